"""Beam parameters of PartBunch::calcBeamParameters (PartBunch.cpp:584-657) as
the engines return them (csrst_beam_stats / csrst_mw_beam_stats), and the
statistics file of DataSinkCommand (DataSink.cpp:39-148)."""
import numpy as np

# order of the CSRST_BEAM_STATS values
FIELDS = ("Ekin", "rrms_x", "rrms_y", "prms_x", "prms_y", "emit_x", "emit_y", "rmean_x", "rmean_y",
          "pmean_x", "pmean_y", "Dy", "DDy", "dE", "rmin_x", "rmin_y", "rmax_x", "rmax_y", "rprms_x", "rprms_y")
N_STATS = len(FIELDS)


def as_dict(values):
    v = np.asarray(values, np.float64)
    return {k: float(v[i]) for i, k in enumerate(FIELDS)}
