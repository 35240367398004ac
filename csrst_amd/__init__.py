"""csrst_amd: the particle-in-cell time step of CSRST on AMD MI355X (gfx950).

Kernels live in libcsrst.so (csrc/, C ABI in include/csrst_hip.h); this
package holds the host-side front end that mirrors the reference's
interfaces for the step.
"""
from . import _lib  # noqa: F401
from .bunch import gaussian_bunch  # noqa: F401
from .pic import Commands, TwoGridStep, bend_radius  # noqa: F401

__all__ = ["Commands", "TwoGridStep", "bend_radius", "gaussian_bunch"]
