"""Extract the known-answer tables of the reference's test/scattergather.cpp
into tests/golden/scattergather.json.

The reference test places one particle at (3.3, 3.2) on a unit mesh with
origin 0 (test/scattergather.cpp:46-88) and lists, for IntCIC, IntS1 and
IntS2, the expected scatter weights on edge-centred (Ex, Ey) and cell-centred
(Bz) patches, with tolerance 1e-8 (test/scattergather.cpp:90-342).  This
script reads those tables as data (the Vcheck/Scheck literals) from the
reference source and writes them as JSON.  It runs only where the reference
tree is mounted; the JSON it writes is committed.
"""
import json
import re
import sys

SRC = "/root/reference/test/scattergather.cpp"


def main():
    text = open(SRC).read()
    # split into the three interpolation sections
    sections = {
        "CIC": text.split("// CHECK CIC")[1].split("// CHECK S1")[0],
        "S1": text.split("// CHECK S1")[1].split("// CHECK S2")[0],
        "S2": text.split("// CHECK S2")[1],
    }
    vre = re.compile(r"Vcheck\((\d+),\s*(\d+),\s*Vector_t\(([-0-9.eE]+)\s*,\s*([-0-9.eE]+)\)\)")
    sre = re.compile(r"Scheck\((\d+),\s*(\d+),\s*([-0-9.eE]+)\)")
    out = {"particle": [3.3, 3.2], "origin": [0.0, 0.0], "spacing": [1.0, 1.0],
           "patch": [21, 21], "tolerance": 1e-8, "source": "test/scattergather.cpp:90-342"}
    for name, sec in sections.items():
        out[name] = {
            "edge": [[int(a), int(b), float(c), float(d)] for a, b, c, d in vre.findall(sec)],
            "cell": [[int(a), int(b), float(c)] for a, b, c in sre.findall(sec)],
        }
    json.dump(out, open(sys.argv[1] if len(sys.argv) > 1 else "scattergather.json", "w"), indent=1)


if __name__ == "__main__":
    main()
