"""Writes tests/golden/cocurve.json from the reference's coexistence tables.

The tables CoCurve/CoCurve_<EOS>_2.txt hold Maxwell-construction coexistence
densities (reduced by rho_c) per reduced temperature Tr; SetMultiphase
(FluidCuda.cu:1356-1384) reads them.  They are data, kept here as the golden
vectors for the EOS / coexistence row of the oracle and the product.

Run in a container that has the reference checkout:
    python tests/golden/make_cocurve_fixture.py /root/reference
"""
import json
import os
import sys


def main(ref):
    out = {}
    for eos in ("CSE", "PRM", "PRW", "RKE", "RKS", "VDW"):
        path = os.path.join(ref, "CoCurve", f"CoCurve_{eos}_2.txt")
        rows = []
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) != 3:
                    continue
                try:
                    rows.append([float(x) for x in parts])
                except ValueError:
                    continue   # header line "T DenGas DenLiquid"
        out[eos] = rows
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cocurve.json")
    with open(dst, "w") as fh:
        json.dump({"source": "CoCurve/CoCurve_<EOS>_2.txt: columns Tr, rhoG/rhoc, rhoL/rhoc",
                   "tables": out}, fh, separators=(",", ":"))
    print(dst, {k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
