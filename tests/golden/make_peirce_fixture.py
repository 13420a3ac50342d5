"""Writes peirce_m14.json: the reference's M = 14 Peirce quadrature table.

Reads the numbers (not the code) of the kPeirceM == 14 branch of
include/alluvion/dg/peirce_quadrature.hpp in the reference checkout and stores
them as a JSON fixture for tests/test_oracle_quadrature.py.  Run once in a
container that has the reference mounted:

    python tests/golden/make_peirce_fixture.py /root/reference
"""
import json
import os
import re
import sys


def main(ref_root):
    path = os.path.join(ref_root, "include", "alluvion", "dg", "peirce_quadrature.hpp")
    text = open(path).read()
    start = text.index("kPeirceM == 14")
    end = text.index("}", start)
    block = text[start:end]
    out = {}
    for name in ("cosphi", "sinphi", "b", "r", "c"):
        vals = {}
        for m in re.finditer(r"\b%s\[(\d+)\]\s*=\s*([-+0-9.eE]+);" % name, block):
            vals[int(m.group(1))] = m.group(2)
        out[name] = [vals[i] for i in range(len(vals))]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "peirce_m14.json")
    with open(dst, "w") as f:
        json.dump({"source": "include/alluvion/dg/peirce_quadrature.hpp (kPeirceM == 14)",
                   "values": out}, f, indent=1)
    print(dst, {k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
