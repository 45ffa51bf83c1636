"""Write the reference-derived golden vectors used by tests/test_oracle.py.

Run where the reference tree is mounted:
    python tests/golden/make_golden.py /root/reference

* gl256.json   -- the 128 positive Gauss-Legendre nodes/weights the
                  reference's Boys table uses
                  (Science/QuantumChemistry/NumericalTools/PointsWeights.hs:25-31).
* mamedov.json -- the Boys-function table quoted in
                  Science/QuantumChemistry/NumericalTools/Boys.hs:27-39:
                  (m, x, Eq(3), Eq(6), reference output).
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _vector(src, name):
    m = re.search(name + r"\s*=\s*U\.fromList\s*\[([^\]]*)\]", src)
    return [float(t.replace("E", "e")) for t in m.group(1).split(",")]


def main(ref):
    nt = os.path.join(ref, "Science", "QuantumChemistry", "NumericalTools")
    with open(os.path.join(nt, "PointsWeights.hs")) as f:
        pw = f.read()
    out = {"points256": _vector(pw, "points256"), "weights256": _vector(pw, "weights256")}
    with open(os.path.join(HERE, "gl256.json"), "w") as f:
        json.dump(out, f)
    with open(os.path.join(nt, "Boys.hs")) as f:
        bz = f.read()
    rows = []
    for m in re.finditer(r"m = ([\d.e-]+)\s+x = ([\d.e-]+)\s+Eq\(3\) = ([\d.e-]+)\s+Eq\(6\) = ([\d.e-]+)\s+ours = ([\d.e-]+)", bz):
        rows.append([float(v) for v in m.groups()])
    with open(os.path.join(HERE, "mamedov.json"), "w") as f:
        json.dump(rows, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
