"""Write the oracle-derived fixtures used by tests/test_oracle.py and
tests/test_gpu_parity.py:

    python tests/golden/make_eri_golden.py

* boys_table.json -- SHA-256 of the oracle's 15 x 1001 Boys table
                     (generateGridBoys, TableBoys.hs:58-63; float64,
                     little-endian, m-major) and a few sampled nodes.
* eri_golden.json -- contracted4Centers with the reference screen and Boys
                     table (IntegralsEvaluation.hs:294-337): every ERI of
                     water/STO-3G in calcIntegrals order (HartreeFock.hs:166-181)
                     and 10^4 seeded canonical quartets of (H2O)_8/6-31G, with
                     the number of primitive quartets the screen keeps.

These pin the oracle against itself over time and give the GPU tests fixed
expected values; the oracle's own checks against the reference and against
the mathematics are in tests/test_oracle.py.
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "haskell-abinitio_amd")]

import numpy as np  # noqa: E402

import hfk_oracle as O  # noqa: E402
from hfk import systems  # noqa: E402

NSAMPLE = 10000
SEED = 1


def boys_fixture():
    t = np.ascontiguousarray(O.boys_table(), dtype="<f8")
    grid = t.reshape(O.NBOYS_M, O.NBOYS_X)
    samples = [[m, i, repr(float(grid[m, i]))] for m in (0, 1, 5, 14) for i in (0, 1, 37, 500, 1000)]
    return {"sha256": hashlib.sha256(t.tobytes()).hexdigest(), "shape": [O.NBOYS_M, O.NBOYS_X],
            "samples": samples}


def eri_fixture():
    water = systems.water("sto-3g")
    flat, kept = O.eri_flat(water)
    mol = systems.water_cluster(8, "6-31g")
    q = O.random_canonical_quartets(mol.nbf, NSAMPLE, seed=SEED)
    vals = O.eri_list(mol, q)
    _, kept8 = O.eri_flat(mol)
    return {
        "water_sto3g": {"nbf": water.nbf, "flat": [repr(float(v)) for v in flat], "kept_primitive_quartets": kept},
        "h2o8_631g": {"nbf": mol.nbf, "seed": SEED, "quartets": q.astype(int).ravel().tolist(),
                      "values": [repr(float(v)) for v in vals], "kept_primitive_quartets_full_array": kept8},
    }


def main():
    with open(os.path.join(HERE, "boys_table.json"), "w") as f:
        json.dump(boys_fixture(), f, indent=1)
    with open(os.path.join(HERE, "eri_golden.json"), "w") as f:
        json.dump(eri_fixture(), f)


if __name__ == "__main__":
    main()
