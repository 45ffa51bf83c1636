"""Run-to-run bitwise reproducibility of the Fock build.

The J/K kernels write their partial sums to per-task slots with plain stores
and fixed-order reductions add them up (no floating-point atomics outside a
workgroup's LDS, whose adds are ordered), so the same density gives the same
bits on every build and on every context: the reference's calcGmatrix
(HartreeFock.hs:225-241) is likewise a fixed-order sum."""
import numpy as np
import pytest

import hfk
from hfk import systems

pytestmark = pytest.mark.gpu


def _density(mol, seed):
    n = mol.nbf
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * 0.1
    return (A + A.T)[np.triu_indices(n)]


@pytest.mark.parametrize("name,opts", [
    ("water8", {}),
    ("water8", {"fp32_threshold": 1e-4}),
    ("water8", {"screen_mode": hfk.HFK_SCREEN_NONE}),
    ("c12h26", {}),
])
def test_fock_build_is_bitwise_reproducible(require_gpu, name, opts):
    print("case: %s %s" % (name, opts), flush=True)
    mol = systems.water_cluster(8, "6-31g") if name == "water8" else systems.alkane(12, "6-31g*")
    D = _density(mol, 3)
    with hfk.Context(mol, hfk.options(**opts)) as ctx:
        builds = [ctx.calcGmatrix(D) for _ in range(4)]   # serial first runs, then concurrent launches
    with hfk.Context(mol, hfk.options(**opts)) as ctx2:
        builds.append(ctx2.calcGmatrix(D))
    for G in builds[1:]:
        assert np.array_equal(G, builds[0])
    assert np.all(np.isfinite(builds[0])) and np.max(np.abs(builds[0])) > 0.1


def test_scf_energy_is_bitwise_reproducible(require_gpu):
    print("case: (H2O)8 SCF twice", flush=True)
    mol = systems.water_cluster(8, "6-31g")
    out = []
    for _ in range(2):
        with hfk.Context(mol) as ctx:
            r = ctx.scfHF(diis_start=float("inf"), diis_size=6, max_steps=60)
        out.append(r)
    assert out[0].getEnergy == out[1].getEnergy and out[0].steps == out[1].steps
    assert np.array_equal(out[0].getDensity, out[1].getDensity)
