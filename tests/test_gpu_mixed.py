"""Mixed precision: shell quartets with Schwarz bound below a threshold are
evaluated and contracted in FP32; everything else stays FP64.  Checked
against the all-FP64 GPU path (itself checked against the oracle)."""
import numpy as np
import pytest

import hfk
from hfk import systems

pytestmark = pytest.mark.gpu

SCF = dict(diis_start=float("inf"), diis_size=6, max_steps=60)


@pytest.fixture(scope="module")
def water8_fp64():
    mol = systems.water_cluster(8, "6-31g")
    with hfk.Context(mol) as ctx:
        r = ctx.scfHF(**SCF)
    return mol, r


@pytest.mark.parametrize("thr", [1e-6, 1e-4])
def test_mixed_energy_within_1uEh(require_gpu, water8_fp64, thr):
    print("case: (H2O)8 mixed thr=%g" % thr, flush=True)
    mol, r64 = water8_fp64
    with hfk.Context(mol, hfk.options(fp32_threshold=thr)) as ctx:
        r32 = ctx.scfHF(**SCF)
        st = ctx.class_stats()
    assert any(c["prec"] == 1 and c["shell_quartets"] > 0 for c in st)
    assert abs(r32.getEnergy - r64.getEnergy) < 1e-6
    assert np.max(np.abs(r32.getOrbE - r64.getOrbE)) < 1e-5


def test_mixed_fock_close_to_fp64(require_gpu, water8_fp64):
    print("case: (H2O)8 mixed Fock", flush=True)
    mol, r64 = water8_fp64
    D = r64.getDensity
    with hfk.Context(mol) as c64:
        G64 = c64.calcGmatrix(D)
    with hfk.Context(mol, hfk.options(fp32_threshold=1e-4)) as c32:
        G32 = c32.calcGmatrix(D)
    assert np.max(np.abs(G32 - G64)) < 1e-5


def test_mixed_parts_partition_the_quartets(require_gpu, water8_fp64):
    """With the precision split, each class's FP64 and FP32 task parts
    together hold exactly the quartets of the all-FP64 plan: the surviving
    shell, primitive and function quartets per class add up, so no quartet
    is evaluated twice or dropped."""
    print("case: (H2O)8 mixed task parts", flush=True)
    mol, _ = water8_fp64
    key = lambda c: (c["la"], c["lb"], c["lc"], c["ld"])
    with hfk.Context(mol) as c64:
        ref = {key(c): c for c in c64.class_stats()}
    with hfk.Context(mol, hfk.options(fp32_threshold=1e-4)) as c32:
        parts = c32.class_stats()
    assert {c["prec"] for c in parts} == {0, 1}
    for k, r in ref.items():
        got = [c for c in parts if key(c) == k]
        for field in ("shell_quartets", "prim_quartets", "func_quartets"):
            assert sum(c[field] for c in got) == r[field], (k, field)


@pytest.mark.parametrize("boys", [hfk.HFK_BOYS_REFERENCE, hfk.HFK_BOYS_ACCURATE])
def test_mixed_gradient_close_to_fp64(require_gpu, water8_fp64, boys):
    """The FP32 part of each task list runs the derivative integrals in FP32
    (both Boys modes): at a fixed density the gradient stays within 1e-7
    Eh/bohr of the FP64 one, and the FP32 parts are really used (the mixed
    gradient differs from the FP64 one)."""
    print("case: (H2O)8 mixed gradient boys=%d" % boys, flush=True)
    mol, r64 = water8_fp64
    D = r64.getDensity
    with hfk.Context(mol, hfk.options(boys_mode=boys)) as c64:
        g64 = c64.gradient(D)
    with hfk.Context(mol, hfk.options(boys_mode=boys, fp32_threshold=1e-4)) as c32:
        g32 = c32.gradient(D)
        st = c32.class_stats()
    assert any(c["prec"] == 1 and c["shell_quartets"] > 0 for c in st)
    d = np.max(np.abs(g32 - g64))
    assert 0.0 < d < 1e-7, d
    assert np.max(np.abs(g32.sum(axis=0))) < 1e-9
