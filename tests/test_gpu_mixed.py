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
