"""Classes with d shells (6-31G*, Cartesian d): the cooperative kernel against
the oracle (flat ERIs, G, SCF energy) and the d-shell gradient against
finite differences; the C12H26/6-31G* workload end to end."""
import sys

import numpy as np
import pytest

import hfk
from hfk import systems
import hfk_oracle as O

pytestmark = pytest.mark.gpu

ACC = dict(screen_mode=hfk.HFK_SCREEN_NONE, boys_mode=hfk.HFK_BOYS_ACCURATE)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def water_star():
    return systems.water("6-31g*")


@pytest.fixture(scope="module")
def ethane():
    return systems.alkane(2, "6-31g*")


def test_water_631gs_eri_flat(require_gpu, water_star):
    _say("case: water 6-31g* eri_flat (reference screen, reference Boys)")
    ref, _ = O.eri_flat(water_star)
    with hfk.Context(water_star) as ctx:
        got = ctx.calcIntegrals()
    err = np.max(np.abs(got - ref))
    _say("  max err %.3e" % err)
    assert err < 1e-12, err


def test_water_631gs_eri_flat_exact(require_gpu, water_star):
    _say("case: water 6-31g* eri_flat (no screen, accurate Boys)")
    ref, _ = O.eri_flat(water_star, exact=True, screen=False)
    with hfk.Context(water_star, hfk.options(**ACC)) as ctx:
        got = ctx.calcIntegrals()
    err = np.max(np.abs(got - ref))
    _say("  max err %.3e" % err)
    assert err < 1e-12, err


def test_ethane_631gs_eri_flat(require_gpu, ethane):
    """Two d shells on different centres: every class up to (dd|dd)."""
    _say("case: C2H6 6-31g* eri_flat")
    ref, _ = O.eri_flat(ethane)
    with hfk.Context(ethane) as ctx:
        got = ctx.calcIntegrals()
    err = np.max(np.abs(got - ref))
    _say("  max err %.3e (n=%d)" % (err, ethane.nbf))
    assert err < 1e-12, err


def test_ethane_631gs_gmatrix(require_gpu, ethane):
    _say("case: C2H6 6-31g* G matrix")
    ref_eri, _ = O.eri_flat(ethane)
    rng = np.random.default_rng(5)
    n = ethane.nbf
    A = rng.standard_normal((n, n)) * 0.3
    D = O.full_to_flat(A + A.T)
    ref = O.gmatrix(n, ref_eri, D)
    with hfk.Context(ethane) as ctx:
        got = ctx.calcGmatrix(D)
    err = np.max(np.abs(got - ref))
    _say("  max err %.3e" % err)
    assert err < 1e-10, err


def test_water_631gs_scf(require_gpu, water_star):
    _say("case: water 6-31g* SCF vs oracle")
    ref_eri, _ = O.eri_flat(water_star)
    ref = O.scf(water_star, lambda D: O.gmatrix(water_star.nbf, ref_eri, D), O.hcore(water_star),
                O.overlap(water_star), max_steps=60, diis_start=float("inf"), ndiis=6)
    with hfk.Context(water_star) as ctx:
        S = ctx.mtxOverlap()
        H = ctx.hcore()
        r = ctx.scfHF(max_steps=60, diis_start=float("inf"), diis_size=6)
    assert np.max(np.abs(S - O.overlap(water_star))) < 1e-13
    assert np.max(np.abs(H - O.hcore(water_star))) < 1e-11
    _say("  E gpu %.12f oracle %.12f" % (r.getEnergy, ref["energy"]))
    assert abs(r.getEnergy - ref["energy"]) < 1e-8


def _scf(mol, opts):
    with hfk.Context(mol, opts) as ctx:
        r = ctx.scfHF(converge=1e-10, max_steps=100, diis_start=float("inf"), diis_size=6)
        assert r.max_error < 1e-9, r.max_error
        g = ctx.gradient(r.getDensity)
    return r, g


def test_water_631gs_gradient_fd(require_gpu, water_star):
    _say("case: water 6-31g* SCF gradient vs finite differences")
    mol = water_star
    opts = hfk.options(**ACC)
    r, g = _scf(mol, opts)
    h = 5e-4
    fd = np.zeros_like(g)
    for a in range(mol.natom):
        for k in range(3):
            R = mol.R.copy()
            R[a, k] += h
            ep = _scf(mol.with_coords(R), opts)[0].getEnergy
            R[a, k] -= 2 * h
            em = _scf(mol.with_coords(R), opts)[0].getEnergy
            fd[a, k] = (ep - em) / (2 * h)
    err = np.max(np.abs(g - fd))
    _say("  max |analytic - fd| = %.3e" % err)
    assert err < 2e-6, (g, fd)
    assert np.max(np.abs(g.sum(axis=0))) < 1e-9
    assert np.max(np.abs(np.cross(mol.R, g).sum(axis=0))) < 1e-7


def test_ethane_631gs_gradient_fd(require_gpu, ethane):
    """d shells on two centres: (dd|dd) derivative integrals enter."""
    _say("case: C2H6 6-31g* gradient vs finite differences (4 coordinates)")
    opts = hfk.options(**ACC)
    r, g = _scf(ethane, opts)
    h = 5e-4
    for a, k in ((0, 0), (1, 2), (3, 1), (6, 0)):
        R = ethane.R.copy()
        R[a, k] += h
        ep = _scf(ethane.with_coords(R), opts)[0].getEnergy
        R[a, k] -= 2 * h
        em = _scf(ethane.with_coords(R), opts)[0].getEnergy
        fd = (ep - em) / (2 * h)
        _say("  atom %d axis %d: analytic %.9f fd %.9f" % (a, k, g[a, k], fd))
        assert abs(g[a, k] - fd) < 2e-6
    assert np.max(np.abs(g.sum(axis=0))) < 1e-9
    assert np.max(np.abs(np.cross(ethane.R, g).sum(axis=0))) < 1e-7


def test_alkane_c12h26_energy_gradient_mixed(require_gpu):
    """The C12H26/6-31G* workload (232 functions, 38 atoms): FP64 and mixed
    precision energies and gradients."""
    mol = systems.alkane(12, "6-31g*")
    assert mol.nbf == 232 and mol.natom == 38
    _say("case: C12H26 6-31g* FP64 SCF + gradient")
    with hfk.Context(mol) as ctx:
        r64 = ctx.scfHF(max_steps=60, diis_start=float("inf"), diis_size=6)
        g64 = ctx.gradient(r64.getDensity)
    _say("  E64 %.10f steps %d err %.1e" % (r64.getEnergy, r64.steps, r64.max_error))
    assert r64.max_error < 1e-7
    assert np.max(np.abs(g64.sum(axis=0))) < 1e-8
    _say("case: C12H26 6-31g* mixed (FP32 below 1e-4) SCF + gradient")
    with hfk.Context(mol, hfk.options(fp32_threshold=1e-4)) as ctx:
        r32 = ctx.scfHF(max_steps=60, diis_start=float("inf"), diis_size=6)
        g32 = ctx.gradient(r32.getDensity)
    dE = abs(r32.getEnergy - r64.getEnergy)
    dg = np.max(np.abs(g32 - g64))
    _say("  |dE| %.2e  max|dg| %.2e" % (dE, dg))
    assert dE < 1e-6
    assert dg < 1e-5
