"""The benchmark workloads as parity tests on the GPU, each through the C ABI:

  * (H2O)_8/6-31G: the SCF energy of the drop-in path against the oracle's
    restatement of the reference driver (scfHF, HartreeFock.hs:70-123) on
    the flat ERI array (calcIntegrals) -- same screen, Boys table and DIIS;
  * (H2O)_16/6-31G: the analytic gradient against central differences of
    the SCF energy along all 144 coordinates (h = 1e-4 bohr);
  * water and (H2O)_2: the analytic gradient against the oracle's FP64 CPU
    gradient (central differences of the oracle's own SCF energy);
  * (H2O)_32/6-31G: mixed precision (FP32 below a Schwarz bound of 1e-4)
    against the all-FP64 GPU path: |dE| <= 1e-6 Eh, orbital energies 1e-5;
  * (H2O)_64/6-31G: mixed-precision energy and gradient against FP64 from
    the fragment guess: |dE| <= 1e-6 Eh, max |dg| <= 1e-5 Eh/bohr.

Finite differences use the accurate Boys function and no screen (the
reference table is piecewise, so its energy is not differentiable across
grid cells)."""
import sys
import time

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


def test_water8_scf_energy_matches_oracle_driver(require_gpu):
    """(H2O)_8/6-31G: 104 functions, 80 electrons; reference screen and Boys
    table, DIIS (6 vectors) from the second step.  The oracle runs the
    reference's scheme: every canonical ERI into the flat array, then the
    SCF loop contracting that array."""
    _say("case: (H2O)8 SCF vs oracle driver")
    mol = systems.water_cluster(8, "6-31g")
    kw = dict(diis_start=float("inf"), max_steps=60)
    with hfk.Context(mol) as ctx:
        r = ctx.scfHF(diis_size=6, **kw)
    flat, _ = O.eri_flat(mol)
    ref = O.scf(mol, lambda D: O.gmatrix(mol.nbf, flat, D), O.hcore(mol), O.overlap(mol), ndiis=6, **kw)
    _say("  GPU E=%.10f (%d steps)  oracle E=%.10f (%d steps)" % (r.getEnergy, r.steps, ref["energy"], ref["steps"]))
    assert abs(r.getEnergy - ref["energy"]) < 1e-8
    assert r.steps == ref["steps"]
    assert np.max(np.abs(np.sort(r.getOrbE) - np.sort(ref["eps"]))) < 1e-7


@pytest.mark.parametrize("name", ["water", "dimer"])
def test_gradient_matches_oracle_cpu_gradient(require_gpu, name):
    """hfk_gradient (analytic derivative integrals on the GPU) against the
    oracle's FP64 CPU gradient (central differences of its SCF energy)."""
    _say("case: %s/6-31G gradient vs oracle FD gradient" % name)
    mol = systems.water("6-31g") if name == "water" else systems.water_cluster(2, "6-31g")
    with hfk.Context(mol, hfk.options(**ACC)) as ctx:
        r = ctx.scfHF(converge=1e-10, max_steps=100, diis_start=float("inf"), diis_size=6)
        assert r.max_error < 1e-9
        g = ctx.gradient(r.getDensity)
    ref = O.fd_gradient(mol, h=1e-4)
    err = float(np.max(np.abs(g - ref)))
    _say("  max|analytic - oracle| %.2e (max|g| %.2e)" % (err, np.max(np.abs(ref))))
    assert err < 1e-6


@pytest.mark.timeout(900)
def test_water16_gradient_every_coordinate_vs_fd(require_gpu):
    """(H2O)_16/6-31G (48 atoms): the analytic gradient (derivative ERIs
    contracted with the converged density plus the Pulay W.dS term)
    against central differences at h = 1e-4 bohr of the variational SCF
    energy 1/2 sum D (H + F(D)) + V_nn along every one of the 144
    coordinates: 288 displaced SCF runs, each started from the undisplaced
    converged density and converged to max|e| < 1e-10."""
    _say("case: (H2O)16 gradient, 144 coordinates vs FD")
    scf = dict(converge=1e-10, diis_start=float("inf"), diis_size=8, max_steps=100)
    opts = hfk.options(**ACC)
    mol = systems.water_cluster(16, "6-31g")
    t0 = time.perf_counter()
    with hfk.Context(mol, opts) as ctx:
        r = ctx.scfHF(guess_density=ctx.fragment_guess(), **scf)
        assert r.max_error < 1e-9
        g = ctx.gradient(r.getDensity)
    D0 = r.getDensity
    h = 1e-4

    def energy(R):
        m = mol.with_coords(R)
        with hfk.Context(m, opts) as c:
            rr = c.scfHF(guess_density=D0, **scf)
            assert rr.max_error < 1e-9
            H = c.hcore()
            F = c.fock(H, rr.getDensity)
        return O.variational_energy(m, rr.getDensity, H, F)

    fd = np.zeros_like(g)
    for a in range(mol.natom):
        for k in range(3):
            R = mol.R.copy()
            R[a, k] += h
            ep = energy(R)
            R[a, k] -= 2 * h
            fd[a, k] = (ep - energy(R)) / (2 * h)
        if a % 8 == 7:
            _say("  atoms 0..%d done, max|analytic - fd| so far %.2e (%.0f s)"
                 % (a, np.max(np.abs(g[:a + 1] - fd[:a + 1])), time.perf_counter() - t0))
    err = float(np.max(np.abs(g - fd)))
    _say("  max|analytic - fd| over %d coordinates %.3e Eh/bohr, max|g| %.3e, sum_A g_A %.1e"
         % (g.size, err, np.max(np.abs(g)), np.max(np.abs(g.sum(axis=0)))))
    assert err < 1e-6
    assert np.max(np.abs(g.sum(axis=0))) < 1e-10


def test_water32_mixed_precision_vs_fp64(require_gpu):
    """(H2O)_32/6-31G (416 functions): FP32 for shell quartets with Schwarz
    bound below 1e-4, FP64 for the rest; integral-direct incremental Fock
    builds; energy and orbital energies against the all-FP64 GPU path."""
    _say("case: (H2O)32 mixed 1e-4 vs FP64")
    mol = systems.water_cluster(32, "6-31g")
    kw = dict(diis_start=float("inf"), diis_size=6, max_steps=60, incremental_tau=1e-12)
    with hfk.Context(mol) as c64:
        r64 = c64.scfHF(**kw)
    with hfk.Context(mol, hfk.options(fp32_threshold=1e-4)) as c32:
        r32 = c32.scfHF(**kw)
        cs = c32.class_stats()
    q32 = sum(c["shell_quartets"] for c in cs if c["prec"] == 1)
    q64 = sum(c["shell_quartets"] for c in cs if c["prec"] == 0)
    dE = r32.getEnergy - r64.getEnergy
    deps = float(np.max(np.abs(r32.getOrbE - r64.getOrbE)))
    _say("  dE %.2e Eh, max|d eps| %.2e, FP32 shell quartets %.1f%%" % (dE, deps, 100.0 * q32 / (q32 + q64)))
    assert q32 > q64 > 0
    assert abs(dE) <= 1e-6
    assert deps <= 1e-5


def test_water64_mixed_energy_and_gradient_vs_fp64(require_gpu):
    """(H2O)_64/6-31G (832 functions, 192 atoms): the mixed-precision
    energy plus analytic gradient against FP64, both SCFs from the same
    fragment guess with incremental Fock builds: |dE| <= 1e-6 Eh and
    max |dg| <= 1e-5 Eh/bohr."""
    _say("case: (H2O)64 mixed energy+gradient vs FP64")
    mol = systems.water_cluster(64, "6-31g")
    kw = dict(diis_start=float("inf"), diis_size=8, max_steps=100, incremental_tau=1e-12)
    with hfk.Context(mol) as ctx:
        Dg = ctx.fragment_guess()
    out = {}
    for name, opts in (("fp64", hfk.options()), ("mixed", hfk.options(fp32_threshold=1e-4))):
        with hfk.Context(mol, opts) as ctx:
            r = ctx.scfHF(guess_density=Dg, **kw)
            g = ctx.gradient(r.getDensity)
        out[name] = (r, g)
        _say("  %s E=%.10f steps=%d scf %.2fs" % (name, r.getEnergy, r.steps, r.seconds))
    (r64, g64), (r32, g32) = out["fp64"], out["mixed"]
    dE = r32.getEnergy - r64.getEnergy
    dg = float(np.max(np.abs(g32 - g64)))
    _say("  dE %.2e Eh, max|dg| %.2e Eh/bohr" % (dE, dg))
    assert abs(dE) <= 1e-6
    assert dg <= 1e-5
    assert np.max(np.abs(g64.sum(axis=0))) < 1e-9
