"""GPU parity: libhfk.so (through the C ABI) against the CPU oracle."""
import sys

import numpy as np
import pytest

import hfk
from hfk import systems
import hfk_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def water():
    return systems.water("sto-3g")


def test_water_sto3g_eri_flat(require_gpu, water):
    _say("case: water sto-3g eri_flat")
    ref, _ = O.eri_flat(water)
    with hfk.Context(water) as ctx:
        got = ctx.calcIntegrals()
    assert got.shape == ref.shape
    err = np.max(np.abs(got - ref))
    assert err < 1e-12, err


def test_water_sto3g_gmatrix(require_gpu, water):
    _say("case: water sto-3g gmatrix")
    ref_eri, _ = O.eri_flat(water)
    rng = np.random.default_rng(1)
    n = water.nbf
    A = rng.standard_normal((n, n))
    D = O.full_to_flat(A + A.T)
    ref = O.gmatrix(n, ref_eri, D)
    with hfk.Context(water) as ctx:
        got = ctx.calcGmatrix(D)
    assert np.max(np.abs(got - ref)) < 1e-11


def test_water_sto3g_onee(require_gpu, water):
    _say("case: water sto-3g S,H")
    with hfk.Context(water) as ctx:
        S = ctx.mtxOverlap()
        H = ctx.hcore()
    assert np.max(np.abs(S - O.overlap(water))) < 1e-13
    assert np.max(np.abs(H - O.hcore(water))) < 1e-11


def test_water_sto3g_scf(require_gpu, water):
    _say("case: water sto-3g scf")
    ref_eri, _ = O.eri_flat(water)
    ref = O.scf(water, lambda D: O.gmatrix(water.nbf, ref_eri, D), O.hcore(water), O.overlap(water))
    with hfk.Context(water) as ctx:
        r = ctx.scfHF()
    assert abs(r.getEnergy - ref["energy"]) < 1e-8
    assert np.max(np.abs(np.sort(r.getOrbE) - np.sort(ref["eps"]))) < 1e-6


def test_scf_delta_energy_is_last_step_change(require_gpu, water):
    """delta_energy after k steps equals E_k - E_{k-1}, read from the result
    of runs stopped after k-1 and k steps (converge = 0 never converges, so
    hfk_scf returns HFK_ESCF with the last step's figures filled in)."""
    _say("case: water sto-3g scf delta_energy")
    import ctypes as C
    from hfk import _lib

    def run(steps):
        o = _lib.HfkScfOptions()
        _lib.lib().hfk_default_scf_options(C.byref(o))
        o.max_steps = steps
        o.converge = 0.0
        r = _lib.HfkScfResult()
        rc = _lib.lib().hfk_scf(ctx._h, C.byref(o), C.byref(r), None, None, None, None)
        assert rc != 0 and r.steps == steps
        return r.energy, r.delta_energy

    with hfk.Context(water) as ctx:
        e1, d1 = run(1)
        e2, d2 = run(2)
        e3, d3 = run(3)
    assert d1 == float("inf")
    assert abs(d2 - (e2 - e1)) < 1e-12
    assert abs(d3 - (e3 - e2)) < 1e-12
    assert abs(d3) > 0.0


def test_water_dimer_631g_eri_flat(require_gpu):
    _say("case: (H2O)2 6-31g eri_flat")
    mol = systems.water_cluster(2, "6-31g")
    ref, _ = O.eri_flat(mol)
    with hfk.Context(mol) as ctx:
        got = ctx.calcIntegrals()
    err = np.max(np.abs(got - ref))
    assert err < 1e-12, err


def test_eri_fixtures_water_and_water8(require_gpu):
    """The flat ERI export against the committed fixtures
    (tests/golden/eri_golden.json): all water/STO-3G ERIs and 10^4 seeded
    (H2O)_8/6-31G quartets, to 1e-12."""
    import json
    import os
    _say("case: ERI fixtures water sto-3g, (H2O)8 6-31g sample")
    eg = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "eri_golden.json")))
    with hfk.Context(systems.water("sto-3g")) as ctx:
        got = ctx.calcIntegrals()
    ref = np.array([float(v) for v in eg["water_sto3g"]["flat"]])
    assert np.max(np.abs(got - ref)) < 1e-12
    s = eg["h2o8_631g"]
    mol = systems.water_cluster(8, "6-31g")
    with hfk.Context(mol) as ctx:
        flat = ctx.calcIntegrals()
    q = np.array(s["quartets"]).reshape(-1, 4)
    idx = np.array([O.flat_rank(mol.nbf, *map(int, r)) for r in q])
    err = np.max(np.abs(flat[idx] - np.array([float(v) for v in s["values"]])))
    _say("  (H2O)8 sample: max|d| %.2e over %d quartets" % (err, len(q)))
    assert err < 1e-12


@pytest.fixture(scope="module")
def water8():
    mol = systems.water_cluster(8, "6-31g")
    flat, _ = O.eri_flat(mol)
    return mol, flat


def test_water8_gmatrix(require_gpu, water8):
    _say("case: (H2O)8 6-31g gmatrix")
    mol, flat = water8
    n = mol.nbf
    rng = np.random.default_rng(7)
    A = rng.standard_normal((n, n)) * 0.1
    D = O.full_to_flat(A + A.T)
    ref = O.gmatrix(n, flat, D)
    with hfk.Context(mol) as ctx:
        got = ctx.calcGmatrix(D)
        got2 = ctx.calcGmatrix(D)
    assert np.max(np.abs(got - ref)) < 1e-11 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(got2 - got)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_water_dimer_scf(require_gpu):
    _say("case: (H2O)2 6-31g scf")
    mol = systems.water_cluster(2, "6-31g")
    flat, _ = O.eri_flat(mol)
    ref = O.scf(mol, lambda D: O.gmatrix(mol.nbf, flat, D), O.hcore(mol), O.overlap(mol))
    with hfk.Context(mol) as ctx:
        r = ctx.scfHF()
    assert abs(r.getEnergy - ref["energy"]) < 1e-8
    assert r.steps == ref["steps"]


def test_scf_energy_converge_rule(require_gpu):
    """energy_converge adds |E_k - E_k-1| < bound to the max|e| rule: the run
    stops no earlier than the reference rule alone, its last step meets the
    bound, and its energy is within the bound of a tightly converged one."""
    _say("case: (H2O)2 6-31g scf energy_converge")
    mol = systems.water_cluster(2, "6-31g")
    kw = dict(diis_start=float("inf"), diis_size=6, max_steps=60)
    with hfk.Context(mol) as ctx:
        r0 = ctx.scfHF(**kw)
        r1 = ctx.scfHF(energy_converge=1e-8, **kw)
        rt = ctx.scfHF(converge=1e-10, **kw)
    print("  steps %d / %d / %d; dE %.2e / %.2e" % (r0.steps, r1.steps, rt.steps, r0.delta_energy,
                                                     r1.delta_energy), flush=True)
    assert r1.steps >= r0.steps
    assert abs(r1.delta_energy) < 1e-8 and r1.max_error < 1e-7
    assert abs(r1.getEnergy - rt.getEnergy) < 1e-8


def test_water_sto3g_scf_reference_harris_guess(require_gpu, water):
    """scfHF from harrisFunctional as written (HFK_GUESS_REFERENCE) against
    the oracle driver with the same guess: same energy, same step count.  For
    (H2O)_8 the guess reads outside the integral array, where the reference
    aborts: an error, no result."""
    print("case: water sto-3g SCF, reference Harris guess", flush=True)
    ref_eri, _ = O.eri_flat(water)
    ref = O.scf(water, lambda D: O.gmatrix(water.nbf, ref_eri, D), O.hcore(water), O.overlap(water),
                guess="reference", eri=ref_eri)
    with hfk.Context(water) as ctx:
        r = ctx.scfHF(guess=hfk.HFK_GUESS_REFERENCE)
    print("  E gpu %.12f oracle %.12f; steps %d / %d" % (r.getEnergy, ref["energy"], r.steps, ref["steps"]), flush=True)
    assert abs(r.getEnergy - ref["energy"]) < 1e-8
    assert r.steps == ref["steps"]
    with hfk.Context(systems.water_cluster(8, "6-31g")) as ctx:
        with pytest.raises(hfk.HfkError):
            ctx.scfHF(guess=hfk.HFK_GUESS_REFERENCE)


def test_textbook_h2_and_heh_through_the_abi(require_gpu):
    """Szabo & Ostlund, STO-3G, exact Boys function and no screen (the
    textbook values are for the exact integrals): H2 at R = 1.4 bohr,
    zeta = 1.24 (S12, four ERIs, E = -1.1167) and HeH+ at R = 1.4632 bohr
    (E = -2.860662; printed to 6 d.p., checked to 5e-6)."""
    _say("case: textbook H2 and HeH+ STO-3G")
    exact = hfk.options(screen_mode=hfk.HFK_SCREEN_NONE, boys_mode=hfk.HFK_BOYS_ACCURATE)
    h2 = systems.h2_szabo()
    with hfk.Context(h2, exact) as ctx:
        S = ctx.mtxOverlap()
        flat = ctx.calcIntegrals()
        r = ctx.scfHF()
    g = lambda *ix: O.lib().ob_flat_get(2, O._dp(np.ascontiguousarray(flat)), *ix)
    assert abs(S[1] - 0.6593) < 5e-5
    assert abs(g(0, 0, 0, 0) - 0.7746) < 5e-5
    assert abs(g(0, 0, 1, 1) - 0.5697) < 5e-5
    assert abs(g(1, 0, 1, 0) - 0.2970) < 5e-5
    assert abs(g(1, 0, 0, 0) - 0.4441) < 5e-5
    assert abs(r.getEnergy - (-1.1167)) < 5e-5
    with hfk.Context(systems.heh_szabo(), exact) as ctx:
        r2 = ctx.scfHF()
    _say("  E(H2) %.6f  E(HeH+) %.6f" % (r.getEnergy, r2.getEnergy))
    assert abs(r2.getEnergy - (-2.860662)) < 5e-6


def test_too_many_primitives_is_an_error(require_gpu):
    """A shell pair holds at most 254 primitive pairs (tasks carry 8-bit
    primitive ranges; 255 marks a whole ket): 16 x 16 is refused at context
    creation, 15 x 15 works.  A pair of 255 = 15 x 17 = 5 x 51 = 3 x 85
    primitive pairs cannot occur alone: the larger shell's own pair (289 or
    more) is refused first."""
    _say("case: primitive-count limit")
    def h2(nprim):
        prims = [(1.0, 0.1 * 1.7 ** k) for k in range(nprim)]
        return systems.custom_molecule([("H", (0, 0, 0), [("S", prims)]), ("H", (0, 0, 1.4), [("S", prims)])])
    with pytest.raises(hfk.HfkError):
        hfk.Context(h2(16))
    with hfk.Context(h2(15)) as ctx:
        assert np.all(np.isfinite(ctx.calcIntegrals()))


def test_scf_option_limits(require_gpu, water):
    """diis_size outside [1, 56] and a negative incremental_tau are refused
    before any work (the DIIS matrix row of the current error is read back
    with max|e| in one 64-double transfer)."""
    _say("case: scf option limits")
    with hfk.Context(water) as ctx:
        for bad in (dict(diis_size=0), dict(diis_size=57), dict(incremental_tau=-1.0)):
            with pytest.raises(hfk.HfkError):
                ctx.scfHF(**bad)
        assert abs(ctx.scfHF(diis_size=56).getEnergy - ctx.scfHF().getEnergy) < 1e-8


def test_one_function_helium_closed_forms(require_gpu):
    """The smallest input: He with one STO-3G s function (zeta = 1.6875; nbf =
    1, occupied = 1, one atom away from the origin).  One orbital fixes the
    density at D = 2/S, so with H = h11, J = (11|11): G(D) = J/S,
    E = 2H/S + J/S^2 (no nuclear repulsion), eps = (H + J/S)/S, and a lone
    atom's gradient is zero.  An empty basis is refused at context creation."""
    import ctypes as C
    from hfk import _lib
    _say("case: He one function, closed forms; empty basis refused")
    mol = systems.custom_molecule([("He", (0.3, -0.2, 0.1), [("S", systems.sto3g_s(1.6875))])])
    ref_eri, _ = O.eri_flat(mol)
    with hfk.Context(mol) as ctx:
        S = float(ctx.mtxOverlap()[0])
        H = float(ctx.hcore()[0])
        flat = ctx.calcIntegrals()
        J = float(flat[0])
        G = ctx.calcGmatrix(np.array([2.0 / S]))
        r = ctx.scfHF()
        g = ctx.gradient(r.getDensity)
    _say("  S %.15f H %.12f J %.12f E %.12f steps %d" % (S, H, J, r.getEnergy, r.steps))
    assert flat.shape == (1,) and abs(J - ref_eri[0]) < 1e-12
    assert abs(S - O.overlap(mol)[0]) < 1e-13 and abs(H - O.hcore(mol)[0]) < 1e-11
    assert abs(G[0] - J / S) < 1e-13
    assert abs(r.getEnergy - (2 * H / S + J / S ** 2)) < 1e-10
    assert abs(float(np.ravel(r.getOrbE)[0]) - (H + J / S) / S) < 1e-10
    assert abs(float(np.ravel(r.getDensity)[0]) - 2.0 / S) < 1e-12
    assert g.shape == (1, 3) and np.max(np.abs(g)) < 1e-12
    b = _lib.HfkBasis()
    b.nbf = 0
    h = C.c_void_p()
    assert _lib.lib().hfk_create(C.byref(b), None, C.byref(h)) == -1 and not h.value   # HFK_EINVAL
