"""Nuclear gradient (hfk_gradient) on the GPU.

The reference has no working gradient (HartreeFock/Derivatives.hs does not
compile), so the gradient is pinned by what the mathematics fixes:
  * term by term at a fixed density D and energy-weighted density W, the
    gradient equals the central finite difference of
        E_fix(R) = sum D H(R) + 1/2 sum D G(D; R) - sum W S(R) + V_nn(R)
    built from the library's own hcore / calcGmatrix / mtxOverlap, whose
    values are pinned against the oracle elsewhere;
  * at SCF convergence it equals the finite difference of the SCF energy;
  * it is translation invariant (sum_A g_A = 0) for any D, W and rotation
    invariant (sum_A R_A x g_A = 0) at convergence.
Finite differences use the accurate Boys function and no screen: the
reference's Boys table is piecewise (a 0.1 grid with 6-term Taylor steps
whose node values carry a 7e-6 relative error), so its energy is not
differentiable across grid cells."""
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


def _full(flat, n):
    return O.flat_to_full(flat, n)


def _e_fixed(mol, D, W):
    n = mol.nbf
    with hfk.Context(mol, hfk.options(**ACC)) as ctx:
        H = _full(ctx.hcore(), n)
        G = _full(ctx.calcGmatrix(D), n)
        S = _full(ctx.mtxOverlap(), n)
    Df, Wf = _full(D, n), _full(W, n)
    return float((Df * H).sum() + 0.5 * (Df * G).sum() - (Wf * S).sum() + mol.nuclear_repulsion())


def _scf(mol, opts=None):
    with hfk.Context(mol, opts if opts is not None else hfk.options(**ACC)) as ctx:
        r = ctx.scfHF(converge=1e-10, max_steps=100, diis_start=float("inf"), diis_size=6)
        assert r.max_error < 1e-9, r.max_error
        g = ctx.gradient(r.getDensity)
    return r, g


def _displaced(mol, a, k, h):
    R = mol.R.copy()
    R[a, k] += h
    return mol.with_coords(R)


def _rand_sym_flat(rng, n, scale):
    A = rng.standard_normal((n, n)) * scale
    return O.full_to_flat(A + A.T)


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g"])
def test_gradient_terms_fixed_density(require_gpu, basis):
    """Each term of the analytic gradient against finite differences at a
    fixed (non-converged) D and W; W = 0 and D = 0 isolate the terms."""
    mol = systems.water(basis)
    n = mol.nbf
    rng = np.random.default_rng(7)
    with hfk.Context(mol, hfk.options(**ACC)) as ctx:
        r = ctx.scfHF(converge=1e-8, max_steps=60, diis_start=float("inf"))
    D = r.getDensity + _rand_sym_flat(rng, n, 0.05)
    W = _rand_sym_flat(rng, n, 0.3)
    zero = np.zeros_like(D)
    h = 1e-4
    for label, Dx, Wx in (("D,W", D, W), ("W=0", D, zero), ("D=0", zero, W)):
        _say("case: water %s fixed-density gradient, %s" % (basis, label))
        with hfk.Context(mol, hfk.options(**ACC)) as ctx:
            g = ctx.gradient(Dx, Wx)
        fd = np.zeros_like(g)
        for a in range(mol.natom):
            for k in range(3):
                ep = _e_fixed(_displaced(mol, a, k, h), Dx, Wx)
                em = _e_fixed(_displaced(mol, a, k, -h), Dx, Wx)
                fd[a, k] = (ep - em) / (2 * h)
        err = np.max(np.abs(g - fd))
        _say("  max |analytic - fd| = %.3e (max |g| %.3e)" % (err, np.max(np.abs(g))))
        assert err < 2e-7 * max(1.0, np.max(np.abs(g))), (label, g, fd)
        assert np.max(np.abs(g.sum(axis=0))) < 1e-9 * max(1.0, np.max(np.abs(g)))


def test_gradient_scf_water_631g(require_gpu):
    _say("case: water 6-31g SCF gradient vs finite differences")
    mol = systems.water("6-31g")
    r, g = _scf(mol)
    h = 5e-4
    fd = np.zeros_like(g)
    for a in range(mol.natom):
        for k in range(3):
            ep = _scf(_displaced(mol, a, k, h))[0].getEnergy
            em = _scf(_displaced(mol, a, k, -h))[0].getEnergy
            fd[a, k] = (ep - em) / (2 * h)
    err = np.max(np.abs(g - fd))
    _say("  max |analytic - fd| = %.3e" % err)
    assert err < 2e-6, (g, fd)
    assert np.max(np.abs(g.sum(axis=0))) < 1e-9
    torque = np.cross(mol.R, g).sum(axis=0)
    assert np.max(np.abs(torque)) < 1e-7, torque


def test_gradient_scf_explicit_w(require_gpu):
    """W passed as 2 sum_occ eps_i C_i C_i^T gives the internal 1/2 D F D result."""
    _say("case: water 6-31g gradient with explicit W")
    mol = systems.water("6-31g")
    n = mol.nbf
    with hfk.Context(mol, hfk.options(**ACC)) as ctx:
        r = ctx.scfHF(converge=1e-10, max_steps=100, diis_start=float("inf"), diis_size=6)
        Cm = r.getCoeff
        occ = mol.occupied
        order = np.argsort(r.getOrbE)
        D = _full(r.getDensity, n)
        for C_cols in (Cm, Cm.T):   # the orientation whose occupied block rebuilds D
            Co = C_cols[:, order[:occ]]
            if np.max(np.abs(2.0 * Co @ Co.T - D)) < 1e-8:
                break
        else:
            pytest.fail("getCoeff does not rebuild the density")
        W = 2.0 * (Co * r.getOrbE[order[:occ]]) @ Co.T
        g_w = ctx.gradient(r.getDensity, O.full_to_flat(W))
        g_int = ctx.gradient(r.getDensity)
    assert np.max(np.abs(g_w - g_int)) < 1e-8


def test_gradient_scf_water_dimer(require_gpu):
    _say("case: (H2O)2 6-31g SCF gradient vs finite differences")
    mol = systems.water_cluster(2, "6-31g")
    r, g = _scf(mol)
    h = 5e-4
    rng = np.random.default_rng(3)
    coords = [(int(a), int(k)) for a, k in zip(rng.integers(0, mol.natom, 5), rng.integers(0, 3, 5))]
    for a, k in coords:
        ep = _scf(_displaced(mol, a, k, h))[0].getEnergy
        em = _scf(_displaced(mol, a, k, -h))[0].getEnergy
        fd = (ep - em) / (2 * h)
        _say("  atom %d axis %d: analytic %.9f fd %.9f" % (a, k, g[a, k], fd))
        assert abs(g[a, k] - fd) < 2e-6
    assert np.max(np.abs(g.sum(axis=0))) < 1e-9
    assert np.max(np.abs(np.cross(mol.R, g).sum(axis=0))) < 1e-7


def test_gradient_reference_modes(require_gpu):
    """With the reference's screen and Boys table the gradient stays within
    the table's accuracy of the exact one."""
    _say("case: (H2O)2 6-31g gradient, reference screen and Boys table")
    mol = systems.water_cluster(2, "6-31g")
    _, g_acc = _scf(mol)
    _, g_ref = _scf(mol, hfk.options())
    err = np.max(np.abs(g_ref - g_acc))
    _say("  max |reference - accurate| = %.3e" % err)
    assert err < 1e-4
    assert np.max(np.abs(g_ref.sum(axis=0))) < 1e-9


def _oracle_energy(mol):
    """FP64 CPU restatement (oracle): exact Boys, no screen, reference SCF driver."""
    n = mol.nbf
    eri, _ = O.eri_flat(mol, exact=True, screen=False)
    r = O.scf(mol, lambda D: O.gmatrix(n, eri, D), O.hcore(mol, exact=True), O.overlap(mol),
              max_steps=100, ndiis=6, converge=1e-10, diis_start=float("inf"))
    return r["energy"]


class _HFEnergy:
    """torch.autograd.Function around the GPU SCF energy and hfk_gradient
    (test harness only)."""

    @staticmethod
    def make():
        import torch

        class F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, R, mol):
                m = mol.with_coords(R.detach().cpu().numpy())
                with hfk.Context(m, hfk.options(**ACC)) as c:
                    r = c.scfHF(converge=1e-10, max_steps=100, diis_start=float("inf"), diis_size=6)
                    g = c.gradient(r.getDensity)
                ctx.save_for_backward(torch.tensor(g, dtype=torch.float64))
                return torch.tensor(r.getEnergy, dtype=torch.float64)

            @staticmethod
            def backward(ctx, grad_out):
                (g,) = ctx.saved_tensors
                return grad_out * g, None

        return F


def test_gradient_autograd_vs_cpu_fd(require_gpu):
    """The GPU gradient through torch.autograd equals the central finite
    difference of the FP64 CPU oracle's SCF energy (water/6-31G)."""
    import torch
    _say("case: water 6-31g autograd gradient vs CPU (oracle) finite differences")
    mol = systems.water("6-31g")
    R = torch.tensor(mol.R, dtype=torch.float64, requires_grad=True)
    E = _HFEnergy.make().apply(R, mol)
    E.backward()
    h = 5e-4
    fd = np.zeros((mol.natom, 3))
    for a in range(mol.natom):
        for k in range(3):
            fd[a, k] = (_oracle_energy(_displaced(mol, a, k, h)) - _oracle_energy(_displaced(mol, a, k, -h))) / (2 * h)
    e_cpu = _oracle_energy(mol)
    _say("  E gpu %.12f cpu %.12f; max |dE/dR gpu - fd cpu| %.3e" % (E.item(), e_cpu, np.max(np.abs(R.grad.numpy() - fd))))
    assert abs(E.item() - e_cpu) < 1e-8
    assert np.max(np.abs(R.grad.numpy() - fd)) < 2e-6


def test_gradient_after_scf_uses_cached_fock(require_gpu):
    """After a converged scfHF the context forms F(D) for W from its last Fock
    build plus G(D - D0); the gradient equals the one from a fresh context
    (full Fock build) at the same density."""
    _say("case: (H2O)2 6-31g gradient after SCF vs fresh context")
    mol = systems.water_cluster(2, "6-31g")
    with hfk.Context(mol) as ctx:
        r = ctx.scfHF(converge=1e-9, max_steps=100, diis_start=float("inf"), diis_size=6)
        g_cached = ctx.gradient(r.getDensity)
        rng = np.random.default_rng(11)
        D2 = r.getDensity + _rand_sym_flat(rng, mol.nbf, 1e-3)
        g_cached2 = ctx.gradient(D2)
    with hfk.Context(mol) as ctx:
        g_full = ctx.gradient(r.getDensity)
        g_full2 = ctx.gradient(D2)
    err = max(np.max(np.abs(g_cached - g_full)), np.max(np.abs(g_cached2 - g_full2)))
    _say("  max |cached - full| = %.3e" % err)
    assert err < 1e-10
