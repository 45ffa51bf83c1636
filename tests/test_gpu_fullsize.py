"""Fock builds at the benchmark sizes, (H2O)_32 and (H2O)_64 in 6-31G (416
and 832 basis functions), where the flat ERI array (28 and 447 GiB) cannot
be formed, and C12H26 in 6-31G* (232 functions, d shells).

- Sampled elements of G(D) against brute-force contractions of the oracle's
  ERIs (contracted4Centers with the reference screen and Boys table):
  G_ab = sum_kl D_lk [(ab|kl) - 1/2 (al|kb)] (HartreeFock.hs:225-241), every
  ERI taken at its canonical index order as calcIntegrals stores it
  (HartreeFock.hs:153-156,166-181).
- Linearity of G in D over the whole matrix.
"""
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


def _canonical(q):
    """sortKeys: i <= j, k <= l, (i, j) <= (k, l) lexicographically."""
    i, j, k, l = (q[:, c].copy() for c in range(4))
    i, j = np.minimum(i, j), np.maximum(i, j)
    k, l = np.minimum(k, l), np.maximum(k, l)
    swap = (i > k) | ((i == k) & (j > l))
    i2 = np.where(swap, k, i)
    j2 = np.where(swap, l, j)
    k2 = np.where(swap, i, k)
    l2 = np.where(swap, j, l)
    return np.stack([i2, j2, k2, l2], axis=1).astype(np.int32)


def _g_element(mol, Dfull, a, b):
    n = mol.nbf
    kk, ll = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    kk, ll = kk.ravel(), ll.ravel()
    coul = np.stack([np.full_like(kk, a), np.full_like(kk, b), kk, ll], axis=1)
    exch = np.stack([np.full_like(kk, a), ll, kk, np.full_like(kk, b)], axis=1)
    vc = O.eri_list(mol, _canonical(coul))
    vx = O.eri_list(mol, _canonical(exch))
    w = Dfull[ll, kk]
    return float(np.dot(vc, w) - 0.5 * np.dot(vx, w))


def _density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * 0.05
    return A + A.T


@pytest.mark.parametrize("system,nsample", [("h2o32", 12), ("h2o64", 6), ("c12h26", 8)])
def test_cluster_gmatrix_sampled_elements(require_gpu, system, nsample):
    _say("case: %s gmatrix, %d sampled elements vs oracle" % (system, nsample))
    if system == "c12h26":   # 6-31G*, six Cartesian d per carbon: the d classes
        mol = systems.alkane(12, "6-31g*")
    else:
        mol = systems.water_cluster(int(system[3:]), "6-31g")
    n = mol.nbf
    Dfull = _density(n, 11)
    with hfk.Context(mol) as ctx:
        G = O.flat_to_full(ctx.calcGmatrix(O.full_to_flat(Dfull)), n)
    rng = np.random.default_rng(n)
    pairs = [(0, 0), (0, 1), (n - 1, n - 1)] + [tuple(sorted(rng.integers(0, n, 2))) for _ in range(nsample - 3)]
    dfun = np.nonzero(mol.lxyz.sum(axis=1) == 2)[0]
    if dfun.size:   # d-d and d-s elements on the first and last carbon
        pairs += [(int(dfun[0]), int(dfun[1])), (int(dfun[2]), int(dfun[-1])), (0, int(dfun[-2]))]
    scale = max(1.0, float(np.max(np.abs(G))))
    worst = 0.0
    for a, b in pairs:
        ref = _g_element(mol, Dfull, a, b)
        err = abs(G[a, b] - ref)
        worst = max(worst, err)
        _say("  G[%d,%d] gpu %.12e oracle %.12e |d| %.2e" % (a, b, G[a, b], ref, err))
    assert worst < 1e-11 * scale, (worst, scale)


@pytest.mark.parametrize("nwat", [32, 64])
def test_cluster_gmatrix_linear_in_density(require_gpu, nwat):
    _say("case: (H2O)%d 6-31g gmatrix linearity" % nwat)
    mol = systems.water_cluster(nwat, "6-31g")
    n = mol.nbf
    D1 = O.full_to_flat(_density(n, 1))
    D2 = O.full_to_flat(_density(n, 2))
    with hfk.Context(mol) as ctx:
        g1 = ctx.calcGmatrix(D1)
        g2 = ctx.calcGmatrix(D2)
        g12 = ctx.calcGmatrix(D1 - 2.0 * D2)
    scale = max(1.0, float(np.max(np.abs(g1))), float(np.max(np.abs(g2))))
    err = float(np.max(np.abs(g12 - (g1 - 2.0 * g2))))
    _say("  max|G(D1-2D2) - G(D1) + 2G(D2)| = %.2e (scale %.2e)" % (err, scale))
    assert err < 1e-12 * scale


def test_h2o32_gradient_fixed_density_fd(require_gpu):
    """hfk_gradient at (H2O)_32 size against central differences of
    E(R) = sum D H(R) + 1/2 sum D G(D; R) - sum W S(R) + V_nn(R) at a fixed
    non-converged D and W, on four coordinates, and translation invariance
    over all 96 atoms.  Accurate Boys function (the reference's table is
    piecewise, so its energy is not differentiable across grid cells) and
    the reference's primitive screen at 1e-20 (a primitive quartet that
    crosses it moves E by < 1e-20)."""
    _say("case: (H2O)32 6-31g fixed-density gradient vs finite differences")
    opts = dict(screen_mode=hfk.HFK_SCREEN_REFERENCE, screen_tau=1e-20, boys_mode=hfk.HFK_BOYS_ACCURATE)
    mol = systems.water_cluster(32, "6-31g")
    n = mol.nbf
    rng = np.random.default_rng(32)
    D = O.full_to_flat(_density(n, 5) + np.eye(n) * 0.5)
    W = O.full_to_flat(_density(n, 6))
    with hfk.Context(mol, hfk.options(**opts)) as ctx:
        g = ctx.gradient(D, W)

    def energy(R):
        m = mol.with_coords(R)
        with hfk.Context(m, hfk.options(**opts)) as c:
            H = O.flat_to_full(c.hcore(), n)
            G = O.flat_to_full(c.calcGmatrix(D), n)
            S = O.flat_to_full(c.mtxOverlap(), n)
        Df, Wf = O.flat_to_full(D, n), O.flat_to_full(W, n)
        return float((Df * H).sum() + 0.5 * (Df * G).sum() - (Wf * S).sum() + m.nuclear_repulsion())

    h = 1e-4
    coords = [(0, 0), (mol.natom - 1, 2)] + [(int(a), int(k)) for a, k in
                                               zip(rng.integers(1, mol.natom - 1, 2), rng.integers(0, 3, 2))]
    scale = max(1.0, float(np.max(np.abs(g))))
    for a, k in coords:
        R = mol.R.copy()
        R[a, k] += h
        ep = energy(R)
        R[a, k] -= 2 * h
        em = energy(R)
        fd = (ep - em) / (2 * h)
        _say("  atom %d axis %d: analytic % .10f fd % .10f |d| %.2e" % (a, k, g[a, k], fd, abs(g[a, k] - fd)))
        assert abs(g[a, k] - fd) < 2e-7 * scale
    resid = float(np.max(np.abs(g.sum(axis=0))))
    _say("  max|sum_A g_A| = %.2e (max|g| %.2e)" % (resid, scale))
    assert resid < 1e-9 * scale
