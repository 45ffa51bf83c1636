"""Pins of the CPU oracle (no GPU).  See oracle/hfk_oracle.py's header for
what each test pins."""
import json
import math
import os

import numpy as np
import pytest

import hfk_oracle as O
from hfk import basis, systems

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_gl256_matches_reference_constants():
    g = json.load(open(os.path.join(GOLD, "gl256.json")))
    x, w = O.gl256()
    np.testing.assert_allclose(x, g["points256"], rtol=2e-16, atol=0)
    np.testing.assert_allclose(w, g["weights256"], rtol=4e-15, atol=0)
    # exactness: the symmetric rule integrates x^(2k) over [-1,1], k < 256
    for k in (0, 1, 10, 100, 255):
        val = 2.0 * np.sum(w * x ** (2 * k))
        assert abs(val - 2.0 / (2 * k + 1)) < 1e-13


def test_boys_table_f0_nodes_are_erf_closed_form():
    t = O.boys_table()
    for i in (1, 7, 100, 499, 500):
        x = 0.1 * i
        ref = 0.5 * math.sqrt(math.pi / x) * math.erf(math.sqrt(x)) if x <= 50 else 0.5 * math.sqrt(math.pi / x)
        assert abs(t[i] - ref) <= 2e-16 * ref
    assert t[0] == 1.0


def test_boys_table_deviation_profile():
    """The reference's m>=1 nodes come from a 256-point Gauss-Legendre
    quadrature of t^(m-1/2) e^(-xt) (Boys.hs:60-66,106-118); its error
    is largest for m=1 (integrable t^(1/2) singularity).  Independent
    accurate values: mpmath (via tests/os_eri.boys)."""
    import os_eri
    t = O.boys_table()
    worst = {}
    for m in (1, 2, 3, 6, 10, 14):
        errs = []
        for i in (0, 3, 17, 150, 400, 999):
            ex = os_eri.boys(m, 0.1 * i)
            errs.append(abs(t[m * 1001 + i] - ex) / ex)
        worst[m] = max(errs)
    assert 1e-6 < worst[1] < 7e-6         # F1 nodes: ~6.9e-6 relative
    assert worst[2] < 1e-8 and worst[3] < 1e-10
    assert max(worst[m] for m in (6, 10, 14)) < 1e-13


def test_boys_taylor_step_matches_formula():
    t = O.boys_table()
    for m, x in ((0, 0.123), (2, 7.77), (4, 99.95), (3, 0.05)):
        k = math.floor(x / 0.1)
        xk = k * 0.1
        dx = x - xk
        f = [t[(m + j) * 1001 + k] for j in range(6)]
        ref = f[0] - f[1] * dx + 0.5 * f[2] * dx ** 2 - f[3] * dx ** 3 / 6 + f[4] * dx ** 4 / 24 - f[5] * dx ** 5 / 120
        assert abs(O.boys_taylor(m, x) - ref) < 1e-15
    # above x_k = 100 the asymptotic formula at the node is expanded
    x = 150.04
    xk = math.floor(x / 0.1) * 0.1
    a = [O.lib().ob_boys_asympt(float(j), xk) for j in range(6)]
    dx = x - xk
    ref = sum((-dx) ** j / math.factorial(j) * a[j] for j in range(6))
    assert abs(O.boys_taylor(0, x) - ref) < 1e-15 * abs(ref) + 1e-18


def test_boys_exact_against_mamedov():
    rows = json.load(open(os.path.join(GOLD, "mamedov.json")))
    for m, x, eq3, eq6, ours in rows:
        if m == 20 and x == 30:
            eq3 = 1.37585444267909e-13   # the comment table prints e-3 (typo); Eq(6) agrees
        v = O.boys_exact(int(m), x)
        assert abs(v - eq3) <= 1e-12 * eq3, (m, x, v, eq3)


def test_boys_exact_downward_recursion():
    for T in (0.01, 1.3, 12.0, 45.0):
        for m in range(0, 12):
            lhs = O.boys_exact(m, T)
            rhs = (2 * T * O.boys_exact(m + 1, T) + math.exp(-T)) / (2 * m + 1)
            assert abs(lhs - rhs) <= 1e-14 * lhs


def _prim_mol(centers, ls, exps):
    atoms = []
    for c, l, e in zip(centers, ls, exps):
        ft = {(0, 0, 0): "S", (1, 0, 0): "Px", (0, 1, 0): "Py", (0, 0, 1): "Pz", (2, 0, 0): "Dxx",
              (1, 1, 0): "Dxy", (1, 0, 1): "Dxz", (0, 2, 0): "Dyy", (0, 1, 1): "Dyz", (0, 0, 2): "Dzz"}[tuple(l)]
        atoms.append(("H", tuple(c), [(ft, [(1.0, e)])]))
    return basis.Molecule(atoms)


def test_ssss_closed_form():
    A, B, C, D = (0, 0, 0), (0.3, -0.2, 0.9), (1.5, 0.4, -0.7), (-0.6, 1.1, 0.2)
    a, b, c, d = 0.8, 1.7, 0.45, 3.1
    mol = _prim_mol([A, B, C, D], [(0, 0, 0)] * 4, [a, b, c, d])
    v, _ = O.eri(mol, 0, 1, 2, 3, exact=True, screen=False)
    p, q = a + b, c + d
    P = np.array([(a * A[i] + b * B[i]) / p for i in range(3)])
    Q = np.array([(c * C[i] + d * D[i]) / q for i in range(3)])
    T = p * q / (p + q) * np.sum((P - Q) ** 2)
    ref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * math.exp(
        -a * b / p * np.sum((np.array(A) - B) ** 2) - c * d / q * np.sum((np.array(C) - D) ** 2)) * O.boys_exact(0, T)
    assert abs(v - ref) < 1e-14


@pytest.mark.parametrize("ls", [
    ((1, 0, 0), (0, 0, 0), (0, 0, 0), (0, 0, 0)),
    ((0, 1, 0), (0, 0, 1), (1, 0, 0), (0, 0, 0)),
    ((1, 0, 0), (1, 0, 0), (0, 1, 0), (0, 1, 0)),
    ((2, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 0)),
    ((1, 1, 0), (0, 0, 2), (0, 1, 1), (2, 0, 0)),
])
def test_md_oracle_vs_independent_obara_saika(ls):
    import os_eri
    A, B, C, D = (0.1, 0, 0), (0.3, -0.2, 0.9), (1.5, 0.4, -0.7), (-0.6, 1.1, 0.2)
    ex = (0.8, 1.7, 0.45, 3.1)
    mol = _prim_mol([A, B, C, D], ls, ex)
    v, _ = O.eri(mol, 0, 1, 2, 3, exact=True, screen=False)
    ref = os_eri.prim_eri(A, ls[0], ex[0], B, ls[1], ex[1], C, ls[2], ex[2], D, ls[3], ex[3])
    assert abs(v - ref) < 1e-13 * max(1.0, abs(ref))


def test_eri_permutation_symmetry_and_translation():
    mol = systems.water("sto-3g")
    flat, _ = O.eri_flat(mol, exact=True, screen=False)
    n = mol.nbf
    T = O.eri_tensor(flat, n)
    # flat lookups are symmetric by construction; check against direct evaluation
    for (i, j, k, l) in ((0, 1, 2, 3), (2, 5, 6, 1), (4, 4, 3, 0)):
        v, _ = O.eri(mol, i, j, k, l, exact=True, screen=False)
        for perm in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (k, l, i, j), (l, k, j, i)):
            w, _ = O.eri(mol, *perm, exact=True, screen=False)
            assert abs(w - v) < 1e-14
        assert abs(T[i, j, k, l] - v) < 1e-14
    shifted = mol.with_coords(mol.R + np.array([0.7, -1.3, 2.1]))
    # the CGF centres move with the atoms
    f2, _ = O.eri_flat(systems.build_molecule(["O", "H", "H"], mol.R + np.array([0.7, -1.3, 2.1]), "sto-3g"),
                       exact=True, screen=False)
    assert np.max(np.abs(f2 - flat)) < 1e-12
    del shifted


def test_p_block_rotation_covariance():
    """(p_i s|s s) transforms as a vector under rotation of all centres."""
    rng = np.random.default_rng(3)
    Rm = systems.quat_matrix(rng.standard_normal(4))
    cen = np.array([[0.1, 0.2, 0.3], [1.0, -0.4, 0.5], [-0.7, 0.9, 0.1]])
    ex = (0.9, 0.6, 1.3)

    def block(cs):
        vals = []
        for l in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
            mol = _prim_mol([cs[0], cs[1], cs[2], cs[2]], [l, (0, 0, 0), (0, 0, 0), (0, 0, 0)],
                            [ex[0], ex[1], ex[2], ex[2]])
            vals.append(O.eri(mol, 0, 1, 2, 3, exact=True, screen=False)[0])
        return np.array(vals)

    v0 = block(cen)
    v1 = block(cen @ Rm.T)
    np.testing.assert_allclose(v1, Rm @ v0, atol=1e-13)


def test_flat_layout_literal_vs_enumeration():
    for n in range(1, 11):
        pos = 0
        for i in range(n):
            for j in range(i, n):
                for k in range(i, n):
                    for l in range(k, n):
                        if O.sort_keys(i, j, k, l) != [i, j, k, l]:
                            continue
                        assert O.flat_rank(n, i, j, k, l) == pos
                        assert O.four_index_ref(n, i, j, k, l) == pos
                        pos += 1
        assert pos == (n * (n + 1) // 2) * (n * (n + 1) // 2 + 1) // 2


def test_gmatrix_against_dense_contraction():
    mol = systems.water("sto-3g")
    n = mol.nbf
    flat, _ = O.eri_flat(mol)
    T = O.eri_tensor(flat, n)
    rng = np.random.default_rng(5)
    A = rng.standard_normal((n, n))
    Df = A + A.T
    G = O.gmatrix(n, flat, O.full_to_flat(Df))
    Gref = np.einsum("abkl,lk->ab", T, Df) - 0.5 * np.einsum("alkb,lk->ab", T, Df)
    np.testing.assert_allclose(O.flat_to_full(G, n), Gref, atol=1e-12)


def test_reference_screen_is_active_for_water():
    mol = systems.water("sto-3g")
    flat, kept = O.eri_flat(mol)
    flat_ns, kept_ns = O.eri_flat(mol, screen=False)
    assert kept_ns == 7 * 0 + 32886            # 406 quartets x 81 primitive quartets
    assert kept == 32274                       # primitive quartets passing !(q q' < 1e-8)
    d = np.max(np.abs(flat - flat_ns))
    assert 1e-8 < d < 1e-6


def _two_terms_closed_form(A, la, c1, a, B, lb, c2, b):
    """twoTermsERI (IntegralsEvaluation.hs:318-337) written out by hand:
    2 c1^2 c2^2 (pi/2p)^(5/2) (sum_tuv E_tuv)(sum_tuv (-1)^(t+u+v) E_tuv).
    E_tuv = E^x_t E^y_u E^z_v and E^x_0 of (0,0) is exp(-mu X_AB^2), so both
    sums are exp(-mu AB^2) times products of per-axis sums S+- = sum_t
    (+-1)^t E^ij_t / E^00_0.  From E^{i+1,j}_t = E^ij_{t-1}/2p + X_PA E^ij_t
    + (t+1) E^ij_{t+1} (and the same in j with X_PB), with h = 1/2p:
    (0,0) 1; (1,0) X_PA +- h; (2,0) (X_PA +- h)^2 + h; (1,1)
    (X_PA +- h)(X_PB +- h) + h; (0,j) as (j,0) with X_PB."""
    p = a + b
    h = 0.5 / p
    P = (a * np.asarray(A) + b * np.asarray(B)) / p
    prod = 1.0
    for x in range(3):
        i, j = la[x], lb[x]
        xa, xb = P[x] - A[x], P[x] - B[x]
        if (i, j) == (0, 0):
            sp = sm = 1.0
        elif i + j == 1:
            xx = xa if i else xb
            sp, sm = xx + h, xx - h
        elif (i, j) in ((2, 0), (0, 2)):
            xx = xa if i else xb
            sp, sm = (xx + h) ** 2 + h, (xx - h) ** 2 + h
        elif (i, j) == (1, 1):
            sp, sm = (xa + h) * (xb + h) + h, (xa - h) * (xb - h) + h
        else:
            raise ValueError("closed forms cover i + j <= 2 per axis")
        prod *= sp * sm
    ab2 = float(np.sum((np.asarray(A) - np.asarray(B)) ** 2))
    return 2.0 * c1 ** 2 * c2 ** 2 * (math.pi / (2 * p)) ** 2.5 * math.exp(-2.0 * a * b / p * ab2) * prod


def _two_terms(A, la, c1, a, B, lb, c2, b):
    import ctypes as C
    ia, ib = (np.ascontiguousarray(l, dtype=np.int32) for l in (la, lb))
    ra, rb = (np.ascontiguousarray(r, dtype=np.float64) for r in (A, B))
    return O.lib().ob_two_terms(O._dp(ra), ia.ctypes.data_as(C.POINTER(C.c_int)), c1, a,
                                O._dp(rb), ib.ctypes.data_as(C.POINTER(C.c_int)), c2, b)


def test_screen_pair_quantity_closed_forms():
    """The oracle's twoTermsERI against the hand-derived per-axis closed
    forms, for s, p and d pairs with i + j <= 2 per axis, on distinct and
    coincident centres (one-centre p-s pairs come out negative, so the
    reference's sqrt gives NaN and `not (NaN < 1e-8)` keeps the quartet).
    For an s-s pair it is a quarter of the exact (ab|ab) of the independent
    Obara-Saika code: the reference's constant 2 (pi/2p)^(5/2) against the
    true 2 pi^(5/2) / (p^2 sqrt(2p))."""
    import os_eri
    rng = np.random.default_rng(318)
    shapes = [((0, 0, 0), (0, 0, 0)), ((1, 0, 0), (0, 0, 0)), ((0, 0, 0), (0, 1, 0)), ((0, 0, 1), (0, 0, 1)),
              ((1, 0, 0), (0, 1, 0)), ((2, 0, 0), (0, 0, 0)), ((0, 0, 0), (0, 2, 0)), ((1, 1, 0), (0, 0, 0)),
              ((1, 0, 1), (0, 1, 0)), ((0, 1, 1), (1, 0, 0))]
    negative = 0
    for la, lb in shapes:
        for same_centre in (False, True):
            A = rng.standard_normal(3)
            B = A.copy() if same_centre else A + rng.standard_normal(3) * 0.8
            a, b = rng.uniform(0.2, 8.0, 2)
            c1, c2 = rng.uniform(0.3, 2.0, 2)
            ref = _two_terms_closed_form(A, la, c1, a, B, lb, c2, b)
            got = _two_terms(A, la, c1, a, B, lb, c2, b)
            assert abs(got - ref) <= 1e-13 * max(abs(ref), 1e-300), (la, lb, same_centre, got, ref)
            negative += ref < 0
    assert negative > 0
    A, B = [0.1, -0.3, 0.2], [0.5, 0.4, -0.6]
    a, b, c1, c2 = 1.3, 0.45, 0.7, 1.1
    exact = os_eri.prim_eri(A, (0, 0, 0), a, B, (0, 0, 0), b, A, (0, 0, 0), a, B, (0, 0, 0), b)
    assert abs(_two_terms(A, (0, 0, 0), c1, a, B, (0, 0, 0), c2, b) - 0.25 * c1 ** 2 * c2 ** 2 * exact) < 1e-14


def test_screen_keeps_32274_water_quartets_by_closed_forms():
    """The reference screen's decision, recomputed from the closed forms
    alone for every primitive quartet of the 406 canonical water/STO-3G
    quartets: keep unless sqrt(q_ab) sqrt(q_cd) < 1e-8, NaN kept
    (IntegralsEvaluation.hs:303-316).  Matches the oracle's count."""
    mol = systems.water("sto-3g")
    n = mol.nbf
    po = mol.prim_off
    q = {}
    for i in range(n):
        for j in range(n):
            for g1 in range(po[i], po[i + 1]):
                for g2 in range(po[j], po[j + 1]):
                    v = _two_terms_closed_form(mol.center[i], mol.lxyz[i], mol.coef[g1], mol.expo[g1],
                                               mol.center[j], mol.lxyz[j], mol.coef[g2], mol.expo[g2])
                    q[g1, g2] = math.sqrt(v) if v >= 0 else float("nan")
    kept = nan_kept = 0
    pairs = [(i, j) for i in range(n) for j in range(i, n)]
    for x, (i, j) in enumerate(pairs):
        for (k, l) in pairs[x:]:
            for g1 in range(po[i], po[i + 1]):
                for g2 in range(po[j], po[j + 1]):
                    for g3 in range(po[k], po[k + 1]):
                        for g4 in range(po[l], po[l + 1]):
                            val = q[g1, g2] * q[g3, g4]
                            if not (val < 1e-8):
                                kept += 1
                                nan_kept += val != val
    assert kept == O.eri_flat(mol)[1] == 32274
    assert nan_kept > 0


def test_textbook_h2_and_heh():
    """Szabo & Ostlund, STO-3G: H2 (R=1.4, zeta=1.24) and HeH+ (R=1.4632)."""
    h2 = systems.h2_szabo()
    S = O.overlap(h2)
    assert abs(S[1] - 0.6593) < 5e-5
    flat, _ = O.eri_flat(h2, exact=True, screen=False)
    g = lambda *ix: O.lib().ob_flat_get(2, O._dp(flat), *ix)
    assert abs(g(0, 0, 0, 0) - 0.7746) < 5e-5
    assert abs(g(0, 0, 1, 1) - 0.5697) < 5e-5
    assert abs(g(1, 0, 1, 0) - 0.2970) < 5e-5
    assert abs(g(1, 0, 0, 0) - 0.4441) < 5e-5
    r = O.scf(h2, lambda D: O.gmatrix(2, flat, D), O.hcore(h2, exact=True), S)
    assert abs(r["energy"] - (-1.1167)) < 5e-5
    heh = systems.heh_szabo()
    fl2, _ = O.eri_flat(heh, exact=True, screen=False)
    r2 = O.scf(heh, lambda D: O.gmatrix(2, fl2, D), O.hcore(heh, exact=True), O.overlap(heh))
    assert abs(r2["energy"] - (-2.860662)) < 5e-6


def test_water_sto3g_energy_and_guess_independence():
    mol = systems.water("sto-3g")
    flat, _ = O.eri_flat(mol)
    S, H = O.overlap(mol), O.hcore(mol)
    g = lambda D: O.gmatrix(mol.nbf, flat, D)
    e1 = O.scf(mol, g, H, S)["energy"]
    e2 = O.scf(mol, g, H, S, guess="core")["energy"]
    assert abs(e1 - e2) < 1e-9
    assert -75.0 < e1 < -74.9


def test_gauss_elem_reference_example():
    # GaussElimination.hs:115-119: q1 x = v1 has the solution (1, -2, 3)
    q1 = [[4.0, -2.0, 1.0], [-2.0, 4.0, -2.0], [1.0, -2.0, 4.0]]
    x = O.gauss_elem(q1, [11.0, -16.0, 17.0])
    np.testing.assert_allclose(x, [1, -2, 3], atol=1e-14)


def test_reference_index_maps_on_valid_indices():
    """The literal restatements of indexFlat2DIM2 / indexDIM2toFlat /
    fourIndex2Flat used by the reference Harris guess agree, on indices inside
    the matrix, with the C restatement and the calcIntegrals enumeration."""
    for dim in range(1, 9):
        for idx in range(dim * (dim + 1) // 2):
            i, j = O.ref_flat_to_dim2(dim, idx)
            assert 0 <= i <= j < dim and O.ref_dim2_to_flat(dim, i, j) == idx
    for n in range(1, 6):
        r = 0
        for i in range(n):
            for j in range(i, n):
                for k in range(n):
                    for l in range(k, n):
                        if (i, j) > (k, l):
                            continue
                        q = [i, j, k, l]
                        assert O.ref_four_index_to_flat(n, q) == O.four_index_ref(n, i, j, k, l) == O.flat_rank(n, i, j, k, l)


def test_reference_index_maps_outside_the_matrix_by_hand():
    """Values of the reference's maps at the arguments the Harris guess hands
    them for water/STO-3G (flat length 28 as the matrix dimension, occupied =
    5 as the basis size), evaluated by hand from LinearAlgebra.hs:194-283:

    indexFlat2DIM2 28 idx: idx < 28 - 0 gives x = 0, y = idx - calcElem 0
      = idx - sum [0] = idx, so every row index is 0.
    fourIndex2Flat 5 [0,0,0,27]: a0 = 0; i = j = k, so l - i = 27.
    fourIndex2Flat 5 [0,0,3,3]: i = j only: (5 - 0) + indexDIM2toFlat 4 2 2
      = 5 + (4 + 3) = 12.
    fourIndex2Flat 5 [0,27,3,3]: 5 + sumIndexTriang 4 (= 10)
      + sum_{m=1..26} (15 - m) (= 39) + (indexDIM2toFlat 5 3 3 (= 12)
      - indexDIM2toFlat 5 0 27 (= 27)) = 39.
    fourIndex2Flat 5 [0,3,3,27]: 5 + 10 + (14 + 13) + ((12 + 24) - 3) = 75.
    fourIndex2Flat 5 [1,1,1,9]: a0 = sumFirstIndex 5 = 5 + 10 + (14 + 13 +
      12 + 11) = 65; i = j = k, so 65 + (9 - 1) = 73."""
    assert [O.ref_flat_to_dim2(28, idx) for idx in range(28)] == [(0, idx) for idx in range(28)]
    for q, want in (([0, 0, 0, 27], 27), ([0, 0, 3, 3], 12), ([0, 27, 3, 3], 39), ([0, 3, 3, 27], 75),
                    ([1, 1, 1, 9], 73)):
        assert O.ref_four_index_to_flat(5, q) == want, (q, want)
    assert O.sort_keys(0, 27, 0, 0) == [0, 0, 0, 27]   # [0,27] > [0,0], so the pairs swap
    assert O.sort_keys(0, 3, 3, 27) == [0, 3, 3, 27]


def test_reference_harris_guess_index_range():
    """computeGuessIntegrals as written: in range for water/STO-3G (n=7,
    occ=5), outside the integral array for (H2O)_8/6-31G (n=104, occ=40),
    where the reference's bounds-checked `!` aborts."""
    n, occ = 7, 5
    M = n * (n + 1) // 2
    eri = np.arange(M * (M + 1) // 2, dtype=float)
    G = O.reference_harris_g(eri, M, occ)
    assert G.shape == (M,) and np.all(np.isfinite(G))
    n, occ = 104, 40
    M = n * (n + 1) // 2
    with pytest.raises(IndexError):
        O.reference_harris_g(np.zeros(M * (M + 1) // 2), M, occ)


def test_oracle_matches_its_committed_fixtures():
    """Regression pins (tests/golden/make_eri_golden.py): the Boys table's
    SHA-256 and sampled nodes; every water/STO-3G ERI and the number of
    primitive quartets the reference screen keeps; 10^4 seeded (H2O)_8/6-31G
    ERIs.  Bit for bit: the oracle is deterministic."""
    import hashlib
    bt = json.load(open(os.path.join(GOLD, "boys_table.json")))
    t = np.ascontiguousarray(O.boys_table(), dtype="<f8")
    assert hashlib.sha256(t.tobytes()).hexdigest() == bt["sha256"]
    grid = t.reshape(bt["shape"])
    for m, i, v in bt["samples"]:
        assert grid[m, i] == float(v)
    eg = json.load(open(os.path.join(GOLD, "eri_golden.json")))
    water = systems.water("sto-3g")
    flat, kept = O.eri_flat(water)
    assert kept == eg["water_sto3g"]["kept_primitive_quartets"]
    np.testing.assert_array_equal(flat, np.array([float(v) for v in eg["water_sto3g"]["flat"]]))
    s = eg["h2o8_631g"]
    mol = systems.water_cluster(8, "6-31g")
    q = np.array(s["quartets"], dtype=np.int32).reshape(-1, 4)
    np.testing.assert_array_equal(q, O.random_canonical_quartets(mol.nbf, len(q), seed=s["seed"]))
    np.testing.assert_array_equal(O.eri_list(mol, q), np.array([float(v) for v in s["values"]]))


def test_oracle_fd_gradient_properties():
    """The oracle's FP64 gradient (central differences of its exact-Boys,
    unscreened SCF energy) on water/6-31G: translation invariant, zero
    normal to the molecular plane (x), the two H atoms mirror images in y,
    and unchanged to 1e-7 when the step is halved."""
    from hfk import systems
    mol = systems.water("6-31g")
    g = O.fd_gradient(mol, h=1e-4)
    assert np.max(np.abs(g.sum(axis=0))) < 1e-7
    assert np.max(np.abs(g[:, 0])) < 1e-7
    assert abs(g[1, 1] + g[2, 1]) < 1e-7 and abs(g[1, 2] - g[2, 2]) < 1e-7
    assert np.max(np.abs(g)) > 1e-3   # not at the equilibrium geometry
    g2 = O.fd_gradient(mol, h=5e-5)
    assert np.max(np.abs(g - g2)) < 1e-7


def _onee_pairs(mol):
    """Function pairs covering s-p, p-p, s-d, p-d and d-d on one centre and
    on two centres (water/6-31G*: O 1s 2s 2p 3s 3p 6d, H 1s 2s)."""
    L = mol.lxyz.sum(axis=1)
    at = mol.atom_of_bf
    pairs = []
    for li, lj in ((0, 1), (1, 1), (0, 2), (1, 2), (2, 2), (0, 0)):
        for same in (True, False):
            cand = [(i, j) for i in range(mol.nbf) for j in range(i, mol.nbf)
                    if {L[i], L[j]} == {li, lj} and (at[i] == at[j]) == same]
            pairs += cand[:: max(1, len(cand) // 4)][:4]
    return pairs


def test_oracle_overlap_kinetic_p_d_against_quadrature():
    """mtxOverlap (IntegralsEvaluation.hs:169-221) and the kinetic part of
    hcore (tab, :263-280) for p and d functions against Gauss-Hermite
    quadrature of the same primitives (tests/os_onee.py, no recursion in
    common).  T is hcore with every nuclear charge set to zero.  The
    reference forces S_ii = 1; the true self-overlaps of the normalised
    contractions differ from 1 by less than 1e-5."""
    import copy
    import os_onee
    from hfk import systems
    mol = systems.water("6-31g*")
    n = mol.nbf
    S = O.flat_to_full(O.overlap(mol), n)
    m0 = copy.copy(mol)
    m0.Z = np.zeros_like(mol.Z)
    T = O.flat_to_full(O.hcore(m0, exact=True), n)
    pairs = _onee_pairs(mol)
    assert len(pairs) >= 30
    for i, j in pairs:
        s_ref = os_onee.contracted(mol, i, j, "S")
        t_ref = os_onee.contracted(mol, i, j, "T")
        if i != j:
            assert abs(S[i, j] - s_ref) < 1e-13, (i, j, S[i, j], s_ref)
        assert abs(T[i, j] - t_ref) < 1e-11 * max(1.0, abs(t_ref)), (i, j, T[i, j], t_ref)
    for i in range(n):
        assert S[i, i] == 1.0
        assert abs(os_onee.contracted(mol, i, i, "S") - 1.0) < 1e-5


def test_oracle_nuclear_attraction_p_d_against_quadrature():
    """The nuclear-attraction part of hcore (vijHermite / mcMurchie,
    IntegralsEvaluation.hs:385-414) for p and d functions against the
    Gaussian-transform quadrature of tests/os_onee.py: with only the first
    H nucleus charged (Z = 1), hcore - T = -V_H."""
    import copy
    import os_onee
    from hfk import systems
    mol = systems.water("6-31g*")
    n = mol.nbf
    m0 = copy.copy(mol)
    m0.Z = np.zeros_like(mol.Z)
    mh = copy.copy(mol)
    mh.Z = np.zeros_like(mol.Z)
    mh.Z[1] = 1.0
    V = -(O.flat_to_full(O.hcore(mh, exact=True), n) - O.flat_to_full(O.hcore(m0, exact=True), n))
    for i, j in _onee_pairs(mol)[::2]:
        v_ref = os_onee.contracted(mol, i, j, "V", C=mol.R[1])
        assert abs(V[i, j] - v_ref) < 1e-10 * max(1.0, abs(v_ref)), (i, j, V[i, j], v_ref)


def test_quadrature_one_electron_closed_forms():
    """The quadrature code itself against closed forms for normalised
    primitives on one centre: <s|s> = 1, <s|T|s> = 3a/2, <px|T|px> = 5a/2,
    <dxy|T|dxy> = 7a/2, <dxx|T|dxx> = 13a/6 (x^2 carries an s admixture:
    7a/6 along x plus a/2 along y and z); and <s_a|1/r|s_a> = 2 sqrt(2a/pi) (nucleus at the
    centre), <s_a|1/r_C|s_a> = erf(sqrt(2a) R)/R (nucleus at distance R)."""
    import math
    import os_onee
    a, O3 = 0.8, (0.0, 0.0, 0.0)

    def nrm(l):
        return 1.0 / math.sqrt(os_onee.prim_overlap(O3, l, a, O3, l, a))

    for l, t in (((0, 0, 0), 1.5), ((1, 0, 0), 2.5), ((1, 1, 0), 3.5), ((2, 0, 0), 13.0 / 6.0)):
        c = nrm(l) ** 2
        assert abs(os_onee.prim_kinetic(O3, l, a, O3, l, a) * c - t * a) < 1e-13
    c = nrm((0, 0, 0)) ** 2
    assert abs(os_onee.prim_nuclear(O3, (0, 0, 0), a, O3, (0, 0, 0), a, O3) * c - 2 * math.sqrt(2 * a / math.pi)) < 1e-12
    R = 1.7
    v = os_onee.prim_nuclear(O3, (0, 0, 0), a, O3, (0, 0, 0), a, (0.0, 0.0, R)) * c
    assert abs(v - math.erf(math.sqrt(2 * a) * R) / R) < 1e-12


def test_water_sto3g_published_energy():
    """Water/STO-3G at the geometry of T. D. Crawford's Hartree-Fock
    programming project #3 (bohr): published E_nuc = 8.002367061810450 and
    E_SCF = -74.942079928192 Eh (exact integrals, normalised contractions).
    With each contraction scaled to unit self-overlap (the quadrature of
    tests/os_onee.py; the stored STO-3G coefficients are normalised to
    ~3e-8) the oracle's exact-Boys, unscreened SCF reproduces the printed
    value.  This pins S, T, V for s and p functions, the ERIs and the SCF
    driver together."""
    import copy
    import math
    import os_onee
    from hfk import basis
    xyz = [(0.0, -0.143225816552, 0.0), (1.638036840407, 1.136548822547, 0.0), (-1.638036840407, 1.136548822547, 0.0)]
    mol = basis.build_molecule(["O", "H", "H"], xyz, "sto-3g")
    assert abs(mol.nuclear_repulsion() - 8.002367061810450) < 1e-11
    e_stored, _ = O.scf_energy_exact(mol)
    mol = copy.deepcopy(mol)
    for i in range(mol.nbf):
        s = os_onee.contracted(mol, i, i, "S")
        assert abs(s - 1.0) < 1e-7
        mol.coef[mol.prim_off[i]:mol.prim_off[i + 1]] /= math.sqrt(s)
    e, r = O.scf_energy_exact(mol, converge=1e-11)
    assert abs(e - (-74.942079928192)) < 1e-10
    # the reference's convention (coefficients as stored, S_ii := 1) moves E by the self-overlap defect only
    assert abs(e_stored - e) < 1e-6


def test_oracle_under_address_and_undefined_behaviour_sanitizers():
    """Every C entry point of the oracle, built with AddressSanitizer and
    UBSan (oracle/Makefile target `sanitize`, single-threaded), runs without
    a report; the same run checks F_m(0) = 1/(2m+1) to the reference nodes'
    error, table Boys against the accurate function, Gauss-Legendre moments,
    the 8-fold ERI symmetry over s/p/d, the flat layout maps and G against
    its dense contraction (oracle/sanitize_check.c)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("needs gcc and make")
    odir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    b = subprocess.run(["make", "-C", odir, "sanitize"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(odir, "_build", "sanitize_check")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert "oracle sanitizer run ok" in r.stdout
