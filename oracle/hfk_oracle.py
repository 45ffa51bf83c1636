"""CPU oracle for the RHF integral path (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product (libhfk.so and the hfk
package) never imports it.

Pieces and what pins them (tests/test_oracle.py):

==========================  ====================================================
oracle piece                pinned by
==========================  ====================================================
GL-256 nodes/weights        golden vector extracted from PointsWeights.hs:25-31;
                            sum w = 2, exact integration of x^510
Boys table (reference)      F0 nodes = erf closed form; F_m(0) = 1/(2m+1);
                            node error vs the accurate Boys function reproduces
                            the deviation profile the reference documents
                            (Boys.hs:27-39 Mamedov table, F1 rel. err ~ 7e-6)
Boys (accurate)             Mamedov reference values (Boys.hs:27-39, Eq(3)),
                            downward recursion identity, mpmath quadrature
primitive ERI (MD)          (ss|ss) closed form; independent Obara-Saika code
                            (tests/os_eri.py) at 1e-13; 8-fold symmetry;
                            translation and rotation covariance
contracted ERI + screen     the screen restates IntegralsEvaluation.hs:
                            303-337; twoTermsERI against hand-derived per-axis
                            closed forms (s, p, d pairs; one-centre p-s pairs
                            are negative, sqrt gives NaN, the quartet is
                            kept); s-s pairs = 1/4 of the exact (ab|ab); the
                            water/STO-3G kept count (32 274 of 32 886)
                            recomputed from the closed forms alone
flat layout                 literal restatement of fourIndex2Flat
                            (LinearAlgebra.hs:266-283) vs the enumeration of
                            calcIntegrals (HartreeFock.hs:166-181), n <= 12
calcGmatrix                 brute force G from a dense 4-index tensor
S, T, V                     textbook H2 / HeH+ values (Szabo & Ostlund)
SCF energy                  Szabo & Ostlund H2 (-1.1167) and HeH+ (-2.860662)
memory and UB safety        every C entry point built with AddressSanitizer and
                            UBSan and run once (oracle/sanitize_check.c, with
                            its own symmetry, layout and G checks)
==========================  ====================================================
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libhfk_oracle.so")
NBOYS_M, NBOYS_X = 15, 1001

_lib = None


class ObBasis(C.Structure):
    _fields_ = [("nbf", C.c_int), ("center", C.POINTER(C.c_double)), ("lxyz", C.POINTER(C.c_int)),
                ("prim_off", C.POINTER(C.c_int)), ("coef", C.POINTER(C.c_double)), ("expo", C.POINTER(C.c_double))]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import subprocess
            subprocess.check_call(["make", "-C", HERE], stdout=subprocess.DEVNULL)
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        bp = C.POINTER(ObBasis)
        L.ob_gl256.argtypes = [dp, dp]
        L.ob_boys_node.argtypes = [C.c_double, C.c_double]
        L.ob_boys_node.restype = C.c_double
        L.ob_boys_table.argtypes = [dp]
        L.ob_boys_taylor.argtypes = [dp, C.c_int, C.c_double]
        L.ob_boys_taylor.restype = C.c_double
        L.ob_boys_exact.argtypes = [C.c_int, C.c_double]
        L.ob_boys_exact.restype = C.c_double
        L.ob_boys_asympt.argtypes = [C.c_double, C.c_double]
        L.ob_boys_asympt.restype = C.c_double
        L.ob_two_terms.argtypes = [dp, C.POINTER(C.c_int), C.c_double, C.c_double, dp, C.POINTER(C.c_int),
                                   C.c_double, C.c_double]
        L.ob_two_terms.restype = C.c_double
        L.ob_eri.argtypes = [bp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_long)]
        L.ob_eri.restype = C.c_double
        L.ob_eri_flat.argtypes = [bp, dp, C.c_int, C.c_int, dp, C.POINTER(C.c_long)]
        L.ob_gmatrix.argtypes = [C.c_long, dp, dp, dp]
        L.ob_flat_rank.argtypes = [C.c_long] * 5
        L.ob_flat_rank.restype = C.c_long
        L.ob_four_index_ref.argtypes = [C.c_long] * 5
        L.ob_four_index_ref.restype = C.c_long
        L.ob_flat_len.argtypes = [C.c_long]
        L.ob_flat_len.restype = C.c_long
        L.ob_index_dim2_to_flat.argtypes = [C.c_long] * 3
        L.ob_index_dim2_to_flat.restype = C.c_long
        L.ob_flat_get.argtypes = [C.c_long, dp, C.c_long, C.c_long, C.c_long, C.c_long]
        L.ob_flat_get.restype = C.c_double
        L.ob_eri_list.argtypes = [bp, dp, C.c_int, C.c_int, C.c_long, C.POINTER(C.c_int), dp]
        L.ob_overlap.argtypes = [bp, dp]
        L.ob_hcore.argtypes = [bp, dp, C.c_int, C.c_int, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


_TABLE = None


def boys_table():
    """generateGridBoys 0.1 (TableBoys.hs:58-63): [15][1001]."""
    global _TABLE
    if _TABLE is None:
        t = np.zeros(NBOYS_M * NBOYS_X)
        lib().ob_boys_table(_dp(t))
        _TABLE = t
    return _TABLE


def gl256():
    x, w = np.zeros(128), np.zeros(128)
    lib().ob_gl256(_dp(x), _dp(w))
    return x, w


def boys_taylor(m, x):
    return lib().ob_boys_taylor(_dp(boys_table()), int(m), float(x))


def boys_exact(m, x):
    return lib().ob_boys_exact(int(m), float(x))


class OracleBasis:
    """Keeps the numpy arrays of a hfk.basis.Molecule alive for the C oracle."""

    def __init__(self, mol):
        self.mol = mol
        self.center = np.ascontiguousarray(mol.center, dtype=np.float64)
        self.lxyz = np.ascontiguousarray(mol.lxyz, dtype=np.int32)
        self.prim_off = np.ascontiguousarray(mol.prim_off, dtype=np.int32)
        self.coef = np.ascontiguousarray(mol.coef, dtype=np.float64)
        self.expo = np.ascontiguousarray(mol.expo, dtype=np.float64)
        b = ObBasis()
        b.nbf = mol.nbf
        b.center = _dp(self.center)
        b.lxyz = self.lxyz.ctypes.data_as(C.POINTER(C.c_int))
        b.prim_off = self.prim_off.ctypes.data_as(C.POINTER(C.c_int))
        b.coef = _dp(self.coef)
        b.expo = _dp(self.expo)
        self.b = b


def eri(mol, i, j, k, l, exact=False, screen=True):
    """contracted4Centers (IntegralsEvaluation.hs:294-306) for CGFs i,j,k,l."""
    ob = OracleBasis(mol)
    kept = C.c_long(0)
    v = lib().ob_eri(C.byref(ob.b), _dp(boys_table()), int(exact), int(screen), i, j, k, l, C.byref(kept))
    return v, kept.value


def eri_flat(mol, exact=False, screen=True):
    """calcIntegrals (HartreeFock.hs:166-181)."""
    ob = OracleBasis(mol)
    n = mol.nbf
    out = np.zeros(lib().ob_flat_len(n))
    kept = C.c_long(0)
    lib().ob_eri_flat(C.byref(ob.b), _dp(boys_table()), int(exact), int(screen), _dp(out), C.byref(kept))
    return out, kept.value


def eri_list(mol, quartets, exact=False, screen=True):
    """contracted4Centers over a list of canonical quartets (OpenMP)."""
    ob = OracleBasis(mol)
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
    out = np.zeros(q.shape[0])
    lib().ob_eri_list(C.byref(ob.b), _dp(boys_table()), int(exact), int(screen), q.shape[0],
                      q.ctypes.data_as(C.POINTER(C.c_int)), _dp(out))
    return out


def random_canonical_quartets(n, count, seed=0):
    """Sample, with replacement, of canonical quartets i<=j, k<=l,
    (i,j)<=(k,l): two pair indices drawn uniformly and ordered, so quartets
    with (i,j) == (k,l) (a fraction 1/M of the list) have half the weight of
    the others."""
    rng = np.random.default_rng(seed)
    iu, ju = np.triu_indices(n)
    m = iu.size
    a = rng.integers(0, m, size=count)
    b = rng.integers(0, m, size=count)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    return np.stack([iu[lo], ju[lo], iu[hi], ju[hi]], axis=1).astype(np.int32)


def flat_rank(n, i, j, k, l):
    return lib().ob_flat_rank(n, i, j, k, l)


def four_index_ref(n, i, j, k, l):
    return lib().ob_four_index_ref(n, i, j, k, l)


def sort_keys(i, j, k, l):
    """sortKeys (HartreeFock.hs:153-156)."""
    a = sorted([i, j])
    b = sorted([k, l])
    return a + b if a <= b else b + a


def gmatrix(n, eri_flat_arr, D_flat):
    """calcGmatrix (HartreeFock.hs:225-241) from a flat ERI array."""
    G = np.zeros(n * (n + 1) // 2)
    lib().ob_gmatrix(n, _dp(np.ascontiguousarray(eri_flat_arr)), _dp(np.ascontiguousarray(D_flat)), _dp(G))
    return G


def overlap(mol):
    ob = OracleBasis(mol)
    S = np.zeros(mol.nbf * (mol.nbf + 1) // 2)
    lib().ob_overlap(C.byref(ob.b), _dp(S))
    return S


def hcore(mol, exact=False):
    ob = OracleBasis(mol)
    H = np.zeros(mol.nbf * (mol.nbf + 1) // 2)
    R = np.ascontiguousarray(mol.R, dtype=np.float64)
    Z = np.ascontiguousarray(mol.Z, dtype=np.float64)
    lib().ob_hcore(C.byref(ob.b), _dp(boys_table()), int(exact), mol.natom, _dp(R), _dp(Z), _dp(H))
    return H


# ----------------------------------------------------------------------
# flat <-> square helpers (LinearAlgebra.hs:200-253)
# ----------------------------------------------------------------------
def flat_to_full(flat, n):
    M = np.zeros((n, n))
    iu = np.triu_indices(n)
    M[iu] = flat
    M[(iu[1], iu[0])] = flat
    return M


def full_to_flat(M):
    n = M.shape[0]
    return M[np.triu_indices(n)].copy()


def eri_tensor(flat, n):
    """Dense 4-index tensor from the flat array (small n only)."""
    T = np.zeros((n, n, n, n))
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for l in range(n):
                    s = sort_keys(i, j, k, l)
                    T[i, j, k, l] = flat[flat_rank(n, *s)]
    return T


# ----------------------------------------------------------------------
# SCF driver restatement (HartreeFock.hs:70-135, DIIS.hs:55-137)
# ----------------------------------------------------------------------
def symm_ortho(S):
    """symmOrtho (BasisOrthogonalization.hs:26-35)."""
    s, U = np.linalg.eigh(S)
    return U @ np.diag(1.0 / np.sqrt(s)) @ U.T


def diagonal_hf(F_flat, X, occ):
    """diagonalHF (HartreeFock.hs:263-273)."""
    n = X.shape[0]
    Fp = X.T @ flat_to_full(F_flat, n) @ X
    e, Cp = np.linalg.eigh(Fp)
    Cm = X @ Cp
    D = 2.0 * Cm[:, :occ] @ Cm[:, :occ].T
    return Cm, full_to_flat(D), e


def gauss_elem(A, b):
    """gaussElem (GaussElimination.hs:42-112): no pivoting."""
    A = np.array(A, dtype=np.float64)
    b = np.array(b, dtype=np.float64)
    m = A.shape[0]
    # a zero pivot makes the rows non-finite; the check below turns that into None
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in range(m):
            A2, b2 = A.copy(), b.copy()
            for i in range(k + 1, m):
                alpha = A[i, k] / A[k, k]
                A2[i, k:] = A[i, k:] - alpha * A[k, k:]
                b2[i] = b[i] - alpha * b[k]
            A, b = A2, b2
    if not all(np.any(np.abs(A[i]) > 1e-6) for i in range(m)) or not np.all(np.isfinite(A)):
        return None
    x = np.zeros(m)
    for i in range(m - 1, -1, -1):
        x[i] = (b[i] - np.dot(x[i + 1:], A[i, i + 1:])) * (1.0 / A[i, i])
    return x


# --- the reference's index maps exactly as written (LinearAlgebra.hs:194-283),
# including indices outside the matrix, as the reference's Harris guess uses them
def ref_calc_elem(dim, w):
    """sum (takeZero w (iterate pred dim)): takeZero 0 = [0]; take of a negative count = []."""
    return sum(dim - t for t in range(w)) if w > 0 else 0


def ref_dim2_to_flat(dim, x, y):
    """indexDIM2toFlat (LinearAlgebra.hs:250-253)."""
    return ref_calc_elem(dim, x) + (y - x) if x <= y else ref_calc_elem(dim, y) + (x - y)


def ref_flat_to_dim2(dim, idx):
    """indexFlat2DIM2 (LinearAlgebra.hs:240-246)."""
    x, n = idx, 0
    while not x < dim - n:
        x, n = x - (dim - n), n + 1
    return n, idx + n - ref_calc_elem(dim, n)


def ref_four_index_to_flat(n, q):
    """fourIndex2Flat with totalFirstIndex/sumFirstIndex/selectPosition (LinearAlgebra.hs:255-283)."""
    tri = lambda m: (m * m + m) // 2
    first = lambda m: m + tri(m - 1) + sum(tri(m) - k for k in range(1, m))
    i, j, k, l = q
    a0 = sum(first(n - m) for m in range(i))
    if i == j == k:
        pos = l - i
    elif i == j:
        pos = (n - i) + ref_dim2_to_flat(n - (i + 1), k - (i + 1), l - (i + 1))
    else:
        pos = (n - i) + tri(n - 1 - i) + sum(tri(n - i) - m for m in range(1, j - i)) + \
            (ref_dim2_to_flat(n, k, l) - ref_dim2_to_flat(n, i, j))
    return a0 + pos


def reference_harris_g(eri, nflat, occ):
    """computeGuessIntegrals as written (HartreeFock.hs:137-149): the flat length
    as the matrix dimension and occ as the basis size.  Raises IndexError where
    the reference's bounds-checked `!` aborts."""
    G = np.zeros(nflat)
    for idx in range(nflat):
        i, j = ref_flat_to_dim2(nflat, idx)
        s = 0.0
        for k in range(occ):
            f1 = ref_four_index_to_flat(occ, sort_keys(i, j, k, k))
            f2 = ref_four_index_to_flat(occ, sort_keys(i, k, k, j))
            if not (0 <= f1 < len(eri) and 0 <= f2 < len(eri)):
                raise IndexError("reference Harris guess indexes the integral array at %d / %d" % (f1, f2))
            s += eri[f1] - 0.5 * eri[f2]
        G[idx] = 2.0 * s
    return G


def scf(mol, fock_G, H_flat, S_flat, guess="core", max_steps=30, ndiis=5, converge=1e-7,
        diis_start=0.1, eri=None):
    """scfHF/scfDIIS.  fock_G(D_flat) -> G_flat.  Returns dict(energy, steps, C, eps, D, F)."""
    n = mol.nbf
    occ = mol.occupied
    vnn = mol.nuclear_repulsion()
    S = flat_to_full(S_flat, n)
    X = symm_ortho(S)
    if guess == "core":
        _, D, _ = diagonal_hf(H_flat, X, occ)
    elif guess == "reference":   # harrisFunctional (HartreeFock.hs:127-134) as written
        _, D, _ = diagonal_hf(H_flat + reference_harris_g(eri, len(H_flat), occ), X, occ)
    else:
        D0 = np.zeros(n * (n + 1) // 2)
        for i in range(occ):
            D0[i * n - i * (i - 1) // 2] = 2.0
        _, D, _ = diagonal_hf(H_flat + fock_G(D0), X, occ)
    es, fs = [], []
    switch = False
    Hf = flat_to_full(H_flat, n)
    for step in range(max_steps):
        F = H_flat + fock_G(D)
        Ff, Df = flat_to_full(F, n), flat_to_full(D, n)
        e = X.T @ (Ff @ Df @ S - S @ Df @ Ff) @ X
        maxe = np.max(np.abs(e))
        Fuse = F
        if not switch:
            if step > 0 and maxe < diis_start:
                es, fs = [e] + es, [F] + fs
                if len(es) >= ndiis:
                    es, fs = es[:-1], fs[:-1]
                switch = True
        else:
            use_e, use_f = list(es), list(fs)
            while True:
                E = [e] + use_e
                Fs = [F] + use_f
                m = len(E)
                B = np.zeros((m + 1, m + 1))
                for i in range(m):
                    for j in range(m):
                        B[i, j] = np.trace(E[i].T @ E[j])
                B[m, :m] = B[:m, m] = -1.0
                rhs = np.zeros(m + 1)
                rhs[m] = -1.0
                cs = gauss_elem(B, rhs)
                if cs is not None:
                    Fuse = sum(c * f for c, f in zip(cs[:m], Fs))
                    break
                if not use_e:
                    break
                use_e, use_f = use_e[:-1], use_f[:-1]
            es, fs = [e] + use_e, [F] + use_f
            if len(es) >= ndiis:
                es, fs = es[:-1], fs[:-1]
        Cm, Dn, eps = diagonal_hf(Fuse, X, occ)
        Dnf = flat_to_full(Dn, n)
        etot = 0.5 * np.sum(Dnf * (Hf + flat_to_full(Fuse, n))) + vnn
        if step != 0 and maxe < converge:
            return dict(energy=etot, steps=step + 1, C=Cm, eps=eps, D=Dn, F=Fuse, max_error=maxe)
        D = Dn
    raise RuntimeError("SCF maxium steps exceeded")


def variational_energy(mol, D_flat, H_flat, F_flat):
    """1/2 sum_ij D_ij (H_ij + F_ij) + V_nn from flat upper triangles
    (variationalE, HartreeFock.hs:277-287, with F = F(D) rebuilt from the
    density itself: second order in the density error)."""
    n = mol.nbf
    Df, Hf, Ff = flat_to_full(D_flat, n), flat_to_full(H_flat, n), flat_to_full(F_flat, n)
    return 0.5 * float(np.sum(Df * (Hf + Ff))) + mol.nuclear_repulsion()


def scf_energy_exact(mol, converge=1e-10):
    """FP64 SCF energy of mol with the accurate Boys function and no screen:
    flat ERIs (calcIntegrals), DIIS from the second step, then the
    variational energy of the converged density with F(D) rebuilt."""
    S, H = overlap(mol), hcore(mol, exact=True)
    flat, _ = eri_flat(mol, exact=True, screen=False)
    n = mol.nbf
    r = scf(mol, lambda D: gmatrix(n, flat, D), H, S, max_steps=100, ndiis=6, converge=converge,
            diis_start=float("inf"))
    D = r["D"]
    return variational_energy(mol, D, H, H + gmatrix(n, flat, D)), r


def fd_gradient(mol, h=1e-4, converge=1e-10):
    """FP64 CPU reference gradient dE/dR (natom x 3, Eh/bohr): central
    differences at +-h bohr of scf_energy_exact along every coordinate.  The
    reference has no working gradient (HartreeFock/Derivatives.hs does not
    compile), so this is the oracle's gradient: the derivative of its own
    SCF energy, whose pieces (S, T, V, ERIs, the SCF driver) are pinned
    separately."""
    g = np.zeros((mol.natom, 3))
    for a in range(mol.natom):
        for k in range(3):
            R = mol.R.copy()
            R[a, k] += h
            ep, _ = scf_energy_exact(mol.with_coords(R), converge)
            R[a, k] -= 2 * h
            em, _ = scf_energy_exact(mol.with_coords(R), converge)
            g[a, k] = (ep - em) / (2 * h)
    return g
