"""Independent one-electron integrals over Cartesian Gaussians by quadrature,
used only to pin the oracle's overlap, kinetic and nuclear-attraction
integrals for p and d functions.  No recursion is shared with the oracle:

  * overlap: per axis, int (x-A)^i (x-B)^j exp(-a(x-A)^2 - b(x-B)^2) dx by
    Gauss-Hermite quadrature about the product centre P (exact for these
    polynomial degrees);
  * kinetic: -1/2 d^2/dx^2 applied to the ket primitive analytically,
    T_x = b(2j+1) S(i,j) - 2 b^2 S(i,j+2) - j(j-1)/2 S(i,j-2), times the
    other two axes' overlaps;
  * nuclear attraction: 1/|r - C| = 2/sqrt(pi) int_0^inf exp(-t^2 |r-C|^2) dt
    with t^2 = p u^2 / (1 - u^2), which gives
        V = 2 pi / p exp(-mu AB^2) int_0^1 exp(-p u^2 PC^2) prod_x M_x(u) du,
    M_x(u) the moment of (x-A)^i (x-B)^j under a normalised Gaussian of
    centre P + u^2 (C - P) and variance (1 - u^2) / (2p), each moment by
    Gauss-Hermite quadrature; the u integral by adaptive quadrature.
"""
import math

import numpy as np
from scipy import integrate

_GH_X, _GH_W = np.polynomial.hermite.hermgauss(24)   # exact to degree 47


def _moment(i, j, A, B, center, var):
    """E[(x-A)^i (x-B)^j] under N(center, var) (Gauss-Hermite)."""
    x = center + math.sqrt(2.0 * var) * _GH_X
    return float(np.sum(_GH_W * (x - A) ** i * (x - B) ** j)) / math.sqrt(math.pi)


def _overlap_1d(i, j, A, B, a, b):
    p = a + b
    P = (a * A + b * B) / p
    return math.exp(-a * b / p * (A - B) ** 2) * math.sqrt(math.pi / p) * _moment(i, j, A, B, P, 1.0 / (2 * p))


def prim_overlap(A, la, a, B, lb, b):
    return math.prod(_overlap_1d(la[k], lb[k], A[k], B[k], a, b) for k in range(3))


def prim_kinetic(A, la, a, B, lb, b):
    s = [[_overlap_1d(la[k], lb[k] + d, A[k], B[k], a, b) if lb[k] + d >= 0 else 0.0 for d in (-2, 0, 2)]
         for k in range(3)]
    t = 0.0
    for k in range(3):
        j = lb[k]
        tk = b * (2 * j + 1) * s[k][1] - 2 * b * b * s[k][2] - 0.5 * j * (j - 1) * s[k][0]
        t += tk * math.prod(s[m][1] for m in range(3) if m != k)
    return t


def prim_nuclear(A, la, a, B, lb, b, C):
    """<a| 1/|r - C| |b> (positive; the Hamiltonian term is -Z times it)."""
    p = a + b
    P = [(a * A[k] + b * B[k]) / p for k in range(3)]
    AB2 = sum((A[k] - B[k]) ** 2 for k in range(3))
    PC2 = sum((P[k] - C[k]) ** 2 for k in range(3))

    def f(u):
        u2 = u * u
        var = (1.0 - u2) / (2.0 * p)
        m = 1.0
        for k in range(3):
            m *= _moment(la[k], lb[k], A[k], B[k], P[k] + u2 * (C[k] - P[k]), var)
        return math.exp(-p * u2 * PC2) * m

    val, _err = integrate.quad(f, 0.0, 1.0, epsabs=1e-15, epsrel=1e-13, limit=200)
    return 2.0 * math.pi / p * math.exp(-a * b / p * AB2) * val


def contracted(mol, i, j, kind, C=None):
    """Contracted integral between basis functions i and j of an hfk Molecule
    (coefficients as stored: primitive norms already folded in)."""
    A, B = mol.center[i], mol.center[j]
    la, lb = tuple(int(v) for v in mol.lxyz[i]), tuple(int(v) for v in mol.lxyz[j])
    tot = 0.0
    for pa in range(mol.prim_off[i], mol.prim_off[i + 1]):
        for pb in range(mol.prim_off[j], mol.prim_off[j + 1]):
            a, b = mol.expo[pa], mol.expo[pb]
            if kind == "S":
                v = prim_overlap(A, la, a, B, lb, b)
            elif kind == "T":
                v = prim_kinetic(A, la, a, B, lb, b)
            else:
                v = prim_nuclear(A, la, a, B, lb, b, C)
            tot += mol.coef[pa] * mol.coef[pb] * v
    return tot
