"""GPU parity of the dense FP64 pieces (MFMA GEMM, rocSOLVER eigen-solve)
against numpy on the same inputs, through the C ABI."""
import numpy as np
import pytest

import hfk
from hfk import systems
import hfk_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx13():
    mol = systems.water_cluster(1, "6-31g")      # 13 basis functions
    c = hfk.Context(mol)
    yield c
    c.close()


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return A @ A.T + n * np.eye(n)


def test_unitary_transf(require_gpu, ctx13):
    print("case: unitaryTransf n=13", flush=True)
    n = ctx13.n
    rng = np.random.default_rng(2)
    X = rng.standard_normal((n, n))
    F = rng.standard_normal((n, n))
    F = F + F.T
    got = ctx13.unitaryTransf(X, O.full_to_flat(F))
    np.testing.assert_allclose(got, X.T @ F @ X, atol=1e-12)


def test_density(require_gpu, ctx13):
    print("case: calcDensity n=13", flush=True)
    n = ctx13.n
    rng = np.random.default_rng(3)
    Cm = rng.standard_normal((n, n))
    for occ in (1, 5, n):
        got = O.flat_to_full(ctx13.calcDensity(Cm, occ), n)
        np.testing.assert_allclose(got, 2 * Cm[:, :occ] @ Cm[:, :occ].T, atol=1e-12)


def test_symm_ortho(require_gpu, ctx13):
    print("case: symmOrtho n=13", flush=True)
    n = ctx13.n
    S = _spd(n, 4)
    X = ctx13.symmOrtho(S)
    np.testing.assert_allclose(X, O.symm_ortho(S), atol=1e-12)
    np.testing.assert_allclose(X.T @ S @ X, np.eye(n), atol=1e-12)


def test_gemm_larger(require_gpu):
    print("case: unitaryTransf n=104", flush=True)
    mol = systems.water_cluster(8, "6-31g")
    with hfk.Context(mol) as c:
        n = c.n
        rng = np.random.default_rng(5)
        X = rng.standard_normal((n, n))
        F = rng.standard_normal((n, n))
        F = F + F.T
        got = c.unitaryTransf(X, O.full_to_flat(F))
        np.testing.assert_allclose(got, X.T @ F @ X, atol=1e-10)
        S = _spd(n, 6)
        Xs = c.symmOrtho(S)
        np.testing.assert_allclose(Xs.T @ S @ Xs, np.eye(n), atol=1e-11)


@pytest.mark.parametrize("system", ["h2", "water", "h2o8", "h2o16"])
def test_symm_ortho_sizes(require_gpu, system):
    """symmOrtho on both sides of the eigen-solver switch: rocSOLVER's Jacobi
    solver up to n = 32 (H2: 2, water/STO-3G: 7) and divide and conquer above
    ((H2O)_8: 104, (H2O)_16: 208).  X = S^-1/2 is unique, so it is compared
    with numpy's directly; S = I (one eigenvalue, n-fold degenerate) gives X = I."""
    mol = {"h2": lambda: systems.h2_szabo(), "water": lambda: systems.water("sto-3g"),
           "h2o8": lambda: systems.water_cluster(8, "6-31g"), "h2o16": lambda: systems.water_cluster(16, "6-31g")}[system]()
    with hfk.Context(mol) as c:
        n = c.n
        print("case: symmOrtho %s n=%d" % (system, n), flush=True)
        S = _spd(n, 8) / n
        X = c.symmOrtho(S)
        ref = O.symm_ortho(S)
        assert np.max(np.abs(X - ref)) < 1e-11 * np.max(np.abs(ref))
        np.testing.assert_allclose(X.T @ S @ X, np.eye(n), atol=1e-11)
        np.testing.assert_allclose(c.symmOrtho(np.eye(n)), np.eye(n), atol=1e-13)
