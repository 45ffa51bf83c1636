"""hfk_eri_flat_noalloc: the calcIntegrals export into caller-owned device
memory with no allocation, against hfk_eri_flat on the same context.  The
same kernels run on the same inputs, so the arrays are compared bit for bit,
on two calls in a row, the second with a smaller basis; a basis above the
fixed scratch size is refused."""
import ctypes as C

import numpy as np
import pytest

import hfk
from hfk import _lib, systems

pytestmark = pytest.mark.gpu


def _noalloc(ctx, n_out):
    import torch
    out = torch.full((n_out,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # the call runs on the context's own stream
    rc = _lib.lib().hfk_eri_flat_noalloc(ctx._h, C.c_void_p(out.data_ptr()), n_out)
    torch.cuda.synchronize()
    return rc, out.cpu().numpy()


def test_noalloc_flat_matches_eri_flat_twice(require_gpu):
    for name, mol in (("(H2O)2/6-31G", systems.water_cluster(2, "6-31g")), ("water/STO-3G", systems.water("sto-3g"))):
        print("case: no-alloc flat export %s (n=%d)" % (name, mol.nbf), flush=True)
        with hfk.Context(mol) as ctx:
            ref = ctx.calcIntegrals()
            rc, got = _noalloc(ctx, ref.size)
        assert rc == 0
        assert np.array_equal(got, ref)


def test_noalloc_refuses_large_basis_and_wrong_length(require_gpu):
    print("case: no-alloc limits", flush=True)
    mol = systems.water_cluster(32, "6-31g")   # 416 functions > HFK_NOALLOC_MAX_NBF
    with hfk.Context(mol) as ctx:
        rc, _ = _noalloc(ctx, 10)
    assert rc == -1   # HFK_EINVAL
    with hfk.Context(systems.water("sto-3g")) as ctx:
        rc, _ = _noalloc(ctx, 10)
    assert rc == -1   # HFK_EINVAL
