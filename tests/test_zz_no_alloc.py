"""The allocation-free GNMFREE entry point against the allocating one
(collected last: it leaves its per-GPU scratch array resident)."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import mesh_case

pytestmark = pytest.mark.gpu


def test_gnmfree_noalloc_matches_gnmfree_dev():
    """Two calls in a row (the second with fewer and shorter Jacobians) give
    bit for bit the z of f25_gnmfree_dev on the same device inputs; a batch
    past the scratch size is refused before any launch."""
    import torch
    from fem25_amd import _lib
    from fem25_amd.api import MeshInfo
    L = _lib.load()
    f = L.f25_gnmfree_dev_noalloc
    m, mod, pml = mesh_case("c1")
    mi = MeshInfo(m, pml)
    mi.set_model(mod)
    inv = mi.inv_setup(b"PP")
    rng = np.random.default_rng(95)
    for njac, lo, hi in ((10, 20, 121), (4, 5, 40)):
        nobs = (3 * rng.integers(lo, hi, size=njac)).astype(np.int32)
        print(f"noalloc gnmfree: njac={njac} obs={int(nobs.sum())} nnpgl={inv.nnpgl}", flush=True)
        tot = int(nobs.sum()) * inv.nnpgl
        jac = torch.from_numpy((rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(np.complex64)).cuda()
        alpha = rng.uniform(0.5, 2.0, size=njac)
        p = torch.from_numpy(rng.standard_normal(inv.na35)).cuda()
        z0 = torch.from_numpy(rng.standard_normal(inv.na35)).cuda()
        za, zb = z0.clone(), z0.clone()
        torch.cuda.synchronize()
        args = (njac, nobs.ctypes.data_as(_lib.ip), alpha.ctypes.data_as(_lib.dp), C.c_void_p(jac.data_ptr()),
                C.c_void_p(p.data_ptr()))
        _lib.check(L.f25_gnmfree_dev(mi.h, *args, C.c_void_p(za.data_ptr())), "f25_gnmfree_dev")
        _lib.check(f(mi.h, *args, C.c_void_p(zb.data_ptr())), "f25_gnmfree_dev_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(za, zb)
        assert not torch.equal(za, z0)
    big = np.full(1, 3 * 2 ** 20, np.int32)   # 3 Mi observations: past 32 MiB
    one = np.ones(1)
    rc = f(mi.h, 1, big.ctypes.data_as(_lib.ip), one.ctypes.data_as(_lib.dp), C.c_void_p(jac.data_ptr()),
           C.c_void_p(p.data_ptr()), C.c_void_p(za.data_ptr()))
    assert rc == 1   # F25_ERR_ARG
    mi.close()
