"""The order-6 benchmark workload at full size: 400 x 200 elements of order 6
(8,650,803 DOFs), the layered model with anomalies.  Size-independent
properties of the batched product and the gradient, with their tolerances
written here:

* the product for three (omega, ky) pairs of the 192-pair batch equals the
  assembled FP64 CSR operator applied pair by pair (1e-12 normwise; the CSR
  path's values are pinned to the compiled ASMBLE_DRIVER on small meshes);
* the gradient is symmetric in its two fields, g(u, lam) = g(lam, u), since
  dS/dm is Hermitian (1e-12), and linear in lam (1e-12);
* graph replays of the product equal the eager product bit for bit."""
import math

import numpy as np
import pytest

from fem25_amd import configs as cf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c5():
    from fem25_amd.api import MeshInfo
    m = cf.structured_mesh(400, 200, 6)
    mod = cf.layered_model(m, seed=2)
    mi = MeshInfo(m)
    mi.set_model(mod)
    yield m, mi
    mi.close()
    import torch
    torch.cuda.empty_cache()   # the 27 GB vectors stay cached by torch otherwise


def test_config5_product_against_assembled_operator(c5):
    import torch
    from fem25_amd import _lib
    m, mi = c5
    om, ky, f, py = cf.freq_ky_batch(6, 32, 0.02, 0.1, 6300.0 / math.sqrt(3.0))
    n, npair = mi.ndof, om.size
    print(f"config5 product: ndof={n} npair={npair}", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(51)
    x = torch.randn(n, npair, dtype=torch.complex128, device="cuda", generator=gen)
    y = torch.empty_like(x)
    mi.apply_batch_dev(om, ky, x, y, _lib.F25_APPLY_F64)
    s = torch.empty(mi.nz, dtype=torch.complex128, device="cuda")
    for b in (0, 77, npair - 1):
        mi.asmble_dev(b"B", float(f[b]), float(py[b]), s)
        yb = torch.empty(n, dtype=torch.complex128, device="cuda")
        mi.csr_apply_dev(s, x[:, b].contiguous(), yb)
        err = float(torch.linalg.norm(y[:, b] - yb) / torch.linalg.norm(yb))
        print(f"pair {b}: rel err {err:.2e}", flush=True)
        assert err < 1e-12, (b, err)
    del s
    yg = torch.empty_like(x)
    for rep in range(2):
        yg.zero_()
        mi.apply_batch_dev(om, ky, x, yg, _lib.F25_APPLY_F64 | _lib.F25_APPLY_GRAPH)
        assert torch.equal(yg, y), rep


def test_config5_gradient_symmetry_and_linearity(c5):
    import torch
    m, mi = c5
    om, ky, f, py = cf.freq_ky_batch(2, 8, 0.02, 0.1, 6300.0 / math.sqrt(3.0))
    n, nt = mi.ndof, om.size
    print(f"config5 gradient: ndof={n} triples={nt}", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(53)
    u = torch.randn(nt, n, dtype=torch.complex128, device="cuda", generator=gen)
    l1 = torch.randn(nt, n, dtype=torch.complex128, device="cuda", generator=gen)
    l2 = torch.randn(nt, n, dtype=torch.complex128, device="cuda", generator=gen)
    g = lambda a, b: (lambda out: (mi.gradient_dev(om, ky, a, b, out), out)[1])(  # noqa: E731
        torch.zeros(m.nnpg, 3, dtype=torch.float64, device="cuda"))
    g_ul, g_lu = g(u, l1), g(l1, u)
    rel = float(torch.linalg.norm(g_ul - g_lu) / torch.linalg.norm(g_ul))
    print(f"symmetry: rel {rel:.2e}", flush=True)
    assert rel < 1e-12
    g_sum = g(u, (l1 + l2).contiguous())
    g_2 = g(u, l2)
    rel = float(torch.linalg.norm(g_sum - g_ul - g_2) / torch.linalg.norm(g_sum))
    print(f"linearity: rel {rel:.2e}", flush=True)
    assert rel < 1e-12
