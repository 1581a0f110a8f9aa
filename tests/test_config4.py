"""Precision and uneven batching at the benchmark size: 200 x 100 elements
of order 4 with two PML rings and one Bielak ring (963,603 DOFs before the
fixed outer DOFs), the layered model with anomalies, 10 frequencies x 24 ky
= 240 (omega, ky) pairs, 12 plane-wave sources with seeded receiver counts
of 20..120.

Errors are reported per (omega, ky) pair for the operator and per gradient
field, against the FP64 path (which the parity tests pin to the oracle on
small meshes and, here, to the assembled FP64 CSR operator).  The bar of
1e-8 normwise relative error holds for the refined FP32 product; plain FP32
(product and gradient) is reported at its FP32 level."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from fem25_amd import configs as cf

pytestmark = pytest.mark.gpu

NSRC = 12


def _case():
    m = cf.structured_mesh(200, 100, 4, npml=2, nbielak=1)
    mod = cf.layered_model(m, seed=2)
    om, ky, f, py = cf.freq_ky_batch(10, 24, 0.02, 0.1, 6300.0 / math.sqrt(3.0))
    return m, mod, om, ky, f, py


def test_config4_product_errors_per_pair():
    import torch
    from fem25_amd import _lib
    from fem25_amd.api import MeshInfo
    m, mod, om, ky, f, py = _case()
    pml = cf.PML(rcoeff=1e-3, kappa=1.0, fcent=0.05)
    mi = MeshInfo(m, pml)
    mi.set_model(mod)
    n, npair = mi.ndof, om.size
    print(f"config4 product: ndof={n} npair={npair}", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(41)
    x = torch.randn(n, npair, dtype=torch.complex128, device="cuda", generator=gen)
    y = torch.empty_like(x)
    mi.apply_batch_dev(om, ky, x, y, _lib.F25_APPLY_F64)
    s = torch.empty(mi.nz, dtype=torch.complex128, device="cuda")
    for b in (0, npair // 2 + 5, npair - 1):
        mi.asmble_dev(b"B", float(f[b]), float(py[b]), s)
        yb = torch.empty(n, dtype=torch.complex128, device="cuda")
        mi.csr_apply_dev(s, x[:, b].contiguous(), yb)
        err = float(torch.linalg.norm(y[:, b] - yb) / torch.linalg.norm(yb))
        assert err < 1e-12, (b, err)
    del s
    yr = torch.empty_like(x)
    mi.apply_batch_dev(om, ky, x, yr, _lib.F25_APPLY_F32_REFINED)
    er = (torch.linalg.norm(yr - y, dim=0) / torch.linalg.norm(y, dim=0)).cpu().numpy()
    del yr
    x32 = x.to(torch.complex64)
    y32 = torch.empty_like(x32)
    mi.apply_batch_dev(om, ky, x32, y32, _lib.F25_APPLY_F32)
    e32 = (torch.linalg.norm(y32.to(torch.complex128) - y, dim=0) / torch.linalg.norm(y, dim=0)).cpu().numpy()
    for i in range(0, npair, 24):
        print(f"pair {i:3d} (f={f[i]:.4f} Hz, ky={ky[i]:.2e}): refined {er[i]:.2e}  fp32 {e32[i]:.2e}")
    print(f"per-pair max: refined {er.max():.2e}, fp32 {e32.max():.2e}", flush=True)
    assert er.max() < 1e-8
    assert e32.max() < 1e-5
    mi.close()


def test_config4_gradient_errors_per_field_and_adjoint_sources():
    """12 sources x 240 pairs = 2880 (omega, ky, source) triples with seeded
    fields: FP32 and FP64 gradients accumulated source by source, errors per
    field and for the pairs of one frequency; and CADJLAM's right-hand sides
    for 10 frequencies x 12 sources whose receivers beyond each source's
    count carry zero weight, against the oracle."""
    import torch
    from fem25_amd import _lib
    from fem25_amd.api import MeshInfo
    m, mod, om, ky, f, py = _case()
    pml = cf.PML(rcoeff=1e-3, kappa=1.0, fcent=0.05)
    mi = MeshInfo(m, pml)
    mi.set_model(mod)
    n, npair = mi.ndof, om.size
    gen = torch.Generator(device="cuda").manual_seed(43)
    g64 = torch.zeros(m.nnpg, 3, dtype=torch.float64, device="cuda")
    g32 = torch.zeros_like(g64)
    for s in range(NSRC):
        print(f"config4 gradient: source {s} ({npair} triples, ndof={n})", flush=True)
        u = torch.randn(npair, n, dtype=torch.complex128, device="cuda", generator=gen)
        lam = torch.randn(npair, n, dtype=torch.complex128, device="cuda", generator=gen)
        mi.gradient_dev(om, ky, u, lam, g64)
        mi.gradient32_dev(om, ky, u.to(torch.complex64), lam.to(torch.complex64), g32)
        if s == 0:   # per (omega, ky) pair for the 24 pairs of the first frequency
            worst = 0.0
            for b in range(24):
                a = torch.zeros_like(g64)
                c = torch.zeros_like(g64)
                mi.gradient_dev(om[b:b + 1], ky[b:b + 1], u[b:b + 1].contiguous(), lam[b:b + 1].contiguous(), a)
                mi.gradient32_dev(om[b:b + 1], ky[b:b + 1], u[b:b + 1].to(torch.complex64),
                                  lam[b:b + 1].to(torch.complex64), c)
                worst = max(worst, float(torch.linalg.norm(c - a) / torch.linalg.norm(a)))
            print(f"fp32 gradient, single pairs of the first frequency: worst rel err {worst:.2e}")
            assert worst < 1e-5
        del u, lam
    for p, name in enumerate(("Vp", "Vs", "rho")):
        e = float(torch.linalg.norm(g32[:, p] - g64[:, p]) / torch.linalg.norm(g64[:, p]))
        print(f"fp32 gradient over 2880 triples, field {name}: rel err {e:.2e}")
        assert e < 1e-5
    # adjoint right-hand sides with uneven receiver counts (zero weights past the count)
    rng = np.random.default_rng(5)
    nrec_max, nf = 120, 10
    counts = rng.integers(20, 121, size=NSRC)
    # receivers: interior free-surface nodes, spread over the surface
    gr = mi.graph()
    cnp = np.frombuffer(gr.cnp, dtype="S1")
    seen, cand = set(), []
    for e in range((m.nz - 1) * m.nx, m.nz * m.nx):
        for a in range(m.nen - m.nl, m.nen):
            d = tuple(int(v) for v in gr.lm[e, a])
            if min(d) > 0 and d not in seen and cnp[d[0] - 1] == b"I":
                seen.add(d)
                cand.append(d)
    mrdof = np.ascontiguousarray(np.array(cand[::len(cand) // nrec_max][:nrec_max], np.int32))
    assert mrdof.shape == (nrec_max, 3)
    yrec = rng.uniform(-4e4, 4e4, nrec_max)
    est = (rng.standard_normal((NSRC, nrec_max, nf, 3)) + 1j * rng.standard_normal((NSRC, nrec_max, nf, 3)))
    obs = (rng.standard_normal((NSRC, nrec_max, nf, 3)) + 1j * rng.standard_normal((NSRC, nrec_max, nf, 3)))
    est, obs = est.astype(np.complex64), obs.astype(np.complex64)
    w = np.ones((NSRC, nrec_max, nf, 3), np.float32)
    for s in range(NSRC):
        w[s, counts[s]:] = 0.0
    ifreq = np.repeat(np.arange(1, nf + 1), NSRC).astype(np.int32)
    isrc = np.tile(np.arange(1, NSRC + 1), nf).astype(np.int32)
    fq = 0.02 + 0.008 * ifreq.astype(np.float64)
    pyv = np.full(nf * NSRC, 0.5e-4)
    rhs = torch.zeros(nf * NSRC * n, dtype=torch.complex64, device="cuda")
    ed, od = torch.from_numpy(est).cuda(), torch.from_numpy(obs).cuda()
    wd = torch.from_numpy(w).cuda()
    iw = C.c_int(-1)
    L = _lib.load()
    ptr = lambda a, t: a.ctypes.data_as(t)   # noqa: E731
    _lib.check(L.f25_cadjlam_rhs_dev(n, nf, nrec_max, NSRC, nf * NSRC, ptr(ifreq, _lib.ip), ptr(isrc, _lib.ip),
                                     ptr(fq, _lib.dp), ptr(pyv, _lib.dp), 0.0, 2, 3, 0.0, 0.0, 0, 0,
                                     ptr(yrec, _lib.dp), ptr(mrdof, _lib.ip), C.c_void_p(ed.data_ptr()),
                                     C.c_void_p(od.data_ptr()), C.cast(C.c_void_p(wd.data_ptr()), _lib.fp),
                                     C.c_void_p(rhs.data_ptr()), C.byref(iw)), "cadjlam_rhs_dev")
    assert iw.value == 0
    rh = rhs.cpu().numpy().reshape(nf * NSRC, n)
    for k in (0, 7, 61, nf * NSRC - 1):
        ref, _ = oracle.cadjlam_rhs(n, int(ifreq[k]), int(isrc[k]), float(fq[k]), float(pyv[k]), 0.0, 2, 3, 0.0, 0.0,
                                    0, False, yrec, mrdof, est, obs, w)
        assert np.abs(rh[k] - ref).max() <= 2e-6 * np.abs(ref).max(), k
        live = np.count_nonzero(ref)
        assert live == 3 * counts[isrc[k] - 1], (k, live)
    mi.close()
