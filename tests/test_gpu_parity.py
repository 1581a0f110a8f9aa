"""HIP path against the oracle (CPU restatement pinned to the compiled
reference): assembled operator, SpMV, batched product, gradient, virtual
forces.  Tolerances: FP64 results 1e-12 normwise relative; COMPLEX*8
results at FP32 level (the reference itself rounds every contribution to
FP32)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fem25_amd import configs as cf
from tests.helpers import TWOPI, csr, mesh_case, oracle_S, oracle_graph, pairs, ps_inv, relerr

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    return torch


def _mi(name):
    from fem25_amd.api import MeshInfo
    m, mod, pml = mesh_case(name)
    mi = MeshInfo(m, pml)
    mi.set_model(mod)
    return mi


def test_mfma_layout():
    import ctypes as C
    from fem25_amd import _lib
    err = C.c_double(-1)
    assert _lib.load().f25_selftest_mfma(C.byref(err)) == 0, err.value
    assert err.value == 0.0


@pytest.mark.parametrize("name", ["tiny", "pml", "p6", "p2_pml"])
def test_graph_bitexact(name):
    mi = _mi(name)
    g = mi.graph()
    go = oracle_graph(name)
    assert (g.ndof, g.nz) == (go.ndof, go.nz)
    np.testing.assert_array_equal(g.ien, go.ien)
    np.testing.assert_array_equal(g.lm, go.lm)
    np.testing.assert_array_equal(g.irptr, go.irptr)
    np.testing.assert_array_equal(g.jcptr, go.jcptr)
    assert g.cnp == go.cnp


@pytest.mark.parametrize("name,freq,py", [("tiny", 0.05, 0.0), ("tiny_layered", 0.08, 1.1e-4),
                                          ("pml", 0.05, 0.7e-4), ("p6", 0.1, 0.5e-4),
                                          ("p2_pml", 0.2, 1e-4)])
def test_asmble_fp64(name, freq, py):
    mi = _mi(name)
    s = mi.asmble(b"B", freq, py, complex128=True)
    m, mod, pml = mesh_case(name)
    ref = oracle.asmble_z(m, mod, oracle_graph(name), freq, py, pml)
    assert relerr(s, ref) < FP64_TOL


@pytest.mark.parametrize("name,freq,py", [("tiny", 0.05, 0.0), ("pml", 0.05, 0.7e-4)])
def test_asmble_cdist_c64(name, freq, py):
    mi = _mi(name)
    s = mi.asmble(b"B", freq, py)
    m, mod, pml = mesh_case(name)
    ref = oracle.asmble_cdist(m, mod, oracle_graph(name), freq, py, pml)
    # entrywise against the row scale: both sides hold FP32-rounded sums
    g = oracle_graph(name)
    rowmax = np.maximum.reduceat(np.abs(ref.astype(np.complex128)), g.irptr[:-1] - 1)
    rows = np.repeat(np.arange(g.ndof), np.diff(g.irptr))
    assert np.all(np.abs(s.astype(np.complex128) - ref) <= 1e-6 * rowmax[rows])


def test_asmble_config1_and_spmv(torch_dev):
    torch = torch_dev
    mi = _mi("c1")
    freq, py = 0.05, 0.0
    s = mi.asmble(b"B", freq, py, complex128=True)
    m, mod, pml = mesh_case("c1")
    g = oracle_graph("c1")
    ref = oracle.asmble_z(m, mod, g, freq, py, pml)
    assert relerr(s, ref) < FP64_TOL
    x = cf.seeded_complex(g.ndof, seed=1)
    ds = torch.from_numpy(s).cuda()
    dx = torch.from_numpy(x).cuda()
    dy = torch.zeros_like(dx)
    mi.csr_apply_dev(ds, dx, dy)
    assert relerr(dy.cpu().numpy(), oracle.spmv_z(g, ref, x)) < FP64_TOL


def test_apply_full_size_against_assembled_csr(torch_dev):
    """At the benchmark size (200 x 100 elements, p = 4, 963,603 DOFs,
    8 frequencies x 16 ky = 128 pairs, layered model with anomalies): the
    batched FP64 product (element by element, one launch per colour) against
    the assembled FP64 CSR operator of the same pair followed by its SpMV,
    for the first, a middle and the last pair; and two runs of the product
    bit for bit equal (the colour scatter has no races)."""
    torch = torch_dev
    from fem25_amd.api import MeshInfo
    m = cf.structured_mesh(200, 100, 4)
    mi = MeshInfo(m)
    mi.set_model(cf.layered_model(m, seed=2))
    om, ky, f, py = cf.freq_ky_batch(8, 16, 0.02, 0.1, 6300.0 / np.sqrt(3.0))
    print(f"full-size product: ndof={mi.ndof} npair={om.size}", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(mi.ndof, om.size, dtype=torch.complex128, device="cuda", generator=gen)
    y = torch.empty_like(x)
    mi.apply_batch_dev(om, ky, x, y)
    y2 = torch.empty_like(x)
    mi.apply_batch_dev(om, ky, x, y2)
    torch.cuda.synchronize()
    assert torch.equal(y, y2)
    del y2
    s = torch.empty(mi.nz, dtype=torch.complex128, device="cuda")
    for b in (0, om.size // 2 + 3, om.size - 1):
        mi.asmble_dev(b"B", float(f[b]), float(py[b]), s)
        xb = x[:, b].contiguous()
        yb = torch.empty_like(xb)
        mi.csr_apply_dev(s, xb, yb)
        err = float(torch.linalg.norm(y[:, b] - yb) / torch.linalg.norm(yb))
        assert err < 1e-13, (b, err)
    mi.close()


def test_gradient_full_size_additive_and_linear(torch_dev):
    """At the benchmark size (200 x 100 elements, p = 4, 128 triples): the
    gradient over all triples equals the sum of the gradients over four
    quarters of them (the kernel takes two triples per pass and loads the
    next pass during the current one), and it is linear in the adjoint
    field: g(u, l1 + 2 l2) = g(u, l1) + 2 g(u, l2)."""
    torch = torch_dev
    from fem25_amd.api import MeshInfo
    m = cf.structured_mesh(200, 100, 4)
    mi = MeshInfo(m)
    mi.set_model(cf.layered_model(m, seed=2))
    om, ky, f, py = cf.freq_ky_batch(8, 16, 0.02, 0.1, 6300.0 / np.sqrt(3.0))
    nt, n = om.size, mi.ndof
    print(f"full-size gradient: ndof={n} ntrip={nt}", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(9)
    u = torch.randn(nt, n, dtype=torch.complex128, device="cuda", generator=gen)
    l1 = torch.randn(nt, n, dtype=torch.complex128, device="cuda", generator=gen)
    l2 = torch.randn(nt, n, dtype=torch.complex128, device="cuda", generator=gen)

    def grad(lam, sl=slice(None)):
        g = torch.zeros(m.nnpg, 3, dtype=torch.float64, device="cuda")
        mi.gradient_dev(om[sl], ky[sl], u[sl].contiguous(), lam[sl].contiguous(), g)
        return g

    g1 = grad(l1)
    parts = sum(grad(l1, slice(q * nt // 4, (q + 1) * nt // 4)) for q in range(4))
    assert float(torch.linalg.norm(parts - g1) / torch.linalg.norm(g1)) < 1e-12
    g2 = grad(l2)
    g12 = grad(l1 + 2 * l2)
    assert float(torch.linalg.norm(g12 - (g1 + 2 * g2)) / torch.linalg.norm(g12)) < 1e-12
    mi.close()


@pytest.mark.parametrize("name", ["tiny_layered", "p5", "p6", "pml", "p2_pml"])
def test_apply_batch_fp64(name):
    mi = _mi(name)
    om, ky, f, py = pairs(2, 3)
    g = oracle_graph(name)
    x = cf.seeded_complex(g.ndof, batch=om.size, seed=3)
    y = mi.apply_batch(om, ky, x)
    for b in range(om.size):
        S = oracle_S(name, f[b], py[b])
        assert relerr(y[b], S @ x[b]) < FP64_TOL, b


def test_apply_batch_fp64_pair_groups():
    """70 pairs: two groups of 64 lanes, the second partly empty.  A pair
    repeated in both groups gives the same bits; pairs at both group edges
    match the oracle operator."""
    mi = _mi("tiny_layered")
    om, ky, f, py = pairs(7, 10)
    assert om.size == 70
    for a, b in ((67, 3), (64, 63)):
        om[a], ky[a], f[a], py[a] = om[b], ky[b], f[b], py[b]
    g = oracle_graph("tiny_layered")
    x = cf.seeded_complex(g.ndof, batch=om.size, seed=21)
    x[67] = x[3]
    x[64] = x[63]
    y = mi.apply_batch(om, ky, x)
    assert np.array_equal(y[67], y[3]) and np.array_equal(y[64], y[63])
    for b in (0, 3, 62, 63, 65, 69):
        S = oracle_S("tiny_layered", f[b], py[b])
        assert relerr(y[b], S @ x[b]) < FP64_TOL, b


@pytest.mark.parametrize("name", ["p5", "p6"])
def test_apply_batch_fp64_row_tiles_pair_groups(name):
    """p5, p6: two 25-row tiles per component (p6's second tile has a padding
    row) and two groups of 64 pairs, the second partly empty."""
    mi = _mi(name)
    om, ky, f, py = pairs(7, 10)
    g = oracle_graph(name)
    x = cf.seeded_complex(g.ndof, batch=om.size, seed=23)
    y = mi.apply_batch(om, ky, x)
    for b in (0, 63, 64, 69):
        S = oracle_S(name, f[b], py[b])
        assert relerr(y[b], S @ x[b]) < FP64_TOL, b


@pytest.mark.parametrize("name,tiles", [("tiny_layered", (1, 2, 3)), ("p5", (2, 3, 4)), ("p6", (2, 3, 4))])
def test_apply_batch_fp64_tile_candidates(name, tiles):
    """Every compiled row-tile count of the FP64 product gives the oracle's
    product; an unknown count is refused."""
    from fem25_amd._lib import Fem25Error
    mi = _mi(name)
    om, ky, f, py = pairs(2, 3)
    g = oracle_graph(name)
    x = cf.seeded_complex(g.ndof, batch=om.size, seed=25)
    refs = [oracle_S(name, f[b], py[b]) @ x[b] for b in range(om.size)]
    for t in tiles:
        mi.set_apply_tiles(t)
        y = mi.apply_batch(om, ky, x)
        for b in range(om.size):
            assert relerr(y[b], refs[b]) < FP64_TOL, (t, b)
    with pytest.raises(Fem25Error):
        mi.set_apply_tiles(7)
    mi.set_apply_tiles(0)


def test_apply_batch_graph_replays(torch_dev):
    """F25_APPLY_GRAPH: captured once, replayed with new (omega, ky) values
    and with eager work between the replays; every replay equals the eager
    product bit for bit, and new buffers are captured anew."""
    torch = torch_dev
    from fem25_amd import _lib
    mi = _mi("tiny_layered")
    g = oracle_graph("tiny_layered")
    x = torch.from_numpy(np.ascontiguousarray(cf.seeded_complex(g.ndof, batch=70, seed=27).T)).cuda()
    yg = torch.zeros_like(x)
    ye = torch.zeros_like(x)
    for rep in range(4):
        om, ky, _, _ = pairs(7, 10, fmin=0.02 + 0.01 * rep)
        mi.apply_batch_dev(om, ky, x, yg, _lib.F25_APPLY_F64 | _lib.F25_APPLY_GRAPH)
        mi.apply_batch_dev(om, ky, x, ye, _lib.F25_APPLY_F64)
        torch.cuda.synchronize()
        assert torch.equal(yg, ye), rep
        x.mul_(1.5)                     # eager work between the replays
    y2 = torch.zeros_like(x)            # other output buffer: a new capture
    mi.apply_batch_dev(om, ky, x, y2, _lib.F25_APPLY_F64 | _lib.F25_APPLY_GRAPH)
    mi.apply_batch_dev(om, ky, x, ye, _lib.F25_APPLY_F64)
    torch.cuda.synchronize()
    assert torch.equal(y2, ye)
    mi.set_apply_tiles(2)               # another kernel: the capture is dropped and redone
    mi.apply_batch_dev(om, ky, x, y2, _lib.F25_APPLY_F64 | _lib.F25_APPLY_GRAPH)
    mi.apply_batch_dev(om, ky, x, ye, _lib.F25_APPLY_F64)
    torch.cuda.synchronize()
    assert torch.equal(y2, ye)
    mi.set_apply_tiles(0)
    from fem25_amd._lib import Fem25Error
    with pytest.raises(Fem25Error):     # graph replay is for the FP64 product only
        mi.apply_batch_dev(om, ky, x, y2, _lib.F25_APPLY_F32_REFINED | _lib.F25_APPLY_GRAPH)


def test_apply_batch_pml_modes_and_graph(torch_dev):
    """PML elements in the batched product (per-omega element matrices):
    uneven pairs per omega (a group of 16 and a partial one), FP64 against the
    oracle operator at 1e-12, the FP32 modes at their precision, a graph
    replayed after the omegas change (the matrices are rebuilt outside the
    graph) equal to the eager product bit for bit."""
    torch = torch_dev
    from fem25_amd import _lib
    name = "pml"
    mi = _mi(name)
    g = oracle_graph(name)
    om, ky, f, py = pairs(3, 7)
    om = np.concatenate([om, np.full(12, om[0])])         # 19 pairs of the first omega
    ky = np.concatenate([ky, np.linspace(0.0, ky[6], 12)])
    f, py = om / TWOPI, ky / om
    x = cf.seeded_complex(g.ndof, batch=om.size, seed=44)
    xd = torch.from_numpy(np.ascontiguousarray(x.T)).cuda()
    y = torch.zeros_like(xd)
    mi.apply_batch_dev(om, ky, xd, y, _lib.F25_APPLY_F64)
    yh = y.cpu().numpy().T
    yref = np.stack([oracle_S(name, f[b], py[b]) @ x[b] for b in range(om.size)])
    for b in range(om.size):
        assert relerr(yh[b], yref[b]) < FP64_TOL, b
    x32 = xd.to(torch.complex64)
    y32 = torch.zeros_like(x32)
    mi.apply_batch_dev(om, ky, x32, y32, _lib.F25_APPLY_F32)
    yr = torch.zeros_like(xd)
    mi.apply_batch_dev(om, ky, xd, yr, _lib.F25_APPLY_F32_REFINED)
    e32 = max(relerr(y32.cpu().numpy().T[b], yref[b]) for b in range(om.size))
    er = max(relerr(yr.cpu().numpy().T[b], yref[b]) for b in range(om.size))
    print(f"pml mesh normwise error: fp32 {e32:.2e}, fp32-refined {er:.2e}")
    assert e32 < 1e-5 and er < 1e-8
    yg = torch.zeros_like(xd)
    for rep in range(3):
        omr = om * (1.0 + 0.1 * rep)
        mi.apply_batch_dev(omr, ky, xd, yg, _lib.F25_APPLY_F64 | _lib.F25_APPLY_GRAPH)
        mi.apply_batch_dev(omr, ky, xd, y, _lib.F25_APPLY_F64)
        torch.cuda.synchronize()
        assert torch.equal(yg, y), rep
        xd.mul_(0.5)                    # eager work between the replays


def test_apply_batch_config1_modes(torch_dev):
    torch = torch_dev
    from fem25_amd import _lib
    mi = _mi("c1")
    om, ky, f, py = pairs(2, 5)
    g = oracle_graph("c1")
    x = cf.seeded_complex(g.ndof, batch=om.size, seed=4)
    xd = torch.from_numpy(np.ascontiguousarray(x.T)).cuda()        # [dof][pair]
    y64 = torch.zeros_like(xd)
    mi.apply_batch_dev(om, ky, xd, y64, _lib.F25_APPLY_F64)
    y64 = y64.cpu().numpy().T
    yref = np.stack([oracle_S("c1", f[b], py[b]) @ x[b] for b in range(om.size)])
    for b in range(om.size):
        assert relerr(y64[b], yref[b]) < FP64_TOL
    x32 = xd.to(torch.complex64)
    y32 = torch.zeros_like(x32)
    mi.apply_batch_dev(om, ky, x32, y32, _lib.F25_APPLY_F32)
    yr = torch.zeros_like(xd)
    mi.apply_batch_dev(om, ky, xd, yr, _lib.F25_APPLY_F32_REFINED)
    e32 = max(relerr(y32.cpu().numpy().T[b], yref[b]) for b in range(om.size))
    er = max(relerr(yr.cpu().numpy().T[b], yref[b]) for b in range(om.size))
    print(f"config-1 normwise error: fp32 {e32:.2e}, fp32-refined {er:.2e}")
    assert e32 < 1e-5
    assert er < 1e-8   # FP32 MFMA with the fixed-point hi/lo split


def test_apply_quadratic_in_ky():
    """S(2k) - 3 S(k) + 3 S(0) - S(-k) = 0 (S is quadratic in ky)."""
    mi = _mi("tiny_layered")
    g = oracle_graph("tiny_layered")
    w, k = TWOPI * 0.07, 1.3e-5
    x = cf.seeded_complex(g.ndof, seed=11)
    om = np.full(4, w)
    ky = np.array([2 * k, k, 0.0, -k])
    y = mi.apply_batch(om, ky, np.stack([x] * 4))
    comb = y[0] - 3 * y[1] + 3 * y[2] - y[3]
    assert np.linalg.norm(comb) < 1e-11 * np.linalg.norm(y[0])


@pytest.mark.parametrize("name,nf,nk", [("tiny_layered", 1, 2), ("pml", 1, 2), ("tiny_layered", 2, 3),
                                         ("p5", 2, 3), ("p6", 1, 2), ("tiny_layered", 1, 3), ("p6", 1, 1),
                                         ("p5", 1, 5)])
def test_gradient_exact(name, nf, nk):
    """The sum-factorised kernel with 1 to 6 triples (passes of 2, the last
    one half empty for an odd count); p6 adds a column of lanes for the 17th
    xi point."""
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    om, ky, f, py = pairs(nf, nk)
    u = cf.seeded_complex(g.ndof, batch=om.size, seed=21)
    lam = cf.seeded_complex(g.ndof, batch=om.size, seed=22)
    gd = mi.gradient(om, ky, u, lam)
    ref = np.zeros((m.nnpg, 3))
    for b in range(om.size):
        ref += oracle.grad_exact(m, mod, g, f[b], py[b], u[b], lam[b], pml)[1]
    for p in range(3):
        assert relerr(gd[:, p], ref[:, p]) < 1e-11, p


@pytest.mark.parametrize("name", ["tiny_layered", "pml", "p5", "p6"])
def test_gradient_fp32(torch_dev, name):
    """The FP32 gradient (complex64 fields, FP32 interpolation and densities,
    FP64 sums) against the FP64 gradient per field: at FP32 level (1e-5
    normwise; the fields alone are rounded to 6e-8)."""
    torch = torch_dev
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    om, ky, f, py = pairs(2, 3)
    u = torch.from_numpy(cf.seeded_complex(g.ndof, batch=om.size, seed=23)).cuda()
    lam = torch.from_numpy(cf.seeded_complex(g.ndof, batch=om.size, seed=24)).cuda()
    g64 = torch.zeros(m.nnpg, 3, dtype=torch.float64, device="cuda")
    g32 = torch.zeros_like(g64)
    mi.gradient_dev(om, ky, u, lam, g64)
    mi.gradient32_dev(om, ky, u.to(torch.complex64), lam.to(torch.complex64), g32)
    for p in range(3):
        e = float(torch.linalg.norm(g32[:, p] - g64[:, p]) / torch.linalg.norm(g64[:, p]))
        print(f"{name} field {p}: fp32 gradient rel err {e:.2e}")
        assert e < 1e-5, (p, e)


@pytest.mark.parametrize("name,cinv", [("tiny_layered", b"PP"), ("pml", b"PP"), ("tiny", b"SS"), ("p5", b"PP"),
                                       ("p6", b"SS")])
def test_cfmat_adjflam(name, cinv):
    """p <= 5: the sum-factorised element products; p6: the direct ones."""
    from fem25_amd import api
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    inv = mi.inv_setup(cinv)
    io = oracle.inv_setup(m, mod, g, pml)
    np.testing.assert_array_equal(inv.maskg, io.maskg)
    np.testing.assert_array_equal(inv.jcsc_fdist, io.jcsc)
    np.testing.assert_array_equal(inv.icsc_fdist, io.icsc)
    np.testing.assert_array_equal(inv.mygrad, io.mygrad)
    np.testing.assert_allclose(inv.elem_wts, io.elem_wts, rtol=1e-15)
    wave = cf.seeded_complex(g.ndof, seed=31).astype(np.complex64)
    freq, py = 0.06, 0.8e-4
    fm, ierr = api.cfmat_dist(inv.nz_fdist, g.ndof, freq, py, wave, mi, inv)
    assert ierr == 0
    ref = oracle.cfmat_dist(m, mod, g, io, freq, py, wave, cinvtype=cinv)
    assert relerr(fm, ref) < 2e-6
    lam = cf.seeded_complex(g.ndof, seed=32).astype(np.complex64)
    gp = api.adjflam_dist(mi, inv, ref, lam)
    gref = oracle.adjflam_dist(io, ref, lam)
    assert relerr(gp, gref) < 1e-5


def test_peff():
    from fem25_amd import api
    mi = _mi("pml")
    m, mod, pml = mesh_case("pml")
    g = oracle_graph("pml")
    cd = oracle.asmble_cdist(m, mod, g, 0.05, 0.0, pml)
    ue = cf.seeded_complex(g.ndof, seed=41).astype(np.complex64)
    pe = api.asmb_peff(mi, cd, ue)
    ref = oracle.asmb_peff(g, cd, ue)
    assert relerr(pe, ref) < 1e-5
    assert np.count_nonzero(ref) > 0


@pytest.mark.parametrize("name", ["tiny_layered", "p5", "p6"])
def test_gn_fv_matches_operator_difference(name):
    """F v = dS/dm[v] u (the batched product on derivative blocks) against a
    central difference of the oracle's FP64 operator (S is cubic in the step
    through rho Vp^2, so the step is small); p4, p5 and p6."""
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    rng = np.random.default_rng(51)
    v = np.stack([rng.standard_normal(m.nnpg) * 50.0, rng.standard_normal(m.nnpg) * 30.0,
                  rng.standard_normal(m.nnpg) * 20.0], axis=1)
    f, py = 0.07, 0.6e-4
    u = cf.seeded_complex(g.ndof, seed=52)
    fv = mi.gn_fv(np.array([TWOPI * f]), np.array([TWOPI * f * py]), u, v)[0]
    h = 1e-3    # truncation (rho Vp^2 is cubic in h) ~ 3e-11, rounding ~ 3e-12
    mp = cf.model_from_velocity(mod.vp + h * v[:, 0], mod.vs + h * v[:, 1], mod.dens + h * v[:, 2])
    mm = cf.model_from_velocity(mod.vp - h * v[:, 0], mod.vs - h * v[:, 1], mod.dens - h * v[:, 2])
    Sp = csr(g, oracle.asmble_z(m, mp, g, f, py, pml))
    Sm = csr(g, oracle.asmble_z(m, mm, g, f, py, pml))
    ref = (Sp @ u - Sm @ u) / (2 * h)
    assert relerr(fv, ref) < 1e-9


@pytest.mark.parametrize("name", ["tiny_layered", "p6", "bench"])
def test_gn_adjoint_identity(name):
    """J v (f25_gn_fv: the batched product on derivative blocks) and J^H w
    (f25_gn_fhw: the gradient kernel) are adjoint:
    sum_t Re <w_t, J_t v> = v . (J^H w), over three (omega, ky) pairs; on
    small meshes and at the benchmark size (200 x 100, p = 4)."""
    if name == "bench":
        from fem25_amd.api import MeshInfo
        m = cf.structured_mesh(200, 100, 4)
        mi = MeshInfo(m)
        mi.set_model(cf.layered_model(m, seed=2))
        ndof = mi.ndof
    else:
        mi = _mi(name)
        m, _, _ = mesh_case(name)
        ndof = oracle_graph(name).ndof
    om, ky, f, py = pairs(1, 3)
    rng = np.random.default_rng(61)
    v = rng.standard_normal((m.nnpg, 3)) * np.array([50.0, 30.0, 20.0])
    u = cf.seeded_complex(ndof, batch=om.size, seed=62)
    w = cf.seeded_complex(ndof, batch=om.size, seed=63)
    fv = mi.gn_fv(om, ky, u, v)
    g = mi.gn_fhw(om, ky, u, w)
    lhs = float(sum(np.vdot(w[t], fv[t]).real for t in range(om.size)))
    rhs = float(np.sum(v * g))
    scale = float(sum(np.linalg.norm(w[t]) * np.linalg.norm(fv[t]) for t in range(om.size)))
    assert abs(lhs) > 1e-6 * scale          # a non-trivial inner product
    assert abs(lhs - rhs) < 1e-11 * scale, (lhs, rhs)


@pytest.mark.parametrize("name", ["p6", "bench"])
def test_gn_diag_against_jv_columns(name):
    """Each pseudo-Hessian diagonal entry (k_fcols_sf + k_fdiag) equals
    sum_t |J_t e_{a,p}|^2 with J e from f25_gn_fv (the batched product on
    derivative blocks): two independent kernels, on a p6 test mesh and at
    the benchmark size (200 x 100, p = 4), for a few anchors."""
    if name == "bench":
        from fem25_amd.api import MeshInfo
        m = cf.structured_mesh(200, 100, 4)
        mi = MeshInfo(m)
        mi.set_model(cf.layered_model(m, seed=2))
        ndof = mi.ndof
        picks = [(m.nnpg // 2 + 7, 0), (1234, 1), (m.nnpg - 3, 2), (0, 0)]
    else:
        mi = _mi(name)
        m, _, _ = mesh_case(name)
        ndof = oracle_graph(name).ndof
        picks = [(a, p) for a in (0, 6, m.nnpg // 2, m.nnpg - 1) for p in range(3)]
    om, ky, f, py = pairs(1, 2)
    u = cf.seeded_complex(ndof, batch=om.size, seed=64)
    d = mi.gn_diag(om, ky, u)
    for a, p in picks:
        v = np.zeros((m.nnpg, 3))
        v[a, p] = 1.0
        fv = mi.gn_fv(om, ky, u, v)
        ref = float(sum(np.vdot(fv[t], fv[t]).real for t in range(om.size)))
        assert ref > 0
        assert abs(d[a, p] - ref) < 1e-10 * ref, (a, p, d[a, p], ref)


@pytest.mark.parametrize("name,stride", [("tiny_layered", 1), ("p5", 1), ("p6", 1), ("pml", 2), ("p2_pml", 1)])
def test_gn_diag_matches_operator_columns(name, stride):
    """Pseudo-Hessian diagonal sum_t |dS_t/dm_{a,p} u_t|^2 against columns
    formed by central differences of the oracle's FP64 operator.  S is
    quadratic in Vp alone and in Vs alone and affine in rho, so a central
    difference is exact up to rounding (h = 1 m/s, 1 kg/m^3).  Every order
    runs the sum-factorised element kernel (p6 with a column of lanes for the
    17th xi point).  PML elements
    are outside the path, so on the PML meshes every stride-th anchor that
    touches no PML element is checked."""
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    om, ky, f, py = pairs(1, 2)
    u = cf.seeded_complex(g.ndof, batch=om.size, seed=71)
    d = mi.gn_diag(om, ky, u)
    in_pml = np.zeros(m.nnpg, dtype=bool)
    for e in range(m.ieng.shape[0]):
        if m.cdomain[e:e + 1] == b"A":
            in_pml[m.ieng[e] - 1] = True
    anchors = [a for a in range(0, m.nnpg, stride) if not in_pml[a]]
    assert len(anchors) >= 6
    ref = np.zeros_like(d)
    for a in anchors:
        for p in range(3):
            dv = np.zeros((3, m.nnpg))
            dv[p, a] = 1.0
            mp = cf.model_from_velocity(mod.vp + dv[0], mod.vs + dv[1], mod.dens + dv[2])
            mm = cf.model_from_velocity(mod.vp - dv[0], mod.vs - dv[1], mod.dens - dv[2])
            for b in range(om.size):
                sp_ = csr(g, oracle.asmble_z(m, mp, g, f[b], py[b], pml))
                sm_ = csr(g, oracle.asmble_z(m, mm, g, f[b], py[b], pml))
                col = (sp_ @ u[b] - sm_ @ u[b]) / 2.0
                ref[a, p] += float(np.vdot(col, col).real)
    for p in range(3):
        assert relerr(d[anchors, p], ref[anchors, p]) < 1e-9, p
    assert np.all(d[anchors] > 0)


def test_pchadj_jacobian():
    """J = G^T F and diag(J^H J) on the GPU against the oracle's COMPLEX*8
    restatement of PCHADJ and the dense complex128 product."""
    name = "pml"
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    mi.inv_setup(b"PP")
    io = oracle.inv_setup(m, mod, g, pml)
    rng = np.random.default_rng(91)
    nrhs, lrhs = 60, g.ndof + 3
    G = (rng.standard_normal((nrhs, lrhs)) + 1j * rng.standard_normal((nrhs, lrhs))).astype(np.complex64)
    F = (rng.standard_normal(io.nz_fdist) + 1j * rng.standard_normal(io.nz_fdist)).astype(np.complex64)
    jac, hd = mi.pchadj_jac(nrhs, G, F)
    jref, href = oracle.pchadj_jac(io, nrhs, lrhs, G, F)
    assert relerr(jac, jref) < 1e-5
    assert relerr(hd, href) < 1e-5
    Fd = np.zeros((g.ndof, io.nnpgl), np.complex128)
    for c in range(io.nnpgl):
        for i in range(io.jcsc[c] - 1, io.jcsc[c + 1] - 1):
            Fd[io.icsc[i] - 1, c] += F[i]
    Jd = G[:, :g.ndof].astype(np.complex128) @ Fd
    assert relerr(jac.T, Jd) < 1e-6      # FP64 sums rounded once to COMPLEX*8


@pytest.mark.parametrize("name", ["tiny_layered", "pml", "p5", "p6"])
@pytest.mark.parametrize("cinv", [b"PP", b"SS"])
def test_qpmfree_element_products(name, cinv):
    """CADJ_NU_ALPHA's right-hand sides and P3, and MET225_MOD's P1 and P2,
    on the GPU (FP64, rounded once) against the oracle's COMPLEX*8
    restatement of qpmfree.f90 (norm-wise 2e-5: the reference's single-
    precision sums are the noise floor)."""
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    mi.inv_setup(cinv)
    io = oracle.inv_setup(m, mod, g, pml)
    rng = np.random.default_rng(96)
    freq, py = 0.07, 1.1e-4
    cvec = lambda: (rng.standard_normal(g.ndof) + 1j * rng.standard_normal(g.ndof)).astype(np.complex64)  # noqa: E731
    u, lam, nu, alpha = cvec(), cvec(), cvec(), cvec()
    v = rng.standard_normal(io.na35).astype(np.float32)
    print(f"qpmfree {name} {cinv}: ndof={g.ndof} nnpgl={io.nnpgl}", flush=True)
    bnu, alp, p3 = mi.qp_nu_alpha(freq, py, v, lam, u)
    rb, ra, r3 = oracle.qp_nu_alpha(m, mod, g, io, freq, py, v, lam, u, cinvtype=cinv, pml=pml)
    assert relerr(alp, ra) < 2e-5
    assert relerr(bnu, rb) < 2e-5
    assert relerr(p3, r3) < 2e-5
    p1, p2 = mi.met225(freq, py, lam, nu, alpha, u)
    r1, r2 = oracle.met225(m, mod, g, io, freq, py, lam, nu, alpha, u, cinvtype=cinv, pml=pml)
    assert relerr(p1, r1) < 2e-5
    assert relerr(p2, r2) < 2e-5
    assert np.count_nonzero(p1) == np.count_nonzero(r1) > 0


@pytest.mark.parametrize("name", ["tiny_layered", "pml", "p6"])
def test_qpmfree_element_products_joint_ps(name):
    """The joint P-S inversion (CINVTYPE 'PS', NVINV = 2): f25_inv_setup's
    structure is the single-variable one with each column copied forward
    (GRADPTRS), and CADJ_NU_ALPHA's right-hand sides and P3 and MET225_MOD's
    P1, P2 on the GPU match the oracle's PS restatement (2e-5 norm-wise, the
    reference's single-precision sums); CFMAT_DIST refuses PS as VFORCE25
    does."""
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    inv = mi.inv_setup(b"PS")
    io = ps_inv(oracle.inv_setup(m, mod, g, pml))
    for k, ko in (("maskg", "maskg"), ("mygrad", "mygrad"), ("jcsc_fdist", "jcsc"), ("icsc_fdist", "icsc"),
                  ("elem_wts", "elem_wts")):
        np.testing.assert_array_equal(getattr(inv, k), getattr(io, ko), err_msg=k)
    rng = np.random.default_rng(97)
    freq, py = 0.07, 1.1e-4
    cvec = lambda: (rng.standard_normal(g.ndof) + 1j * rng.standard_normal(g.ndof)).astype(np.complex64)  # noqa: E731
    u, lam, nu, alpha = cvec(), cvec(), cvec(), cvec()
    v = rng.standard_normal(io.na35).astype(np.float32)
    print(f"qpmfree PS {name}: ndof={g.ndof} nnpgl={io.nnpgl}", flush=True)
    bnu, alp, p3 = mi.qp_nu_alpha(freq, py, v, lam, u)
    rb, ra, r3 = oracle.qp_nu_alpha(m, mod, g, io, freq, py, v, lam, u, cinvtype=b"PS", pml=pml)
    assert relerr(alp, ra) < 2e-5
    assert relerr(bnu, rb) < 2e-5
    assert relerr(p3, r3) < 2e-5
    p1, p2 = mi.met225(freq, py, lam, nu, alpha, u)
    r1, r2 = oracle.met225(m, mod, g, io, freq, py, lam, nu, alpha, u, cinvtype=b"PS", pml=pml)
    assert relerr(p1, r1) < 2e-5
    assert relerr(p2, r2) < 2e-5
    assert np.count_nonzero(p1[0::2]) > 0 and np.count_nonzero(p1[1::2]) > 0
    from fem25_amd import api
    _, code = api.cfmat_dist(inv.nz_fdist, g.ndof, freq, py, u, mi, inv)
    assert code == 7          # F25_ERR_INVTYPE


def _peff_fp64(m, mod, g, pml, freq, py, ue):
    """ASMB_PEFF in FP64 from the oracle's assembled operator (pinned to the
    compiled ASMBLE_DRIVER): -S_BE u_E on 'B' rows, +S_EB u_E on 'E' rows."""
    S = csr(g, oracle.asmble_z(m, mod, g, freq, py, pml)).tocoo()
    cnp = np.frombuffer(g.cnp, dtype="S1")
    sign = (np.where((cnp[S.row] == b"B") & (cnp[S.col] == b"E"), -1.0, 0.0)
            + np.where((cnp[S.row] == b"E") & (cnp[S.col] == b"B"), 1.0, 0.0))
    import scipy.sparse as sps
    return sps.csr_matrix((S.data * sign, (S.row, S.col)), shape=S.shape) @ ue


@pytest.mark.parametrize("name", ["pml", "p2_pml"])
def test_bielak_sources_batched(name, torch_dev):
    """The Bielak sources of a batch on the GPU: FILL_UWAVE_QMESH for four
    plane waves (two back-azimuths, P-type polarisation) against the oracle
    (itself bit-exact with the compiled routine) to 1e-15; ASMB_PEFF's
    effective force for six (omega, ky) pairs at once against the FP64
    assembled operator restricted by DOF classes (1e-12 per pair), and equal
    to -M_B S M_E u + M_E S M_B u from the full batched product (1e-13)."""
    torch = torch_dev
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    n = g.ndof
    waves = [(2.1e-4, 3.3e-4, 0.0, 20.0), (-1.5e-4, 2.2e-4, 30.0, 35.0), (1.0e-4, 4.0e-4, 0.0, 10.0),
             (3.0e-4, 1.0e-4, -45.0, 60.0)]
    kx, kz, bz, ai = (np.array([w[i] for w in waves]) for i in range(4))
    ue = torch.empty(n * len(waves), dtype=torch.complex128, device="cuda")
    print(f"bielak {name}: ndof={n} waves={len(waves)}", flush=True)
    mi.fill_uwave_dev(kx, kz, bz, ai, ue)
    got = ue.cpu().numpy().reshape(n, len(waves))
    for w, (a, b, c, d) in enumerate(waves):
        ref = oracle.fill_uwave(m, g, a, b, c, d, model=mod, pml=pml)
        assert np.abs(got[:, w] - ref).max() <= 1e-15, w
    # six pairs: each of the first three waves at two (omega, ky)
    freqs = np.array([0.05, 0.05, 0.08, 0.08, 0.11, 0.11])
    pys = np.array([0.0, 1.2e-4, 0.0, 0.8e-4, 0.5e-4, 0.0])
    om = TWOPI * freqs
    ky = om * pys
    src = [0, 1, 1, 2, 0, 2]
    u = torch.stack([ue.view(n, len(waves))[:, k] for k in src], dim=1).contiguous()
    peff = torch.empty_like(u)
    mi.peff_batch_dev(om, ky, u, peff)
    pe = peff.cpu().numpy()
    uh = u.cpu().numpy()
    for b in range(len(src)):
        want = _peff_fp64(m, mod, g, pml, freqs[b], pys[b], uh[:, b])
        assert relerr(pe[:, b], want) < 1e-12, b
    cnp = np.frombuffer(g.cnp, dtype="S1")
    mb = torch.from_numpy((cnp == b"B").astype(np.float64)).cuda()[:, None]
    me = torch.from_numpy((cnp == b"E").astype(np.float64)).cuda()[:, None]
    y1, y2 = torch.empty_like(u), torch.empty_like(u)
    mi.apply_batch_dev(om, ky, (u * me).contiguous(), y1)
    mi.apply_batch_dev(om, ky, (u * mb).contiguous(), y2)
    alt = (-mb * y1 + me * y2).cpu().numpy()
    assert relerr(pe, alt) < 1e-13
    assert np.count_nonzero(pe[cnp == b"B"]) > 0 and np.count_nonzero(pe[cnp == b"E"]) > 0
    assert not pe[(cnp != b"B") & (cnp != b"E")].any()


def test_bielak_rhs_graph_replay(torch_dev):
    """The plane-wave sources of a batch in one call (FILL_UWAVE_QMESH per
    pair, then ASMB_PEFF): the eager launches equal the separate entry
    points bit for bit, and the HIP-graph replays equal the eager result on
    every replay, with eager products and synchronisations in between and
    with new wave and frequency parameters staged into the same graph."""
    torch = torch_dev
    name = "pml"
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    n = g.ndof
    rng = np.random.default_rng(98)
    npair = 5

    def params():
        f = rng.uniform(0.03, 0.12, npair)
        om = TWOPI * f
        ky = om * rng.uniform(0.0, 1.5e-4, npair)
        kx = om * rng.uniform(-1.0, 1.0, npair) / 6000.0
        kz = om * rng.uniform(0.2, 1.0, npair) / 6000.0
        return om, ky, kx, kz, rng.uniform(-60, 60, npair), rng.uniform(5, 60, npair)

    ue = torch.empty(n * npair, dtype=torch.complex128, device="cuda")
    pe = torch.empty_like(ue)
    ue2, pe2 = torch.empty_like(ue), torch.empty_like(ue)
    x = torch.randn(n * npair, dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)
    for trial in range(2):
        om, ky, kx, kz, bz, ai = params()
        print(f"bielak graph trial {trial}: ndof={n} npair={npair}", flush=True)
        mi.bielak_rhs_dev(om, ky, kx, kz, bz, ai, ue, pe)
        mi.fill_uwave_dev(kx, kz, bz, ai, ue2)
        mi.peff_batch_dev(om, ky, ue2, pe2)
        assert torch.equal(ue, ue2) and torch.equal(pe, pe2)
        want_u, want_p = ue.clone(), pe.clone()
        for rep in range(3):
            ue.zero_()
            pe.fill_(complex(float("nan"), 0.0))
            mi.apply_batch_dev(om, ky, x, y)     # eager work between replays
            torch.cuda.synchronize()
            mi.bielak_rhs_dev(om, ky, kx, kz, bz, ai, ue, pe, graph=True)
            assert torch.equal(ue, want_u), (trial, rep)
            assert torch.equal(pe, want_p), (trial, rep)


def test_gnmfree_batched_uneven():
    """GNMFREE's sum_j alpha_j Re(J_j^H J_j p) over 12 Jacobians with 20..120
    observations each, against the oracle restatement (FP64 both)."""
    name = "pml"
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    inv = mi.inv_setup(b"PP")
    rng = np.random.default_rng(92)
    nobs = rng.integers(20, 121, size=12).astype(np.int32)
    tot = int(nobs.sum()) * inv.nnpgl
    jac = (rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(np.complex64)
    alpha = rng.uniform(0.5, 2.0, size=12)
    p = rng.standard_normal(inv.na35)
    z = mi.gnmfree(nobs, alpha, jac, p)
    zref = oracle.gnmfree(nobs, alpha, inv.nnpgl, inv.mygrad, inv.na35, jac, p)
    assert relerr(z, zref) < 1e-12


def test_gnmfree_repeatable():
    """On the 40 x 20 mesh (861 gradient columns, 14 column chunks of J p):
    two calls give z bit for bit (the chunks are added in a fixed order, no
    atomics), and z matches the oracle."""
    name = "c1"
    mi = _mi(name)
    inv = mi.inv_setup(b"PP")
    rng = np.random.default_rng(93)
    nobs = (3 * rng.integers(20, 121, size=10)).astype(np.int32)
    tot = int(nobs.sum()) * inv.nnpgl
    jac = (rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(np.complex64)
    alpha = rng.uniform(0.5, 2.0, size=nobs.size)
    p = rng.standard_normal(inv.na35)
    z1 = mi.gnmfree(nobs, alpha, jac, p)
    z2 = mi.gnmfree(nobs, alpha, jac, p)
    np.testing.assert_array_equal(z1, z2)
    zref = oracle.gnmfree(nobs, alpha, inv.nnpgl, inv.mygrad, inv.na35, jac, p)
    assert relerr(z1, zref) < 1e-12


def test_gn_hessian(torch_dev):
    """GENHLOC's local J^T J over the 861 columns of the 40 x 20 mesh for 10
    Jacobians with uneven observation counts: FP64 MFMA (SYRK) against the
    oracle, exactly symmetric, accumulated by a second call; the Jacobian
    file written from device memory equals the one written from host
    memory."""
    torch = torch_dev
    from fem25_amd import _lib
    L = _lib.load()
    mi = _mi("c1")
    inv = mi.inv_setup(b"PP")
    n = inv.nnpgl
    rng = np.random.default_rng(94)
    nobs = (3 * rng.integers(20, 121, size=10)).astype(np.int32)
    tot = int(nobs.sum()) * n
    jac = (rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(np.complex64)
    alpha = rng.uniform(0.5, 2.0, size=nobs.size)
    jd = torch.from_numpy(jac).cuda()
    h = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    args = (mi.h, nobs.size, nobs.ctypes.data_as(_lib.ip), alpha.ctypes.data_as(_lib.dp), C.c_void_p(jd.data_ptr()),
            C.c_void_p(h.data_ptr()))
    _lib.check(L.f25_gn_hessian_dev(*args), "f25_gn_hessian_dev")
    hh = h.cpu().numpy()
    ref = oracle.gn_hessian(nobs, alpha, n, jac)
    assert np.abs(hh - ref).max() <= 1e-12 * np.abs(ref).max()
    np.testing.assert_array_equal(hh, hh.T)
    _lib.check(L.f25_gn_hessian_dev(*args), "f25_gn_hessian_dev")
    assert np.abs(h.cpu().numpy() - 2 * ref).max() <= 1e-12 * np.abs(ref).max()
    import os
    import tempfile
    no = int(nobs[0])
    lobs = np.ones((no // 3, 3), np.int32)
    with tempfile.TemporaryDirectory() as d:
        pa, pb = os.path.join(d, "dev.dat"), os.path.join(d, "host.dat")
        for path, ptr, ondev in ((pa, C.c_void_p(jd.data_ptr()), 1), (pb, jac.ctypes.data_as(C.c_void_p), 0)):
            _lib.check(L.f25_write_jac(path.encode(), no, n, 3, no // 3, lobs.ctypes.data_as(_lib.ip), ptr, ondev,
                                       inv.mygrad.ctypes.data_as(_lib.ip)), "f25_write_jac")
        assert open(pa, "rb").read() == open(pb, "rb").read()


def test_pml_anchors_fixed():
    """Anchors of PML elements hold fixed parameters in the gradient-type
    products: their gradient and pseudo-Hessian entries stay zero and a
    perturbation placed only there gives F v = 0."""
    name = "pml"
    mi = _mi(name)
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    in_pml = np.zeros(m.nnpg, dtype=bool)
    for e in range(m.nelem):
        if m.cdomain[e:e + 1] == b"A":
            in_pml[m.ieng[e] - 1] = True
    om, ky, f, py = pairs(1, 2)
    u = cf.seeded_complex(g.ndof, batch=om.size, seed=31)
    lam = cf.seeded_complex(g.ndof, batch=om.size, seed=32)
    gd = mi.gradient(om, ky, u, lam)
    assert np.all(gd[in_pml] == 0.0) and np.all(gd[~in_pml] != 0.0)
    dg = mi.gn_diag(om, ky, u)
    assert np.all(dg[in_pml] == 0.0) and np.all(dg[~in_pml] > 0.0)
    v = np.zeros((m.nnpg, 3))
    v[in_pml] = np.random.default_rng(33).standard_normal((int(in_pml.sum()), 3))
    fv = mi.gn_fv(om, ky, u, v)
    assert np.all(fv == 0.0)


@pytest.mark.parametrize("name", ["tiny_layered", "pml"])
def test_asmble_dispersive(name):
    """Dispersive medium (KMDAMP/CMDAMP, VCMPLX): FP64 operator against the
    oracle at 1e-12, COMPLEX*8 CDIST against the compiled reference's
    fixture at FP32 level; the batched product (every element as per-omega
    matrices) against the oracle's dispersive operator at 1e-12, eager and
    replayed from a graph; the gradient (element blocks) refuses the
    medium."""
    import os
    from fem25_amd import _lib
    from fem25_amd.api import MeshInfo
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", f"asmq_{name}.npz"))
    md = cf.with_attenuation(mod, f["qp"], f["qs"], float(f["fref"]))
    mi = MeshInfo(m, pml)
    mi.set_model(md)
    for k in range(2):
        freq, py = float(f[f"freq{k}"]), float(f[f"py{k}"])
        s = mi.asmble(b"B", freq, py, complex128=True)
        assert relerr(s, oracle.asmble_z(m, md, g, freq, py, pml)) < FP64_TOL
        c = mi.asmble(b"B", freq, py)
        ref = f[f"cdist{k}"]
        rowmax = np.maximum.reduceat(np.abs(ref.astype(np.complex128)), g.irptr[:-1] - 1)
        rows = np.repeat(np.arange(g.ndof), np.diff(g.irptr))
        assert np.all(np.abs(c.astype(np.complex128) - ref) <= 1e-6 * rowmax[rows])
    om, ky, fq, py = pairs(2, 3)
    x = cf.seeded_complex(g.ndof, batch=om.size, seed=1)
    y = mi.apply_batch(om, ky, x)
    for b in range(om.size):
        S = csr(g, oracle.asmble_z(m, md, g, fq[b], py[b], pml))
        assert relerr(y[b], S @ x[b]) < FP64_TOL, b
    import torch
    xd = torch.from_numpy(np.ascontiguousarray(x.T)).cuda()
    yg, ye = torch.zeros_like(xd), torch.zeros_like(xd)
    for rep in range(2):
        mi.apply_batch_dev(om * (1 + 0.05 * rep), ky, xd, yg, _lib.F25_APPLY_F64 | _lib.F25_APPLY_GRAPH)
        mi.apply_batch_dev(om * (1 + 0.05 * rep), ky, xd, ye, _lib.F25_APPLY_F64)
        torch.cuda.synchronize()
        assert torch.equal(yg, ye), rep
    with pytest.raises(_lib.Fem25Error):
        mi.gradient(om[:1], ky[:1], x[:1], x[:1])
    mi.set_model(mod)        # back to the elastic operator
    y2 = mi.apply_batch(om[:1], ky[:1], x[:1])
    assert relerr(y2[0], oracle_S(name, fq[0], py[0]) @ x[0]) < FP64_TOL
    s = mi.asmble(b"B", 0.05, 0.7e-4, complex128=True)
    assert relerr(s, oracle.asmble_z(m, mod, g, 0.05, 0.7e-4, pml)) < FP64_TOL


@pytest.mark.parametrize("name", ["p5", "p6"])
def test_apply_batch_modes_high_order(name):
    """FP32 and refined products at orders 5 and 6 (remainder rows on the
    vector ALUs, per-component launches at p6) against the FP64 oracle."""
    import torch
    from fem25_amd import _lib
    mi = _mi(name)
    om, ky, f, py = pairs(2, 4)
    g = oracle_graph(name)
    x = cf.seeded_complex(g.ndof, batch=om.size, seed=6)
    yref = np.stack([oracle_S(name, f[b], py[b]) @ x[b] for b in range(om.size)])
    xd = torch.from_numpy(np.ascontiguousarray(x.T)).cuda()
    x32 = xd.to(torch.complex64)
    y32 = torch.zeros_like(x32)
    mi.apply_batch_dev(om, ky, x32, y32, _lib.F25_APPLY_F32)
    yr = torch.zeros_like(xd)
    mi.apply_batch_dev(om, ky, xd, yr, _lib.F25_APPLY_F32_REFINED)
    e32 = max(relerr(y32.cpu().numpy().T[b], yref[b]) for b in range(om.size))
    er = max(relerr(yr.cpu().numpy().T[b], yref[b]) for b in range(om.size))
    print(f"{name}: fp32 {e32:.2e}, refined {er:.2e}")
    assert e32 < 1e-5
    assert er < 1e-8


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _adj_case(seed, ndof=3000, nfreq=4, nrec=60, nsrc=3):
    rng = np.random.default_rng(seed)
    mrdof = rng.integers(1, ndof + 1, size=(nrec, 3)).astype(np.int32)
    mrdof[4, 1] = mrdof[9, 0]                   # two receivers on one DOF
    yrec = rng.uniform(-3e4, 3e4, nrec)
    cv = lambda *s: (rng.standard_normal(s) + 1j * rng.standard_normal(s)).astype(np.complex64)  # noqa: E731
    est = cv(nsrc, nrec, nfreq, 3)
    obs = (est + 0.4 * cv(nsrc, nrec, nfreq, 3)).astype(np.complex64)
    obs[0, ::7, :, 2] = 0                       # missing data
    w = rng.uniform(0.3, 2.0, (nsrc, nrec, nfreq, 3)).astype(np.float32)
    return rng, mrdof, yrec, est, obs, w


def test_exresp25d(torch_dev):
    """EXRESP25D on the GPU (host and device forms) against the oracle's
    restatement (itself bit-exact with the compiled routine): 2e-7 of the
    largest value (single-precision sin/cos of the receiver phase)."""
    from fem25_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(111)
    ndof, nsrc, nrec = 4000, 3, 45
    mrdof = rng.integers(0, ndof + 1, size=(nrec, 3)).astype(np.int32)
    recv = (rng.standard_normal((nrec, 3)) + 1j * rng.standard_normal((nrec, 3))).astype(np.complex64)
    yrec = rng.uniform(-3e4, 3e4, nrec)
    py = rng.uniform(0, 1.4e-4, nsrc)
    sol = (rng.standard_normal(nsrc * ndof) + 1j * rng.standard_normal(nsrc * ndof)).astype(np.complex64)
    ref = oracle.exresp25d(ndof, 0.083, py, yrec, sol, mrdof, recv)
    est = np.zeros_like(ref)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    _lib.check(L.f25_exresp25d(ndof, nsrc, nrec, 0.083, _ptr(py, _lib.dp), _ptr(yrec, _lib.dp), vp(sol),
                               _ptr(mrdof, _lib.ip), vp(recv), vp(est)), "exresp25d")
    assert np.abs(est - ref).max() <= 2e-7 * np.abs(ref).max()
    assert np.array_equal(est == 0, ref == 0)
    torch = torch_dev
    sd = torch.from_numpy(sol).cuda()
    ed = torch.zeros(est.size, dtype=torch.complex64, device="cuda")
    _lib.check(L.f25_exresp25d_dev(ndof, nsrc, nrec, 0.083, _ptr(py, _lib.dp), _ptr(yrec, _lib.dp),
                                   C.c_void_p(sd.data_ptr()), _ptr(mrdof, _lib.ip), vp(recv),
                                   C.c_void_p(ed.data_ptr())), "exresp25d_dev")
    np.testing.assert_array_equal(ed.cpu().numpy().reshape(est.shape), est)


@pytest.mark.parametrize("norm,dtmax", [(2, 0.0), (2, 2.0), (1, 0.0), (3, 0.0), (4, 0.0)])
def test_cadjlam_rhs(norm, dtmax):
    """CADJLAM's right-hand side on the GPU for a batch of (frequency,
    source) pairs covering every residual type, both data forms (complex and
    magnitude/phase), every HDIFFER option and a rotated frame, against the
    oracle (whose parts are bit-exact with the compiled BPRESID4, CRESID,
    CROTATE, CPHM2CM): 2e-6 of the pair's largest value; NaN exactly where
    the reference's EPS = 0 combination norm gives NaN.  The L1 and Huber
    residuals dif/|dif| magnify an ulp of sin, cos or atan2 (which no two
    math libraries round alike) by 1/|dif|; for them each entry may also
    differ by 10x the largest change the oracle itself shows when its data
    and estimates move by four parts in 2^23 (an ulp of a phase near pi)."""
    from fem25_amd import _lib
    L = _lib.load()
    rng, mrdof, yrec, est, obs, w = _adj_case(112 + norm)
    nsrc, nrec, nfreq, _ = est.shape
    ndof = 3000
    pairs = [(f + 1, s + 1) for f in range(nfreq) for s in range(nsrc)]
    ifreq = np.array([p[0] for p in pairs], np.int32)
    isrc = np.array([p[1] for p in pairs], np.int32)
    freq = 0.02 + 0.02 * ifreq.astype(np.float64)
    py = rng.uniform(0, 1.3e-4, len(pairs))
    vp = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    for irestp in (1, 2, 3):
        for lunwrap in (0, 1):
            ib = (irestp + 2 * lunwrap + norm) % 7
            az = 17.5 * irestp - 30.0
            print(f"cadjlam norm={norm} dtmax={dtmax} irestp={irestp} lunwrap={lunwrap} ibphase={ib}", flush=True)
            rhs = np.zeros((len(pairs), ndof), np.complex64)
            iw = C.c_int(-1)
            _lib.check(L.f25_cadjlam_rhs(ndof, nfreq, nrec, nsrc, len(pairs), _ptr(ifreq, _lib.ip),
                                         _ptr(isrc, _lib.ip), _ptr(freq, _lib.dp), _ptr(py, _lib.dp), az,
                                         norm, irestp, 0.05, dtmax, ib, lunwrap, _ptr(yrec, _lib.dp),
                                         _ptr(mrdof, _lib.ip), vp(est), vp(obs), _ptr(w, _lib.fp), vp(rhs),
                                         C.byref(iw)), "cadjlam_rhs")
            assert iw.value == 0
            for k, (f, s) in enumerate(pairs):
                args = (ndof, f, s, freq[k], py[k], az, norm, irestp, 0.05, dtmax, ib, bool(lunwrap), yrec, mrdof)
                ref, iwr = oracle.cadjlam_rhs(*args, est, obs, w)
                assert np.array_equal(np.isnan(rhs[k]), np.isnan(ref))
                ok = ~np.isnan(ref)
                scale = np.abs(ref[ok]).max() if ok.any() else 0.0
                tol = np.full(ndof, 2e-6 * max(scale, 1e-30))
                if norm in (1, 3):
                    sens = np.zeros(ndof)
                    for t in range(3):
                        sg = lambda: rng.choice([-1.0, 1.0], est.shape)   # noqa: E731
                        pe = 1 + 2.0 ** -21 * (sg() + 1j * sg())      # moves magnitude and phase
                        po = 1 + 2.0 ** -21 * (sg() + 1j * sg())
                        alt, _ = oracle.cadjlam_rhs(*args, (est * pe).astype(np.complex64),
                                                    (obs * po).astype(np.complex64), w)
                        sens = np.maximum(sens, np.abs(alt - ref))
                    tol += 10 * sens
                err = np.abs(rhs[k] - ref)
                worst = int(np.argmax(np.where(ok, err / tol, 0)))
                assert np.all(err[ok] <= tol[ok]), (k, irestp, lunwrap, worst, err[worst], tol[worst], scale)
    # an invalid receiver DOF is skipped and reported
    bad = mrdof.copy()
    bad[3, 2] = ndof + 5
    rhs = np.zeros((1, ndof), np.complex64)
    iw = C.c_int(0)
    _lib.check(L.f25_cadjlam_rhs(ndof, nfreq, nrec, nsrc, 1, _ptr(ifreq, _lib.ip), _ptr(isrc, _lib.ip),
                                 _ptr(freq, _lib.dp), _ptr(py, _lib.dp), 0.0, 2, 3, 0.05, 0.0, 0, 0,
                                 _ptr(yrec, _lib.dp), _ptr(bad, _lib.ip), vp(est), vp(obs),
                                 _ptr(w, _lib.fp), vp(rhs), C.byref(iw)), "cadjlam_rhs")
    assert iw.value == 1


def _covd_case(seed, nrec, nfreq=3, nsrc=3, ndof=5000):
    rng = np.random.default_rng(seed)
    mrdof = rng.integers(1, ndof + 1, size=(nrec, 3)).astype(np.int32)
    mrdof[4, 1] = mrdof[9, 0]                   # two receivers on one DOF
    yrec = rng.uniform(-3e4, 3e4, nrec)
    x = np.cumsum(rng.uniform(0.5, 3.0, nrec))
    dx, ierr = oracle.rec_space1d(x)
    assert ierr == 0
    cv = lambda *s: (rng.standard_normal(s) + 1j * rng.standard_normal(s)).astype(np.complex64)  # noqa: E731
    est = cv(nsrc, nrec, nfreq, 3)
    obs = (est + 0.4 * cv(nsrc, nrec, nfreq, 3)).astype(np.complex64)
    obs[rng.random(obs.shape) < 0.2] = 0        # gaps in every component
    obs[0, :, 1, 0] = 0                         # a dead component
    obs[1, :, 0, :] = 0                         # a dead frequency
    obs[2, :, 2, 1] = 0
    obs[2, 7, 2, 1] = 1 + 1j                    # a component with one observed receiver
    w = rng.uniform(0.3, 2.0, (nsrc, nrec, nfreq, 3)).astype(np.float32)
    return rng, ndof, mrdof, yrec, dx, est, obs, w


@pytest.mark.parametrize("ldwght,dtmax,nrec", [(1, 0.0, 60), (0, 0.0, 60), (1, 1.5, 60), (1, 0.0, 700)])
def test_cadjlam_rhs_covd(ldwght, dtmax, nrec):
    """CADJLAM's right-hand side with the data covariance (LAPLACE_1DV2,
    PERM_OBS, BPRESID_COVD4) on the GPU for every (frequency, source) pair,
    every residual type, both data forms and a rotated frame, against the
    oracle (whose parts are pinned to the compiled routines): 2e-6 of the
    pair's largest value.  The data hold gaps in every component (the
    reference's roughening matrix is then numbered apart from PERM_OBS's
    positions), a dead component, a dead frequency and a component with a
    single observed receiver; 700 receivers take the position scan over
    several blocks of 256."""
    from fem25_amd import _lib
    L = _lib.load()
    rng, ndof, mrdof, yrec, dx, est, obs, w = _covd_case(120 + nrec + 2 * ldwght, nrec)
    nsrc, _, nfreq, _ = est.shape
    pairs = [(f + 1, s + 1) for f in range(nfreq) for s in range(nsrc)]
    ifreq = np.array([p[0] for p in pairs], np.int32)
    isrc = np.array([p[1] for p in pairs], np.int32)
    freq = 0.02 + 0.02 * ifreq.astype(np.float64)
    py = rng.uniform(0, 1.3e-4, len(pairs))
    vp = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    for irestp in (1, 2, 3):
        for lunwrap in (0, 1):
            ib = (irestp + 2 * lunwrap) % 7
            az = 17.5 * irestp - 30.0
            print(f"cadjlam_covd ldwght={ldwght} dtmax={dtmax} nrec={nrec} irestp={irestp} lunwrap={lunwrap}",
                  flush=True)
            rhs = np.zeros((len(pairs), ndof), np.complex64)
            iw = C.c_int(-1)
            _lib.check(L.f25_cadjlam_rhs_covd(ndof, nfreq, nrec, nsrc, len(pairs), _ptr(ifreq, _lib.ip),
                                              _ptr(isrc, _lib.ip), _ptr(freq, _lib.dp), _ptr(py, _lib.dp), az,
                                              irestp, dtmax, ib, lunwrap, ldwght, _ptr(dx, _lib.dp),
                                              _ptr(yrec, _lib.dp), _ptr(mrdof, _lib.ip), vp(est), vp(obs),
                                              _ptr(w, _lib.fp), vp(rhs), C.byref(iw)), "cadjlam_rhs_covd")
            assert iw.value == 0
            for k, (f, s) in enumerate(pairs):
                ref, iwr, ierr = oracle.cadjlam_rhs_covd(ndof, f, s, freq[k], py[k], az, irestp, dtmax, ib,
                                                         bool(lunwrap), ldwght, dx, yrec, mrdof, est, obs, w)
                assert ierr == 0 and iwr == 0
                scale = np.abs(ref).max()
                err = np.abs(rhs[k] - ref)
                assert err.max() <= 2e-6 * max(scale, 1e-30), (k, irestp, lunwrap, err.max(), scale)
                if not obs[s - 1, :, f - 1].any():
                    assert not rhs[k].any()


def test_cadjlam_rhs_covd_dev_matches_host_form(torch_dev):
    """The device-pointer form of the data-covariance right-hand side gives
    the host form's values bit for bit."""
    from fem25_amd import _lib
    torch = torch_dev
    L = _lib.load()
    rng, ndof, mrdof, yrec, dx, est, obs, w = _covd_case(131, 90)
    nsrc, nrec, nfreq, _ = est.shape
    ifreq = np.repeat(np.arange(1, nfreq + 1), nsrc).astype(np.int32)
    isrc = np.tile(np.arange(1, nsrc + 1), nfreq).astype(np.int32)
    npair = ifreq.size
    freq = 0.03 * ifreq.astype(np.float64)
    py = np.full(npair, 0.7e-4)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    rhs = np.zeros((npair, ndof), np.complex64)
    iw = C.c_int(-1)
    _lib.check(L.f25_cadjlam_rhs_covd(ndof, nfreq, nrec, nsrc, npair, _ptr(ifreq, _lib.ip), _ptr(isrc, _lib.ip),
                                      _ptr(freq, _lib.dp), _ptr(py, _lib.dp), 12.0, 3, 0.0, 1, 0, 1,
                                      _ptr(dx, _lib.dp), _ptr(yrec, _lib.dp), _ptr(mrdof, _lib.ip), vp(est), vp(obs),
                                      _ptr(w, _lib.fp), vp(rhs), C.byref(iw)), "cadjlam_rhs_covd")
    ed, od, wd = (torch.from_numpy(a).cuda() for a in (est, obs, w))
    rd = torch.full((npair * ndof,), 7 + 7j, dtype=torch.complex64, device="cuda")
    _lib.check(L.f25_cadjlam_rhs_covd_dev(ndof, nfreq, nrec, nsrc, npair, _ptr(ifreq, _lib.ip), _ptr(isrc, _lib.ip),
                                          _ptr(freq, _lib.dp), _ptr(py, _lib.dp), 12.0, 3, 0.0, 1, 0, 1,
                                          _ptr(dx, _lib.dp), _ptr(yrec, _lib.dp), _ptr(mrdof, _lib.ip),
                                          C.c_void_p(ed.data_ptr()), C.c_void_p(od.data_ptr()),
                                          C.cast(C.c_void_p(wd.data_ptr()), _lib.fp), C.c_void_p(rd.data_ptr()),
                                          C.byref(iw)), "cadjlam_rhs_covd_dev")
    np.testing.assert_array_equal(rd.cpu().numpy().reshape(npair, ndof), rhs)


def test_qp_rhs_full_size_against_operator_difference():
    """At the benchmark size (200 x 100 elements, p = 4, 963,603 DOFs): with
    every anchor of a uniform mesh inverted (ELEM_WTS = 1), CADJ_NU_ALPHA's
    RHSALP is S[lambda = f, mu = 0, rho = 0] u, f = 2 Vp V at the anchors, and
    RHSBNU the same (Hermitian) operator on lambda; S is linear in lambda, so
    both equal S(model with lambda + f) x - S(model) x from the batched FP64
    product -- two independent kernels (norm-wise 1e-5: the outputs are
    COMPLEX*8)."""
    import dataclasses
    from fem25_amd.api import MeshInfo
    m = cf.structured_mesh(200, 100, 4)
    mod = cf.layered_model(m, seed=2)
    mi = MeshInfo(m)
    mi.set_model(mod)
    inv = mi.inv_setup(b"PP")
    assert np.all(inv.elem_wts == 1.0)
    rng = np.random.default_rng(97)
    n = mi.ndof
    u = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    lam = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    v = rng.standard_normal(inv.na35).astype(np.float32)
    freq, py = 0.07, 1.1e-4
    print(f"qp full size: ndof={n} nnpgl={inv.nnpgl}", flush=True)
    bnu, alp, _ = mi.qp_nu_alpha(freq, py, v, lam, u)
    vp = np.sqrt((mod.ecoeff[:, 0] + 2 * mod.ecoeff[:, 1]) / mod.dens)
    f = np.zeros(m.nnpg)
    sel = inv.maskg > 0
    f[sel] = 2 * vp[sel] * v[inv.maskg[sel] - 1].astype(np.float64)
    om = np.full(2, 2 * np.pi * freq)
    ky = om * py
    x = np.stack([u.astype(np.complex128), lam.astype(np.complex128)])
    y0 = mi.apply_batch(om, ky, x)
    ec = np.asfortranarray(np.stack([mod.ecoeff[:, 0] + f, mod.ecoeff[:, 1]], axis=1))
    mi.set_model(dataclasses.replace(mod, ecoeff=ec))
    y1 = mi.apply_batch(om, ky, x)
    mi.set_model(mod)
    d = y1 - y0
    assert relerr(alp, d[0]) < 1e-5
    assert relerr(bnu, d[1]) < 1e-5
