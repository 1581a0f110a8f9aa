"""The adjoint-gradient workload with Bielak plane-wave sources (tests/config3.py:
40 x 20 elements of order 4 with two PML rings and one Bielak ring, three
P/SV plane waves at 0.05 Hz, 60 free-surface receivers, observed data from
the model with doubled anomalies, L2 misfit).

CPU: the oracle's exact gradient against central finite differences of the
misfit at ten seeded inversion anchors (tests/golden/c3_fd.npz, written by
tests/golden/make_config3_fd.py), and the plane waves against the
homogeneous operator.  GPU: every step of the path on the device against
the oracle: the COMPLEX*8 operator, the Bielak effective forces, the batched
FP64 product with the PML elements, receiver extraction, the adjoint
right-hand sides, the exact gradient (also against the finite differences)
and the reference-parity gradient (CFMAT_DIST + ADJFLAM_DIST)."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from tests import config3 as c3
from tests.helpers import csr, relerr

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fd_fixture():
    return np.load(os.path.join(GOLD, "c3_fd.npz"), allow_pickle=False)


def _oracle_gradient(st):
    g = np.zeros((st.mesh.nnpg, 3))
    for s in range(st.u.shape[0]):
        g += oracle.grad_exact(st.mesh, st.model, st.g, c3.FREQ, c3.PY, st.u[s], st.lam[s], c3.PML)[1]
    return g


def _fd_check(g, fx):
    anchors, fd = fx["anchors"], fx["fd"]
    scale = np.abs(fd).max()
    for i, a in enumerate(anchors):
        for p in range(3):
            assert abs(g[a, p] - fd[i, p]) <= 1e-5 * max(abs(fd[i, p]), 1e-3 * scale), (a, p, g[a, p], fd[i, p])


def test_config3_oracle_gradient_vs_finite_differences():
    """The setup is the one the fixture was made on (same misfit to 1e-12) and
    the oracle's exact gradient matches the finite differences."""
    st = c3.setup()
    fx = _fd_fixture()
    assert abs(st.chi - float(fx["chi"])) <= 1e-12 * st.chi
    np.testing.assert_array_equal(c3.fd_anchors(st), fx["anchors"])
    _fd_check(_oracle_gradient(st), fx)


def test_config3_plane_waves_solve_homogeneous_operator():
    """On the homogeneous model each plane wave satisfies the elastic
    equations: S u_pw vanishes on the rows of interior nodes (away from the
    free surface and the PML), to the discretisation error of order 4."""
    st = c3.setup()
    m, g = st.mesh, st.g
    hom = c3.anomaly_model(m, scale=0.0)
    S = csr(g, oracle.asmble_z(m, hom, g, c3.FREQ, c3.PY, c3.PML))
    xd, zd, _, _ = c3.node_xz(m, g)
    inner = (st.cnp == b"I") & (zd < -6e3)
    for s in range(st.ue.shape[0]):
        r = S @ st.ue[s]
        ref = abs(S).multiply(1.0).dot(np.abs(st.ue[s]))
        assert np.abs(r[inner]).max() < 1e-3 * ref[inner].max(), s


@pytest.mark.gpu
def test_config3_pipeline_on_gpu():
    import torch
    from fem25_amd import _lib, api
    from fem25_amd.api import MeshInfo
    assert torch.cuda.is_available()
    L = _lib.load()
    st = c3.setup()
    m, g, nsrc = st.mesh, st.g, st.ue.shape[0]
    print(f"config3: ndof={g.ndof} sources={nsrc} receivers={c3.NREC}", flush=True)
    mi = MeshInfo(m, c3.PML)
    mi.set_model(st.model)
    gg = mi.graph()
    np.testing.assert_array_equal(gg.lm, g.lm)
    assert gg.cnp == g.cnp
    # COMPLEX*8 operator (entrywise against the row scale) and P_eff
    cd = mi.asmble(b"B", c3.FREQ, c3.PY)
    cref = oracle.asmble_cdist(m, st.model, g, c3.FREQ, c3.PY, c3.PML)
    rowmax = np.maximum.reduceat(np.abs(cref.astype(np.complex128)), g.irptr[:-1] - 1)
    rows = np.repeat(np.arange(g.ndof), np.diff(g.irptr))
    assert np.all(np.abs(cd.astype(np.complex128) - cref) <= 1e-6 * rowmax[rows])
    S = csr(g, oracle.asmble_z(m, st.model, g, c3.FREQ, c3.PY, c3.PML))
    for s in range(nsrc):
        ue8 = st.ue[s].astype(np.complex64)
        pe = api.asmb_peff(mi, cref, ue8)
        assert relerr(pe, oracle.asmb_peff(g, cref, ue8)) < 1e-5, s
        assert relerr(pe, c3.peff_fp64(S, st.cnp, st.ue[s])) < 1e-5, s
    # batched FP64 product with the PML elements on the three forward fields
    om = np.full(nsrc, 2 * np.pi * c3.FREQ)
    ky = om * c3.PY
    y = mi.apply_batch(om, ky, st.u)
    for s in range(nsrc):
        assert relerr(y[s], S @ st.u[s]) < 1e-12, s
    # receiver extraction (EXRESP25D), unit response
    mrdof = np.ascontiguousarray(st.rec_dofs, np.int32)
    recv = np.ones((c3.NREC, 3), np.complex64)
    yrec = np.zeros(c3.NREC)
    py = np.full(nsrc, c3.PY)
    sol = st.u.astype(np.complex64).ravel()
    est = np.zeros((nsrc, c3.NREC, 3), np.complex64)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    ptr = lambda a, t: a.ctypes.data_as(t)        # noqa: E731
    _lib.check(L.f25_exresp25d(g.ndof, nsrc, c3.NREC, c3.FREQ, ptr(py, _lib.dp), ptr(yrec, _lib.dp), vp(sol),
                               ptr(mrdof, _lib.ip), vp(recv), vp(est)), "exresp25d")
    eref = oracle.exresp25d(g.ndof, c3.FREQ, py, yrec, sol, mrdof, recv)
    assert np.abs(est - eref).max() <= 2e-7 * np.abs(eref).max()
    assert relerr(est, st.u[:, st.rec_dofs - 1]) < 1e-6
    # adjoint right-hand sides (CADJLAM: L2 norm, full complex residual, unit weights)
    e4 = est[:, :, None, :]
    o4 = st.d.astype(np.complex64)[:, :, None, :]
    w4 = np.ones_like(e4, dtype=np.float32)
    ifreq = np.ones(nsrc, np.int32)
    isrc = np.arange(1, nsrc + 1, dtype=np.int32)
    fq = np.full(nsrc, c3.FREQ)
    rhs = np.zeros((nsrc, g.ndof), np.complex64)
    iw = C.c_int(-1)
    _lib.check(L.f25_cadjlam_rhs(g.ndof, 1, c3.NREC, nsrc, nsrc, ptr(ifreq, _lib.ip), ptr(isrc, _lib.ip),
                                 ptr(fq, _lib.dp), ptr(py, _lib.dp), 0.0, 2, 3, 0.0, 0.0, 0, 0,
                                 ptr(yrec, _lib.dp), ptr(mrdof, _lib.ip), vp(np.ascontiguousarray(e4)),
                                 vp(np.ascontiguousarray(o4)), ptr(w4, _lib.fp), vp(rhs), C.byref(iw)),
               "cadjlam_rhs")
    assert iw.value == 0
    for s in range(nsrc):
        ref, _ = oracle.cadjlam_rhs(g.ndof, 1, s + 1, c3.FREQ, c3.PY, 0.0, 2, 3, 0.0, 0.0, 0, False, yrec, mrdof,
                                    e4, o4, w4)
        assert np.abs(rhs[s] - ref).max() <= 2e-6 * np.abs(ref).max(), s
        want = np.zeros(g.ndof, np.complex128)
        np.add.at(want, st.rec_dofs.ravel() - 1, (st.d[s] - st.u[s, st.rec_dofs - 1]).ravel())
        assert np.abs(rhs[s] - np.conj(want)).max() <= 1e-5 * np.abs(want).max(), s
    # exact gradient over the three (omega, ky, source) triples
    gd = mi.gradient(om, ky, st.u, st.lam)
    gref = _oracle_gradient(st)
    for p in range(3):
        assert relerr(gd[:, p], gref[:, p]) < 1e-11, p
    _fd_check(gd, _fd_fixture())
    # reference-parity gradient: CFMAT_DIST then ADJFLAM_DIST per source
    inv = mi.inv_setup(b"PP")
    io = oracle.inv_setup(m, st.model, g, c3.PML)
    np.testing.assert_array_equal(inv.maskg, io.maskg)
    gsum = np.zeros(inv.na35, np.float64)
    rsum = np.zeros(inv.na35, np.float64)
    for s in range(nsrc):
        wave = st.u[s].astype(np.complex64)
        fm, ierr = api.cfmat_dist(inv.nz_fdist, g.ndof, c3.FREQ, c3.PY, wave, mi, inv)
        assert ierr == 0
        fref = oracle.cfmat_dist(m, st.model, g, io, c3.FREQ, c3.PY, wave)
        assert relerr(fm, fref) < 2e-6, s
        lam8 = st.lam[s].astype(np.complex64)
        gsum += api.adjflam_dist(mi, inv, fref, lam8)
        rsum += oracle.adjflam_dist(io, fref, lam8)
    assert relerr(gsum, rsum) < 1e-5
    mi.close()


@pytest.mark.gpu
def test_config3_refinement_of_single_precision_solves():
    """The forward solves factorised in single precision (as CMUMPS does) and
    refined with the FP64 residual of the batched product (with the PML
    elements): the residual falls by orders of magnitude per step and the
    refined fields match the FP64 solutions to 1e-10 per source."""
    import scipy.sparse.linalg as sla
    import torch
    from fem25_amd.api import MeshInfo
    st = c3.setup()
    m, g, nsrc = st.mesh, st.g, st.ue.shape[0]
    mi = MeshInfo(m, c3.PML)
    mi.set_model(st.model)
    S = csr(g, oracle.asmble_z(m, st.model, g, c3.FREQ, c3.PY, c3.PML))
    lu32 = sla.splu(S.astype(np.complex64).tocsc())
    f = np.stack([c3.peff_fp64(S, st.cnp, st.ue[s]) for s in range(nsrc)])
    u = np.stack([lu32.solve(f[s].astype(np.complex64)).astype(np.complex128) for s in range(nsrc)])
    om = np.full(nsrc, 2 * np.pi * c3.FREQ)
    ky = om * c3.PY
    fd = torch.from_numpy(np.ascontiguousarray(f.T)).cuda()
    hist = []
    for it in range(6):
        ud = torch.from_numpy(np.ascontiguousarray(u.T)).cuda()
        rd = torch.empty_like(ud)
        rn = mi.residual_batch_dev(om, ky, ud, fd, rd)
        hist.append(rn)
        r = rd.cpu().numpy().T
        if it == 0:
            for s in range(nsrc):   # the residual itself against the oracle operator
                want = f[s] - S @ u[s]   # cancels to ~1e-6 |f|: compare against |f|
                assert np.linalg.norm(r[s] - want) < 1e-13 * np.linalg.norm(f[s]), s
        print(f"refinement step {it}: |r|/|f| per source {np.array2string(rn, precision=2)}", flush=True)
        if rn.max() < 1e-13:
            break
        u = u + np.stack([lu32.solve(r[s].astype(np.complex64)).astype(np.complex128) for s in range(nsrc)])
    assert hist[0].max() > 1e-8          # the single-precision solve alone misses the bar
    assert hist[-1].max() < 1e-12
    for s in range(nsrc):
        assert relerr(u[s], st.u[s]) < 1e-10, s
    mi.close()
