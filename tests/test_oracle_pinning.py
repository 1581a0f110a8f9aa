"""Pins the oracle (oracle.c) to the reference's own outputs (golden fixtures
from the compiled Fortran, tests/golden/make_golden.py) and to what the
mathematics fixes (closed forms, invariants, finite differences)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
from numpy.polynomial import legendre as npleg
from scipy.sparse.linalg import spsolve

import oracle
from fem25_amd import configs as cf
from tests.helpers import TWOPI, csr, mesh_case, oracle_graph, ps_inv, relerr

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# -- quadrature and shape functions ------------------------------------------------

def test_gaussq_bitexact_vs_reference():
    z = gold("gaussq.npz")
    for n in range(3, 20):
        x, w = oracle.gll(n)
        np.testing.assert_array_equal(x, z[f"x{n}"])
        np.testing.assert_array_equal(w, z[f"w{n}"])


@pytest.mark.parametrize("n", [3, 5, 7, 13, 17])
def test_gll_closed_form(n):
    """Interior nodes are the roots of P'_{n-1}; w = 2 / (n (n-1) P_{n-1}(x)^2)."""
    x, w = oracle.gll(n)
    c = np.zeros(n)
    c[-1] = 1.0
    roots = np.sort(npleg.legroots(npleg.legder(c)))
    np.testing.assert_allclose(x[1:-1], roots, atol=2e-15)
    np.testing.assert_allclose(x[[0, -1]], [-1, 1], atol=1e-15)
    wc = 2.0 / (n * (n - 1) * npleg.legval(x, c) ** 2)
    np.testing.assert_allclose(w, wc, rtol=1e-13)
    assert abs(w.sum() - 2.0) < 1e-14


@pytest.mark.parametrize("p", [4, 5, 6])
def test_dshl_vs_reference(p):
    z = gold("dshl.npz")
    xi, xq, wq, shl = oracle.dshl(p)
    np.testing.assert_array_equal(xq, z[f"xq{p}"])
    np.testing.assert_array_equal(wq, z[f"wq{p}"])
    assert np.abs(shl - z[f"shl{p}"]).max() < 1e-13


@pytest.mark.parametrize("p", [2, 4, 6])
def test_dshl_polynomial_exactness(p):
    """Partition of unity and exact interpolation/differentiation of x^i z^j, i, j <= p."""
    xi, xq, wq, shl = oracle.dshl(p)
    nl = p + 1
    ax = np.tile(xi, nl)            # node a = ie*nl + ix
    az = np.repeat(xi, nl)
    qx, qz = np.meshgrid(xq, xq)    # shl[iz][ix]
    assert np.abs(shl[..., 2].sum(-1) - 1).max() < 1e-13
    assert np.abs(shl[..., 0].sum(-1)).max() < 1e-12
    for i in range(p + 1):
        for j in range(p + 1 - i):
            f = ax ** i * az ** j
            val = shl[..., 2] @ f
            dx = shl[..., 0] @ f
            dz = shl[..., 1] @ f
            np.testing.assert_allclose(val, qx ** i * qz ** j, atol=1e-12)
            np.testing.assert_allclose(dx, i * qx ** max(i - 1, 0) * qz ** j * (i > 0), atol=1e-11)
            np.testing.assert_allclose(dz, j * qx ** i * qz ** max(j - 1, 0) * (j > 0), atol=1e-11)


# -- numbering ---------------------------------------------------------------------

@pytest.mark.parametrize("case", ["tiny", "p2_pml", "pml"])
def test_graph_bitexact_vs_reference(case):
    z = gold(f"graph_{case}.npz")
    g = oracle_graph(case)
    assert (g.nnp, g.ndof) == (int(z["nnp"]), int(z["ndof"]))
    for k in ("ien", "lm", "irptr", "jcptr"):
        np.testing.assert_array_equal(getattr(g, k), z[k])
    assert g.cnp == z["cnp"].tobytes()


def test_graph_counts_and_bruteforce_pattern():
    m, _, _ = mesh_case("p2_pml")
    g = oracle_graph("p2_pml")
    assert g.nnp == (m.nx * m.nord + 1) * (m.nz * m.nord + 1)
    # brute-force element adjacency (the GENCRS_DUMB idea, graph25.f:270-311)
    rows, cols = [], []
    for e in range(m.nelem):
        d = g.lm[e].ravel()
        d = d[d > 0]
        rr, cc = np.meshgrid(d, d, indexing="ij")
        rows.append(rr.ravel())
        cols.append(cc.ravel())
    A = sp.csr_matrix((np.ones(sum(r.size for r in rows)), (np.concatenate(rows) - 1,
                       np.concatenate(cols) - 1)), shape=(g.ndof, g.ndof))
    A.sum_duplicates()
    A.sort_indices()
    np.testing.assert_array_equal(A.indptr + 1, g.irptr)
    np.testing.assert_array_equal(A.indices + 1, g.jcptr)


def test_config1_numbering_checksum():
    import hashlib
    m, _, _ = mesh_case("c1")
    g = oracle_graph("c1")
    h = hashlib.sha256()
    for a in (g.ien, g.lm, g.irptr, g.jcptr):
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    h.update(g.cnp)
    assert h.hexdigest() == str(gold("c1_graph_sha.npz")["sha"])
    assert g.ndof == 3 * (40 * 4 + 1) * (20 * 4 + 1)


# -- element operators and assembly -----------------------------------------------------

@pytest.mark.parametrize("case", ["tiny", "p2_pml", "pml"])
def test_cdist_vs_reference(case):
    z = gold(f"asm_{case}.npz")
    m, mod, pml = mesh_case(case)
    g = oracle_graph(case)
    cd = oracle.asmble_cdist(m, mod, g, float(z["freq"]), float(z["py"]), pml)
    ref = z["cdist"]
    rowmax = np.maximum.reduceat(np.abs(ref.astype(np.complex128)), g.irptr[:-1] - 1)
    rows = np.repeat(np.arange(g.ndof), np.diff(g.irptr))
    assert np.all(np.abs(cd.astype(np.complex128) - ref) <= 1e-6 * rowmax[rows])
    assert np.mean(cd == ref) > 0.9      # mostly bit-identical
    zz = oracle.asmble_z(m, mod, g, float(z["freq"]), float(z["py"]), pml)
    assert relerr(zz, ref) < 1e-6


def test_rigid_translations_in_stiffness_null_space():
    m, mod, pml = mesh_case("tiny_layered")
    g = oracle_graph("tiny_layered")
    S = csr(g, oracle.asmble_z(m, mod, g, 1e-9, 0.0, pml))   # omega ~ 0, ky = 0
    comp = np.zeros(g.ndof, int)
    comp[g.lm.reshape(-1, 3)[:, 0] - 1] = 0
    comp[g.lm.reshape(-1, 3)[:, 1] - 1] = 1
    comp[g.lm.reshape(-1, 3)[:, 2] - 1] = 2
    scale = abs(S).sum(1).max()
    for c in range(3):
        t = (comp == c).astype(float)
        assert np.abs(S @ t).max() < 1e-12 * scale


def test_operator_hermitian_and_mass_identity():
    m, mod, pml = mesh_case("tiny")
    g = oracle_graph("tiny")
    f, ky = 0.05, 1.2e-5
    w = TWOPI * f
    S = csr(g, oracle.asmble_z(m, mod, g, f, ky / w, pml))
    assert abs(S - S.conj().T).max() < 1e-15 * abs(S).max()
    # x-translation: K0 t = 0, so t^T S t = -w^2 int rho + ky^2 int mu
    comp = np.zeros(g.ndof, int)
    for c in range(3):
        comp[g.lm.reshape(-1, 3)[:, c] - 1] = c
    t = (comp == 0).astype(float)
    area = 12e3 * 8e3
    rho, muv = mod.dens[0], mod.ecoeff[0, 1]
    ref = -w * w * rho * area + ky * ky * muv * area
    assert abs(t @ (S @ t) - ref) < 1e-12 * abs(ref)


def test_quadratic_in_ky_affine_in_omega2():
    m, mod, pml = mesh_case("tiny_layered")
    g = oracle_graph("tiny_layered")
    def S(w, k):
        return csr(g, oracle.asmble_z(m, mod, g, w / TWOPI, k / w, pml))
    w, k = TWOPI * 0.07, 2e-5
    comb = S(w, 2 * k) - 3 * S(w, k) + 3 * S(w, 0.0) - S(w, -k)
    assert abs(comb).max() < 1e-12 * abs(S(w, k)).max()
    w1, w2, w3 = TWOPI * 0.03, TWOPI * 0.06, TWOPI * 0.09
    kk = 0.0
    d1 = (S(w2, kk) - S(w1, kk)) / (w2 ** 2 - w1 ** 2)
    d2 = (S(w3, kk) - S(w2, kk)) / (w3 ** 2 - w2 ** 2)
    assert abs(d1 - d2).max() < 1e-9 * abs(d1).max()


# -- virtual forces and gradient ---------------------------------------------------------

@pytest.mark.parametrize("case", ["tiny", "p2_pml", "pml"])
def test_inversion_structures_and_cfmat_vs_reference(case):
    z = gold(f"inv_{case}.npz")
    m, mod, pml = mesh_case(case)
    g = oracle_graph(case)
    inv = oracle.inv_setup(m, mod, g, pml)
    for k in ("maskg", "mconn", "mygrad", "jcsc", "icsc", "elem_wts"):
        np.testing.assert_array_equal(getattr(inv, k), z[k])
    wave = cf.seeded_complex(g.ndof, seed=int(z["wave_seed"])).astype(np.complex64)
    for cinv in ("PP", "SS"):
        fm = oracle.cfmat_dist(m, mod, g, inv, float(z["freq"]), float(z["py"]), wave,
                               cinvtype=cinv.encode(), pml=pml)
        ref = z[f"fmat_{cinv}"]
        assert np.abs(fm - ref).max() <= 1e-6 * np.abs(ref).max()


def test_adjflam_hand_example():
    inv = oracle.Inv(nnpinv=2, ncon=1, nnpgl=2, nz_fdist=3, mbufrhs=2, na35=2,
                     maskg=np.array([1, 2], np.int32), mconn=np.zeros((1, 2), np.int32),
                     mygrad=np.array([2, 1], np.int32), jcsc=np.array([1, 3, 4], np.int32),
                     icsc=np.array([1, 3, 2], np.int32), elem_wts=np.ones(1))
    f = np.array([1 + 2j, 3 - 1j, 0.5 + 0.5j], np.complex64)
    lam = np.array([2 + 1j, -1 + 4j, 1 - 1j], np.complex64)
    g = oracle.adjflam_dist(inv, f, lam)
    # column 1 -> gradient entry 2: -Re(conj(1+2i)(2+i) + conj(3-i)(1-i)) = -(4 + 4) = -8
    # column 2 -> gradient entry 1: -Re(conj(.5+.5i)(-1+4i)) = -1.5
    np.testing.assert_allclose(g, [-1.5, -8.0], rtol=1e-6)


def _objective_setup():
    m, mod, pml = mesh_case("tiny_layered")
    g = oracle_graph("tiny_layered")
    rng = np.random.default_rng(61)
    f = 1e10 * (rng.standard_normal(g.ndof) + 1j * rng.standard_normal(g.ndof))
    # receivers: the three components at the top-surface nodes
    top = g.lm[:, -(m.nord + 1):, :][(np.arange(m.nelem) // m.nx) == m.nz - 1].ravel() - 1
    rec = np.unique(top)
    # data: the response of a model with doubled anomalies, so that the
    # residual is of the order of the response
    m2 = cf.layered_model(m, seed=7, n_anom=4, amp=0.1, anomaly_scale=2.0)
    S2 = csr(g, oracle.asmble_z(m, m2, g, 0.06, 0.5e-4, pml)).tocsc()
    d = spsolve(S2, f)[rec]
    return m, mod, pml, g, f, rec, d


def _chi(m, mod, pml, g, freq, py, f, rec, d):
    S = csr(g, oracle.asmble_z(m, mod, g, freq, py, pml)).tocsc()
    u = spsolve(S, f)
    r = u[rec] - d
    return 0.5 * np.vdot(r, r).real, S, u, r


def test_exact_gradient_matches_finite_differences():
    """dchi/dm = Re(lam^H dS/dm u) with S^H lam = R^T (d - R u)."""
    m, mod, pml, g, f, rec, d = _objective_setup()
    freq, py = 0.06, 0.5e-4
    chi, S, u, r = _chi(m, mod, pml, g, freq, py, f, rec, d)
    rhs = np.zeros(g.ndof, complex)
    rhs[rec] = -r
    lam = spsolve(S.conj().T.tocsc(), rhs)
    _, gv = oracle.grad_exact(m, mod, g, freq, py, u, lam, pml)
    rng = np.random.default_rng(62)
    nodes = rng.choice(m.nnpg, 6, replace=False)
    for a in nodes:
        for p, base in ((0, mod.vp), (1, mod.vs), (2, mod.dens)):
            h = 1e-4 * base[a]
            vals = []
            for s in (+1, -1):
                vp, vs, rho = mod.vp.copy(), mod.vs.copy(), mod.dens.copy()
                (vp, vs, rho)[p][a] += s * h
                mm = cf.model_from_velocity(vp, vs, rho)
                vals.append(_chi(m, mm, pml, g, freq, py, f, rec, d)[0])
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(gv[a, p] - fd) <= 1e-5 * max(abs(fd), 1e-3 * abs(gv).max()), (a, p, gv[a, p], fd)


# the compiled reference (oracle/_ref, built from the reference sources by
# oracle/Makefile) pins routines that have no golden fixture
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built")


def _jac_batch(seed, nfreq, nsrc, nnpgl, na35):
    rng = np.random.default_rng(seed)
    mygrad = np.sort(rng.choice(np.arange(1, na35 + 1), size=nnpgl, replace=False)).astype(np.int32)
    nobs = rng.integers(20, 121, size=(nfreq, nsrc)).astype(np.int32)   # uneven receiver counts
    tot = int(nobs.sum()) * nnpgl
    jac = (rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(np.complex64)
    p = rng.standard_normal(na35)
    src = (rng.standard_normal((nfreq, nsrc)) + 1j * rng.standard_normal((nfreq, nsrc))).astype(np.complex64)
    return mygrad, nobs, jac, p, src


def _gnmfree_dense(nobs, alpha, nnpgl, mygrad, na35, jac, p):
    z = np.zeros(na35)
    off = 0
    for j, no in enumerate(np.ravel(nobs)):
        J = jac[off:off + no * nnpgl].reshape(nnpgl, no).T.astype(np.complex128)
        off += no * nnpgl
        np.add.at(z, mygrad - 1, alpha[j] * np.real(J.conj().T @ (J @ p[mygrad - 1])))
    return z


@needs_ref
@pytest.mark.parametrize("case,freq,py", [("pml", 0.05, 0.0), ("p2_pml", 0.2, 1e-4), ("pml", 0.08, 0.6e-4)])
def test_asmb_peff_vs_reference(case, freq, py):
    """orc_asmb_peff against the compiled ASMB_PEFF (asmble.f90:318-377) on
    the same CDIST and u_E: bit for bit (same COMPLEX*8 products and the
    same row order)."""
    m, mod, pml = mesh_case(case)
    g = oracle_graph(case)
    cd = oracle.asmble_cdist(m, mod, g, freq, py, pml)
    ue = cf.seeded_complex(g.ndof, seed=41).astype(np.complex64)
    mine = oracle.asmb_peff(g, cd, ue)
    ref = oracle.ref_asmb_peff(g, cd, ue)
    assert np.count_nonzero(ref) > 0
    np.testing.assert_array_equal(mine, ref)


@pytest.mark.parametrize("case,freq,py", [("pml", 0.05, 0.0), ("p2_pml", 0.2, 1e-4)])
def test_asmb_peff_closed_form(case, freq, py):
    """P_eff = -S_BE u_E on 'B' rows and +S_EB u_E on 'E' rows, 0 elsewhere,
    with S the FP64 operator (orc_asmble_z, pinned above) restricted by the
    DOF classes: the COMPLEX*8 result within FP32 rounding of it."""
    m, mod, pml = mesh_case(case)
    g = oracle_graph(case)
    cd = oracle.asmble_cdist(m, mod, g, freq, py, pml)
    S = csr(g, oracle.asmble_z(m, mod, g, freq, py, pml)).tocoo()
    cnp = np.frombuffer(g.cnp, dtype="S1")
    ue = cf.seeded_complex(g.ndof, seed=42)
    rb, cb = cnp[S.row] == b"B", cnp[S.col] == b"B"
    re_, ce = cnp[S.row] == b"E", cnp[S.col] == b"E"
    sign = np.where(rb & ce, -1.0, 0.0) + np.where(re_ & cb, 1.0, 0.0)
    Sr = sp.csr_matrix((S.data * sign, (S.row, S.col)), shape=S.shape)
    want = Sr @ ue
    got = oracle.asmb_peff(g, cd, ue.astype(np.complex64))
    assert np.count_nonzero(sign) > 0
    assert relerr(got, want) < 2e-6
    other = (cnp != b"B") & (cnp != b"E")
    assert np.all(got[other] == 0)


def test_gn_hessian_vs_dense():
    """orc_gn_hessian against the dense product sum_j alpha_j Re(J_j^H J_j)."""
    rng = np.random.default_rng(71)
    nnpgl, nobs = 9, np.array([4, 7, 1, 5], np.int32)
    jac = (rng.standard_normal(int(nobs.sum()) * nnpgl)
           + 1j * rng.standard_normal(int(nobs.sum()) * nnpgl)).astype(np.complex64)
    alpha = rng.uniform(0.2, 3.0, nobs.size)
    h = oracle.gn_hessian(nobs, alpha, nnpgl, jac)
    want = np.zeros((nnpgl, nnpgl))
    off = 0
    for j, no in enumerate(nobs):
        J = jac[off:off + no * nnpgl].astype(np.complex128).reshape(nnpgl, no).T   # (obs, column)
        want += alpha[j] * (J.conj().T @ J).real
        off += no * nnpgl
    assert np.abs(h - want).max() <= 1e-13 * np.abs(want).max()
    np.testing.assert_array_equal(h, h.T)


@needs_ref
def test_jacobian_writer_vs_reference_statements(tmp_path):
    """f25_write_jac's bytes equal those of PCHADJ's WRITE statements
    (pchadj.f90:371-395) compiled with the reference's compiler, names
    aside; LOBS mixed true/false."""
    import ctypes as C
    from fem25_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(72)
    ndim, nrec, nnpgl = 3, 5, 4
    nrhsf = ndim * nrec
    lobs = (rng.random((nrec, ndim)) < 0.7).astype(np.int32)
    jac = (rng.standard_normal((nnpgl, nrhsf)) + 1j * rng.standard_normal((nnpgl, nrhsf))).astype(np.complex64)
    mygrad = rng.permutation(np.arange(1, 12))[:nnpgl].astype(np.int32)
    a, b = str(tmp_path / "mine.dat"), str(tmp_path / "ref.dat")
    rc = L.f25_write_jac(a.encode(), nrhsf, nnpgl, ndim, nrec, lobs.ctypes.data_as(_lib.ip),
                         jac.ctypes.data_as(C.c_void_p), 0, mygrad.ctypes.data_as(_lib.ip))
    assert rc == 0
    oracle.ref_write_jac(b, nrhsf, nnpgl, ndim, nrec, lobs, jac, mygrad)
    assert open(a, "rb").read() == open(b, "rb").read()


@needs_ref
def test_gnmfree_vs_reference(tmp_path):
    """GNMFREE restated: bit-exact against the compiled reference, which
    reads Jacobian files written in its own record layout; and against the
    dense closed form Re(J^H J p).  The reference clears its running sum
    with DSCAL(NA35, 0, Z2) on freshly allocated memory (gnmfree.f90:60),
    which keeps any NaN the heap held there; such entries are skipped."""
    nfreq, nsrc, nnpgl, na35 = 2, 3, 40, 55
    mygrad, nobs, jac, p, src = _jac_batch(81, nfreq, nsrc, nnpgl, na35)
    alpha = np.ones(nfreq * nsrc)
    z = oracle.gnmfree(nobs.ravel(), alpha, nnpgl, mygrad, na35, jac, p)
    zr = oracle.ref_gnmfree([0.05, 0.0625], nsrc, src, 0, nnpgl, mygrad, na35, nobs, jac, p,
                            str(tmp_path))
    ok = np.isfinite(zr)
    assert ok.sum() >= na35 // 2       # how many the heap spoils varies run to run
    np.testing.assert_array_equal(z[ok], zr[ok])
    zd = _gnmfree_dense(nobs, alpha, nnpgl, mygrad, na35, jac, p)
    assert np.abs(z - zd).max() <= 1e-12 * np.abs(zd).max()
    # migration weights |source|^2 (computed by the reference itself in REAL*4)
    a2 = (np.abs(src.ravel()) ** 2).astype(np.float32).astype(np.float64)
    z2 = oracle.gnmfree(nobs.ravel(), a2, nnpgl, mygrad, na35, jac, p)
    zr2 = oracle.ref_gnmfree([0.05, 0.0625], nsrc, src, 1, nnpgl, mygrad, na35, nobs, jac, p,
                             str(tmp_path))
    ok = np.isfinite(zr2)
    assert ok.sum() >= na35 // 2
    assert np.abs(z2 - zr2)[ok].max() <= 1e-6 * np.abs(zr2[ok]).max()


def test_pchadj_vs_dense():
    """PCHADJ's J = G^T F (COMPLEX*8 sums) against the dense product in
    complex128, and diag(J^H J) against numpy, on the CSC structures of the
    PML mesh."""
    m, mod, pml = mesh_case("pml")
    g = oracle_graph("pml")
    inv = oracle.inv_setup(m, mod, g, pml)
    rng = np.random.default_rng(82)
    nrhs, lrhs = 36, g.ndof + 5
    G = (rng.standard_normal((nrhs, lrhs)) + 1j * rng.standard_normal((nrhs, lrhs))).astype(np.complex64)
    F = (rng.standard_normal(inv.nz_fdist) + 1j * rng.standard_normal(inv.nz_fdist)).astype(np.complex64)
    jac, hd = oracle.pchadj_jac(inv, nrhs, lrhs, G, F)
    Fd = np.zeros((g.ndof, inv.nnpgl), np.complex128)
    for c in range(inv.nnpgl):
        for i in range(inv.jcsc[c] - 1, inv.jcsc[c + 1] - 1):
            Fd[inv.icsc[i] - 1, c] += F[i]
    Jd = G[:, :g.ndof].astype(np.complex128) @ Fd          # (nrhs, nnpgl)
    assert np.linalg.norm(jac.T - Jd) <= 1e-6 * np.linalg.norm(Jd)
    hdd = np.zeros(inv.na35)
    hdd[inv.mygrad - 1] = np.sum(np.abs(Jd) ** 2, axis=0)
    assert np.linalg.norm(hd - hdd) <= 1e-6 * np.linalg.norm(hdd)


@pytest.mark.parametrize("case", ["tiny_layered", "pml"])
def test_dispersive_cdist_vs_reference(case):
    """KMDAMP/CMDAMP + VCMPLX through ASMBLE_DRIVER with LDISP: the oracle's
    COMPLEX*8 CDIST equals the compiled reference's (fixture) bit for bit."""
    m, mod, pml = mesh_case(case)
    g = oracle_graph(case)
    f = np.load(os.path.join(GOLD, f"asmq_{case}.npz"))
    md = cf.with_attenuation(mod, f["qp"], f["qs"], float(f["fref"]))
    for k in range(2):
        cd = oracle.asmble_cdist(m, md, g, float(f[f"freq{k}"]), float(f[f"py{k}"]), pml)
        np.testing.assert_array_equal(cd, f[f"cdist{k}"])


@pytest.mark.parametrize("case", ["tiny", "pml"])
def test_dispersive_without_attenuation_is_density_normalised(case):
    """With Q = 9999 VCMPLX returns the real velocity, and on a homogeneous
    model (velocities interpolated in V equal those interpolated in lambda,
    mu) the dispersive operator is the elastic one divided by rho."""
    m, _, pml = mesh_case(case)
    mod = cf.homogeneous_model(m)
    g = oracle_graph(case)
    md = cf.with_attenuation(mod, 9999.0, 9999.0, 1.0)
    for freq, py in ((0.05, 0.7e-4), (0.13, 0.0)):
        s_el = oracle.asmble_z(m, mod, g, freq, py, pml)
        s_q = oracle.asmble_z(m, md, g, freq, py, pml)
        assert relerr(s_q * mod.dens[0], s_el) < 1e-13


# ---- second-order (QPMFREE) element routines ---------------------------------
_ELEM = dict(xlocs=[0.0, 1000.0, 1130.0, -60.0], zlocs=[0.0, 40.0, 910.0, 1000.0],
             dens=[2600.0, 2750.0, 2900.0, 2680.0],
             ecoeff=np.array([[3.1e10, 2.9e10], [3.4e10, 3.0e10], [3.9e10, 3.3e10], [3.3e10, 3.1e10]]))


@needs_ref
@pytest.mark.parametrize("nord", [2, 4])
def test_djacab2_dabm_vs_reference(nord):
    """DJACAB2, DABMKISOB25 and DABM2KISOB25 restated: bit-exact against the
    compiled reference on a distorted element, collocated on each corner,
    for P, S and joint (PS) derivatives.  Both sides use the compiled DSHL
    tables (the oracle's own agree with them to 1e-13, pinned above)."""
    omega, py = 2 * np.pi * 0.07, 1.1e-4
    xg, wg, shl = oracle.ref_dshl(nord)
    for inpg in range(1, 5):
        o = oracle.djacab2(nord, inpg, tables=(xg, shl), **_ELEM)
        r = oracle.ref_djacab2(nord, inpg, **_ELEM)
        for a, b in zip(o, r):
            np.testing.assert_array_equal(a, b)
        det, dab, dab2, shg = r
        assert np.count_nonzero(dab[0]) and np.count_nonzero(dab2[1])
        for lisp, liss in ((1, 0), (0, 1), (1, 1)):
            mo = oracle.dabmkisob25(nord, lisp, liss, omega, py, det, dab, shg, wq=wg)
            mr = oracle.dabmkisob25(nord, lisp, liss, omega, py, det, dab, shg, reference=True)
            for a, b in zip(mo, mr):
                np.testing.assert_array_equal(a, b)
            assert np.count_nonzero(mr[0 if lisp else 1]) and np.count_nonzero(mr[2 if lisp else 3])
            m2o = oracle.dabmkisob25(nord, lisp, liss, omega, py, det, dab, shg, dab2=dab2, wq=wg)
            m2r = oracle.dabmkisob25(nord, lisp, liss, omega, py, det, dab, shg, reference=True, dab2=dab2)
            for a, b in zip(m2o, m2r):
                np.testing.assert_array_equal(a, b)


def _perturbation_operator(m, mod, g, pml, fl, fm, freq, py):
    """The assembled operator of the model (lambda, mu, rho) = (fl, fm, 0) at
    the anchors: S is linear in the three fields, so this is the sum over
    elements of the element matrices whose DAB fields are fl, fm."""
    import dataclasses
    pert = dataclasses.replace(mod, dens=np.zeros_like(mod.dens),
                               ecoeff=np.asfortranarray(np.stack([fl, fm], axis=1)))
    return csr(g, oracle.asmble_z(m, pert, g, freq, py, pml))


def _dab_fields(mod, inv, v, kind):
    """Anchor fields (lambda, mu) of sum_c V_c dS/dm_c in DABMKISOB25's
    density-normalised form: PP lambda = 2 Vp V; SS lambda = -4 Vs V, mu = 2 Vs V."""
    lam, mu, rho = mod.ecoeff[:, 0], mod.ecoeff[:, 1], mod.dens
    vp, vs = np.sqrt((lam + 2 * mu) / rho), np.sqrt(mu / rho)
    w = np.zeros(mod.dens.size)
    sel = inv.maskg > 0
    w[sel] = np.asarray(v, np.float64)[inv.maskg[sel] - 1]
    if kind == b"PP":
        return 2 * vp * w, np.zeros_like(w)
    return -4 * vs * w, 2 * vs * w


@pytest.mark.parametrize("kind", [b"PP", b"SS"])
def test_qp_nu_alpha_against_assembled_operator(kind):
    """CADJ_NU_ALPHA's right-hand sides restated, against an independent
    route: summed over every column and connected element, RHSALP is
    S[lambda = f_L, mu = f_M, rho = 0] u and RHSBNU is that matrix's conjugate
    transpose on lambda, with S assembled by orc_asmble_z (pinned against the
    compiled ASMBLE_DRIVER) on a uniform interior mesh (ELEM_WTS = 1).  P3 on
    anchors that touch one element is -Re((S_a u)^H lambda) with S_a the
    operator of that anchor's field alone; MET225_MOD's P1, P2 likewise, and
    P3 = V P1 wherever nu = lambda.  Tolerance: the reference's COMPLEX*8
    accumulation."""
    name = "tiny_layered"
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    inv = oracle.inv_setup(m, mod, g, pml)
    assert np.all(inv.elem_wts == 1.0)
    rng = np.random.default_rng(95)
    freq, py = 0.07, 1.1e-4
    v = rng.standard_normal(inv.na35).astype(np.float32)
    u = (rng.standard_normal(g.ndof) + 1j * rng.standard_normal(g.ndof)).astype(np.complex64)
    lam = (rng.standard_normal(g.ndof) + 1j * rng.standard_normal(g.ndof)).astype(np.complex64)
    bnu, alp, p3 = oracle.qp_nu_alpha(m, mod, g, inv, freq, py, v, lam, u, cinvtype=kind, pml=pml)
    fl, fm = _dab_fields(mod, inv, v, kind)
    S = _perturbation_operator(m, mod, g, pml, fl, fm, freq, py)
    assert relerr(alp, S @ u.astype(np.complex128)) < 1e-5
    assert relerr(bnu, S.conj().T @ lam.astype(np.complex128)) < 1e-5
    # P3 / P1 / P2 on anchors with a single element
    cnt = np.bincount(m.ieng.ravel() - 1, minlength=m.nnpg)
    corners = np.nonzero((cnt == 1) & (inv.maskg > 0))[0]
    assert corners.size >= 2
    nu = (rng.standard_normal(g.ndof) + 1j * rng.standard_normal(g.ndof)).astype(np.complex64)
    alpha = (rng.standard_normal(g.ndof) + 1j * rng.standard_normal(g.ndof)).astype(np.complex64)
    p1, p2 = oracle.met225(m, mod, g, inv, freq, py, lam, nu, alpha, u, cinvtype=kind, pml=pml)
    for a in corners:
        one = np.zeros(inv.na35, np.float32)
        one[inv.maskg[a] - 1] = 1.0
        f1l, f1m = _dab_fields(mod, inv, one, kind)
        Sa = _perturbation_operator(m, mod, g, pml, f1l, f1m, freq, py)
        col = inv.maskg[a] - 1
        for got, x, y, wgt in ((p3[col], u, lam, v[col]), (p1[col], u, nu, 1.0), (p2[col], alpha, lam, 1.0)):
            t = Sa @ x.astype(np.complex128)
            terms = np.conj(t) * y
            want = -wgt * np.real(terms.sum())
            assert abs(got - want) <= 1e-5 * abs(wgt) * np.abs(terms).sum(), (a, got, want)
    _, _, p3n = oracle.qp_nu_alpha(m, mod, g, inv, freq, py, v, nu, u, cinvtype=kind, pml=pml)
    p1n, _ = oracle.met225(m, mod, g, inv, freq, py, lam, nu, alpha, u, cinvtype=kind, pml=pml)
    assert np.abs(p3n - v * p1n).max() <= 1e-5 * np.abs(p3n).max()


def _ps_split(inv_ps, v):
    """V of a PS inversion (P and S of each anchor interleaved) as the V of
    the PP and of the SS inversion on the same anchors."""
    v = np.asarray(v, np.float32)
    return v[0::2].copy(), v[1::2].copy()


@needs_ref
@pytest.mark.parametrize("case", ["pml", "p2_pml", "tiny"])
def test_ps_inversion_structures_vs_reference(case):
    """Joint P-S inversion (CINVTYPE 'PS', NVINV = 2): the restated GENGMASK,
    GRADPTRS and ELEM_WTS against the compiled routines, and equal to the
    single-variable structure copied forward; where every anchor of the
    interior is inverted (case tiny) the compiled GRADPTRS stops with
    'ICSC_FDIST too small' (its bound NNPG*NCON*NEE) and so does the oracle."""
    m, mod, pml = mesh_case(case)
    g = oracle_graph(case)
    one = oracle.inv_setup(m, mod, g, pml)
    if case == "tiny":
        with pytest.raises(RuntimeError):
            oracle.inv_setup(m, mod, g, pml, nvinv=2)
        with pytest.raises(RuntimeError):
            oracle.ref_inv_setup(m, mod, g, cinvtype=b"PS")
        return
    a = oracle.inv_setup(m, mod, g, pml, nvinv=2)
    b = oracle.ref_inv_setup(m, mod, g, cinvtype=b"PS")
    c = ps_inv(one)
    for k in ("maskg", "mconn", "mygrad", "jcsc", "icsc", "elem_wts"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
        np.testing.assert_array_equal(getattr(c, k), getattr(b, k), err_msg=k)
    assert a.na35 == 2 * a.nnpinv and a.nnpgl == 2 * a.nnpinv


def test_qp_nu_alpha_joint_ps():
    """CADJ_NU_ALPHA and MET225_MOD for the joint P-S inversion: the right-hand
    sides are those of the PP inversion weighted by the P entries of V plus
    those of the SS inversion weighted by the S entries (each entry adds its
    P term, then its S term, so the sums differ by COMPLEX*8 rounding), and
    against the assembled operator with lambda = 2 Vp V_P - 4 Vs V_S, mu = 2 Vs
    V_S; P3, P1, P2 of each variable equal the single-variable values bit for
    bit (the same element matrices and the same order)."""
    name = "tiny_layered"
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    ione = oracle.inv_setup(m, mod, g, pml)
    ips = ps_inv(ione)
    rng = np.random.default_rng(96)
    freq, py = 0.07, 1.1e-4
    v = rng.standard_normal(ips.na35).astype(np.float32)
    vp_, vs_ = _ps_split(ips, v)
    u, lam, nu, alpha = ((rng.standard_normal(g.ndof) + 1j * rng.standard_normal(g.ndof)).astype(np.complex64)
                         for _ in range(4))
    bnu, alp, p3 = oracle.qp_nu_alpha(m, mod, g, ips, freq, py, v, lam, u, cinvtype=b"PS", pml=pml)
    bp, ap, p3p = oracle.qp_nu_alpha(m, mod, g, ione, freq, py, vp_, lam, u, cinvtype=b"PP", pml=pml)
    bs, as_, p3s = oracle.qp_nu_alpha(m, mod, g, ione, freq, py, vs_, lam, u, cinvtype=b"SS", pml=pml)
    assert relerr(alp, ap.astype(np.complex128) + as_) < 1e-6
    assert relerr(bnu, bp.astype(np.complex128) + bs) < 1e-6
    np.testing.assert_array_equal(p3[0::2], p3p)
    np.testing.assert_array_equal(p3[1::2], p3s)
    flp, fmp = _dab_fields(mod, ione, vp_, b"PP")
    fls, fms = _dab_fields(mod, ione, vs_, b"SS")
    S = _perturbation_operator(m, mod, g, pml, flp + fls, fmp + fms, freq, py)
    assert relerr(alp, S @ u.astype(np.complex128)) < 1e-5
    assert relerr(bnu, S.conj().T @ lam.astype(np.complex128)) < 1e-5
    p1, p2 = oracle.met225(m, mod, g, ips, freq, py, lam, nu, alpha, u, cinvtype=b"PS", pml=pml)
    p1p, p2p = oracle.met225(m, mod, g, ione, freq, py, lam, nu, alpha, u, cinvtype=b"PP", pml=pml)
    p1s, p2s = oracle.met225(m, mod, g, ione, freq, py, lam, nu, alpha, u, cinvtype=b"SS", pml=pml)
    np.testing.assert_array_equal(p1[0::2], p1p)
    np.testing.assert_array_equal(p1[1::2], p1s)
    np.testing.assert_array_equal(p2[0::2], p2p)
    np.testing.assert_array_equal(p2[1::2], p2s)
    # VFORCE25 has no joint derivatives (vforce25.f:81-85): CFMAT_DIST refuses PS
    with pytest.raises(RuntimeError):
        oracle.cfmat_dist(m, mod, g, ips, freq, py, u, cinvtype=b"PS", pml=pml)


# ---- adjoint-source pipeline --------------------------------------------------
def _receivers(seed, nrec, ndof, zero_frac=0.1):
    rng = np.random.default_rng(seed)
    mrdof = rng.integers(1, ndof + 1, size=(nrec, 3)).astype(np.int32)
    mrdof[rng.random((nrec, 3)) < zero_frac] = 0
    recv = (rng.standard_normal((nrec, 3)) + 1j * rng.standard_normal((nrec, 3))).astype(np.complex64)
    yrec = rng.uniform(-3e4, 3e4, nrec)
    return mrdof, recv, yrec


def _cvals(rng, shape, scale=1.0):
    return (scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))).astype(np.complex64)


@needs_ref
def test_exresp25d_vs_reference():
    """EXRESP25D restated: bit-exact against the compiled routine (several
    sources with their own slowness, receivers off the y = 0 plane, unused
    components)."""
    ndof, nsrc, nrec = 500, 4, 37
    mrdof, recv, yrec = _receivers(101, nrec, ndof)
    rng = np.random.default_rng(102)
    sol = _cvals(rng, nsrc * ndof)
    py = rng.uniform(0, 1.4e-4, nsrc)
    e = oracle.exresp25d(ndof, 0.083, py, yrec, sol, mrdof, recv)
    r = oracle.exresp25d(ndof, 0.083, py, yrec, sol, mrdof, recv, reference=True)
    np.testing.assert_array_equal(e, r)
    assert np.count_nonzero(e) == nsrc * np.count_nonzero(mrdof)


@needs_ref
def test_residual_routines_vs_reference():
    """CRESID, CPHM2CM, CROTATE and BPRESID4 restated: bit-exact against the
    compiled routines for every residual type, every norm (the Huber and
    combination branches with the reference's EPS = 0 included, NaN where the
    reference gives NaN), the phase gate of the L2 norm, and receivers
    without data."""
    rng = np.random.default_rng(103)
    n = 300
    obs = _cvals(rng, n)
    est = obs + _cvals(rng, n, 0.3)
    obs[::17] = 0
    for irestp in (1, 2, 3):
        np.testing.assert_array_equal(oracle.cresid(irestp, obs, est),
                                      oracle.cresid(irestp, obs, est, reference=True))
    mag = rng.uniform(0, 3, n).astype(np.float32)
    ph = rng.uniform(-7, 7, n).astype(np.float32)
    np.testing.assert_array_equal(oracle.cphm2cm(mag, ph), oracle.cphm2cm(mag, ph, reference=True))
    for th in (0.0, 33.7, -121.25):
        a, b = oracle.crotate(th, obs, est), oracle.crotate(th, obs, est, reference=True)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
    nrec = 100
    o3, e3 = obs.reshape(nrec, 3), est.reshape(nrec, 3)
    w = rng.uniform(0.2, 2.0, (nrec, 3)).astype(np.float32)
    nan_seen = False
    for norm in (1, 2, 3, 4):
        for irestp in (1, 2, 3):
            for dtmax in (0.0, 1.5):
                a = oracle.bpresid4(norm, irestp, 0.05, 0.071, dtmax, w, o3, e3)
                b = oracle.bpresid4(norm, irestp, 0.05, 0.071, dtmax, w, o3, e3, reference=True)
                np.testing.assert_array_equal(a, b)
                nan_seen |= bool(np.isnan(a).any())
    assert nan_seen    # combination norm with EPS = 0
    gated = oracle.bpresid4(2, 3, 0.05, 0.071, 1.5, w, o3, e3)
    assert 0 < np.count_nonzero(gated) < np.count_nonzero(oracle.bpresid4(2, 3, 0.05, 0.071, 0.0, w, o3, e3))


def test_hdiffer_closed_form():
    """HDIFFER against its closed forms: sqrt(-i w) = sqrt(w/2)(1 - i),
    sqrt(i w) = sqrt(w/2)(1 + i), i w, i, 1 (one REAL*4 rounding)."""
    f = 0.137
    w = np.float32(2 * np.pi * f)
    s = np.sqrt(w / 2)
    want = {0: 1, 1: s * (1 - 1j), 2: -s * (1 - 1j), 3: s * (1 + 1j), 4: -s * (1 + 1j), 5: 1j * w, 6: 1j, 7: 1}
    for k, v in want.items():
        assert abs(oracle.hdiffer(k, f) - v) <= 2e-7 * abs(v), k


def test_resid_drhs_scatter():
    """RESID_DRHS by brute force: repeated DOFs sum in receiver order,
    out-of-range DOFs are skipped with IWARN = 1, the result is conjugated."""
    rng = np.random.default_rng(104)
    ndof, nrec = 40, 25
    mrdof = rng.integers(1, ndof + 1, size=(nrec, 3)).astype(np.int32)
    mrdof[3, 1] = mrdof[7, 2]                   # a repeated DOF
    resid = _cvals(rng, (nrec, 3))
    rhs, iw = oracle.resid_drhs(ndof, 5, 0.09, mrdof, resid)
    h = complex(oracle.hdiffer(5, 0.09))
    want = np.zeros(ndof, np.complex128)
    for r in range(nrec):
        for i in range(3):
            want[mrdof[r, i] - 1] += h * complex(resid[r, i])
    assert iw == 0
    assert np.abs(rhs - np.conj(want)).max() <= 1e-6 * np.abs(want).max()
    mrdof[5, 0] = 0
    mrdof[6, 2] = ndof + 1
    _, iw = oracle.resid_drhs(ndof, 0, 0.09, mrdof, resid)
    assert iw == 1


def test_cadjlam_rhs_closed_form():
    """CADJLAM's right-hand side: with the L2 norm, full residuals and equal
    weights on the two horizontal components, the residual rotated back to
    the model frame is W^2 (R(-az) OBS - EST) after the receiver phase
    shift, placed at the receiver DOFs and conjugated; magnitude/phase input
    (LUNWRAP) gives the same as the equivalent complex input."""
    rng = np.random.default_rng(105)
    ndof, nfreq, nrec, nsrc = 120, 3, 20, 2
    mrdof = np.arange(1, 3 * nrec + 1, dtype=np.int32).reshape(nrec, 3) * 2
    yrec = rng.uniform(-2e4, 2e4, nrec)
    est = _cvals(rng, (nsrc, nrec, nfreq, 3))
    obs = _cvals(rng, (nsrc, nrec, nfreq, 3))
    wv = rng.uniform(0.5, 2.0, (nsrc, nrec, nfreq))
    w = np.repeat(wv[..., None], 3, axis=-1).astype(np.float32)
    ifreq, isrc, freq, py, az = 2, 2, 0.061, 0.9e-4, 28.0
    rhs, iw = oracle.cadjlam_rhs(ndof, ifreq, isrc, freq, py, az, 2, 3, 0.0, 0.0, 0, False, yrec, mrdof, est,
                                 obs, w)
    assert iw == 0
    th = np.deg2rad(az)
    want = np.zeros(ndof, np.complex128)
    for r in range(nrec):
        e = est[isrc - 1, r, ifreq - 1].astype(np.complex128)
        o = obs[isrc - 1, r, ifreq - 1].astype(np.complex128)
        ob = np.array([np.cos(th) * o[0] + np.sin(th) * o[1], -np.sin(th) * o[0] + np.cos(th) * o[1], o[2]])
        sh = np.exp(-1j * 2 * np.pi * freq * py * yrec[r])
        res = float(w[isrc - 1, r, ifreq - 1, 0]) ** 2 * (ob - e) * sh
        want[mrdof[r] - 1] += res
    assert np.abs(rhs - np.conj(want)).max() <= 2e-6 * np.abs(want).max()
    # magnitude/phase form of the same data
    mp = lambda a: (np.abs(a) + 1j * np.angle(a)).astype(np.complex64)   # noqa: E731
    rhs2, _ = oracle.cadjlam_rhs(ndof, ifreq, isrc, freq, py, az, 2, 3, 0.0, 0.0, 0, True, yrec, mrdof, mp(est),
                                 mp(obs), w)
    assert np.abs(rhs2 - rhs).max() <= 2e-6 * np.abs(rhs).max()


# -- CADJLAM's data covariance (LCOVD) ---------------------------------------------

def _covd_obs(rng, nrec, pz, dead=()):
    raw = _cvals(rng, (nrec, 3))
    raw[rng.random((nrec, 3)) < pz] = 0
    for c in dead:
        raw[:, c] = 0
    return raw


@needs_ref
def test_rec_space1d_vs_reference():
    """REC_SPACE1D: ascending and descending lines, a duplicate, a reversal
    (at receivers 2..nrec-1, the pairs the reference checks, and at the last
    pair, which it does not), one receiver."""
    rng = np.random.default_rng(201)
    x = np.cumsum(rng.uniform(0.5, 3.0, 15))
    cases = [x, x[::-1].copy(), np.r_[x[:5], x[4], x[5:]], np.r_[x[:6], x[3], x[6:]],
             np.r_[x[:-1], x[-3]], x[:1], x[:2]]
    for xr in cases:
        a, ea = oracle.rec_space1d(xr)
        b, eb = oracle.rec_space1d(xr, reference=True)
        assert ea == eb, xr
        np.testing.assert_array_equal(a, b)
        if ea == 0:
            d = np.abs(np.diff(xr))
            np.testing.assert_allclose(a, d / d.min(), rtol=0, atol=0)


@needs_ref
@pytest.mark.parametrize("nrec,pz,ldwght,irestp,dtmax,dead", [
    (12, 0.0, 0, 3, 0.0, ()), (12, 0.3, 1, 3, 0.0, ()), (40, 0.3, 1, 1, 0.0, ()),
    (40, 0.5, 1, 2, 0.5, ()), (40, 0.2, 0, 3, 2.0, ()), (7, 0.2, 1, 3, 0.0, (2,)),
    (7, 0.2, 1, 3, 0.0, (0, 1)), (5, 0.0, 1, 3, 0.0, (0, 1, 2)), (60, 0.9, 1, 3, 0.0, ())])
def test_covd_resid_vs_reference(nrec, pz, ldwght, irestp, dtmax, dead):
    """LAPLACE_1DV2, PERM_OBS and BPRESID_COVD4 against the compiled
    routines: the roughening matrix and the permutation bit for bit; the
    residual to 1e-6 of its largest value (the oracle's SGEMV adds column by
    column as the reference BLAS does; the linked BLAS orders and fuses its
    single-precision sums its own way).  Cases: every receiver observed,
    missing observations in every component (the first component taken,
    the vertical, included), a dead component, two dead components, a dead
    frequency, nine of ten receivers unobserved."""
    rng = np.random.default_rng(202 + nrec + 7 * ldwght + irestp)
    x = np.cumsum(rng.uniform(1.0, 3.0, nrec))
    dx, _ = oracle.rec_space1d(x)
    raw = _covd_obs(rng, nrec, pz, dead)
    obs = (raw * np.complex64(np.exp(0.3j))).astype(np.complex64)
    est = _cvals(rng, (nrec, 3))
    w = rng.uniform(0.5, 2.0, (nrec, 3)).astype(np.float32)
    a = oracle.covd_resid(nrec, irestp, ldwght, dtmax, 0.25, dx, raw, obs, est, w)
    b = oracle.covd_resid(nrec, irestp, ldwght, dtmax, 0.25, dx, raw, obs, est, w, reference=True)
    assert a[3] == b[3] == 0
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
    scale = np.abs(b[0]).max()
    assert np.abs(a[0] - b[0]).max() <= 1e-6 * max(scale, 1e-30)
    if not raw.any():
        assert not b[0].any() and not b[1].any()


def test_covd_closed_form_chain_laplacian():
    """With the vertical and second components observed at every receiver
    (so that the roughening matrix's numbering and PERM_OBS's positions
    agree) and unit weights, each component's residual is the 1-D finite
    element Laplacian of the data residual over its observed receivers:
    r_k = (d_k - d_k-1)/h_k-1 + (d_k - d_k+1)/h_k, h the distance to the
    next observed receiver divided by the smallest spacing (1 without
    ldwght), and zero at unobserved receivers; rows sum to zero (a constant
    residual along a component gives none) and <d, r> = sum (d_k+1 - d_k)^2
    / h_k."""
    rng = np.random.default_rng(203)
    nrec = 25
    x = np.cumsum(rng.uniform(1.0, 4.0, nrec))
    dx, _ = oracle.rec_space1d(x)
    raw = _cvals(rng, (nrec, 3))
    raw[[2, 3, 9, 17, 24], 0] = 0            # gaps in the last component taken
    est = _cvals(rng, (nrec, 3))
    w = np.ones((nrec, 3), np.float32)
    for ldwght in (0, 1):
        r, _, _, ierr = oracle.covd_resid(nrec, 3, ldwght, 0.0, 0.2, dx, raw, raw, est, w)
        assert ierr == 0
        d = (raw - est).astype(np.complex128)
        for c in range(3):
            act = np.flatnonzero(raw[:, c] != 0)
            h = np.diff(x[act]) / np.diff(x).min() if ldwght else np.ones(act.size - 1)
            want = np.zeros(nrec, np.complex128)
            dd = d[act, c]
            want[act[:-1]] += (dd[:-1] - dd[1:]) / h
            want[act[1:]] += (dd[1:] - dd[:-1]) / h
            assert np.abs(r[:, c] - want).max() <= 2e-6 * np.abs(want).max(), (ldwght, c)
            energy = np.vdot(dd, r[act, c].astype(np.complex128)).real
            assert abs(energy - np.sum(np.abs(np.diff(dd)) ** 2 / h)) <= 1e-5 * energy
    const = np.ones((nrec, 3), np.complex64) * np.complex64(1 + 2j)
    r, _, _, _ = oracle.covd_resid(nrec, 3, 1, 0.0, 0.2, dx, const, const, np.zeros_like(const), w)
    assert np.abs(r).max() <= 1e-5


@needs_ref
def test_cadjlam_rhs_covd_composition():
    """CADJLAM with LCOVD: with no rotation, no receiver shift and no
    HDIFFER factor, the right-hand side is the conjugate of the compiled
    BPRESID_COVD4 residual (on the compiled LAPLACE_1DV2 and PERM_OBS)
    placed at the receiver DOFs; a rotation by az with equal horizontal
    weights and no horizontal gaps commutes with the component-wise
    roughening, so it equals rotating the data instead."""
    rng = np.random.default_rng(204)
    ndof, nfreq, nrec, nsrc = 200, 2, 30, 2
    mrdof = np.arange(1, 3 * nrec + 1, dtype=np.int32).reshape(nrec, 3) * 2
    yrec = rng.uniform(-2e4, 2e4, nrec)
    x = np.cumsum(rng.uniform(1.0, 2.0, nrec))
    dx, _ = oracle.rec_space1d(x)
    est = _cvals(rng, (nsrc, nrec, nfreq, 3))
    obs = _cvals(rng, (nsrc, nrec, nfreq, 3))
    obs[1, ::4, 0, 2] = 0
    w = rng.uniform(0.5, 2.0, (nsrc, nrec, nfreq, 3)).astype(np.float32)
    ifreq, isrc = 1, 2
    rhs, iw, ierr = oracle.cadjlam_rhs_covd(ndof, ifreq, isrc, 0.07, 0.0, 0.0, 3, 0.0, 0, False, True, dx, yrec,
                                            mrdof, est, obs, w)
    assert iw == 0 and ierr == 0
    o, e, ww = obs[isrc - 1, :, ifreq - 1], est[isrc - 1, :, ifreq - 1], w[isrc - 1, :, ifreq - 1]
    r, _, _, _ = oracle.covd_resid(nrec, 3, 1, 0.0, 0.07, dx, o, o, e, ww, reference=True)
    want = np.zeros(ndof, np.complex64)
    want[mrdof.ravel() - 1] = np.conj(r.ravel())
    assert np.abs(rhs - want).max() <= 1e-6 * np.abs(want).max()
    # rotation: equal horizontal weights, every component observed everywhere
    # (a gap in an earlier component shifts the roughening matrix's
    # numbering against PERM_OBS's positions, as in the reference)
    wr = w.copy()
    wr[..., 1] = wr[..., 0]
    full = _cvals(rng, obs.shape)
    az = 33.0
    rot, _, _ = oracle.cadjlam_rhs_covd(ndof, ifreq, isrc, 0.07, 0.0, az, 3, 0.0, 0, False, True, dx, yrec, mrdof,
                                        est, full, wr)
    th = np.deg2rad(az)
    o = full[isrc - 1, :, ifreq - 1]
    o2 = o.astype(np.complex128)
    ob = np.stack([np.cos(th) * o2[:, 0] + np.sin(th) * o2[:, 1], -np.sin(th) * o2[:, 0] + np.cos(th) * o2[:, 1],
                   o2[:, 2]], axis=1).astype(np.complex64)
    rr, _, _, _ = oracle.covd_resid(nrec, 3, 1, 0.0, 0.07, dx, o, ob, e, wr[isrc - 1, :, ifreq - 1])
    # R(-az) C W (OBS - R(az) EST) = C W (R(-az) OBS - EST): no rotation left over
    want = np.zeros(ndof, np.complex64)
    want[mrdof.ravel() - 1] = np.conj(rr.ravel())
    assert np.abs(rot - want).max() <= 5e-6 * np.abs(want).max()


def test_covd_entry_point_argument_checks():
    """The data-covariance entry point refuses non-positive or non-finite
    receiver spacings (LAPLACE_1DV2 divides by their sums) before any device
    work; REC_SPACE1D's host form reports its three errors."""
    import ctypes as C
    from fem25_amd import _lib
    L = _lib.load()
    nrec, ndof = 4, 30
    ip, dp, fp = _lib.ip, _lib.dp, _lib.fp
    one = np.ones(1, np.int32)
    fq, py = np.full(1, 0.1), np.zeros(1)
    yrec = np.zeros(nrec)
    mrdof = np.arange(1, 3 * nrec + 1, dtype=np.int32)
    est = np.zeros(3 * nrec, np.complex64)
    w = np.ones(3 * nrec, np.float32)
    rhs = np.zeros(ndof, np.complex64)
    iw = C.c_int(0)
    p = lambda a, t: a.ctypes.data_as(t)   # noqa: E731
    for dx in (np.array([1.0, 0.0, 2.0]), np.array([1.0, -1.0, 2.0]), np.array([1.0, np.inf, 2.0])):
        rc = L.f25_cadjlam_rhs_covd(ndof, 1, nrec, 1, 1, p(one, ip), p(one, ip), p(fq, dp), p(py, dp), 0.0, 3,
                                    C.c_float(0.0), 0, 0, 1, p(dx, dp), p(yrec, dp), p(mrdof, ip),
                                    est.ctypes.data_as(C.c_void_p), est.ctypes.data_as(C.c_void_p), p(w, fp),
                                    rhs.ctypes.data_as(C.c_void_p), C.byref(iw))
        assert rc == _lib.F25_ERR_ARG
        assert b"spacings" in L.f25_last_error()
    out = np.zeros(3)
    x = np.array([0.0, 1.0, 3.0, 6.0])
    assert L.f25_rec_space1d(4, p(x, dp), p(out, dp)) == 0
    np.testing.assert_array_equal(out, [1.0, 2.0, 3.0])
    for bad in (np.array([0.0, 1.0, 1.0, 2.0]), np.array([0.0, 1.0, 0.5, 2.0])):
        assert L.f25_rec_space1d(4, p(bad, dp), p(out, dp)) == _lib.F25_ERR_ARG
        assert not out.any()
    assert L.f25_rec_space1d(1, p(x, dp), p(out, dp)) == _lib.F25_ERR_ARG


# -- Bielak plane-wave sources ------------------------------------------------------

@needs_ref
@pytest.mark.parametrize("case", ["pml", "p2_pml", "p6"])
def test_fill_uwave_vs_reference_and_closed_form(case):
    """FILL_UWAVE_QMESH: the restatement bit for bit against the compiled
    routine (back-azimuths 0 and 30 degrees, P-type polarisation), and the
    plane wave pol * exp(-i (kx x + kz z)) at the GLL nodes placed by the
    bilinear map, computed here independently (from the last element holding
    a node, where the routine takes the first: the phase kx x + kz z of up
    to ~10 rad moves by a few ulps), to 4e-15."""
    m, mod, pml = mesh_case(case)
    g = oracle_graph(case)
    nl = m.nord + 1
    xi, _ = oracle.gll(nl)
    xyz = np.full((g.ndof, 2), np.nan)
    comp = np.zeros(g.ndof, int)
    lm = np.asarray(g.lm).reshape(m.nelem, nl * nl, 3)
    for e in range(m.nelem):
        cx = m.xlocs[m.ieng[e] - 1]
        cz = m.zlocs[m.ieng[e] - 1]
        for ie in range(nl):
            for ix in range(nl):
                a, b = xi[ix], xi[ie]
                sf = 0.25 * np.array([(1 - a) * (1 - b), (1 + a) * (1 - b), (1 + a) * (1 + b), (1 - a) * (1 + b)])
                for c in range(3):
                    d = lm[e, ie * nl + ix, c]
                    if d > 0:
                        xyz[d - 1] = (sf @ cx, sf @ cz)
                        comp[d - 1] = c
    assert not np.isnan(xyz).any()
    for kx, kz, baz, aoi in ((2.1e-4, 3.3e-4, 0.0, 20.0), (-1.5e-4, 2.2e-4, 30.0, 35.0)):
        a = oracle.fill_uwave(m, g, kx, kz, baz, aoi, model=mod, pml=pml)
        b = oracle.fill_uwave(m, g, kx, kz, baz, aoi, reference=True)
        np.testing.assert_array_equal(a, b)
        pol = np.array([np.cos(np.radians(baz)) * np.sin(np.radians(aoi)),
                        np.sin(np.radians(baz)) * np.sin(np.radians(aoi)), np.cos(np.radians(aoi))])
        want = pol[comp] * np.exp(-1j * (kx * xyz[:, 0] + kz * xyz[:, 1]))
        assert np.abs(a - want).max() <= 4e-15


@pytest.mark.parametrize("kind", [b"PP", b"SS", b"PS"])
def test_qp_p3_p1_p2_on_anchors_with_several_elements(kind):
    """CADJ_NU_ALPHA's P3 and MET225_MOD's P1, P2 on anchors held by two or
    four elements: the reference clears TEMP, T1 and T2 per element, so the
    column is the last element of MCONN(INPG,:) alone.  Checked against
    that element's operator assembled on its own (a one-element mesh with
    the same anchors, orc_asmble_z, lambda/mu the anchor's derivative field,
    rho = 0): P = -w Re sum over the element's DOFs of conj(S_e x) y."""
    name = "tiny_layered"
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    one = oracle.inv_setup(m, mod, g, pml)
    io = ps_inv(one) if kind == b"PS" else one
    nv = 2 if kind == b"PS" else 1
    rng = np.random.default_rng(99)
    freq, py = 0.07, 1.1e-4
    cv = lambda: (rng.standard_normal(g.ndof) + 1j * rng.standard_normal(g.ndof)).astype(np.complex64)  # noqa: E731
    u, lam, nu, alpha = cv(), cv(), cv(), cv()
    v = rng.standard_normal(io.na35).astype(np.float32)
    _, _, p3 = oracle.qp_nu_alpha(m, mod, g, io, freq, py, v, lam, u, cinvtype=kind, pml=pml)
    p1, p2 = oracle.met225(m, mod, g, io, freq, py, lam, nu, alpha, u, cinvtype=kind, pml=pml)
    nl = m.nord + 1
    lm = np.asarray(g.lm).reshape(m.nelem, nl * nl, 3)
    cnt = np.bincount(m.ieng.ravel() - 1, minlength=m.nnpg)
    anchors = [a for a in range(m.nnpg) if one.maskg[a] > 0 and cnt[a] >= 2][:6]
    assert len(anchors) >= 3
    lam_, mu_, rho_ = mod.ecoeff[:, 0], mod.ecoeff[:, 1], mod.dens
    vp, vs = np.sqrt((lam_ + 2 * mu_) / rho_), np.sqrt(mu_ / rho_)
    for a in anchors:
        e = one.mconn[:, a][one.mconn[:, a] > 0][-1] - 1
        ie = m.ieng[e] - 1
        m1 = cf.Mesh(nord=m.nord, nx=1, nz=1, ieng=np.array([[1, 2, 3, 4]], np.int32), xlocs=m.xlocs[ie].copy(),
                     zlocs=m.zlocs[ie].copy(), cdomain=b"I", cnnpg=b"II" * 4, xd=np.zeros(4), zd=np.zeros(4))
        g1 = oracle.graph(m1)
        lm1 = np.asarray(g1.lm).reshape(1, nl * nl, 3)
        k = int(np.nonzero(ie == a)[0][0])
        dofs = lm[e].ravel() - 1
        dofs1 = lm1[0].ravel() - 1
        assert np.all(dofs >= 0) and np.all(dofs1 >= 0)
        for iv in range(nv):
            p_var = (kind == b"PP") or (kind == b"PS" and iv == 0)
            fl, fm = np.zeros(4), np.zeros(4)
            if p_var:
                fl[k] = 2 * vp[a]
            else:
                fl[k], fm[k] = -4 * vs[a], 2 * vs[a]
            mod1 = cf.Model(dens=np.zeros(4), ecoeff=np.asfortranarray(np.stack([fl, fm], axis=1)),
                            vp=np.zeros(4), vs=np.zeros(4))
            S = csr(g1, oracle.asmble_z(m1, mod1, g1, freq, py, cf.PML())).toarray()
            Se = np.zeros((g1.ndof, g1.ndof), np.complex128)
            Se[np.ix_(dofs1, dofs1)] = S[np.ix_(dofs1, dofs1)]
            col = (one.maskg[a] - 1) * nv + iv
            w = float(np.float32(io.elem_wts[e]))
            for got, x, y, wv in ((p3[col], u, lam, w * float(v[col])), (p1[col], u, nu, w), (p2[col], alpha, lam, w)):
                t = Se @ x[dofs].astype(np.complex128)
                terms = np.conj(t) * y[dofs]
                want = -wv * np.real(terms.sum())
                assert abs(got - want) <= 2e-5 * abs(wv) * np.abs(terms).sum() + 1e-30, (kind, a, iv, got, want)
