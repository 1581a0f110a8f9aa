"""Host-side interface of the MI355X path, mirroring the reference's calls.

`MeshInfo` holds what the reference keeps in TYPE(MESH_INFO) for this path
(fwd_struc.h:10-141) plus the device-resident copies the kernels read.  The
module-level functions keep the reference routines' names and argument
meaning:

    gen_graph25(mesh)                               GEN_GRAPH25 numbering (graph25_driver.f90:2)
    asmble_driver(nzloc, cftype, freq, py, msh)     ASMBLE_DRIVER (asmble.f90:1)  -> (cdist, ierr)
    asmb_peff(msh, cdist, ue)                       ASMB_PEFF (asmble.f90:318)
    cfmat_dist(nwork, n, freq, py, wave, msh, inv)  CFMAT_DIST (hessloc.f90:1)     -> (fmat, ierr)
    adjflam_dist(msh, inv, fmat, adjlam)            ADJFLAM_DIST (adjgrad.f90:88) -> gprod

Every call goes to libfem25hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from .configs import PML, Mesh, Model
from ._lib import check, dp, fp, ip


def _p(a, t):
    return a.ctypes.data_as(t)


def _ready(*tensors) -> None:
    """Device inputs are produced on PyTorch's stream; wait for them."""
    import torch
    for t in tensors:
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("device arguments must be contiguous GPU tensors")
    torch.cuda.current_stream().synchronize()


def _expect(t, dtype, numel: int, name: str) -> None:
    """A device argument must have the element type and the exact number of
    values the kernels index (a short buffer would be read or written past
    its end on the device)."""
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if t.numel() != numel:
        raise ValueError(f"{name}: {t.numel()} values, expected {numel}")


def _model_arrays(model: Model, nnpg: int):
    """DENS(NNPG) and ECOEFF(NNPG,2) as contiguous FP64 arrays of the mesh's size."""
    dens = np.ascontiguousarray(model.dens, np.float64)
    ec = np.asfortranarray(np.asarray(model.ecoeff, np.float64))
    if dens.shape != (nnpg,) or ec.shape != (nnpg, 2):
        raise ValueError(f"model arrays must be dens ({nnpg},) and ecoeff ({nnpg}, 2); got "
                         f"{dens.shape} and {ec.shape}")
    return dens, ec


@dataclasses.dataclass
class Graph:
    """GLL numbering in the reference's conventions (1-based)."""
    nnp: int
    ndof: int
    nz: int
    ien: np.ndarray    # (nelem, nen)      IEN(NEN,NELEM)
    lm: np.ndarray     # (nelem, nen, 3)   LM(3,NEN,NELEM), 0 = fixed
    irptr: np.ndarray  # (ndof+1,)
    jcptr: np.ndarray  # (nz,)
    cnp: bytes


@dataclasses.dataclass
class InvInfo:
    """The INV_INFO fields of the virtual-force path (fwd_struc.h:249-368)."""
    nnpinv: int
    ncon: int
    nnpgl: int
    nz_fdist: int
    mbufrhs: int
    na35: int
    maskg: np.ndarray
    mygrad: np.ndarray
    jcsc_fdist: np.ndarray
    icsc_fdist: np.ndarray
    elem_wts: np.ndarray
    cinvtype: bytes = b"PP"
    vpvs: float = 3 ** 0.5


class MeshInfo:
    """Mesh + model on the current HIP device."""

    def __init__(self, mesh: Mesh, pml: PML | None = None, pml_srf: PML | None = None,
                 graph: Graph | None = None):
        L = _lib.load()
        self.mesh = mesh
        d, self._keep = _mesh_desc(mesh, pml, pml_srf)
        self._desc = d
        h = C.c_void_p()
        if graph is None:
            check(L.f25_mesh_create(C.byref(d), C.byref(h)), "f25_mesh_create")
        else:
            lm = np.ascontiguousarray(graph.lm, np.int32)
            irp = np.ascontiguousarray(graph.irptr, np.int32)
            jcp = np.ascontiguousarray(graph.jcptr, np.int32)
            check(L.f25_mesh_create_with_graph(C.byref(d), graph.ndof, _p(lm, ip), _p(irp, ip),
                                               _p(jcp, ip), C.byref(h)),
                  "f25_mesh_create_with_graph")
        self.h = h
        s = np.zeros(6, np.int64)
        check(L.f25_mesh_sizes(h, s.ctypes.data_as(_lib.i64p)), "f25_mesh_sizes")
        self.nnp, self.ndof, self.nz, self.ncolour, self.nen, self.nq1 = (int(v) for v in s)
        self.inv: InvInfo | None = None

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            _lib.load().f25_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- numbering ----------------------------------------------------------
    def graph(self) -> Graph:
        m = self.mesh
        ien = np.zeros((m.nelem, self.nen), np.int32)
        lm = np.zeros((m.nelem, self.nen, 3), np.int32)
        irptr = np.zeros(self.ndof + 1, np.int32)
        jcptr = np.zeros(self.nz, np.int32)
        cnp = C.create_string_buffer(self.ndof)
        check(_lib.load().f25_mesh_graph(self.h, _p(ien, ip), _p(lm, ip), _p(irptr, ip),
                                         _p(jcptr, ip), cnp), "f25_mesh_graph")
        return Graph(self.nnp, self.ndof, self.nz, ien, lm, irptr, jcptr, cnp.raw[:self.ndof])

    # -- model --------------------------------------------------------------
    def set_model(self, model: Model):
        self.model = model
        dens, ec = _model_arrays(model, self.mesh.nnpg)
        check(_lib.load().f25_set_model(self.h, _p(dens, dp), ec.ctypes.data_as(dp)),
              "f25_set_model")
        if getattr(model, "qp", None) is not None:    # dispersive medium (LDISP)
            qp = np.ascontiguousarray(model.qp, np.float64)
            qs = np.ascontiguousarray(model.qs, np.float64)
            if qp.shape != (self.mesh.nnpg,) or qs.shape != (self.mesh.nnpg,):
                raise ValueError("qp and qs must hold one value per anchor")
            check(_lib.load().f25_set_attenuation(self.h, _p(qp, dp), _p(qs, dp), float(model.freq0)),
                  "f25_set_attenuation")
        else:
            check(_lib.load().f25_set_attenuation(self.h, None, None, 0.0), "f25_set_attenuation")

    # -- assembled operator -------------------------------------------------
    def asmble(self, cftype: bytes, freq: float, py: float, complex128: bool = False) -> np.ndarray:
        out = np.zeros(self.nz, np.complex128 if complex128 else np.complex64)
        check(_lib.load().f25_asmble(self.h, cftype, freq, py,
                                     _lib.F25_C128 if complex128 else _lib.F25_C64,
                                     out.ctypes.data_as(C.c_void_p)), "f25_asmble")
        return out

    def asmble_dev(self, cftype: bytes, freq: float, py: float, out) -> None:
        """out: torch complex64/complex128 tensor of NZ values on the GPU."""
        import torch
        _ready(out)
        if out.dtype not in (torch.complex64, torch.complex128):
            raise ValueError(f"out: dtype {out.dtype}, expected complex64 or complex128")
        _expect(out, out.dtype, self.nz, "out")
        prec = _lib.F25_C128 if out.dtype == torch.complex128 else _lib.F25_C64
        check(_lib.load().f25_asmble_dev(self.h, cftype, freq, py, prec, C.c_void_p(out.data_ptr())),
              "f25_asmble_dev")

    def csr_apply_dev(self, s, x, y) -> None:
        import torch
        _ready(s, x, y)
        if s.dtype not in (torch.complex64, torch.complex128):
            raise ValueError(f"s: dtype {s.dtype}, expected complex64 or complex128")
        _expect(s, s.dtype, self.nz, "s")
        _expect(x, s.dtype, self.ndof, "x")
        _expect(y, s.dtype, self.ndof, "y")
        prec = _lib.F25_C128 if s.dtype == torch.complex128 else _lib.F25_C64
        check(_lib.load().f25_csr_apply_dev(self.h, prec, C.c_void_p(s.data_ptr()),
                                            C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())),
              "f25_csr_apply_dev")

    # -- batched element-block product ----------------------------------------
    def fill_uwave_dev(self, kappax, kappaz, baz, aoi, ue) -> None:
        """FILL_UWAVE_QMESH for a batch of plane waves into ue, a torch
        complex128 tensor [ndof, nwave] on the device."""
        _ready(ue)
        kx, kz, bz, ai = (np.ascontiguousarray(np.atleast_1d(a), np.float64) for a in (kappax, kappaz, baz, aoi))
        if not (kx.shape == kz.shape == bz.shape == ai.shape) or kx.ndim != 1:
            raise ValueError("kappax, kappaz, baz, aoi must be 1-D arrays of one length")
        import torch
        _expect(ue, torch.complex128, self.ndof * kx.size, "ue")
        check(_lib.load().f25_fill_uwave_dev(self.h, kx.size, _p(kx, dp), _p(kz, dp), _p(bz, dp), _p(ai, dp),
                                             C.c_void_p(ue.data_ptr())), "f25_fill_uwave_dev")

    def peff_batch_dev(self, omega, ky, ue, peff) -> None:
        """ASMB_PEFF for every (omega, ky) pair of the batch: ue, peff torch
        complex128 tensors [ndof, npair] on the device."""
        import torch
        _ready(ue, peff)
        om = np.ascontiguousarray(omega, np.float64)
        kv = np.ascontiguousarray(ky, np.float64)
        if om.shape != kv.shape or om.ndim != 1:
            raise ValueError("omega and ky must be 1-D arrays of one length")
        _expect(ue, torch.complex128, self.ndof * om.size, "ue")
        _expect(peff, torch.complex128, self.ndof * om.size, "peff")
        check(_lib.load().f25_peff_batch_dev(self.h, om.size, _p(om, dp), _p(kv, dp), C.c_void_p(ue.data_ptr()),
                                             C.c_void_p(peff.data_ptr())), "f25_peff_batch_dev")

    def bielak_rhs_dev(self, omega, ky, kappax, kappaz, baz, aoi, ue, peff, graph: bool = False) -> None:
        """Plane-wave sources of a batch, one per pair: ue[:, b] and its
        effective force peff[:, b] (torch complex128 [ndof, npair])."""
        import torch
        _ready(ue, peff)
        arrs = [np.ascontiguousarray(np.atleast_1d(a), np.float64) for a in (omega, ky, kappax, kappaz, baz, aoi)]
        if any(a.shape != arrs[0].shape for a in arrs) or arrs[0].ndim != 1:
            raise ValueError("the per-pair arrays must be 1-D arrays of one length")
        npair = arrs[0].size
        _expect(ue, torch.complex128, self.ndof * npair, "ue")
        _expect(peff, torch.complex128, self.ndof * npair, "peff")
        check(_lib.load().f25_bielak_rhs_dev(self.h, npair, *[_p(a, dp) for a in arrs], C.c_void_p(ue.data_ptr()),
                                             C.c_void_p(peff.data_ptr()), _lib.F25_APPLY_GRAPH if graph else 0),
              "f25_bielak_rhs_dev")

    def apply_batch_dev(self, omega, ky, x, y, mode: int = _lib.F25_APPLY_F64) -> None:
        """x, y: torch tensors [ndof, npair] (complex128, or complex64 for F32)."""
        import torch
        _ready(x, y)
        om = np.ascontiguousarray(omega, np.float64)
        kv = np.ascontiguousarray(ky, np.float64)
        if om.shape != kv.shape or om.ndim != 1:
            raise ValueError("omega and ky must be 1-D arrays of one length")
        dt = torch.complex64 if (mode & ~_lib.F25_APPLY_GRAPH) == _lib.F25_APPLY_F32 else torch.complex128
        _expect(x, dt, self.ndof * om.size, "x")
        _expect(y, dt, self.ndof * om.size, "y")
        check(_lib.load().f25_apply_batch_dev(self.h, om.size, _p(om, dp), _p(kv, dp), mode,
                                              C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())),
              "f25_apply_batch_dev")

    def residual_batch_dev(self, omega, ky, u, f, r) -> np.ndarray:
        """r = f - S u per pair in FP64 ([ndof, npair] complex128 device
        tensors); returns |r_b| / |f_b| per pair."""
        import torch
        _ready(u, f, r)
        om = np.ascontiguousarray(omega, np.float64)
        kv = np.ascontiguousarray(ky, np.float64)
        if om.shape != kv.shape or om.ndim != 1:
            raise ValueError("omega and ky must be 1-D arrays of one length")
        for t, nm in ((u, "u"), (f, "f"), (r, "r")):
            _expect(t, torch.complex128, self.ndof * om.size, nm)
        rn = np.zeros(om.size)
        check(_lib.load().f25_residual_batch_dev(self.h, om.size, _p(om, dp), _p(kv, dp), C.c_void_p(u.data_ptr()),
                                                 C.c_void_p(f.data_ptr()), C.c_void_p(r.data_ptr()), _p(rn, dp)),
              "f25_residual_batch_dev")
        return rn

    def set_apply_tiles(self, nrt: int = 0) -> None:
        """Row tiles per output component of the FP64 product (0: the order's default)."""
        check(_lib.load().f25_set_apply_tiles(self.h, int(nrt)), "f25_set_apply_tiles")

    def apply_batch(self, omega, ky, x: np.ndarray, mode: int = _lib.F25_APPLY_F64) -> np.ndarray:
        """x: [npair, ndof] complex128 host array; returns S_b x_b per row."""
        om = np.ascontiguousarray(omega, np.float64)
        kv = np.ascontiguousarray(ky, np.float64)
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros_like(x)
        check(_lib.load().f25_apply_batch(self.h, om.size, _p(om, dp), _p(kv, dp), mode,
                                          x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p)),
              "f25_apply_batch")
        return y

    # -- gradient -----------------------------------------------------------
    def gradient(self, omega, ky, u: np.ndarray, lam: np.ndarray) -> np.ndarray:
        """sum_t Re(lam_t^H dS_t/dm u_t) for m = (Vp, Vs, rho) at anchors: (nnpg, 3)."""
        om = np.ascontiguousarray(np.atleast_1d(omega), np.float64)
        kv = np.ascontiguousarray(np.atleast_1d(ky), np.float64)
        u = np.ascontiguousarray(u, np.complex128).reshape(om.size, -1)
        lam = np.ascontiguousarray(lam, np.complex128).reshape(om.size, -1)
        g = np.zeros((self.mesh.nnpg, 3))
        check(_lib.load().f25_gradient(self.h, om.size, _p(om, dp), _p(kv, dp),
                                       u.ctypes.data_as(C.c_void_p), lam.ctypes.data_as(C.c_void_p),
                                       _p(g, dp)), "f25_gradient")
        return g

    def gradient_dev(self, omega, ky, u, lam, g) -> None:
        """u, lam: torch complex128 [ntrip, ndof] on the GPU; g: torch float64 (nnpg, 3), accumulated."""
        import torch
        _ready(u, lam, g)
        om = np.ascontiguousarray(omega, np.float64)
        kv = np.ascontiguousarray(ky, np.float64)
        if om.shape != kv.shape or om.ndim != 1:
            raise ValueError("omega and ky must be 1-D arrays of one length")
        _expect(u, torch.complex128, self.ndof * om.size, "u")
        _expect(lam, torch.complex128, self.ndof * om.size, "lam")
        _expect(g, torch.float64, self.mesh.nnpg * 3, "g")
        check(_lib.load().f25_gradient_dev(self.h, om.size, _p(om, dp), _p(kv, dp),
                                           C.c_void_p(u.data_ptr()), C.c_void_p(lam.data_ptr()),
                                           C.c_void_p(g.data_ptr())), "f25_gradient_dev")

    def gradient32_dev(self, omega, ky, u, lam, g) -> None:
        """FP32 gradient: u, lam torch complex64 [ntrip, ndof]; g torch float64 (nnpg, 3), accumulated."""
        import torch
        _ready(u, lam, g)
        om = np.ascontiguousarray(omega, np.float64)
        kv = np.ascontiguousarray(ky, np.float64)
        if om.shape != kv.shape or om.ndim != 1:
            raise ValueError("omega and ky must be 1-D arrays of one length")
        _expect(u, torch.complex64, self.ndof * om.size, "u")
        _expect(lam, torch.complex64, self.ndof * om.size, "lam")
        _expect(g, torch.float64, self.mesh.nnpg * 3, "g")
        check(_lib.load().f25_gradient32_dev(self.h, om.size, _p(om, dp), _p(kv, dp),
                                             C.c_void_p(u.data_ptr()), C.c_void_p(lam.data_ptr()),
                                             C.c_void_p(g.data_ptr())), "f25_gradient32_dev")

    # -- Gauss-Newton element products -----------------------------------------
    def gn_fv(self, omega, ky, u: np.ndarray, v: np.ndarray) -> np.ndarray:
        om = np.ascontiguousarray(np.atleast_1d(omega), np.float64)
        kv = np.ascontiguousarray(np.atleast_1d(ky), np.float64)
        u = np.ascontiguousarray(u, np.complex128).reshape(om.size, -1)
        v = np.ascontiguousarray(v, np.float64)
        out = np.zeros_like(u)
        check(_lib.load().f25_gn_fv(self.h, om.size, _p(om, dp), _p(kv, dp),
                                    u.ctypes.data_as(C.c_void_p), _p(v, dp),
                                    out.ctypes.data_as(C.c_void_p)), "f25_gn_fv")
        return out

    def gn_fhw(self, omega, ky, u, w) -> np.ndarray:
        om = np.ascontiguousarray(np.atleast_1d(omega), np.float64)
        kv = np.ascontiguousarray(np.atleast_1d(ky), np.float64)
        u = np.ascontiguousarray(u, np.complex128).reshape(om.size, -1)
        w = np.ascontiguousarray(w, np.complex128).reshape(om.size, -1)
        g = np.zeros((self.mesh.nnpg, 3))
        check(_lib.load().f25_gn_fhw(self.h, om.size, _p(om, dp), _p(kv, dp),
                                     u.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                                     _p(g, dp)), "f25_gn_fhw")
        return g

    def gn_diag(self, omega, ky, u) -> np.ndarray:
        """Pseudo-Hessian diagonal sum_t |dS_t/dm_{a,p} u_t|^2, (NNPG, 3)."""
        om = np.ascontiguousarray(np.atleast_1d(omega), np.float64)
        kv = np.ascontiguousarray(np.atleast_1d(ky), np.float64)
        u = np.ascontiguousarray(u, np.complex128).reshape(om.size, -1)
        d = np.zeros((self.mesh.nnpg, 3))
        check(_lib.load().f25_gn_diag(self.h, om.size, _p(om, dp), _p(kv, dp),
                                      u.ctypes.data_as(C.c_void_p), _p(d, dp)), "f25_gn_diag")
        return d

    def pchadj_jac(self, nrhs: int, g: np.ndarray, fmat: np.ndarray, hdiag=None):
        """PCHADJ's J = G^T F: g (nrhs, lrhs) complex64, fmat (NZ_FDIST,);
        returns jac (nnpgl, nrhs) complex64 and the accumulated hdiag."""
        inv = self.inv
        g = np.ascontiguousarray(g, np.complex64)
        fmat = np.ascontiguousarray(fmat, np.complex64)
        jac = np.zeros((inv.nnpgl, nrhs), np.complex64)
        hd = np.zeros(inv.na35, np.float32) if hdiag is None else np.ascontiguousarray(hdiag, np.float32)
        check(_lib.load().f25_pchadj_jac(self.h, nrhs, g.shape[1], g.ctypes.data_as(C.c_void_p),
                                         fmat.ctypes.data_as(C.c_void_p), jac.ctypes.data_as(C.c_void_p),
                                         _p(hd, fp)), "f25_pchadj_jac")
        return jac, hd

    def gnmfree(self, nobs, alpha, jac, p, z=None) -> np.ndarray:
        """GNMFREE's z += sum_j alpha_j Re(J_j^H J_j p); jac packed (j, c, obs)."""
        inv = self.inv
        nobs = np.ascontiguousarray(nobs, np.int32)
        alpha = np.ascontiguousarray(alpha, np.float64)
        jac = np.ascontiguousarray(jac, np.complex64)
        p = np.ascontiguousarray(p, np.float64)
        z = np.zeros(inv.na35) if z is None else np.ascontiguousarray(z, np.float64)
        check(_lib.load().f25_gnmfree(self.h, nobs.size, _p(nobs, ip), _p(alpha, dp),
                                      jac.ctypes.data_as(C.c_void_p), _p(p, dp), _p(z, dp)), "f25_gnmfree")
        return z

    def qp_nu_alpha(self, freq: float, py: float, v, adjlam, wave):
        """CADJ_NU_ALPHA's element loops (qpmfree.f90:622-912): returns
        (RHSBNU, RHSALP) COMPLEX*8 (NDOF) and P3 REAL*4 (NA35), before the
        solves for nu and alpha."""
        inv = self.inv
        v = np.ascontiguousarray(v, np.float32)
        adjlam = np.ascontiguousarray(adjlam, np.complex64)
        wave = np.ascontiguousarray(wave, np.complex64)
        bnu = np.zeros(self.ndof, np.complex64)
        alp = np.zeros(self.ndof, np.complex64)
        p3 = np.zeros(inv.na35, np.float32)
        vp = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
        check(_lib.load().f25_qp_nu_alpha(self.h, freq, py, _p(v, fp), vp(adjlam), vp(wave), vp(bnu), vp(alp),
                                          _p(p3, fp)), "f25_qp_nu_alpha")
        return bnu, alp, p3

    def met225(self, freq: float, py: float, adjlam, nu, alpha, wave):
        """MET225_MOD's element loops (qpmfree.f90:1000-1215): (P1, P2)."""
        inv = self.inv
        vecs = [np.ascontiguousarray(x, np.complex64) for x in (adjlam, nu, alpha, wave)]
        p1 = np.zeros(inv.na35, np.float32)
        p2 = np.zeros(inv.na35, np.float32)
        check(_lib.load().f25_met225(self.h, freq, py, *[x.ctypes.data_as(C.c_void_p) for x in vecs],
                                     _p(p1, fp), _p(p2, fp)), "f25_met225")
        return p1, p2

    # -- inversion bookkeeping -------------------------------------------------
    def inv_setup(self, cinvtype: bytes = b"PP", vpvs: float = 3 ** 0.5, linrec: bool = True,
                  lkbdry: bool = True, mrdof1=None) -> InvInfo:
        sizes = np.zeros(6, np.int32)
        nrec = 0 if mrdof1 is None else len(mrdof1)
        mr = None if mrdof1 is None else np.ascontiguousarray(mrdof1, np.int32)
        check(_lib.load().f25_inv_setup(self.h, cinvtype, vpvs, int(linrec), int(lkbdry), nrec,
                                        None if mr is None else _p(mr, ip), _p(sizes, ip)),
              "f25_inv_setup")
        nnpinv, ncon, nnpgl, nzf, mbuf, na35 = (int(v) for v in sizes)
        maskg = np.zeros(self.mesh.nnpg, np.int32)
        mygrad = np.zeros(nnpgl, np.int32)
        jcsc = np.zeros(nnpgl + 1, np.int32)
        icsc = np.zeros(nzf, np.int32)
        ew = np.zeros(self.mesh.nelem)
        check(_lib.load().f25_inv_get(self.h, _p(maskg, ip), _p(mygrad, ip), _p(jcsc, ip),
                                      _p(icsc, ip), _p(ew, dp)), "f25_inv_get")
        self.inv = InvInfo(nnpinv, ncon, nnpgl, nzf, mbuf, na35, maskg, mygrad, jcsc, icsc, ew,
                           cinvtype, vpvs)
        return self.inv


# ---------------------------------------------------------------------------
# Reference-named entry points.

def _mesh_desc(mesh: Mesh, pml: PML | None = None, pml_srf: PML | None = None):
    """f25_mesh_desc of a mesh and the arrays it points into."""
    pml = pml or PML()
    pml_srf = pml_srf or pml
    keep = [np.ascontiguousarray(mesh.ieng, np.int32), np.ascontiguousarray(mesh.xlocs),
            np.ascontiguousarray(mesh.zlocs), np.ascontiguousarray(mesh.xd),
            np.ascontiguousarray(mesh.zd)]
    d = _lib.MeshDesc()
    d.nord, d.nelem, d.nnpg = mesh.nord, mesh.nelem, mesh.nnpg
    d.ieng = _p(keep[0], ip)
    d.xlocs, d.zlocs = _p(keep[1], dp), _p(keep[2], dp)
    d.xd, d.zd = _p(keep[3], dp), _p(keep[4], dp)
    d.cdomain, d.cnnpg = mesh.cdomain, mesh.cnnpg
    d.xwidth, d.zwidth = mesh.xwidth, mesh.zwidth
    d.pml[:] = [pml.rcoeff, pml.dmax, pml.kappa, pml.fcent]
    d.pml_srf[:] = [pml_srf.rcoeff, pml_srf.dmax, pml_srf.kappa, pml_srf.fcent]
    return d, keep


def gen_graph25(mesh: Mesh, pml: PML | None = None) -> Graph:
    """GLL numbering and CSR pattern with GEN_GRAPH25's semantics, on the
    host (f25_graph_build: no device needed)."""
    L = _lib.load()
    d, keep = _mesh_desc(mesh, pml)
    h = C.c_void_p()
    s = np.zeros(3, np.int64)
    check(L.f25_graph_build(C.byref(d), 1, C.byref(h), s.ctypes.data_as(_lib.i64p)), "f25_graph_build")
    try:
        nnp, ndof, nz = (int(v) for v in s)
        ien = np.zeros((mesh.nelem, mesh.nen), np.int32)
        lm = np.zeros((mesh.nelem, mesh.nen, 3), np.int32)
        irptr = np.zeros(ndof + 1, np.int32)
        jcptr = np.zeros(nz, np.int32)
        cnp = C.create_string_buffer(ndof + 1)
        check(L.f25_graph_get(h, _p(ien, ip), _p(lm, ip), _p(irptr, ip), _p(jcptr, ip), cnp),
              "f25_graph_get")
    finally:
        L.f25_graph_destroy(h)
    return Graph(nnp, ndof, nz, ien, lm, irptr, jcptr, cnp.raw[:ndof])


def asmble_driver(nzloc: int, cftype: bytes, freq: float, py: float, msh: MeshInfo):
    """ASMBLE_DRIVER(NZLOC,CFTYPE,FREQ,PY,MSH,CDIST,IERR): returns (CDIST COMPLEX*8, IERR)."""
    if nzloc != msh.nz:
        return np.zeros(max(nzloc, 0), np.complex64), 1
    try:
        return msh.asmble(cftype, freq, py), 0
    except _lib.Fem25Error as e:
        return np.zeros(nzloc, np.complex64), e.code


def asmb_peff(msh: MeshInfo, cdist: np.ndarray, ue: np.ndarray) -> np.ndarray:
    """ASMB_PEFF for one rank: the Bielak effective force, COMPLEX*8."""
    cdist = np.ascontiguousarray(cdist, np.complex64)
    ue = np.ascontiguousarray(ue, np.complex64)
    out = np.zeros(msh.ndof, np.complex64)
    check(_lib.load().f25_asmb_peff(msh.h, cdist.ctypes.data_as(C.c_void_p),
                                    ue.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)),
          "f25_asmb_peff")
    return out


def cfmat_dist(nwork: int, n: int, freq: float, py: float, wave: np.ndarray, msh: MeshInfo,
               inv: InvInfo):
    """CFMAT_DIST(NWORK,N,FREQ,PY,WAVE,MSH,INV,FMAT_DIST,IERR) -> (FMAT_DIST, IERR)."""
    if nwork != inv.nz_fdist or n != msh.ndof:
        return np.zeros(max(nwork, 0), np.complex64), 1
    wave = np.ascontiguousarray(wave, np.complex64)
    out = np.zeros(nwork, np.complex64)
    code = _lib.load().f25_cfmat_dist(msh.h, freq, py, wave.ctypes.data_as(C.c_void_p),
                                      out.ctypes.data_as(C.c_void_p))
    return out, int(code)


def adjflam_dist(msh: MeshInfo, inv: InvInfo, fmat: np.ndarray, adjlam: np.ndarray) -> np.ndarray:
    """ADJFLAM_DIST for one rank: GPROD(NA35) = -Re(F^H lambda), REAL*4."""
    fmat = np.ascontiguousarray(fmat, np.complex64)
    adjlam = np.ascontiguousarray(adjlam, np.complex64)
    g = np.zeros(inv.na35, np.float32)
    check(_lib.load().f25_adjflam_dist(msh.h, fmat.ctypes.data_as(C.c_void_p),
                                       adjlam.ctypes.data_as(C.c_void_p), _p(g, fp)),
          "f25_adjflam_dist")
    return g
