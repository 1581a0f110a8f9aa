"""Writes the golden fixtures in this directory from the reference's own
Fortran, compiled in place by oracle/Makefile (oracle/_ref/libfem25ref.so).

    python tests/golden/make_golden.py

Inputs are regenerated deterministically from fem25_amd.configs (seeded);
each fixture stores what it was made from (mesh case, frequency, slowness,
seeds) and the reference's outputs:
  gaussq.npz       GAUSSQ(kind=5, kpts=2) nodes and weights, n = 3..19
  dshl.npz         DSHL tables for NORD = 4, 5, 6
  graph_<case>.npz GENIEN/GENID/GENLM/GENCRS_QUICK/CNMESH outputs
  asm_<case>.npz   ASMBLE_DRIVER COMPLEX*8 CDIST at one (freq, py)
  inv_<case>.npz   GENGMASK/ELEM_WTS/GRADPTRS structures and CFMAT_DIST
                   columns for a seeded COMPLEX*8 wavefield, PP and SS
  asmq_<case>.npz  ASMBLE_DRIVER CDIST on a dispersive medium (LDISP, seeded
                   QP/QS, FREQ0 = 1 Hz) at two (freq, py)
  c1_graph_sha.npz SHA-256 of the config-1 (40 x 20, p = 4) numbering
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from fem25_amd import configs as cf  # noqa: E402
from tests.helpers import mesh_case  # noqa: E402

CASES = {"tiny": (0.05, 0.7e-4), "p2_pml": (0.2, 1e-4), "pml": (0.05, 0.7e-4)}


def graph_sha(g) -> str:
    h = hashlib.sha256()
    for a in (g.ien, g.lm, g.irptr, g.jcptr):
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    h.update(g.cnp)
    return h.hexdigest()


def main() -> None:
    if not oracle.ref_available():
        oracle.build(ref=True)
    R = oracle.ref()
    i = lambda v: C.byref(C.c_int(v))  # noqa: E731
    xs, ws = {}, {}
    for n in range(3, 20):
        x = np.zeros(n)
        w = np.zeros(n)
        ep = np.array([-1.0, 1.0])
        R.gaussq_(i(5), i(n), C.byref(C.c_double(0)), C.byref(C.c_double(0)), i(2),
                  ep.ctypes.data_as(oracle.dp), x.ctypes.data_as(oracle.dp),
                  w.ctypes.data_as(oracle.dp))
        xs[f"x{n}"], ws[f"w{n}"] = x, w
    np.savez_compressed(os.path.join(HERE, "gaussq.npz"), **xs, **ws)
    d = {}
    for p in (4, 5, 6):
        xq, wq, shl = oracle.ref_dshl(p)
        d[f"xq{p}"], d[f"wq{p}"], d[f"shl{p}"] = xq, wq, shl
    np.savez_compressed(os.path.join(HERE, "dshl.npz"), **d)
    for case, (freq, py) in CASES.items():
        m, mod, pml = mesh_case(case)
        g = oracle.ref_graph(m)
        np.savez_compressed(os.path.join(HERE, f"graph_{case}.npz"), ien=g.ien, lm=g.lm,
                            irptr=g.irptr, jcptr=g.jcptr, cnp=np.frombuffer(g.cnp, np.uint8),
                            nnp=g.nnp, ndof=g.ndof)
        cd = oracle.ref_asmble(m, mod, g, freq, py, pml)
        np.savez_compressed(os.path.join(HERE, f"asm_{case}.npz"), cdist=cd, freq=freq, py=py)
        out = {}
        wave = cf.seeded_complex(g.ndof, seed=31).astype(np.complex64)
        for cinv in (b"PP", b"SS"):
            inv = oracle.ref_inv_setup(m, mod, g, cinvtype=cinv)
            fm = oracle.ref_cfmat(0.06, 0.8e-4, wave, inv.nz_fdist)
            out[f"fmat_{cinv.decode()}"] = fm
        out.update(maskg=inv.maskg, mconn=inv.mconn, mygrad=inv.mygrad, jcsc=inv.jcsc,
                   icsc=inv.icsc, elem_wts=inv.elem_wts, freq=0.06, py=0.8e-4, wave_seed=31)
        np.savez_compressed(os.path.join(HERE, f"inv_{case}.npz"), **out)
        print(case, g.ndof, g.nz, "written")
    for case in ("tiny_layered", "pml"):
        m, mod, pml = mesh_case(case)
        g = oracle.ref_graph(m)
        md = q_model(mod, m.nnpg)
        out = dict(qp=md.qp, qs=md.qs, fref=md.freq0)
        for k, (freq, py) in enumerate(((0.05, 0.7e-4), (0.2, 0.0))):
            out[f"cdist{k}"] = oracle.ref_asmble(m, md, g, freq, py, pml)
            out[f"freq{k}"], out[f"py{k}"] = freq, py
        np.savez_compressed(os.path.join(HERE, f"asmq_{case}.npz"), **out)
        print(case, "dispersive written")
    m, mod, pml = mesh_case("c1")
    np.savez(os.path.join(HERE, "c1_graph_sha.npz"), sha=graph_sha(oracle.ref_graph(m)))


def q_model(mod, nnpg):
    """Seeded quality factors (QP 60-100, QS 30-50) with FREQ0 = 1 Hz."""
    rng = np.random.default_rng(5)
    return cf.with_attenuation(mod, 60 + 40 * rng.random(nnpg), 30 + 20 * rng.random(nnpg), 1.0)


if __name__ == "__main__":
    main()
