"""The library's host numbering (sort-based GENIEN, GENID, GENLM,
GENCRS_QUICK, CNMESH) against the reference's outputs; no device needed."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

from fem25_amd import _lib
from tests.helpers import mesh_case

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def host_graph(mesh, with_csr=True):
    L = _lib.load()
    keep = [np.ascontiguousarray(mesh.ieng, np.int32), mesh.xlocs, mesh.zlocs, mesh.xd, mesh.zd]
    d = _lib.MeshDesc()
    d.nord, d.nelem, d.nnpg = mesh.nord, mesh.nelem, mesh.nnpg
    d.ieng = keep[0].ctypes.data_as(_lib.ip)
    d.xlocs, d.zlocs = keep[1].ctypes.data_as(_lib.dp), keep[2].ctypes.data_as(_lib.dp)
    d.xd, d.zd = keep[3].ctypes.data_as(_lib.dp), keep[4].ctypes.data_as(_lib.dp)
    d.cdomain, d.cnnpg = mesh.cdomain, mesh.cnnpg
    h = C.c_void_p()
    s = np.zeros(3, np.int64)
    _lib.check(L.f25_graph_build(C.byref(d), int(with_csr), C.byref(h), s.ctypes.data_as(_lib.i64p)),
               "f25_graph_build")
    nnp, ndof, nz = (int(v) for v in s)
    ien = np.zeros((mesh.nelem, mesh.nen), np.int32)
    lm = np.zeros((mesh.nelem, mesh.nen, 3), np.int32)
    irptr = np.zeros(ndof + 1, np.int32)
    jcptr = np.zeros(max(nz, 1), np.int32)
    cnp = C.create_string_buffer(ndof)
    _lib.check(L.f25_graph_get(h, ien.ctypes.data_as(_lib.ip), lm.ctypes.data_as(_lib.ip),
                               irptr.ctypes.data_as(_lib.ip) if with_csr else None,
                               jcptr.ctypes.data_as(_lib.ip) if with_csr else None, cnp),
               "f25_graph_get")
    L.f25_graph_destroy(h)
    return dict(nnp=nnp, ndof=ndof, nz=nz, ien=ien, lm=lm, irptr=irptr, jcptr=jcptr[:nz],
                cnp=cnp.raw[:ndof])


@pytest.mark.parametrize("case", ["tiny", "p2_pml", "pml"])
def test_host_numbering_bitexact_vs_reference(case):
    m, _, _ = mesh_case(case)
    g = host_graph(m)
    z = np.load(os.path.join(GOLD, f"graph_{case}.npz"))
    assert (g["nnp"], g["ndof"]) == (int(z["nnp"]), int(z["ndof"]))
    for k in ("ien", "lm", "irptr", "jcptr"):
        np.testing.assert_array_equal(g[k], z[k])
    assert g["cnp"] == z["cnp"].tobytes()


def test_host_numbering_config1_checksum():
    m, _, _ = mesh_case("c1")
    g = host_graph(m)
    h = hashlib.sha256()
    for k in ("ien", "lm", "irptr", "jcptr"):
        h.update(np.ascontiguousarray(g[k], np.int32).tobytes())
    h.update(g["cnp"])
    assert h.hexdigest() == str(np.load(os.path.join(GOLD, "c1_graph_sha.npz"))["sha"])


def test_host_numbering_rejects_bad_input():
    L = _lib.load()
    m, _, _ = mesh_case("tiny")
    bad = np.ascontiguousarray(m.ieng, np.int32).copy()
    bad[0, 0] = m.nnpg + 5
    d = _lib.MeshDesc()
    d.nord, d.nelem, d.nnpg = m.nord, m.nelem, m.nnpg
    d.ieng = bad.ctypes.data_as(_lib.ip)
    d.xlocs, d.zlocs = m.xlocs.ctypes.data_as(_lib.dp), m.zlocs.ctypes.data_as(_lib.dp)
    d.cdomain, d.cnnpg = m.cdomain, m.cnnpg
    h = C.c_void_p()
    assert L.f25_graph_build(C.byref(d), 1, C.byref(h), None) == 1
    assert b"IENG" in L.f25_last_error()
    d.ieng = np.ascontiguousarray(m.ieng, np.int32).ctypes.data_as(_lib.ip)
    d.nord = 9
    assert L.f25_graph_build(C.byref(d), 1, C.byref(h), None) == 1
