"""The Fortran drop-in entry points (fem25_amd/fortran/f25_dropin.f90):
ASMBLE_DRIVER, CFMAT_DIST, ADJFLAM_DIST and ASMB_PEFF with the reference's
names and argument lists, called from a Fortran driver that fills the
reference's MESH_INFO / INV_INFO (tests/fortran/dropin_check.f90), checked
against the oracle at the tolerances of the C-ABI parity tests.

The drop-in is compiled against the reference's own fwd_struc.h, so both
binaries are built where the reference sources are available
(__graft_entry__.build()); these tests skip when they were never built."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from fem25_amd import configs as cf
from tests.helpers import mesh_case, oracle_graph, relerr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "fem25_amd", "fortran", "_build", "libfem25dropin.so")
CHECK = os.path.join(ROOT, "tests", "fortran", "_build", "dropin_check")
SYMBOLS = ("asmble_driver_", "asmb_peff_", "cfmat_dist_", "adjflam_dist_")

needs_build = pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(CHECK)),
                                 reason="Fortran drop-in not built (needs the reference's fwd_struc.h)")


@needs_build
def test_dropin_exports_reference_names():
    out = subprocess.run(["nm", "-D", "--defined-only", DROPIN], capture_output=True, text=True,
                         check=True).stdout
    names = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    for s in SYMBOLS:
        assert s in names, s


def _write_input(path, m, mod, pml, g, io, freq, py, wave, lam, ue, cftype=b"B", cinv=b"PP",
                 vpvs=3 ** 0.5):
    i4 = lambda a: np.ascontiguousarray(a, dtype="<i4").tobytes()  # noqa: E731
    f8 = lambda a: np.ascontiguousarray(a, dtype="<f8").tobytes()  # noqa: E731
    c8 = lambda a: np.ascontiguousarray(a, dtype="<c8").tobytes()  # noqa: E731
    hdr = [m.nord, m.nelem, m.nnpg, g.ndof, g.nz, io.ncon, io.nnpgl, io.nz_fdist, io.na35]
    pa = pml.as_array(m)
    sc = [pa[0], pa[1], pa[2], pa[3], pa[4], pa[5], pa[2], pa[3], pa[4], pa[5], freq, py]
    with open(path, "wb") as f:
        f.write(i4(hdr))
        f.write(f8(sc))
        f.write(f8([vpvs]))
        f.write(cftype[:1] + cinv[:2] + b" ")
        f.write(i4(m.ieng))                       # IENG(4,NELEM): element-major rows
        for a in (m.xlocs, m.zlocs, m.xd, m.zd, mod.dens):
            f.write(f8(a))
        f.write(f8(np.asfortranarray(mod.ecoeff).ravel(order="F")))
        f.write(m.cdomain)
        f.write(m.cnnpg)
        f.write(i4(g.lm))                         # LM(3,NEN,NELEM)
        f.write(i4(g.irptr))
        f.write(i4(g.jcptr))
        f.write(g.cnp)
        f.write(i4(io.mconn))                     # MCONN(NNPG,NCON) memory order
        f.write(i4(io.maskg))
        f.write(i4(io.mygrad))
        f.write(i4(io.jcsc))
        f.write(i4(io.icsc))
        f.write(f8(io.elem_wts))
        for v in (wave, lam, ue):
            f.write(c8(v))


@needs_build
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["pml", "tiny_layered"])
def test_dropin_against_oracle(name, tmp_path):
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    io = oracle.inv_setup(m, mod, g, pml)
    freq, py = 0.06, 0.8e-4
    wave = cf.seeded_complex(g.ndof, seed=61).astype(np.complex64)
    lam = cf.seeded_complex(g.ndof, seed=62).astype(np.complex64)
    ue = cf.seeded_complex(g.ndof, seed=63).astype(np.complex64)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, m, mod, pml, g, io, freq, py, wave, lam, ue)
    print(f"dropin_check {name}: ndof={g.ndof} nz={g.nz} nz_fdist={io.nz_fdist}", flush=True)
    r = subprocess.run([CHECK, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a

    st = take("<i4", 5)
    cdist = take("<c8", g.nz)
    fmat = take("<c8", io.nz_fdist)
    gprod = take("<f4", io.na35)
    peff = take("<c8", g.ndof)
    assert off == len(raw)
    assert st[0] == 0 and st[1] == 0 and st[2] == 0, st
    assert st[3] != 0, "a wrong NZLOC must be rejected"
    assert st[4] == 0, "the cached second call must reproduce the first"
    ref = oracle.asmble_cdist(m, mod, g, freq, py, pml)
    rowmax = np.maximum.reduceat(np.abs(ref.astype(np.complex128)), g.irptr[:-1] - 1)
    rows = np.repeat(np.arange(g.ndof), np.diff(g.irptr))
    assert np.all(np.abs(cdist.astype(np.complex128) - ref) <= 1e-6 * rowmax[rows])
    fref = oracle.cfmat_dist(m, mod, g, io, freq, py, wave, cinvtype=b"PP")
    assert relerr(fmat, fref) < 2e-6
    assert relerr(gprod, oracle.adjflam_dist(io, fmat, lam)) < 1e-5
    pref = oracle.asmb_peff(g, cdist, ue)
    if np.count_nonzero(pref):
        assert relerr(peff, pref) < 1e-5
    else:
        assert np.count_nonzero(peff) == 0
