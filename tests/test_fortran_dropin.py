"""The Fortran drop-in entry points (fem25_amd/fortran/f25_dropin.f90):
ASMBLE_DRIVER, CFMAT_DIST, ADJFLAM_DIST, ASMB_PEFF and GNMFREE with the
reference's names and argument lists, called from Fortran drivers that fill
the reference's MESH_INFO / INV_INFO / FRQ_INFO / SRC_INFO
(tests/fortran/dropin_check.f90, gnmfree_check.f90), checked against the
oracle at the tolerances of the C-ABI parity tests.

The drop-in is compiled against the reference's own fwd_struc.h, so both
binaries are built where the reference sources are available
(__graft_entry__.build()); these tests skip when they were never built."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from fem25_amd import configs as cf
from tests.helpers import mesh_case, oracle_graph, ps_inv, relerr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "fem25_amd", "fortran", "_build", "libfem25dropin.so")
CHECK = os.path.join(ROOT, "tests", "fortran", "_build", "dropin_check")
GNCHECK = os.path.join(ROOT, "tests", "fortran", "_build", "gnmfree_check")
SYMBOLS = ("asmble_driver_", "asmb_peff_", "cfmat_dist_", "adjflam_dist_", "gnmfree_", "met225_mod_",
           "f25_cadj_nu_alpha_rhs_")

needs_build = pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(CHECK)
                                      and os.path.exists(GNCHECK)),
                                 reason="Fortran drop-in not built (needs the reference's fwd_struc.h)")
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built")


@needs_build
def test_dropin_exports_reference_names():
    out = subprocess.run(["nm", "-D", "--defined-only", DROPIN], capture_output=True, text=True,
                         check=True).stdout
    names = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    for s in SYMBOLS:
        assert s in names, s


def _write_input(path, m, mod, pml, g, io, freq, py, wave, lam, ue, nu, alpha, v, cftype=b"B", cinv=b"PP",
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
        for x in (wave, lam, ue, nu, alpha):
            f.write(c8(x))
        f.write(np.ascontiguousarray(v, "<f4").tobytes())


@needs_build
@pytest.mark.gpu
@pytest.mark.parametrize("name,cinv", [("pml", b"PP"), ("tiny_layered", b"PP"), ("tiny_layered", b"SS"),
                                       ("pml", b"PS")])
def test_dropin_against_oracle(name, cinv, tmp_path):
    """PS (NVINV = 2): MET225_MOD and CADJ_NU_ALPHA's element part against
    the oracle's PS restatement; CFMAT_DIST returns IERR /= 0 as VFORCE25
    does, so FMAT and GPROD are not compared."""
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    io = oracle.inv_setup(m, mod, g, pml)
    if cinv == b"PS":
        io = ps_inv(io)
    freq, py = 0.06, 0.8e-4
    wave = cf.seeded_complex(g.ndof, seed=61).astype(np.complex64)
    lam = cf.seeded_complex(g.ndof, seed=62).astype(np.complex64)
    ue = cf.seeded_complex(g.ndof, seed=63).astype(np.complex64)
    nu = cf.seeded_complex(g.ndof, seed=64).astype(np.complex64)
    alpha = cf.seeded_complex(g.ndof, seed=65).astype(np.complex64)
    v = np.random.default_rng(66).standard_normal(io.na35).astype(np.float32)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, m, mod, pml, g, io, freq, py, wave, lam, ue, nu, alpha, v, cinv=cinv)
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

    st = take("<i4", 7)
    cdist = take("<c8", g.nz)
    fmat = take("<c8", io.nz_fdist)
    gprod = take("<f4", io.na35)
    peff = take("<c8", g.ndof)
    p1 = take("<f4", io.na35)
    p2 = take("<f4", io.na35)
    rhsbnu = take("<c8", g.ndof)
    rhsalp = take("<c8", g.ndof)
    p3 = take("<f4", io.na35)
    assert off == len(raw)
    assert st[0] == 0 and st[1] == 0 and st[5] == 0 and st[6] == 0, st
    assert (st[2] != 0) == (cinv == b"PS"), st
    assert st[3] != 0, "a wrong NZLOC must be rejected"
    assert st[4] == 0, "the cached second call must reproduce the first"
    r1, r2 = oracle.met225(m, mod, g, io, freq, py, lam, nu, alpha, wave, cinvtype=cinv, pml=pml)
    assert relerr(p1, r1) < 2e-5 and relerr(p2, r2) < 2e-5
    rb, ra, r3 = oracle.qp_nu_alpha(m, mod, g, io, freq, py, v, lam, wave, cinvtype=cinv, pml=pml)
    assert relerr(rhsbnu, rb) < 2e-5 and relerr(rhsalp, ra) < 2e-5 and relerr(p3, r3) < 2e-5
    ref = oracle.asmble_cdist(m, mod, g, freq, py, pml)
    rowmax = np.maximum.reduceat(np.abs(ref.astype(np.complex128)), g.irptr[:-1] - 1)
    rows = np.repeat(np.arange(g.ndof), np.diff(g.irptr))
    assert np.all(np.abs(cdist.astype(np.complex128) - ref) <= 1e-6 * rowmax[rows])
    if cinv != b"PS":
        fref = oracle.cfmat_dist(m, mod, g, io, freq, py, wave, cinvtype=cinv)
        assert relerr(fmat, fref) < 2e-6
        assert relerr(gprod, oracle.adjflam_dist(io, fmat, lam)) < 1e-5
    pref = oracle.asmb_peff(g, cdist, ue)
    if np.count_nonzero(pref):
        assert relerr(peff, pref) < 1e-5
    else:
        assert np.count_nonzero(peff) == 0


def _fortran_record(f, a):
    """One Fortran sequential unformatted record: 4-byte length, data, length."""
    b = np.ascontiguousarray(a).tobytes()
    n = np.array([len(b)], "<i4").tobytes()
    f.write(n + b + n)


def _jac_name(freq, isrc, rank=0):
    """./scratch/jac_<rank>-<freq F12.5>-<isrc>.dat (gnmfree.f90:88-97)."""
    return f"jac_{rank}-{format(freq, '12.5f').strip()}-{isrc}.dat"


def _write_jacobians(workdir, freqs, nsrc, nobs, nnpgl, jac):
    """Jacobian files in the record layout pchadj.f90:371-395 writes: header
    (NOBS, NNPGL, NDIM, NREC), LOBS(NDIM, NREC), Re J and Im J (NOBS x NNPGL
    REAL*4, observations fastest); here NREC = NOBS and every LOBS true."""
    os.makedirs(os.path.join(workdir, "scratch"), exist_ok=True)
    off = 0
    for fi, fr in enumerate(freqs):
        for s in range(nsrc):
            no = int(nobs[fi, s])
            J = jac[off:off + no * nnpgl]
            off += no * nnpgl
            with open(os.path.join(workdir, "scratch", _jac_name(fr, s + 1)), "wb") as f:
                _fortran_record(f, np.array([no, nnpgl, 3, no], "<i4"))
                _fortran_record(f, np.ones(3 * no, "<i4"))
                _fortran_record(f, J.real.astype("<f4"))
                _fortran_record(f, J.imag.astype("<f4"))


def _jac_case(seed, nfreq, nsrc, nnpgl, na35):
    rng = np.random.default_rng(seed)
    mygrad = rng.permutation(np.arange(1, na35 + 1))[:nnpgl].astype(np.int32)
    nobs = rng.integers(20, 121, size=(nfreq, nsrc)).astype(np.int32)
    tot = int(nobs.sum()) * nnpgl
    jac = (rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(np.complex64)
    p = rng.standard_normal(na35)
    src = (rng.standard_normal((nfreq, nsrc)) + 1j * rng.standard_normal((nfreq, nsrc))).astype(np.complex64)
    freqs = 0.05 + 0.0125 * np.arange(nfreq)
    return freqs, mygrad, nobs, jac, p, src


@needs_ref
def test_jacobian_files_match_reference_writer(tmp_path):
    """The Jacobian files written here are byte for byte those the compiled
    reference's Fortran writes for the same Jacobians (names included)."""
    freqs, mygrad, nobs, jac, p, src = _jac_case(91, 2, 2, 7, 11)
    oracle.ref_gnmfree(freqs, 2, src, 0, 7, mygrad, 11, nobs, jac, p, str(tmp_path / "ref"))
    _write_jacobians(str(tmp_path / "py"), freqs, 2, nobs, 7, jac)
    names = sorted(os.listdir(tmp_path / "ref" / "scratch"))
    assert names == sorted(os.listdir(tmp_path / "py" / "scratch")) and len(names) == 4
    for n in names:
        assert (tmp_path / "ref" / "scratch" / n).read_bytes() == (tmp_path / "py" / "scratch" / n).read_bytes(), n


@needs_build
@pytest.mark.gpu
@pytest.mark.parametrize("case", [dict(nfreq=3, nsrc=2, npg=1, ipg=0, lmigr=1, nnpgl=150, na35=210, seed=92),
                                  dict(nfreq=4, nsrc=3, npg=2, ipg=1, lmigr=0, nnpgl=97, na35=97, seed=93)],
                         ids=["one_group_migration", "second_of_two_groups"])
def test_gnmfree_dropin(case, tmp_path):
    """GNMFREE drop-in: reads the Jacobian files of its frequency group
    (frequencies IPGROUP+1, IPGROUP+1+NPGROUPS, ...), z = sum alpha
    Re(J^H J p) on the GPU, against the oracle in FP64 (relative 1e-12 of
    max |z|; the Jacobians are COMPLEX*8, the sums FP64 in both); a
    frequency without a file must fail with IERR /= 0."""
    nfreq, nsrc, nnpgl, na35 = case["nfreq"], case["nsrc"], case["nnpgl"], case["na35"]
    freqs, mygrad, nobs, jac, p, src = _jac_case(case["seed"], nfreq, nsrc, nnpgl, na35)
    # purely real or purely imaginary source spectra: |s| is then exact in
    # any complex ABS (hypot(x, 0) = |x|) and |s|^2 the REAL*4 square
    rng = np.random.default_rng(case["seed"] + 100)
    x = rng.standard_normal((nfreq, nsrc)).astype(np.float32)
    src = np.where(rng.random((nfreq, nsrc)) < 0.5, x + 0j, 1j * x).astype(np.complex64)
    _write_jacobians(str(tmp_path), freqs, nsrc, nobs, nnpgl, jac)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array([na35, nnpgl, nfreq, nsrc, case["lmigr"], case["npg"], case["ipg"]], "<i4").tobytes())
        f.write(np.asarray(freqs, "<f8").tobytes())
        f.write(np.asfortranarray(src).ravel(order="F").astype("<c8").tobytes())
        f.write(mygrad.astype("<i4").tobytes())
        f.write(p.astype("<f8").tobytes())
    print(f"gnmfree_check {case}", flush=True)
    r = subprocess.run([GNCHECK, str(fin), str(fout)], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    st = np.frombuffer(raw, "<i4", count=2)
    z = np.frombuffer(raw, "<f8", offset=8)
    assert z.size == na35
    assert st[0] == 0, r.stdout
    assert st[1] != 0, "a frequency without a Jacobian file must be rejected"
    # this group's Jacobians, in the reference's (frequency, source) order
    sel, off, js = [], 0, []
    for fi in range(nfreq):
        for s in range(nsrc):
            n = int(nobs[fi, s]) * nnpgl
            if fi % case["npg"] == case["ipg"]:
                js.append(jac[off:off + n])
                sel.append((fi, s))
            off += n
    no = np.array([nobs[fi, s] for fi, s in sel], np.int32)
    if case["lmigr"]:
        alpha = np.array([x[fi, s] * x[fi, s] for fi, s in sel], np.float32).astype(np.float64)
    else:
        alpha = np.ones(len(sel))
    zref = oracle.gnmfree(no, alpha, nnpgl, mygrad, na35, np.concatenate(js), p)
    assert np.abs(z - zref).max() <= 1e-12 * np.abs(zref).max()
