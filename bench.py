"""Throughput of the batched K(omega, ky) u product (and the adjoint gradient)
on one MI355X.  Default workload: the order-6 Gauss-Newton configuration,
400 x 200 elements, GLL order 6 (8,650,803 DOFs), 6 frequencies x 32
out-of-plane wavenumbers = 192 operator applications per step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c5p6|c2|...] [--skip-cpu]

Prints ONE JSON line on stdout; progress goes to stderr.  A step is one
application of all 192 S(omega_b, ky_b) to their vectors (inputs resident in
HBM, layout [dof][pair]).  The CPU baseline is the reference's own Fortran
(ASMBLE_DRIVER, compiled from source into oracle/_ref) plus the CSR product,
one single-threaded process per host core on disjoint (omega, ky) pairs of a
40 x 20 element piece of the same mesh (same element size and order), the
per-element cost extrapolated linearly to the full mesh; it runs before the
GPU is touched.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# the headline metric; value = FP64 product Gdof/s (the mode that meets the
# 1e-8 accuracy bar), gradients/s and the FP32 modes are in "extra"
METRIC = "K(\u03c9,ky)\u00b7u Gdof/s + gradients/s, FP32+refine vs FP64 at 1e-8 rel err, 1 GPU"
PEAK_FP64 = 78.6e12
PEAK_FP32_MFMA = 157.3e12
PEAK_HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c5p6")
    ap.add_argument("--skip-cpu", action="store_true", help="omit the CPU baseline leg")
    ap.add_argument("--skip-extras", action="store_true", help="only the timed FP64 product")
    ap.add_argument("--only", choices=("grad", "grad32", "gn_diag", "cfmat", "qp", "met225", "pchadj", "gnmfree"),
                    default=None,
                    help="time one secondary kernel family alone (for profiling); prints its own JSON line")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one plain device copy of the input (known bytes) for counter scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the batch of "
              "(omega, ky) problems would be split across GPUs with one gradient all-reduce)",
              file=sys.stderr)
        return 2

    # the CPU leg first, while no GPU state exists in this process (its
    # worker processes are started from here)
    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(args.config)
        log(f"cpu baseline: {cpu}")

    import torch

    from fem25_amd import _lib
    from fem25_amd import configs as cf
    from fem25_amd.api import MeshInfo

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    L = _lib.load()

    shape = cf.CONFIGS[args.config]
    mesh = cf.structured_mesh(shape["nx"], shape["nz"], shape["nord"])
    model = cf.layered_model(mesh, seed=2)
    vs_min = 6300.0 / math.sqrt(3.0)
    omega, ky, freq, py = cf.freq_ky_batch(shape["nfreq"], shape["nky"], 0.02, 0.1, vs_min)
    npair = omega.size
    t0 = time.time()
    mi = MeshInfo(mesh)
    mi.set_model(model)
    ndof, nen = mi.ndof, mi.nen
    log(f"setup: {mesh.nelem} elements, p={mesh.nord}, {ndof} DOF, {mi.nz} CSR nonzeros, "
        f"{npair} pairs, {time.time() - t0:.1f} s")

    g = torch.Generator(device="cuda").manual_seed(3)
    if args.only:
        return only_kernel(args, mi, mesh, L, torch, omega, ky, freq, py, g)
    x = torch.randn(ndof, npair, dtype=torch.complex128, device="cuda", generator=g)
    y = torch.empty_like(x)

    def step():
        mi.apply_batch_dev(omega, ky, x, y, _lib.F25_APPLY_F64)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kern_ms = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
        kern_ms.append(L.f25_last_kernel_ms(b"apply"))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    ms_step = 1e3 * wall / args.steps
    gdofs = ndof * npair / (ms_step * 1e-3) / 1e9
    k_ms = float(np.mean(kern_ms))
    # launches per step: one per element colour
    ncol = mi.ncolour
    flops = 60.0 * nen * nen * mesh.nelem * npair          # useful block flops per step
    bytes_alg = 2 * ndof * npair * 16 + 9 * nen * nen * 8 * mesh.nelem
    achieved = flops / (k_ms * 1e-3)
    log(f"FP64 product: {ms_step:.3f} ms/step, kernels {k_ms:.3f} ms, {gdofs:.2f} Gdof/s, "
        f"{achieved / 1e12:.2f} TFLOP/s")

    if args.copy_calib:
        yc = torch.empty_like(x)
        yc.copy_(x)
        torch.cuda.synchronize()
        log(f"copy calibration: {x.numel() * 16} bytes read, {x.numel() * 16} bytes written")
        del yc

    extra = {}
    if not args.skip_extras:
        extra = extras(mi, mesh, model, L, torch, x, y, omega, ky, freq, py, npair, gdofs, g)
        extra["bielak_rhs"] = bielak_extra(L, torch)
        log(json.dumps(extra))

    traffic, traffic_src = profiled_traffic(args.config)
    out = {
        "metric": METRIC,
        "value": round(gdofs, 3),
        "unit": "Gdof/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded layered model with 20 Gaussian anomalies; seeded N(0,1) vectors)",
        "config": {"workload": f"{shape['nx']}x{shape['nz']} elements, GLL order {shape['nord']}, "
                               f"{shape['nfreq']} freq x {shape['nky']} ky = {npair} K(omega,ky) "
                               f"products per step",
                   "nelem": mesh.nelem, "nord": mesh.nord, "ndof": ndof, "npair": npair},
        "roofline": {"bound": "mfma", "achieved": round(achieved / 1e12, 3), "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": f"k_apply_v<{mesh.nord + 1}>",
                     "compute": "FP64 FMA on the vector ALUs; on gfx950 their peak equals the FP64 "
                                "MFMA peak (78.6 TFLOP/s)",
                     "launch_avg_ms": round(k_ms / ncol, 4),
                     "algorithmic_flops_per_launch": flops / ncol,
                     "algorithmic_bytes_per_launch": bytes_alg / ncol,
                     "launches_per_step": ncol, "flops_per_step": flops,
                     "algorithmic_bytes_per_step": bytes_alg,
                     "kernel_ms_per_step": round(k_ms, 4)},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    mi.close()
    return 0


def roof(bound, achieved, unit):
    peak = {"hbm": PEAK_HBM, "fp64": PEAK_FP64, "fp32_mfma": PEAK_FP32_MFMA}[bound]
    scale = 1e9 if unit == "GB/s" else 1e12
    return {"bound": bound, "achieved": round(achieved / scale, 3), "peak": peak / scale,
            "unit": unit, "frac": round(achieved / peak, 4)}


def timed(fn, L, family, reps):
    """Runs fn reps times after one warm-up; mean kernel ms of the family
    (HIP events inside the library) and mean wall ms per call."""
    import torch
    fn()
    torch.cuda.synchronize()
    ks = []
    t = time.perf_counter()
    for _ in range(reps):
        fn()
        ks.append(L.f25_last_kernel_ms(family))
    torch.cuda.synchronize()
    return float(np.mean(ks)), 1e3 * (time.perf_counter() - t) / reps


def extras(mi, mesh, model, L, torch, x, y, omega, ky, freq, py, npair, gdofs, g):
    """Accuracy of the product, the other precision modes, and every other
    kernel of the path with its own roofline."""
    import ctypes as C
    from fem25_amd import _lib
    ndof, nen, nq = mi.ndof, mi.nen, mi.nq1 * mi.nq1
    out = {}
    # accuracy of the MFMA product against the assembled CSR operator (FP64)
    b = npair - 1
    s = torch.empty(mi.nz, dtype=torch.complex128, device="cuda")
    mi.asmble_dev(b"B", float(freq[b]), float(py[b]), s)
    xb = x[:, b].contiguous()
    yb = torch.empty_like(xb)
    mi.csr_apply_dev(s, xb, yb)
    out["fp64_rel_err_vs_assembled_csr"] = float(torch.linalg.norm(y[:, b] - yb) / torch.linalg.norm(yb))
    # plain FP64 GPU path: assemble S(omega, ky) in CSR, then SpMV, per pair
    nplain = 8
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for i in range(nplain):
        mi.asmble_dev(b"B", float(freq[i]), float(py[i]), s)
        mi.csr_apply_dev(s, x[:, i].contiguous(), yb)
    torch.cuda.synchronize()
    plain = ndof * nplain / (time.perf_counter() - tp) / 1e9
    out["plain_fp64_csr_gdof_per_s"] = round(plain, 3)
    out["speedup_vs_plain_fp64"] = round(gdofs / plain, 2)
    # the same FP64 product replayed from a captured HIP graph (F25_APPLY_GRAPH)
    yg = torch.empty_like(y)
    k_g, wall_g = timed(lambda: mi.apply_batch_dev(omega, ky, x, yg, _lib.F25_APPLY_F64 | _lib.F25_APPLY_GRAPH),
                        L, b"apply", 5)
    out["fp64_graph_replay"] = {"ms_per_step": round(wall_g, 3), "gdof_per_s": round(ndof * npair / wall_g * 1e-6, 3),
                                "bit_equal_to_eager": bool(torch.equal(yg, y))}
    del yg
    # FP32 and FP32 refined products
    x32 = x.to(torch.complex64)
    y32 = torch.empty_like(x32)
    yr = torch.empty_like(x)
    for name, mode, xx, yy in (("fp32", _lib.F25_APPLY_F32, x32, y32),
                               ("fp32_refined", _lib.F25_APPLY_F32_REFINED, x, yr)):
        k_ms, wall = timed(lambda: mi.apply_batch_dev(omega, ky, xx, yy, mode), L, b"apply", 3)
        d = yy.to(torch.complex128) - y
        per = (torch.linalg.norm(d, dim=0) / torch.linalg.norm(y, dim=0)).cpu().numpy()
        err = float(per.max())
        out[name] = {"gdof_per_s": round(ndof * npair / wall * 1e-6, 3), "max_rel_err_vs_fp64": err,
                     "median_rel_err_vs_fp64": float(np.median(per)),
                     "rel_err_per_pair": [float(f"{v:.3e}") for v in per],
                     "kernel_ms": round(k_ms, 3),
                     "roofline": roof("fp32_mfma", 60.0 * nen * nen * mesh.nelem * npair * (4 if "ref" in name else 1)
                                      / (k_ms * 1e-3), "TFLOP/s")}
    del x32, y32, yr
    kern = {}
    # assembled FP64 operator: one pass over the CSR basis (HBM)
    k_ms, wall = timed(lambda: mi.asmble_dev(b"B", float(freq[0]), float(py[0]), s), L, b"asmble", 5)
    # bytes of one pass: per entry its kind (1), V0 (8) and the value (16);
    # V1, V2 (16) for the same-component entries only; a base per 64 entries
    gr = mi.graph()
    comp = np.zeros(ndof, np.int8)
    flat = gr.lm.reshape(-1)
    comp[flat[flat > 0] - 1] = (np.nonzero(flat > 0)[0] % 3).astype(np.int8)
    rows = np.repeat(np.arange(ndof), np.diff(gr.irptr.astype(np.int64)))
    n1 = int(np.count_nonzero(comp[rows] == comp[gr.jcptr - 1]))
    del rows, flat, gr
    nzb = mi.nz * (1 + 8 + 16) + n1 * 16 + (mi.nz // 64 + 1) * 8
    kern["asmble_fp64"] = {"kernel_ms": round(k_ms, 3), "calls_per_s": round(1e3 / wall, 1),
                           "roofline": roof("hbm", nzb / (k_ms * 1e-3), "GB/s")}
    del s
    # adjoint gradient over npair (omega, ky) wavefield pairs (FP64 VALU)
    u = torch.randn(npair, ndof, dtype=torch.complex128, device="cuda")
    lam = torch.randn(npair, ndof, dtype=torch.complex128, device="cuda")
    gg = torch.zeros(mesh.nnpg, 3, dtype=torch.float64, device="cuda")
    k_ms, wall = timed(lambda: mi.gradient_dev(omega, ky, u, lam, gg), L, b"grad", 2)
    # flops per (element, triple), sum factorised: two 1-D passes per field
    # and component, 4 (12 N1^2 Q1 + 18 N1 Q1^2), plus ~150 per quadrature point
    n1, q1 = mesh.nord + 1, 2 * mesh.nord + 5
    per = 4.0 * (12 * n1 * n1 * q1 + 18 * n1 * q1 * q1)
    gfl = (per + 150.0 * nq) * mesh.nelem * npair
    kern["gradient"] = {"kernel_ms": round(k_ms, 3), "gradients_per_s": round(npair / (wall * 1e-3), 1),
                        "roofline": roof("fp64", gfl / (k_ms * 1e-3), "TFLOP/s")}
    out["gradients_per_s"] = kern["gradient"]["gradients_per_s"]
    # the FP32 gradient (complex64 fields, FP32 densities, FP64 sums) and its
    # error per field against the FP64 gradient
    u32, l32 = u.to(torch.complex64), lam.to(torch.complex64)
    g64 = torch.zeros_like(gg)
    g32 = torch.zeros_like(gg)
    mi.gradient_dev(omega, ky, u, lam, g64)
    k32, wall32 = timed(lambda: mi.gradient32_dev(omega, ky, u32, l32, g32), L, b"grad", 2)
    g32.zero_()
    mi.gradient32_dev(omega, ky, u32, l32, g32)
    ferr = {nm: float(torch.linalg.norm(g32[:, p] - g64[:, p]) / torch.linalg.norm(g64[:, p]))
            for p, nm in enumerate(("vp", "vs", "rho"))}
    kern["gradient_fp32"] = {"kernel_ms": round(k32, 3), "gradients_per_s": round(npair / (wall32 * 1e-3), 1),
                             "rel_err_vs_fp64_per_field": ferr,
                             "roofline": roof("fp64", gfl / (k32 * 1e-3), "TFLOP/s")}
    del u32, l32, g64, g32
    # pseudo-Hessian diagonal (host arrays, 8 wavefields)
    nt = 8
    uh = u[:nt].cpu().numpy()
    t0 = time.perf_counter()
    mi.gn_diag(omega[:nt], ky[:nt], uh)
    wall = time.perf_counter() - t0
    k_ms = L.f25_last_kernel_ms(b"gn_diag")
    # flops per (element, wavefield), sum factorised: fields and gradients
    # at the points 24 N1^2 Q1 + 36 N1 Q1^2, 17 row sums per row and output
    # lane 136 N1 Q1^2, the eta contraction 224 N1^2 Q1, about 100 per point
    dper = 248.0 * n1 * n1 * q1 + 172.0 * n1 * q1 * q1 + 100.0 * nq
    dfl = dper * mesh.nelem * nt
    kern["gn_diag"] = {"kernel_ms": round(k_ms, 3), "wavefields_per_s": round(nt / (k_ms * 1e-3), 1),
                       "roofline": roof("fp64", dfl / (k_ms * 1e-3), "TFLOP/s")}
    del u, lam, gg, uh
    # inversion structures of the mesh (all anchors of interior elements)
    inv = mi.inv_setup(b"PP")
    wave = (np.random.default_rng(7).standard_normal(ndof) + 0j).astype(np.complex64)
    t0 = time.perf_counter()
    fm, _ = cfmat_host(mi, inv, float(freq[0]), float(py[0]), wave)
    k_ms = L.f25_last_kernel_ms(b"cfmat")
    # one element's DKM_VPVS product per column; per (item, vector), sum
    # factorised: 4 (15 N1^2 Q1 + 18 N1 Q1^2) + 84 Q
    n1, q1 = mesh.nord + 1, 2 * mesh.nord + 5
    per_item = 4.0 * (15 * n1 * n1 * q1 + 18 * n1 * q1 * q1) + 84.0 * nq
    cfl = per_item * inv.nnpgl
    kern["cfmat_dist"] = {"kernel_ms": round(k_ms, 3), "columns_per_s": round(inv.nnpgl / (k_ms * 1e-3), 1),
                          "roofline": roof("fp64", cfl / (k_ms * 1e-3), "TFLOP/s")}
    # PCHADJ's J = G^T F: 60 receivers x 3 components
    nrhs = 180
    gt = torch.randn(ndof, nrhs, dtype=torch.complex64, device="cuda")
    fd = torch.from_numpy(fm).cuda()
    jac = torch.empty(inv.nnpgl, nrhs, dtype=torch.complex64, device="cuda")
    hd = torch.zeros(inv.na35, dtype=torch.float32, device="cuda")
    args = (mi.h, nrhs, C.c_void_p(gt.data_ptr()), C.c_void_p(fd.data_ptr()), C.c_void_p(jac.data_ptr()),
            C.c_void_p(hd.data_ptr()))
    k_ms, wall = timed(lambda: _lib.check(L.f25_pchadj_jac_dev(*args), "pchadj"), L, b"pchadj", 3)
    pb = ndof * nrhs * 8 + inv.nz_fdist * 12 + inv.nnpgl * nrhs * 8
    kern["pchadj_jac"] = {"kernel_ms": round(k_ms, 3), "nrhs": nrhs,
                          "roofline": roof("hbm", pb / (k_ms * 1e-3), "GB/s")}
    del gt, fd, jac, hd
    # GNMFREE over 10 frequencies x 12 sources with 20..120 receivers each
    rng = np.random.default_rng(8)
    nobs = (3 * rng.integers(20, 121, size=120)).astype(np.int32)
    alpha = np.ones(120)
    jt = torch.randn(int(nobs.sum()) * inv.nnpgl, dtype=torch.complex64, device="cuda")
    pv = torch.randn(inv.na35, dtype=torch.float64, device="cuda")
    zv = torch.zeros(inv.na35, dtype=torch.float64, device="cuda")
    args = (mi.h, 120, nobs.ctypes.data_as(_lib.ip), alpha.ctypes.data_as(_lib.dp), C.c_void_p(jt.data_ptr()),
            C.c_void_p(pv.data_ptr()), C.c_void_p(zv.data_ptr()))
    k_ms, wall = timed(lambda: _lib.check(L.f25_gnmfree_dev(*args), "gnmfree"), L, b"gnmfree", 3)
    jb = 2 * jt.numel() * 8
    kern["gnmfree"] = {"kernel_ms": round(k_ms, 3), "jacobians": 120, "jacobian_bytes": jt.numel() * 8,
                       "roofline": roof("hbm", jb / (k_ms * 1e-3), "GB/s")}
    del jt, pv, zv
    # QPMFREE's second-order element products: CADJ_NU_ALPHA (every element
    # holding local columns, two vectors, + one column per anchor) and
    # MET225_MOD (one column per anchor, two vectors); FP64 VALU, per
    # flops per (item, vector) as for CFMAT_DIST
    cv = lambda s: (np.random.default_rng(s).standard_normal(ndof)   # noqa: E731
                    + 1j * np.random.default_rng(s + 1).standard_normal(ndof)).astype(np.complex64)
    lamv, nuv, alv = cv(11), cv(13), cv(15)
    vq = rng.standard_normal(inv.na35).astype(np.float32)
    mi.qp_nu_alpha(float(freq[0]), float(py[0]), vq, lamv, wave)
    k_ms = L.f25_last_kernel_ms(b"qp")
    qfl = per_item * (2 * mesh.nelem + inv.nnpgl)
    kern["qp_nu_alpha"] = {"kernel_ms": round(k_ms, 3), "roofline": roof("fp64", qfl / (k_ms * 1e-3), "TFLOP/s")}
    mi.met225(float(freq[0]), float(py[0]), lamv, nuv, alv, wave)
    k_ms = L.f25_last_kernel_ms(b"met225")
    mfl = per_item * 2 * inv.nnpgl
    kern["met225"] = {"kernel_ms": round(k_ms, 3), "roofline": roof("fp64", mfl / (k_ms * 1e-3), "TFLOP/s")}
    # adjoint right-hand sides of 10 frequencies x 12 sources, 60 receivers:
    # one dense RHS of NDOF COMPLEX*8 per pair (HBM: the zeroed RHS writes)
    nf, ns, nr = 10, 12, 60
    ar = np.random.default_rng(9)
    mr = ar.integers(1, ndof + 1, size=(nr, 3)).astype(np.int32)
    yr = ar.uniform(-4e4, 4e4, nr)
    ee = torch.randn(ns * nr * nf * 3, dtype=torch.complex64, device="cuda")
    oo = torch.randn(ns * nr * nf * 3, dtype=torch.complex64, device="cuda")
    ww = torch.rand(ns * nr * nf * 3, dtype=torch.float32, device="cuda")
    rr = torch.empty(nf * ns * ndof, dtype=torch.complex64, device="cuda")
    ifq = np.repeat(np.arange(1, nf + 1), ns).astype(np.int32)
    isr = np.tile(np.arange(1, ns + 1), nf).astype(np.int32)
    fq = 0.02 + 0.008 * ifq.astype(np.float64)
    pyy = np.full(nf * ns, 0.5e-4)
    iw = C.c_int(0)
    args = (ndof, nf, nr, ns, nf * ns, ifq.ctypes.data_as(_lib.ip), isr.ctypes.data_as(_lib.ip),
            fq.ctypes.data_as(_lib.dp), pyy.ctypes.data_as(_lib.dp), 20.0, 2, 3, 0.05, 0.0, 1, 0,
            yr.ctypes.data_as(_lib.dp), mr.ctypes.data_as(_lib.ip), C.c_void_p(ee.data_ptr()),
            C.c_void_p(oo.data_ptr()), C.cast(C.c_void_p(ww.data_ptr()), _lib.fp), C.c_void_p(rr.data_ptr()),
            C.byref(iw))
    k_ms, wall = timed(lambda: _lib.check(L.f25_cadjlam_rhs_dev(*args), "cadjlam"), L, b"cadjlam", 3)
    kern["cadjlam_rhs"] = {"kernel_ms": round(k_ms, 3), "pairs": nf * ns,
                           "roofline": roof("hbm", nf * ns * ndof * 8 / (k_ms * 1e-3), "GB/s")}
    # the same with the data covariance (LAPLACE_1DV2 + BPRESID_COVD4),
    # element lengths from the receivers' spacing along the line
    xr = np.sort(ar.uniform(0.0, 8e4, nr)) + np.arange(nr)
    dx = np.zeros(nr - 1)
    _lib.check(L.f25_rec_space1d(nr, xr.ctypes.data_as(_lib.dp), dx.ctypes.data_as(_lib.dp)), "rec_space1d")
    cargs = (ndof, nf, nr, ns, nf * ns, ifq.ctypes.data_as(_lib.ip), isr.ctypes.data_as(_lib.ip),
             fq.ctypes.data_as(_lib.dp), pyy.ctypes.data_as(_lib.dp), 20.0, 3, 0.0, 1, 0, 1,
             dx.ctypes.data_as(_lib.dp), yr.ctypes.data_as(_lib.dp), mr.ctypes.data_as(_lib.ip),
             C.c_void_p(ee.data_ptr()), C.c_void_p(oo.data_ptr()), C.cast(C.c_void_p(ww.data_ptr()), _lib.fp),
             C.c_void_p(rr.data_ptr()), C.byref(iw))
    k_ms, wall = timed(lambda: _lib.check(L.f25_cadjlam_rhs_covd_dev(*cargs), "cadjlam_covd"), L,
                       b"cadjlam_covd", 3)
    kern["cadjlam_rhs_covd"] = {"kernel_ms": round(k_ms, 3), "pairs": nf * ns,
                                "roofline": roof("hbm", nf * ns * ndof * 8 / (k_ms * 1e-3), "GB/s")}
    del ee, oo, ww, rr
    out["kernels"] = kern
    return out


def bielak_extra(L, torch):
    """The Bielak sources of a batch on the configuration-4 mesh (200 x 100
    elements of order 4, two PML rings and one Bielak ring, 963,603 DOFs),
    10 frequencies x 24 ky = 240 pairs, one plane wave per pair: the wave
    (FILL_UWAVE_QMESH) and its effective force (ASMB_PEFF), eager and as one
    HIP graph.  HBM: the two dense [dof][pair] complex128 results written
    (the ring's product is ~1.5% of a full one)."""
    import math
    from fem25_amd import configs as cf
    from fem25_amd.api import MeshInfo
    m = cf.structured_mesh(200, 100, 4, npml=2, nbielak=1)
    mi = MeshInfo(m, cf.PML(rcoeff=1e-3, kappa=1.0, fcent=0.05))
    mi.set_model(cf.layered_model(m, seed=2))
    om, ky, f, py = cf.freq_ky_batch(10, 24, 0.02, 0.1, 6300.0 / math.sqrt(3.0))
    n, npair = mi.ndof, om.size
    rng = np.random.default_rng(12)
    aoi = rng.uniform(10.0, 40.0, npair)
    baz = rng.uniform(-30.0, 30.0, npair)
    s = np.sin(np.radians(aoi)) / 6000.0
    kx, kz = om * s, om * np.cos(np.radians(aoi)) / 6000.0
    ue = torch.empty(n * npair, dtype=torch.complex128, device="cuda")
    pe = torch.empty_like(ue)
    res = {"workload": "200x100 p4, 2 PML + 1 Bielak rings, 240 pairs", "ndof": n}
    for name, graph in (("eager", False), ("graph", True)):
        k_ms, wall = timed(lambda: mi.bielak_rhs_dev(om, ky, kx, kz, baz, aoi, ue, pe, graph=graph), L,
                           b"bielak_rhs", 5)
        res[name] = {"ms": round(k_ms, 3), "wall_ms": round(wall, 3),
                     "roofline": roof("hbm", 2 * 16.0 * n * npair / (k_ms * 1e-3), "GB/s")}
    mi.close()
    return res


def only_kernel(args, mi, mesh, L, torch, omega, ky, freq, py, g):
    """One kernel family alone, warm-up then --steps calls: its mean kernel
    ms (HIP events inside the library) as one JSON line."""
    ndof, npair = mi.ndof, omega.size
    fam = args.only
    if fam in ("grad", "grad32", "gn_diag"):
        u = torch.randn(npair, ndof, dtype=torch.complex128, device="cuda", generator=g)
        lam = torch.randn(npair, ndof, dtype=torch.complex128, device="cuda", generator=g)
        gg = torch.zeros(mesh.nnpg, 3, dtype=torch.float64, device="cuda")
        if fam == "grad":
            fn, lib_fam = (lambda: mi.gradient_dev(omega, ky, u, lam, gg)), b"grad"
        elif fam == "grad32":
            u32, l32 = u.to(torch.complex64), lam.to(torch.complex64)
            del u, lam
            fn, lib_fam = (lambda: mi.gradient32_dev(omega, ky, u32, l32, gg)), b"grad"
        else:
            uh = u[:8].cpu().numpy()
            del u, lam
            fn, lib_fam = (lambda: mi.gn_diag(omega[:8], ky[:8], uh)), b"gn_diag"
    else:
        import ctypes as C
        inv = mi.inv_setup(b"PP")
        rng = np.random.default_rng(7)
        cv = lambda: (rng.standard_normal(ndof) + 1j * rng.standard_normal(ndof)).astype(np.complex64)  # noqa: E731
        wave, lamv, nuv, alv = cv(), cv(), cv(), cv()
        if fam == "cfmat":
            fn, lib_fam = (lambda: cfmat_host(mi, inv, float(freq[0]), float(py[0]), wave)), b"cfmat"
        elif fam == "gnmfree":
            # GNMFREE over 10 frequencies x 12 sources with 20..120 receivers each
            from fem25_amd import _lib
            nobs = (3 * rng.integers(20, 121, size=120)).astype(np.int32)
            alpha = np.ones(120)
            jt = torch.randn(int(nobs.sum()) * inv.nnpgl, dtype=torch.complex64, device="cuda", generator=g)
            pv = torch.randn(inv.na35, dtype=torch.float64, device="cuda", generator=g)
            zv = torch.zeros(inv.na35, dtype=torch.float64, device="cuda")
            gargs = (mi.h, 120, nobs.ctypes.data_as(_lib.ip), alpha.ctypes.data_as(_lib.dp),
                     C.c_void_p(jt.data_ptr()), C.c_void_p(pv.data_ptr()), C.c_void_p(zv.data_ptr()))
            fn, lib_fam = (lambda: _lib.check(L.f25_gnmfree_dev(*gargs), "gnmfree")), b"gnmfree"
        elif fam == "pchadj":
            # J = G^T F for 60 receivers x 3 components, F from CFMAT_DIST
            from fem25_amd import _lib
            fm, _ = cfmat_host(mi, inv, float(freq[0]), float(py[0]), wave)
            nrhs = 180
            gt = torch.randn(ndof, nrhs, dtype=torch.complex64, device="cuda", generator=g)
            fd = torch.from_numpy(fm).cuda()
            jac = torch.empty(inv.nnpgl, nrhs, dtype=torch.complex64, device="cuda")
            hd = torch.zeros(inv.na35, dtype=torch.float32, device="cuda")
            pargs = (mi.h, nrhs, C.c_void_p(gt.data_ptr()), C.c_void_p(fd.data_ptr()),
                     C.c_void_p(jac.data_ptr()), C.c_void_p(hd.data_ptr()))
            fn, lib_fam = (lambda: _lib.check(L.f25_pchadj_jac_dev(*pargs), "pchadj")), b"pchadj"
        elif fam == "qp":
            vq = rng.standard_normal(inv.na35).astype(np.float32)
            fn, lib_fam = (lambda: mi.qp_nu_alpha(float(freq[0]), float(py[0]), vq, lamv, wave)), b"qp"
        else:
            fn, lib_fam = (lambda: mi.met225(float(freq[0]), float(py[0]), lamv, nuv, alv, wave)), b"met225"
    for _ in range(max(args.warmup, 1)):
        fn()
    torch.cuda.synchronize()
    ks = []
    for _ in range(args.steps):
        fn()
        ks.append(L.f25_last_kernel_ms(lib_fam))
    torch.cuda.synchronize()
    print(json.dumps({"only": fam, "config": args.config, "kernel_ms": float(np.mean(ks)), "calls": args.steps}),
          flush=True)
    mi.close()
    return 0


def cfmat_host(mi, inv, freq, py, wave):
    from fem25_amd import api
    return api.cfmat_dist(inv.nz_fdist, mi.ndof, freq, py, wave, mi, inv)


# rocprofv3 FETCH_SIZE / WRITE_SIZE summaries of the product, per workload
TRAFFIC_PROFILES = {
    "c2": "k_apply_v_f64_p4_200x100_128pairs.md",
    "c5p6": "k_apply_v_f64_p6_400x200_192pairs.md",
}


def profiled_traffic(config: str):
    """HBM bytes per k_apply_v launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (None when there is none)."""
    name = TRAFFIC_PROFILES.get(config)
    if name is None:
        return None, None
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1]), "profiles/" + name
    except OSError:
        pass
    return None, None


# the CPU sample: a piece of the workload's mesh with the same element size
# and order, each worker process one (omega, ky) pair after another
CPU_SAMPLE = (40, 20)
CPU_PAIRS_PER_WORKER = 2


def _cpu_worker(args):
    """One single-threaded process: ASMBLE_DRIVER (compiled reference Fortran,
    or the C restatement where it is not built) and the CSR product for its
    (omega, ky) pairs on the sample mesh; returns seconds and counts."""
    nx, nz, nord, width, depth, pairs, seed = args
    import oracle
    from fem25_amd import configs as cf
    mesh = cf.structured_mesh(nx, nz, nord, width=width, depth=depth)
    model = cf.layered_model(mesh, seed=2)
    g = oracle.graph(mesh)
    x = cf.seeded_complex(g.ndof, seed=seed).astype(np.complex64)
    kind = "reference" if oracle.ref_available() else "port"
    ta = tm = 0.0
    for freq, py in pairs:
        t = time.perf_counter()
        if kind == "reference":
            s = oracle.ref_asmble(mesh, model, g, freq, py, cf.PML())
        else:
            s = oracle.asmble_cdist(mesh, model, g, freq, py, cf.PML())
        ta += time.perf_counter() - t
        t = time.perf_counter()
        oracle.spmv_c(g, s, x)
        tm += time.perf_counter() - t
    return {"kind": kind, "asm_s": ta, "spmv_s": tm, "npair": len(pairs), "nelem": mesh.nelem,
            "ndof": g.ndof}


def cpu_baseline(config: str):
    """The reference path (ASMBLE_DRIVER + CSR SpMV per (omega, ky) pair) on
    the host cores, one single-threaded process per core over disjoint pairs
    (the reference's frequency-group parallelism, bielak25.f90:360), on a
    40 x 20 element piece of the workload's mesh; the cost of a pair is linear
    in the element count, so the full mesh's rate is extrapolated from the
    per-element time."""
    import multiprocessing as mp
    from fem25_amd import configs as cf
    shape = cf.CONFIGS[config]
    nx, nz = CPU_SAMPLE
    nx, nz = min(nx, shape["nx"]), min(nz, shape["nz"])
    width, depth = 80e3 * nx / shape["nx"], 40e3 * nz / shape["nz"]
    vs_min = 6300.0 / math.sqrt(3.0)
    _, _, freq, py = cf.freq_ky_batch(shape["nfreq"], shape["nky"], 0.02, 0.1, vs_min)
    cores = max(1, min(16, os.cpu_count() or 1))
    jobs = []
    for w in range(cores):
        idx = [(w * CPU_PAIRS_PER_WORKER + k) % freq.size for k in range(CPU_PAIRS_PER_WORKER)]
        jobs.append((nx, nz, shape["nord"], width, depth, [(float(freq[i]), float(py[i])) for i in idx], 5 + w))
    try:
        t = time.perf_counter()
        with mp.get_context("spawn").Pool(cores) as pool:
            res = pool.map(_cpu_worker, jobs)
        wall = time.perf_counter() - t
    except Exception as e:  # pragma: no cover
        return {"error": str(e)}
    # per (element, pair) seconds of one core, averaged over the workers
    per = float(np.mean([(r["asm_s"] + r["spmv_s"]) / (r["npair"] * r["nelem"]) for r in res]))
    n1 = shape["nord"] + 1
    ndof_full = 3 * ((shape["nx"] * shape["nord"] + 1) * (shape["nz"] * shape["nord"] + 1))
    nelem_full = shape["nx"] * shape["nz"]
    t_pair = per * nelem_full                       # one core, one pair, full mesh
    value = cores * ndof_full / t_pair / 1e9
    asm = float(np.mean([r["asm_s"] / r["npair"] for r in res]))
    spmv = float(np.mean([r["spmv_s"] / r["npair"] for r in res]))
    return {"value": round(value, 7), "unit": "Gdof/s", "cores": cores, "kind": res[0]["kind"],
            "sample": f"{cores} single-threaded processes x {CPU_PAIRS_PER_WORKER} (omega,ky) pairs on a "
                      f"{nx}x{nz}-element piece of the mesh (p={shape['nord']}, {res[0]['ndof']} DOFs): "
                      f"ASMBLE_DRIVER {asm:.2f} s + CSR SpMV {spmv:.3f} s per pair, {per * 1e3:.3f} ms per "
                      f"element and pair, extrapolated to {nelem_full} elements "
                      f"({t_pair:.0f} s per pair per core)",
            "seconds_wall": round(wall, 1), "nen": n1 * n1}


if __name__ == "__main__":
    sys.exit(main())
