"""Times the row-tile candidates of the FP64 batched product per element order
on the GPU and reports the fastest.

    python -m fem25_amd.tune_apply [--orders 4 5 6] [--nx 400] [--nz 200]
                                   [--nfreq 6] [--nky 32] [--reps 5]

Every candidate runs on the same mesh, model and vectors; the time is the
library's HIP-event figure for the product's launches (one per element
colour).  The results of the candidates are compared with each other (they
differ only in the order of floating-point sums).  The per-order defaults
compiled into the library (apply_nrt_default in csrc/capi.hip) come from
this script; f25_set_apply_tiles selects another candidate at run time.
One JSON line per (order, candidate), then one per order with the fastest.
"""
from __future__ import annotations

import argparse
import json
import math
import sys

# row-tile counts compiled into the library, by element order p
CANDIDATES = {2: (1, 2), 3: (1, 2), 4: (1, 2, 3), 5: (2, 3, 4), 6: (2, 3, 4)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--orders", type=int, nargs="+", default=[4, 5, 6])
    ap.add_argument("--nx", type=int, default=400)
    ap.add_argument("--nz", type=int, default=200)
    ap.add_argument("--nfreq", type=int, default=6)
    ap.add_argument("--nky", type=int, default=32)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)

    import numpy as np
    import torch
    from fem25_amd import _lib
    from fem25_amd import configs as cf
    from fem25_amd.api import MeshInfo

    assert torch.cuda.is_available(), "tune_apply needs a HIP device"
    torch.cuda.set_device(0)
    L = _lib.load()
    vs_min = 6300.0 / math.sqrt(3.0)
    omega, ky, _, _ = cf.freq_ky_batch(args.nfreq, args.nky, 0.02, 0.1, vs_min)
    npair = omega.size
    for nord in args.orders:
        if nord not in CANDIDATES:
            print(json.dumps({"order": nord, "error": "no candidates compiled"}), flush=True)
            continue
        mesh = cf.structured_mesh(args.nx, args.nz, nord)
        mi = MeshInfo(mesh)
        mi.set_model(cf.layered_model(mesh, seed=2))
        g = torch.Generator(device="cuda").manual_seed(3)
        x = torch.randn(mi.ndof, npair, dtype=torch.complex128, device="cuda", generator=g)
        y = torch.empty_like(x)
        yref = None
        best = None
        for nrt in CANDIDATES[nord]:
            mi.set_apply_tiles(nrt)
            mi.apply_batch_dev(omega, ky, x, y, _lib.F25_APPLY_F64)
            torch.cuda.synchronize()
            ms = []
            for _ in range(args.reps):
                mi.apply_batch_dev(omega, ky, x, y, _lib.F25_APPLY_F64)
                ms.append(L.f25_last_kernel_ms(b"apply"))
            torch.cuda.synchronize()
            if yref is None:
                yref = y.clone()
                diff = 0.0
            else:
                diff = float(torch.linalg.norm(y - yref) / torch.linalg.norm(yref))
            t = float(np.median(ms))
            rec = {"order": nord, "row_tiles": nrt, "rows_per_tile": -(-(nord + 1) ** 2 // nrt),
                   "kernel_ms_per_step": round(t, 4),
                   "gdof_per_s": round(mi.ndof * npair / (t * 1e-3) / 1e9, 3), "rel_diff_vs_first": diff}
            print(json.dumps(rec), flush=True)
            if best is None or t < best[1]:
                best = (nrt, t)
        print(json.dumps({"order": nord, "best_row_tiles": best[0], "kernel_ms_per_step": round(best[1], 4),
                          "mesh": f"{args.nx}x{args.nz}", "npair": npair}), flush=True)
        mi.set_apply_tiles(0)
        del x, y, yref
        mi.close()
        torch.cuda.empty_cache()
    return 0


if __name__ == "__main__":
    sys.exit(main())
