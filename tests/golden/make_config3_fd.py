"""Writes tests/golden/c3_fd.npz: central finite differences (relative step
1e-4) of the config-3 misfit (tests/config3.py: 40 x 20 elements with PML and
Bielak rings, three plane-wave sources, 60 receivers) with respect to Vp, Vs
and rho at ten seeded inversion anchors, from FP64 sparse solves of the
oracle's operator.  Test infrastructure: about six minutes on one core.

    python tests/golden/make_config3_fd.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests import config3 as c3  # noqa: E402


def main():
    st = c3.setup()
    anchors = c3.fd_anchors(st)
    fd = c3.fd_gradient(st, anchors)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c3_fd.npz")
    np.savez(out, anchors=anchors, fd=fd, chi=np.float64(st.chi), rel_step=np.float64(1e-4))
    print(f"chi {st.chi:.12e}; anchors {anchors.tolist()}; wrote {out}")


if __name__ == "__main__":
    main()
