"""MI355X (gfx950) implementation of fem25's element-operator, assembly and
adjoint-gradient hot path.  The compute lives in libfem25hip.so (C ABI in
include/fem25hip.h); this package is its host-side interface."""
from .configs import PML, Mesh, Model  # noqa: F401


def __getattr__(name):
    # the HIP-backed API is imported lazily so that mesh/config helpers work
    # on machines without the built library
    if name in ("MeshInfo", "Graph", "InvInfo", "gen_graph25", "asmble_driver", "asmb_peff",
                "cfmat_dist", "adjflam_dist"):
        from . import api
        return getattr(api, name)
    raise AttributeError(name)
