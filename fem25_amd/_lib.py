"""ctypes binding of libfem25hip.so (C ABI in include/fem25hip.h).

The library is loaded from the package directory; there is no fallback: if
it is missing the import of this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfem25hip.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)


class MeshDesc(C.Structure):
    _fields_ = [("nord", C.c_int), ("nelem", C.c_int), ("nnpg", C.c_int), ("ieng", ip),
                ("xlocs", dp), ("zlocs", dp), ("xd", dp), ("zd", dp), ("cdomain", C.c_char_p),
                ("cnnpg", C.c_char_p), ("xwidth", C.c_double), ("zwidth", C.c_double),
                ("pml", C.c_double * 4), ("pml_srf", C.c_double * 4)]


F25_C64, F25_C128 = 0, 1
F25_APPLY_F64, F25_APPLY_F32, F25_APPLY_F32_REFINED = 0, 1, 2
F25_APPLY_GRAPH = 16   # or'ed with F25_APPLY_F64: replay a captured HIP graph
F25_ERR_ARG, F25_ERR_HIP, F25_ERR_STATE = 1, 4, 8    # status codes of fem25hip.h

EXPORTS = {
    "f25_mesh_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    "f25_mesh_create_with_graph": (C.c_int, [C.POINTER(MeshDesc), C.c_int, ip, ip, ip,
                                             C.POINTER(C.c_void_p)]),
    "f25_mesh_destroy": (C.c_int, [C.c_void_p]),
    "f25_mesh_sizes": (C.c_int, [C.c_void_p, i64p]),
    "f25_mesh_graph": (C.c_int, [C.c_void_p, ip, ip, ip, ip, C.c_char_p]),
    "f25_set_model": (C.c_int, [C.c_void_p, dp, dp]),
    "f25_set_attenuation": (C.c_int, [C.c_void_p, dp, dp, C.c_double]),
    "f25_asmble": (C.c_int, [C.c_void_p, C.c_char, C.c_double, C.c_double, C.c_int, C.c_void_p]),
    "f25_asmble_dev": (C.c_int, [C.c_void_p, C.c_char, C.c_double, C.c_double, C.c_int, C.c_void_p]),
    "f25_csr_apply_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "f25_asmb_peff": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "f25_apply_batch_dev": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_int, C.c_void_p,
                                      C.c_void_p]),
    "f25_apply_batch": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_int, C.c_void_p, C.c_void_p]),
    "f25_set_apply_tiles": (C.c_int, [C.c_void_p, C.c_int]),
    "f25_residual_batch_dev": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_void_p, C.c_void_p, C.c_void_p, dp]),
    "f25_gradient": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_void_p, C.c_void_p, dp]),
    "f25_gradient_dev": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "f25_gradient32_dev": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_void_p, C.c_void_p, C.c_void_p]),
    "f25_inv_setup": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_int, C.c_int, C.c_int,
                                ip, ip]),
    "f25_inv_get": (C.c_int, [C.c_void_p, ip, ip, ip, ip, dp]),
    "f25_inv_attach": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_int, C.c_int, ip, ip,
                                 C.c_int, ip, ip, C.c_int, ip, C.c_int, dp]),
    "f25_asmb_peff_csr": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_char_p, ip, ip, ip, C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
    "f25_adjflam_csc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, ip, C.c_void_p,
                                  C.c_void_p, fp]),
    "f25_cfmat_dist": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p]),
    "f25_adjflam_dist": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, fp]),
    "f25_pchadj_jac": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, fp]),
    "f25_pchadj_jac_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "f25_gnmfree": (C.c_int, [C.c_void_p, C.c_int, ip, dp, C.c_void_p, dp, dp]),
    "f25_gnmfree_dev": (C.c_int, [C.c_void_p, C.c_int, ip, dp, C.c_void_p, C.c_void_p, C.c_void_p]),
    "f25_gn_hessian_dev": (C.c_int, [C.c_void_p, C.c_int, ip, dp, C.c_void_p, C.c_void_p]),
    "f25_write_jac": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, ip, C.c_void_p, C.c_int, ip]),
    "f25_gnmfree_dev_noalloc": (C.c_int, [C.c_void_p, C.c_int, ip, dp, C.c_void_p, C.c_void_p, C.c_void_p]),
    "f25_gn_jp": (C.c_int, [C.c_int, C.c_int, ip, C.c_int, ip, C.c_void_p, dp, C.c_void_p]),
    "f25_gn_jhz": (C.c_int, [C.c_int, C.c_int, ip, C.c_int, ip, dp, C.c_void_p, C.c_void_p, dp]),
    "f25_qp_nu_alpha": (C.c_int, [C.c_void_p, C.c_double, C.c_double, fp, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, fp]),
    "f25_met225": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                             fp, fp]),
    "f25_exresp25d": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, dp, dp, C.c_void_p, ip, C.c_void_p,
                                C.c_void_p]),
    "f25_exresp25d_dev": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, dp, dp, C.c_void_p, ip, C.c_void_p,
                                    C.c_void_p]),
    "f25_cadjlam_rhs": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, dp, dp, C.c_double, C.c_int,
                                  C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, dp, ip, C.c_void_p, C.c_void_p,
                                  fp, C.c_void_p, ip]),
    "f25_cadjlam_rhs_dev": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, dp, dp, C.c_double,
                                      C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, dp, ip, C.c_void_p,
                                      C.c_void_p, fp, C.c_void_p, ip]),
    "f25_cadjlam_rhs_covd": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, dp, dp, C.c_double,
                                       C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, dp, dp, ip, C.c_void_p,
                                       C.c_void_p, fp, C.c_void_p, ip]),
    "f25_cadjlam_rhs_covd_dev": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, dp, dp, C.c_double,
                                           C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, dp, dp, ip, C.c_void_p,
                                           C.c_void_p, fp, C.c_void_p, ip]),
    "f25_rec_space1d": (C.c_int, [C.c_int, dp, dp]),
    "f25_fill_uwave_dev": (C.c_int, [C.c_void_p, C.c_int, dp, dp, dp, dp, C.c_void_p]),
    "f25_peff_batch_dev": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_void_p, C.c_void_p]),
    "f25_bielak_rhs_dev": (C.c_int, [C.c_void_p, C.c_int, dp, dp, dp, dp, dp, dp, C.c_void_p, C.c_void_p,
                                     C.c_int]),
    "f25_gn_fv": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_void_p, dp, C.c_void_p]),
    "f25_gn_fhw": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_void_p, C.c_void_p, dp]),
    "f25_gn_diag": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_void_p, dp]),
    "f25_graph_build": (C.c_int, [C.POINTER(MeshDesc), C.c_int, C.POINTER(C.c_void_p), i64p]),
    "f25_graph_get": (C.c_int, [C.c_void_p, ip, ip, ip, ip, C.c_char_p]),
    "f25_graph_destroy": (C.c_int, [C.c_void_p]),
    "f25_last_error": (C.c_char_p, []),
    "f25_last_kernel_ms": (C.c_double, [C.c_char_p]),
    "f25_selftest_mfma": (C.c_int, [dp]),
    "f25_selftest_residency": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
}

_lib = None


def load():
    """Loads and types the library (raises OSError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built; run python -m fem25_amd.build")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class Fem25Error(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = load().f25_last_error().decode(errors="replace")
        super().__init__(f"{where}: error {code}: {msg}")
        self.code = code


def check(code: int, where: str) -> None:
    if code:
        raise Fem25Error(code, where)
