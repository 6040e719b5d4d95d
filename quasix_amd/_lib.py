"""ctypes binding of libqxgw.so (the C ABI declared in include/qxgw.h).

The library is built in-tree (``python -m quasix_amd.build``) and loaded from this package
directory.  There is no fallback: if the library is missing or the GPU is absent, calls
raise, so a result can only come from the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libqxgw.so"

QX_OK = 0
STATUS_NAMES = {
    0: "QX_OK", 1: "QX_ERR_INVALID_INPUT", 2: "QX_ERR_DIMENSION_MISMATCH",
    3: "QX_ERR_NUMERICAL", 4: "QX_ERR_CONVERGENCE", 5: "QX_ERR_NOT_IMPLEMENTED",
    6: "QX_ERR_DEVICE", 7: "QX_ERR_ALLOC",
}
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2, "mixed": 2}
SOLVER = {"newton": 0, "linearized": 1}
PHASES = ["prep", "pi", "potrf", "trtri", "wnm", "newton", "residue", "spare"]

EXPORTED = [
    "qx_create", "qx_destroy", "qx_last_error", "qx_set_stream", "qx_synchronize",
    "qx_release_workspace", "qx_newton_config_default", "qx_g0w0_desc_default",
    "qx_gauss_legendre_grid", "qx_scaled_legendre_roots", "qx_rho_response_dev",
    "qx_screened_wc_dev", "qx_wmn_imag_dev", "qx_sigma_cd_dev", "qx_g0w0_cd_dev", "qx_g0w0_cd_ov_dev", "qx_g0w0_cd",
    "qx_set_profiling", "qx_get_profile", "qx_g0w0_adjoint_ov_dev", "qx_g0w0_batch_ov_dev",
    "qx_sigma_ac_dev", "qx_thiele", "qx_pade_thiele", "qx_ac_pade_thiele_diag", "qx_qp_solve_ac",
    "qx_exchange_diag_dev", "qx_df_mo_transform_dev", "qx_screening_w_dev",
    "qx_sigma_c_pyscf_cd_dev", "qx_solve_qp_linearized", "qx_g0w0_cd_ov_noalloc_dev",
    "qx_assemble_blocks_dev",
]


class NewtonConfig(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int), ("energy_tolerance", C.c_double),
        ("initial_damping", C.c_double), ("min_damping", C.c_double),
        ("max_damping", C.c_double), ("max_step", C.c_double),
        ("derivative_delta", C.c_double), ("use_richardson", C.c_int),
        ("z_min", C.c_double), ("z_max", C.c_double),
        ("use_bisection_fallback", C.c_int), ("max_bisection_iterations", C.c_int),
    ]


class G0W0Desc(C.Structure):
    _fields_ = [
        ("nocc", C.c_int), ("nvir", C.c_int), ("naux", C.c_int), ("nfreq", C.c_int),
        ("x0", C.c_double), ("eta", C.c_double), ("precision", C.c_int),
        ("ntarget", C.c_int), ("npair", C.c_int), ("newton", NewtonConfig),
        ("qp_solver", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("chol_fallbacks", C.c_int), ("singular_fallbacks", C.c_int),
        ("residue_targets", C.c_int), ("residue_jobs", C.c_int),
        ("newton_failures", C.c_int), ("seconds", C.c_double), ("pi_variant", C.c_int),
        ("waves", C.c_int), ("graph_replay", C.c_int), ("resolved_instances", C.c_int),
    ]


class BatchInstance(C.Structure):
    _fields_ = [
        ("nocc", C.c_int), ("nvir", C.c_int), ("naux", C.c_int), ("ntarget", C.c_int),
        ("mo_energy", C.c_void_p), ("B_ov", C.c_void_p), ("targets", C.c_void_p),
        ("static_terms", C.c_void_p),
    ]


class QxError(RuntimeError):
    """A non-OK qx_status; ``code`` holds the status name (QuasixError variant)."""

    def __init__(self, code: int, msg: str):
        self.code = STATUS_NAMES.get(code, str(code))
        super().__init__(f"{self.code}: {msg}")


_lib = None
_lock = threading.Lock()


def load():
    """Load libqxgw.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing; build it with `python -m quasix_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        vp, dp, ip = C.c_void_p, C.c_void_p, C.c_void_p
        lib.qx_create.argtypes = [C.POINTER(vp), C.c_int]
        lib.qx_destroy.argtypes = [vp]
        lib.qx_destroy.restype = None
        lib.qx_last_error.argtypes = [vp]
        lib.qx_last_error.restype = C.c_char_p
        lib.qx_set_stream.argtypes = [vp, vp]
        lib.qx_synchronize.argtypes = [vp]
        lib.qx_release_workspace.argtypes = [vp]
        lib.qx_newton_config_default.argtypes = [C.POINTER(NewtonConfig)]
        lib.qx_newton_config_default.restype = None
        lib.qx_g0w0_desc_default.argtypes = [C.POINTER(G0W0Desc)]
        lib.qx_g0w0_desc_default.restype = None
        lib.qx_gauss_legendre_grid.argtypes = [C.c_int, dp, dp]
        lib.qx_scaled_legendre_roots.argtypes = [C.c_int, C.c_double, dp, dp]
        lib.qx_rho_response_dev.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp,
                                            C.c_int, dp, dp]
        lib.qx_screened_wc_dev.argtypes = [vp, C.c_int, C.c_int, dp, dp, C.POINTER(C.c_int)]
        lib.qx_wmn_imag_dev.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, C.c_int,
                                        dp, C.c_int, dp, dp]
        lib.qx_sigma_cd_dev.argtypes = [vp, C.POINTER(G0W0Desc), dp, dp, dp, dp, ip, ip, ip, dp,
                                        dp, dp, C.c_int, dp, dp]
        lib.qx_g0w0_cd_dev.argtypes = [vp, C.POINTER(G0W0Desc), dp, dp, dp, dp, ip, ip, ip, dp,
                                       dp, dp, dp, dp, dp, ip, C.POINTER(Stats)]
        lib.qx_g0w0_cd.argtypes = lib.qx_g0w0_cd_dev.argtypes
        lib.qx_g0w0_cd_ov_dev.argtypes = [vp, C.POINTER(G0W0Desc), dp, dp, dp, ip, dp, dp, dp, dp,
                                          dp, dp, ip, C.POINTER(Stats)]
        lib.qx_g0w0_cd_ov_noalloc_dev.argtypes = lib.qx_g0w0_cd_ov_dev.argtypes
        lib.qx_assemble_blocks_dev.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp,
                                               C.c_int, ip, dp, dp, C.c_int, dp, ip, ip,
                                               C.POINTER(C.c_int)]
        lib.qx_g0w0_adjoint_ov_dev.argtypes = [vp, C.POINTER(G0W0Desc), dp, dp, ip, dp, dp, dp, dp]
        lib.qx_g0w0_batch_ov_dev.argtypes = [vp, C.POINTER(G0W0Desc), C.c_int,
                                             C.POINTER(BatchInstance), dp, dp, dp, dp, ip,
                                             C.POINTER(Stats)]
        lib.qx_sigma_ac_dev.argtypes = [vp, C.POINTER(G0W0Desc), dp, dp, dp, ip, ip, ip,
                                        C.c_double, dp, dp, C.POINTER(C.c_int)]
        lib.qx_thiele.argtypes = [C.c_int, dp, dp, dp]
        lib.qx_pade_thiele.argtypes = [C.c_int, C.c_double, dp, dp, dp]
        lib.qx_ac_pade_thiele_diag.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp,
                                               C.POINTER(C.c_int)]
        lib.qx_qp_solve_ac.argtypes = [C.c_int, C.c_int, dp, dp, ip, C.c_int, C.c_int, dp, dp, dp,
                                       C.c_int, dp, C.POINTER(NewtonConfig), dp, dp, dp, dp, ip]
        lib.qx_exchange_diag_dev.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
        lib.qx_df_mo_transform_dev.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp,
                                               dp, dp]
        lib.qx_screening_w_dev.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp, dp]
        lib.qx_sigma_c_pyscf_cd_dev.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, dp,
                                                dp, dp, dp, dp, dp]
        lib.qx_solve_qp_linearized.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp]
        lib.qx_set_profiling.argtypes = [vp, C.c_int]
        lib.qx_get_profile.argtypes = [vp, dp, dp]
        non_status = ("qx_destroy", "qx_last_error", "qx_newton_config_default",
                      "qx_g0w0_desc_default")
        for name in EXPORTED:
            if name not in non_status:
                getattr(lib, name).restype = C.c_int
        _lib = lib
        return lib


class Handle:
    """Owns one qx_handle (one per device and thread)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = C.c_void_p()
        rc = self.lib.qx_create(C.byref(self.h), device)
        if rc != QX_OK:
            raise QxError(rc, f"qx_create(device={device}) failed (no GPU or HIP runtime error)")
        self.device = device

    def check(self, rc: int):
        if rc != QX_OK:
            raise QxError(rc, self.lib.qx_last_error(self.h).decode(errors="replace"))

    def close(self):
        if self.h:
            self.lib.qx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_handles: dict = {}


def handle(device: int = 0) -> Handle:
    key = (threading.get_ident(), device)
    h = _handles.get(key)
    if h is None:
        h = _handles[key] = Handle(device)
    return h


def newton_config(**overrides) -> NewtonConfig:
    c = NewtonConfig()
    load().qx_newton_config_default(C.byref(c))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c
