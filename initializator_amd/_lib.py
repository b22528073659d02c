"""ctypes binding of libnmr_hip.so (the C ABI in include/nmr_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
initializator_amd/csrc``).  Loading fails loudly when it is missing: there is
no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_uint8, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnmr_hip.so")

NMR_OK = 0
NMR_ERR_INVALID = -1
NMR_ERR_NULL = -2
NMR_ERR_DEVICE = -3
NMR_ERR_NOMEM = -4
NMR_ERR_STATE = -5
NMR_ERR_UNSUPPORTED = -6

NMR_F_DIFFUSION = 1
NMR_F_VELOCITY = 2
NMR_F_PRECESSION = 4
NMR_F_ADVECT_MASK = 8

NMR_LAYOUT_COMPACT = 0
NMR_LAYOUT_PADDED = 1

NMR_SCHEME_UPWIND = 0
NMR_SCHEME_MACCORMACK = 1

NMR_BC_NEUMANN = 0
NMR_BC_PERIODIC = 1
NMR_BC_DIRICHLET = 2

NMR_CONV_ZMARCH = 0
NMR_CONV_TILES = 1
NMR_CONV_ROWS = 2
NMR_CONV_ZMARCH_SPARSE = 3

NMR_PHASE_LINEAR = 0
NMR_PHASE_EXACT = 1

# Functions every build of the library must export (the declarations of
# include/nmr_hip.h).
EXPORTS = (
    "nmr_hip_create",
    "nmr_hip_destroy",
    "nmr_hip_upload_mask",
    "nmr_hip_upload_velocity",
    "nmr_hip_upload_magnetization",
    "nmr_hip_download_magnetization",
    "nmr_hip_upload_coefficients",
    "nmr_hip_upload_field",
    "nmr_hip_upload_phase_weights",
    "nmr_hip_run",
    "nmr_hip_signal",
    "nmr_hip_copy_bandwidth",
    "nmr_hip_mfma_rate",
    "nmr_hip_batch_set_layout",
    "nmr_hip_run_static",
    "nmr_hip_pulse",
    "nmr_hip_device_count",
    "nmr_hip_set_device",
    "nmr_hip_last_run_ms",
    "nmr_hip_set_tile",
    "nmr_hip_set_layout",
    "nmr_hip_num_variants",
    "nmr_hip_variant_info",
    "nmr_hip_device_bytes",
    "nmr_hip_batch_create",
    "nmr_hip_batch_destroy",
    "nmr_hip_batch_upload",
    "nmr_hip_batch_download",
    "nmr_hip_batch_run",
    "nmr_hip_batch_set_tile",
    "nmr_hip_batch_autotune",
    "nmr_hip_batch_last_run_ms",
    "nmr_hip_corrector_create",
    "nmr_hip_corrector_destroy",
    "nmr_hip_corrector_num_params",
    "nmr_hip_corrector_get_params",
    "nmr_hip_corrector_set_params",
    "nmr_hip_corrector_set_scales",
    "nmr_hip_corrector_get_scales",
    "nmr_hip_corrector_set_conv",
    "nmr_hip_corrector_forward",
    "nmr_hip_corrector_backward",
    "nmr_hip_corrector_train",
    "nmr_hip_run_learned",
    "nmr_hip_strerror",
    "nmr_hip_last_error",
    "nmr_hip_abi_version",
)

NMR_CORRECTOR_PARAMS = 10402


class NmrProblemC(ctypes.Structure):
    """Mirror of ``nmr_problem`` in include/nmr_hip.h."""

    _fields_ = [
        ("nx", c_int32),
        ("ny", c_int32),
        ("nz", c_int32),
        ("ncomp", c_int32),
        ("flags", c_int32),
        ("scheme", c_int32),
        ("phase_mode", c_int32),
        ("buffer", c_int32 * 3),
        ("bc", c_int32 * 3),
        ("reserved", c_int32),
        ("dx", c_double),
        ("dy", c_double),
        ("dz", c_double),
        ("dt", c_double),
        ("diffusivity", c_double),
        ("r1_bulk", c_double),
        ("r2_bulk", c_double),
        ("rho1", c_double),
        ("rho2", c_double),
        ("gamma", c_double),
        ("p_eq", c_double),
        ("b0_dir", c_double * 3),
        ("grad", c_double * 3),
    ]


class NmrBatchInstanceC(ctypes.Structure):
    """Mirror of ``nmr_batch_instance`` in include/nmr_hip.h."""

    _fields_ = [
        ("nx", c_int32),
        ("ny", c_int32),
        ("nz", c_int32),
        ("reserved", c_int32),
        ("rho1", c_double),
        ("rho2", c_double),
        ("grad_amp", c_double),
        ("delta", c_double),
        ("big_delta", c_double),
    ]


class NmrTrainOptsC(ctypes.Structure):
    """Mirror of ``nmr_train_opts`` in include/nmr_hip.h."""

    _fields_ = [
        ("factor", c_int32),
        ("rounds", c_int32),
        ("stride", c_int32),
        ("reserved", c_int32),
        ("lr", c_double),
    ]


class NmrError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, code: int, where: str, detail: str):
        super().__init__(f"{where}: {code} ({_strerror(code)}): {detail}")
        self.code = code


_lib = None


def _strerror(code: int) -> str:
    if _lib is None:
        return "?"
    return _lib.nmr_hip_strerror(code).decode()


def load(path: str | None = None) -> ctypes.CDLL:
    """Load libnmr_hip.so once and declare its signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"{p} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(p)
    vp = c_void_p
    dp = POINTER(c_double)
    sig = {
        "nmr_hip_create": (c_int, [POINTER(NmrProblemC), POINTER(vp)]),
        "nmr_hip_destroy": (c_int, [vp]),
        "nmr_hip_upload_mask": (c_int, [vp, POINTER(c_uint8), c_int32]),
        "nmr_hip_upload_velocity": (c_int, [vp, dp, dp, dp, c_int32]),
        "nmr_hip_upload_magnetization": (c_int, [vp, dp, dp, dp, c_int32]),
        "nmr_hip_download_magnetization": (c_int, [vp, dp, dp, dp, c_int32]),
        "nmr_hip_upload_coefficients": (c_int, [vp, dp, dp, dp, c_int32]),
        "nmr_hip_upload_field": (c_int, [vp, dp, c_int32]),
        "nmr_hip_upload_phase_weights": (c_int, [vp, dp, dp, c_int32]),
        # the hot call takes plain addresses (HipSolver.run passes integers)
        "nmr_hip_run": (c_int, [vp, c_int64, vp, vp]),
        "nmr_hip_signal": (c_int, [vp, dp]),
        "nmr_hip_copy_bandwidth": (c_int, [c_int64, c_int32, dp]),
        "nmr_hip_mfma_rate": (c_int, [c_int64, c_int32, dp]),
        "nmr_hip_batch_set_layout": (c_int, [vp, c_int32]),
        "nmr_hip_run_static": (c_int, [vp, c_int64, dp, dp]),
        "nmr_hip_pulse": (c_int, [vp, c_int32, c_double, c_double, c_double, c_double]),
        "nmr_hip_device_count": (c_int, [POINTER(c_int32)]),
        "nmr_hip_set_device": (c_int, [c_int32]),
        "nmr_hip_last_run_ms": (c_int, [vp, dp, POINTER(c_int64)]),
        "nmr_hip_set_tile": (c_int, [vp, c_int32, c_int32]),
        "nmr_hip_set_layout": (c_int, [vp, c_int32]),
        "nmr_hip_num_variants": (c_int, []),
        "nmr_hip_variant_info": (
            c_int,
            [c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)],
        ),
        "nmr_hip_device_bytes": (c_int, [vp, POINTER(c_int64)]),
        "nmr_hip_batch_create": (c_int, [POINTER(NmrProblemC), c_int32, POINTER(NmrBatchInstanceC), POINTER(vp)]),
        "nmr_hip_batch_destroy": (c_int, [vp]),
        "nmr_hip_batch_upload": (c_int, [vp, c_int32, POINTER(c_uint8), dp, dp, dp, dp, dp, dp]),
        "nmr_hip_batch_download": (c_int, [vp, c_int32, dp, dp, dp]),
        "nmr_hip_batch_run": (c_int, [vp, c_int64, dp]),
        "nmr_hip_batch_set_tile": (c_int, [vp, c_int32, c_int32, c_int32, c_int32]),
        "nmr_hip_batch_autotune": (c_int, [vp, c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
        "nmr_hip_batch_last_run_ms": (c_int, [vp, dp, POINTER(c_int64)]),
        "nmr_hip_corrector_create": (c_int, [ctypes.c_uint64, POINTER(vp)]),
        "nmr_hip_corrector_destroy": (c_int, [vp]),
        "nmr_hip_corrector_num_params": (c_int, []),
        "nmr_hip_corrector_get_params": (c_int, [vp, POINTER(ctypes.c_float)]),
        "nmr_hip_corrector_set_params": (c_int, [vp, POINTER(ctypes.c_float)]),
        "nmr_hip_corrector_set_scales": (c_int, [vp, c_double, c_double]),
        "nmr_hip_corrector_get_scales": (c_int, [vp, dp, dp]),
        "nmr_hip_corrector_set_conv": (c_int, [vp, c_int32, c_int32]),
        "nmr_hip_corrector_forward": (c_int, [vp, c_int32, c_int32, c_int32, POINTER(ctypes.c_float),
                                              POINTER(ctypes.c_float)]),
        "nmr_hip_corrector_backward": (c_int, [vp, c_int32, c_int32, c_int32, POINTER(ctypes.c_float),
                                               POINTER(ctypes.c_float), POINTER(ctypes.c_float)]),
        "nmr_hip_corrector_train": (c_int, [vp, vp, POINTER(NmrTrainOptsC), dp]),
        "nmr_hip_run_learned": (c_int, [vp, vp, c_int64, c_int32, dp, c_double, dp, POINTER(c_int64)]),
        "nmr_hip_strerror": (c_char_p, [c_int]),
        "nmr_hip_last_error": (c_char_p, []),
        "nmr_hip_abi_version": (c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code: int, where: str) -> None:
    """Raise NmrError for a non-zero status."""
    if code != NMR_OK:
        lib = load()
        raise NmrError(code, where, lib.nmr_hip_last_error().decode())
