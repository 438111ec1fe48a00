"""ctypes binding of the gfx950 C-ABI library ``libtenet_amd.so`` (declared in
``include/tenet_amd.h``).

The library is loaded lazily on the first device call.  There is no fallback: if it is
missing, or no HIP device is present, every compute entry point raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

__all__ = ["lib", "check", "stream_ptr", "LIB_PATH", "TN_F64", "TN_F32", "PRECISIONS"]

LIB_PATH = Path(__file__).resolve().parent / "libtenet_amd.so"

TN_F64, TN_F32 = 0, 1
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2}
JOB_WORDS = 24
MAX_RANK = 6
PROB_WORDS = 8
SEG_WORDS = 6
MPO_JOB_WORDS = 8
MPO_TILE = 1024  # positions per tn_mpo_apply tile (256 threads x 4)
LQ_PROB_WORDS = 6
SVD_PROB_WORDS = 11
SVD_MAX_W = 64
SVD_MAX_K = 4096

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_d = ctypes.c_double

_SIGNATURES = {
    "tn_last_error": (ctypes.c_char_p, []),
    "tn_version": (_i32, []),
    "tn_device_cus": (_i32, [ctypes.POINTER(_i32)]),
    "tn_recouple": (_i32, [_i32, _p, _p, _p, _p, _p, _p, _i64, _p]),
    "tn_recouple_cast": (_i32, [_i32, _i32, _p, _p, _p, _p, _p, _p, _i64, _p]),
    "tn_recouple_plan": (_i32, [_i64, _i64, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _i64, _i64,
                                _p, _p, _p]),
    "tn_gemm_grouped": (_i32, [_i32, _i32, _i32, _i32, _p, _p, _p, _p, _p, _p, _p, _i64, _i32, _i32, _i32, _p]),
    "tn_gemm_grouped_static": (_i32, [_i32, _i32, _i32, _i32, _p, _p, _p, _p, _p, _p, _i64, _i32, _i32, _i32,
                                      _i64, _p, _p, _p, _p, _i64, _p]),
    "tn_static_scratch_bytes": (_i64, []),
    "tn_gemm_grouped_rows": (_i32, [_i32, _i32, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _i32, _i32, _p]),
    "tn_gemm_skinny": (_i32, [_i32, _i32, _i32, _i32, _p, _p, _p, _p, _p, _p, _p, _i64, _p]),
    "tn_dot": (_i32, [_i32, _p, _p, _p, _i64, _p, _p, _p]),
    "tn_krylov_orth": (_i32, [_i32, _i64, _i32, _p, _p, _p, _i32, _p, _p, _p, _p, _p]),
    "tn_krylov_orth_work": (_i64, []),
    "tn_lq": (_i32, [_p, _i64, _i64, _p, _p, _p, _p]),
    "tn_axpby": (_i32, [_i32, _i64, _d, _p, _i32, _p, _d, _p, _i32, _p, _p]),
    "tn_scale": (_i32, [_i32, _i64, _d, _p, _i32, _p, _p]),
    "tn_cast": (_i32, [_i32, _p, _i32, _p, _i64, _p]),
    "tn_mpo_apply": (_i32, [_i32, _p, _p, _p, _i64, _p, _p, _p, _p, _i64, _p]),
    "tn_svd_jacobi_norms": (_i32, [_p, _i64, _p, _p, _p]),
    "tn_svd_jacobi_round": (_i32, [_p, _p, _i64, _i64, _p, _p, _p, _i32, _i32, _i32, _d, _p, _p, _d, _p]),
    "tn_svd_jacobi_converged": (_i32, [_p, _i64, _p, _p]),
    "tn_svd_jacobi_finish": (_i32, [_p, _i64, _p, _p, _p, _p, _p, _p, _d, _i32, _p]),
    "tn_tridiag_ground": (_i32, [_p, _p, _i32, _d, _p, _p, _p]),
    "tn_tridiag_ground_host": (_i32, [_p, _p, _i32, _d, _p, _p]),
    "tn_tridiag_ground_batched": (_i32, [_p, _p, _i32, _i64, _i64, _d, _p, _p, _p]),
    "tn_dot_segmented": (_i32, [_i32, _p, _p, _p, _p, _i64, _p, _p, _i64, _p, _i64, _p]),
    "tn_axpby_segmented": (_i32, [_i32, _p, _i64, _d, _p, _i64, _i32, _p, _d, _p, _i64, _i32, _p, _p]),
}
TRIDIAG_MAX = 64

_lib = None


def load_library(path: Path | None = None) -> ctypes.CDLL:
    """Open the shared library and declare every exported signature (no device needed)."""
    path = Path(path or os.environ.get("TENET_AMD_LIB", LIB_PATH))
    if not path.exists():
        raise RuntimeError(
            f"the gfx950 library {path} is not built; run `python tenet-py_amd/build.py` "
            "(there is no CPU fallback for tenet's contraction engine)"
        )
    cdll = ctypes.CDLL(str(path))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(cdll, name)
        fn.restype = res
        fn.argtypes = args
    return cdll


def lib() -> ctypes.CDLL:
    """The loaded library; raises when it is missing or no GPU is visible."""
    global _lib
    if _lib is None:
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError(
                "tenet's contraction engine runs only on a HIP device (MI355X / gfx950); "
                "no GPU is visible to torch"
            )
        _lib = load_library()
    return _lib


_host = None


def host_lib() -> ctypes.CDLL:
    """The loaded library for its host-side entry points (plan compilation), which need no
    device; raises when the library is missing."""
    global _host
    if _lib is not None:
        return _lib
    if _host is None:
        _host = load_library()
    return _host


def check_host(rc: int, exc=ValueError) -> None:
    """Raise ``exc`` with the library's message when a host-side entry point failed."""
    if rc != 0:
        raise exc(host_lib().tn_last_error().decode(errors="replace"))


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().tn_last_error().decode(errors="replace")
        raise RuntimeError(f"tenet HIP kernel error: {msg}")


def stream_ptr() -> int:
    """The raw ``hipStream_t`` of torch's current stream on the current device (the
    direct query; ``torch.cuda.current_stream()`` builds a Python Stream object per call)."""
    import torch

    return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())


def loaded() -> bool:
    return _lib is not None
