"""ctypes binding of libsumcheck_hip.so (C ABI in include/sumcheck_hip.h).

The library is the only compute path: if it is missing or cannot be loaded, every entry
point raises; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_NAME = "libsumcheck_hip.so"
LIB_PATH = Path(os.environ.get("SUMCHECK_HIP_LIB", Path(__file__).with_name(LIB_NAME)))

SC_OK = 0
SC_ERR_ARG = 1
SC_ERR_UNSUPPORTED = 2
SC_ERR_HIP = 3
SC_ERR_MEMORY = 4

SC_INPUT_HOST_CANONICAL = 0
SC_INPUT_HOST_MONT = 1
SC_INPUT_DEVICE_MONT = 2
SC_INPUT_AFFINE = 3
SC_INPUT_UNIFORM = 4
SC_INPUT_HOST_MONT_LIMB = 5
SC_INPUT_DEVICE_MONT_LIMB = 6

SC_CHALLENGE_FIXED = 0
SC_CHALLENGE_SHA3 = 1
SC_CHALLENGE_EXPLICIT = 2

CHALLENGE_MODES = {"fixed": SC_CHALLENGE_FIXED, "sha3": SC_CHALLENGE_SHA3, "explicit": SC_CHALLENGE_EXPLICIT}

u32p = ctypes.POINTER(ctypes.c_uint32)
i32p = ctypes.POINTER(ctypes.c_int32)
u64p = ctypes.POINTER(ctypes.c_uint64)


class ScComposition(ctypes.Structure):
    _fields_ = [
        ("num_tables", ctypes.c_int32),
        ("num_terms", ctypes.c_int32),
        ("term_lens", i32p),
        ("term_offsets", i32p),
        ("term_vars", i32p),
        ("coeffs", u32p),
    ]


class ScInput(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("instance", ctypes.c_uint32),
        ("data", ctypes.c_void_p),
        ("seed", ctypes.c_uint64),
        ("selector_mask", ctypes.c_uint64),
    ]


# (name, restype, argtypes) of every exported function; tests check the library exports
# exactly what include/sumcheck_hip.h declares.
SIGNATURES = [
    ("sc_last_error", ctypes.c_char_p, []),
    ("sc_version", ctypes.c_char_p, []),
    ("sc_context_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("sc_context_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("sc_context_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("sc_proof_create", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ScComposition), ctypes.c_int32,
      ctypes.POINTER(ScInput), ctypes.POINTER(ctypes.c_void_p)]),
    ("sc_proof_set_schedule", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32]),
    ("sc_proof_set_point_parallel", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    ("sc_proof_set_fold_chunk", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("sc_proof_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, u32p]),
    ("sc_proof_capture", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, u32p]),
    ("sc_proof_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("sc_proof_result", ctypes.c_int, [ctypes.c_void_p, u32p, u32p, u64p]),
    ("sc_proof_run_async", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, u32p, ctypes.POINTER(ctypes.c_int32)]),
    ("sc_proof_slot_result", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, u32p, u32p, u64p]),
    ("sc_proof_tables", ctypes.c_int, [ctypes.c_void_p, u32p]),
    ("sc_proof_info", ctypes.c_int, [ctypes.c_void_p, i32p, i32p, u64p, ctypes.POINTER(ctypes.c_int64)]),
    ("sc_proof_use_compiled", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("sc_proof_set_profiling", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("sc_proof_kernel_times", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    ("sc_proof_round_times", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int32,
      ctypes.POINTER(ctypes.c_float)]),
    ("sc_proof_final", ctypes.c_int, [ctypes.c_void_p, u32p, u32p]),
    ("sc_proof_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("sc_prove", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ScComposition), ctypes.c_int32,
      ctypes.POINTER(ScInput), ctypes.c_int32, u32p, u32p, u32p, u64p]),
    ("sc_prove_noalloc", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ScComposition), ctypes.c_int32,
      ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("sc_field_op", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, u32p, u32p, u32p]),
    ("sc_batch_create", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ScComposition), i32p,
      ctypes.POINTER(ScInput), ctypes.POINTER(ctypes.c_void_p)]),
    ("sc_batch_set_tile", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("sc_batch_set_point_parallel", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    ("sc_batch_set_schedules", ctypes.c_int,
     [ctypes.c_void_p, i32p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    ("sc_batch_instance_schedule", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, i32p, ctypes.POINTER(ctypes.c_int64), i32p]),
    ("sc_batch_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("sc_batch_capture", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("sc_batch_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("sc_batch_result", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, u32p, u32p, u64p]),
    ("sc_batch_info", ctypes.c_int, [ctypes.c_void_p, u64p, i32p, u64p]),
    ("sc_batch_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("sc_server_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    ("sc_server_submit", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ScComposition), ctypes.c_int32, ctypes.POINTER(ScInput),
      ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    ("sc_server_step", ctypes.c_int, [ctypes.c_void_p, i32p, i32p]),
    ("sc_server_try_step", ctypes.c_int, [ctypes.c_void_p, i32p, i32p, i32p]),
    ("sc_server_set_policy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("sc_server_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, i32p]),
    ("sc_server_set_request_schedule", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64]),
    ("sc_server_drain", ctypes.c_int, [ctypes.c_void_p]),
    ("sc_server_result", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, u32p, u32p, u64p]),
    ("sc_server_info", ctypes.c_int, [ctypes.c_void_p, u64p, u64p, u64p]),
    ("sc_server_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("sc_calibrate", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int32]),
    ("sc_transcript_challenge", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, u32p, ctypes.c_int32, u32p]),
]

_lib: ctypes.CDLL | None = None


class LibraryUnavailable(RuntimeError):
    """libsumcheck_hip.so is missing or failed to load."""


def load() -> ctypes.CDLL:
    """Load the library once, binding every signature. Raises LibraryUnavailable."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise LibraryUnavailable(
            f"{LIB_PATH} not found: build it with `make -C sumcheck-triton_amd` "
            "(or __graft_entry__.build())"
        )
    try:
        lib = ctypes.CDLL(str(LIB_PATH))
    except OSError as exc:
        raise LibraryUnavailable(f"cannot load {LIB_PATH}: {exc}") from exc
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


class SumcheckError(RuntimeError):
    pass


def check(rc: int) -> None:
    """Raise the Python exception matching a C ABI status code."""
    if rc == SC_OK:
        return
    msg = (load().sc_last_error() or b"").decode(errors="replace")
    if rc in (SC_ERR_ARG, SC_ERR_UNSUPPORTED):
        raise ValueError(msg)
    if rc == SC_ERR_MEMORY:
        raise MemoryError(msg)
    raise SumcheckError(f"HIP error: {msg}")


def as_u32p(arr) -> ctypes.POINTER(ctypes.c_uint32):
    return arr.ctypes.data_as(u32p) if arr is not None else None
