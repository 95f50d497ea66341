"""ctypes binding of libwaltz.so (the gfx950 HIP library).

There is no CPU fallback: if the library is missing or fails to load, every
call raises. Build it with ``python -m waltz_amd.build``.
"""
import ctypes as C
import os
from functools import lru_cache
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libwaltz.so"
_P, _SZ, _U32, _U = C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint


class WaltzError(RuntimeError):
    pass


class _HipConfig(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32), ("width", C.c_int), ("bound", C.c_double)]


class BatchShape(C.Structure):
    _fields_ = [(name, C.c_uint32) for name in ("x", "y", "z", "fx", "fy", "fz", "samples", "sides", "width")]


_SIGS = {
    "waltz_amd_last_error": (C.c_char_p, []),
    "waltz_amd_version": (C.c_char_p, []),
    "waltz_amd_max_compressed_bytes": (_SZ, [_U32, _U32, _U32, C.c_int]),
    "waltz_amd_compress": (C.c_int, [C.c_int, _P, _U32, _U32, _U32, C.c_double, _P, _SZ, C.POINTER(_SZ), _P]),
    "waltz_amd_decompress": (C.c_int, [C.c_int, _P, _SZ, _U32, _U32, _U32, _P, _P]),
    "waltz_amd_transform": (C.c_int, [C.c_int, C.c_int, _P, _P, _U32, _U32, _U32, _U, _U, C.c_int, _P]),
    "waltz_amd_quantize": (C.c_int, [C.c_int, _P, _U32, _U32, _U32, _U32, C.c_double, _P, _P, _P, _P,
                                     C.POINTER(_U32), _P]),
    "waltz_amd_wzp_encode": (C.c_int, [_P, _P, _U32, _P, _SZ, C.POINTER(_SZ), _P]),
    "waltz_amd_wzp_decode": (C.c_int, [_P, _SZ, _U32, _P, _P, _P]),
    "waltz_hip_create": (_P, [C.c_int]),
    "waltz_hip_destroy": (None, [_P]),
    "waltz_hip_error": (C.c_char_p, [_P]),
    "waltz_hip_max_bytes": (_SZ, [C.POINTER(_HipConfig)]),
    "waltz_hip_encode": (C.c_int, [_P, C.POINTER(_HipConfig), _P, _P, _SZ, C.POINTER(_SZ)]),
    "waltz_hip_decode": (C.c_int, [_P, C.POINTER(_HipConfig), _P, _SZ, _P, _SZ]),
    "waltz_cuda_decode_chunk": (C.c_char_p, [_U32, _U32, _U32, C.c_int, _P, _SZ, _P, _SZ, _P]),
    "waltz_amd_profile_enable": (C.c_int, [C.c_int]),
    "waltz_amd_profile_report": (_SZ, [C.c_char_p, _SZ]),
}
_SIGS.update({
    "waltz_amd_batch_create": (_P, [C.c_int, _U32, C.POINTER(_U32)]),
    "waltz_amd_batch_error": (C.c_char_p, []),
    "waltz_amd_batch_destroy": (None, [_P]),
    "waltz_amd_batch_workspace_bytes": (_SZ, [_P]),
    "waltz_amd_batch_output_elements": (_SZ, [_P]),
    "waltz_amd_batch_decode": (C.c_int, [_P, _P, _SZ, _P, _P, _P, _SZ, _P]),
    "waltz_amd_batch_check": (C.c_int, [_P, _P]),
    "waltz_amd_batch_compress_workspace_bytes": (_SZ, [_P]),
    "waltz_amd_batch_archive_capacity": (_SZ, [_P]),
    "waltz_amd_batch_compress": (C.c_int, [_P, _P, C.c_double, _P, _SZ, _P, _P, _SZ, _P]),
    "waltz_amd_batch_compress_check": (C.c_int, [_P, _P]),
    "waltz_amd_transform_adjoint": (C.c_int, [C.c_int, C.c_int, _P, _P, _U32, _U32, _U32, _U, _U, C.c_int, _P]),
    "waltz_amd_transform_async": (C.c_int, [C.c_int, C.c_int, _P, _P, _P, _P, _U32, _U32, _U32, _U, _U, C.c_int,
                                            _P]),
    "waltz_amd_transform_adjoint_async": (C.c_int, [C.c_int, C.c_int, _P, _P, _P, _P, _U32, _U32, _U32, _U, _U,
                                                    C.c_int, _P]),
    "waltz_amd_autotune": (C.c_int, [C.c_int, _U32, _U32, _U32, _P]),
    "waltz_amd_set_tuning": (C.c_int, [C.c_int, _U32, _U32, _U32, C.c_int, _U, C.POINTER(_U32)]),
    "waltz_amd_clear_tuning": (None, []),
    "waltz_amd_set_pass_tuning": (C.c_int, [C.c_int, _U32, _U32, _U32, C.c_int, C.POINTER(_U32), _U32]),
    "waltz_amd_tuning_report": (_SZ, [C.c_char_p, _SZ]),
})
OPTIONAL = {
    "waltz_gpu_decode_batch": (C.c_char_p, [C.POINTER(BatchShape), _P, _P, _SZ, _P, _SZ, _P, _SZ, C.c_double, _P]),
}


@lru_cache(maxsize=1)
def lib():
    path = os.environ.get("WALTZ_AMD_LIBRARY") or str(LIB_PATH)
    if not Path(path).exists():
        raise WaltzError(f"{path} is not built; run `python -m waltz_amd.build` (no CPU fallback exists)")
    handle = C.CDLL(path)
    for name, (res, args) in {**_SIGS, **OPTIONAL}.items():
        fn = getattr(handle, name, None)
        if fn is None:
            if name in OPTIONAL:
                continue
            raise WaltzError(f"{path} lacks {name}")
        fn.restype, fn.argtypes = res, args
    return handle


def profile(enable):
    """Turn per-stage HIP-event timing on/off; returns the previous state."""
    return bool(lib().waltz_amd_profile_enable(int(enable)))


def profile_report():
    """{stage: (total_ms, launches)} since the last report (then reset)."""
    import json
    cap = 1 << 16
    buf = C.create_string_buffer(cap)
    need = lib().waltz_amd_profile_report(buf, cap)  # the call also resets the totals
    if need >= cap:
        raise WaltzError("profile report truncated")
    return {k: tuple(v) for k, v in json.loads(buf.value.decode()).items()}


def check(rc, what):
    if rc:
        raise WaltzError(f"{what}: {lib().waltz_amd_last_error().decode(errors='replace')}")
