"""ORACLE — test infrastructure only.

NumPy-facing ctypes bindings for:
  * ``liboracle.so``: this project's C restatement of the reference CPU scalar
    codec (``oracle/waltz_oracle.c``);
  * ``_ref/libwaltz_cpu_ref.so``: the reference CPU codec compiled from its own
    sources by ``oracle/Makefile`` (present only when it was built).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module, and only to check results against. The product never imports it.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_P, _SZ, _U32, _U = C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint


def _load(path):
    return C.CDLL(os.fspath(path)) if Path(path).exists() else None


def build():
    """Compile liboracle.so (and the reference lib when its sources exist)."""
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", os.fspath(HERE)], check=True)  # -B: always recompile
    if Path("/root/reference/cpu/src/codec.cpp").exists():
        subprocess.run(["make", "-s", "-B", "-C", os.fspath(HERE), "ref"], check=True)


class OracleError(RuntimeError):
    pass


class Oracle:
    """The C restatement (always required)."""

    def __init__(self):
        lib = _load(HERE / "liboracle.so")
        if lib is None:
            raise OracleError("oracle/liboracle.so is missing; run make -C oracle")
        self.lib = lib
        sig = {
            "wo_last_error": (C.c_char_p, []),
            "wo_dwt_levels": (_U, [_U32]),
            "wo_zband": (_U32, [_U32, _U32, _U32, _U, C.c_int]),
            "wo_morton_rank": (_U32, [_U32] * 6),
            "wo_transform_f32": (C.c_int, [_P, _U32, _U32, _U32, _U, _U, C.c_int, C.c_int]),
            "wo_transform_f64": (C.c_int, [_P, _U32, _U32, _U32, _U, _U, C.c_int, C.c_int]),
            "wo_quantize_f32": (C.c_int, [_P, _U32, _U32, _U32, _U32, C.c_double, _P, _P, _P, _P, _P]),
            "wo_quantize_f64": (C.c_int, [_P, _U32, _U32, _U32, _U32, C.c_double, _P, _P, _P, _P, _P]),
            "wo_dequantize_f32": (None, [_P, _P, _U32, _U32, _U32, _U32, C.c_double, _P]),
            "wo_dequantize_f64": (None, [_P, _P, _U32, _U32, _U32, _U32, C.c_double, _P]),
            "wo_repairs_f32": (C.c_int, [_P, _P, _U32, C.c_double, _P, _P, _P]),
            "wo_repairs_f64": (C.c_int, [_P, _P, _U32, C.c_double, _P, _P, _P]),
            "wo_within_bound": (C.c_int, [C.c_double, C.c_double, C.c_double]),
            "wo_wzp_max_bytes": (_SZ, [_U32]),
            "wo_wzp_encode": (C.c_int64, [_P, _P, _U32, _P, _SZ]),
            "wo_wzp_decode": (C.c_int, [_P, _SZ, _U32, _P, _P]),
            "wo_max_compressed_bytes": (_SZ, [_U32, _U32, _U32, C.c_int]),
            "wo_compress_f32": (C.c_int, [_P, _U32, _U32, _U32, C.c_double, _P, _SZ, C.POINTER(_SZ)]),
            "wo_compress_f64": (C.c_int, [_P, _U32, _U32, _U32, C.c_double, _P, _SZ, C.POINTER(_SZ)]),
            "wo_decompress_f32": (C.c_int, [_P, _SZ, _U32, _U32, _U32, _P]),
            "wo_decompress_f64": (C.c_int, [_P, _SZ, _U32, _U32, _U32, _P]),
            "wo_fnv1a64": (C.c_uint64, [_P, _SZ]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args

    def _err(self, what):
        raise OracleError(f"{what}: {self.lib.wo_last_error().decode()}")

    @staticmethod
    def _sfx(dtype):
        return "f32" if np.dtype(dtype) == np.float32 else "f64"

    def fnv1a64(self, array):
        a = np.ascontiguousarray(array)
        return "%016x" % self.lib.wo_fnv1a64(a.ctypes.data, a.nbytes)

    # geometry -------------------------------------------------------------
    def levels(self, n):
        return self.lib.wo_dwt_levels(n)

    def geometry(self, nx, ny, nz):
        lx, lz = self.levels(min(nx, ny)), self.levels(nz)
        dyadic = lx == lz
        return lx, lz, dyadic, self.lib.wo_zband(nx, ny, nz, lz, int(dyadic))

    def morton_rank(self, x, y, z, nx, ny, nz):
        return self.lib.wo_morton_rank(x, y, z, nx, ny, nz)

    # stages ---------------------------------------------------------------
    def transform(self, volume, lxy, lz, dyadic, inverse):
        """volume: (nz, ny, nx) array; returns a transformed copy."""
        a = np.ascontiguousarray(volume).copy()
        nz, ny, nx = a.shape
        fn = getattr(self.lib, "wo_transform_" + self._sfx(a.dtype))
        if fn(a.ctypes.data, nx, ny, nz, lxy, lz, int(dyadic), int(inverse)):
            self._err("transform")
        return a

    def quantize(self, coeff, bz, quantum):
        """Returns (dequantized, mag[n] u16, sign[ceil(n/32)] u32, qidx, qval)."""
        c = np.ascontiguousarray(coeff).copy()
        nz, ny, nx = c.shape
        n = c.size
        mag = np.zeros(n, np.uint16)
        sign = np.zeros((n + 31) // 32, np.uint32)
        qidx = np.zeros(max(n, 1), np.uint32)
        qval = np.zeros(max(n, 1), np.uint32)
        nq = _U32()
        fn = getattr(self.lib, "wo_quantize_" + self._sfx(c.dtype))
        if fn(c.ctypes.data, nx, ny, nz, bz, quantum, mag.ctypes.data, sign.ctypes.data,
              qidx.ctypes.data, qval.ctypes.data, C.byref(nq)):
            self._err("quantize")
        return c, mag, sign, qidx[:nq.value].copy(), qval[:nq.value].copy()

    def dequantize(self, mag, sign, shape, bz, quantum, dtype):
        out = np.empty(shape, dtype)
        nz, ny, nx = shape
        fn = getattr(self.lib, "wo_dequantize_" + self._sfx(dtype))
        fn(np.ascontiguousarray(mag, np.uint16).ctypes.data, np.ascontiguousarray(sign, np.uint32).ctypes.data,
           nx, ny, nz, bz, quantum, out.ctypes.data)
        return out

    def repairs(self, original, reconstructed, bound):
        o = np.ascontiguousarray(original).ravel()
        r = np.ascontiguousarray(reconstructed, dtype=o.dtype).ravel()
        idx = np.zeros(max(o.size, 1), np.uint32)
        err = np.zeros(max(o.size, 1), o.dtype)
        cnt = _U32()
        fn = getattr(self.lib, "wo_repairs_" + self._sfx(o.dtype))
        if fn(o.ctypes.data, r.ctypes.data, o.size, bound, idx.ctypes.data, err.ctypes.data, C.byref(cnt)):
            self._err("repairs")
        return idx[:cnt.value].copy(), err[:cnt.value].copy()

    def within_bound(self, value, original, bound):
        return bool(self.lib.wo_within_bound(value, original, bound))

    def wzp_encode(self, mag, sign):
        mag = np.ascontiguousarray(mag, np.uint16)
        sign = np.ascontiguousarray(sign, np.uint32)
        cap = self.lib.wo_wzp_max_bytes(mag.size)
        out = np.zeros(max(cap, 1), np.uint8)
        r = self.lib.wo_wzp_encode(mag.ctypes.data, sign.ctypes.data, mag.size, out.ctypes.data, cap)
        if r < 0:
            self._err("wzp_encode")
        return out[:r].tobytes()

    def wzp_decode(self, data, n):
        src = np.frombuffer(bytes(data), np.uint8)
        mag = np.zeros(max(n, 1), np.uint16)
        sign = np.zeros(max((n + 31) // 32, 1), np.uint32)
        if self.lib.wo_wzp_decode(src.ctypes.data, src.size, n, mag.ctypes.data, sign.ctypes.data):
            self._err("wzp_decode")
        return mag[:n], sign[:(n + 31) // 32]

    # codec ----------------------------------------------------------------
    def max_compressed_bytes(self, shape, dtype):
        nz, ny, nx = shape
        return self.lib.wo_max_compressed_bytes(nx, ny, nz, np.dtype(dtype).itemsize)

    def compress(self, volume, abs_bound):
        a = np.ascontiguousarray(volume)
        nz, ny, nx = a.shape
        cap = self.max_compressed_bytes(a.shape, a.dtype)
        out = np.zeros(cap, np.uint8)
        written = _SZ()
        fn = getattr(self.lib, "wo_compress_" + self._sfx(a.dtype))
        if fn(a.ctypes.data, nx, ny, nz, abs_bound, out.ctypes.data, cap, C.byref(written)):
            self._err("compress")
        return out[:written.value].tobytes()

    def decompress(self, archive, shape, dtype):
        src = np.frombuffer(bytes(archive), np.uint8)
        out = np.empty(shape, dtype)
        nz, ny, nx = shape
        fn = getattr(self.lib, "wo_decompress_" + self._sfx(dtype))
        if fn(src.ctypes.data, src.size, nx, ny, nz, out.ctypes.data):
            self._err("decompress")
        return out


class _RefConfig(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
                ("type", C.c_int), ("absolute_error", C.c_double), ("backend", C.c_int)]


class Reference:
    """The reference CPU codec built by ``make -C oracle ref`` (optional)."""

    SCALAR, AVX2 = 1, 2

    def __init__(self):
        lib = _load(HERE / "_ref" / "libwaltz_cpu_ref.so")
        if lib is None:
            raise OracleError("oracle/_ref/libwaltz_cpu_ref.so is not built")
        self.lib = lib
        cfg = C.POINTER(_RefConfig)
        for name, res, args in [
            ("waltz_cpu_create", _P, []), ("waltz_cpu_destroy", None, [_P]),
            ("waltz_cpu_last_error", C.c_char_p, [_P]), ("waltz_cpu_has_avx2", C.c_int, []),
            ("waltz_cpu_max_compressed_bytes", _SZ, [cfg]),
            ("waltz_cpu_compress", C.c_int, [_P, cfg, _P, _P, _SZ, C.POINTER(_SZ)]),
            ("waltz_cpu_decompress", C.c_int, [_P, cfg, _P, _SZ, _P, _SZ]),
            ("ref_workspace_elements", _SZ, [_U32, _U32, _U32]),
            ("ref_transform", C.c_int, [C.c_int, C.c_int, _P, _P, _U32, _U32, _U32, _U, _U, C.c_int, C.c_int]),
            ("ref_wzp_max_bytes", _SZ, [_U32]),
            ("ref_wzp_encode", _SZ, [_P, _P, _U32, _P, _SZ, C.c_int]),
            ("ref_wzp_decode", C.c_int, [_P, _SZ, _U32, _P, _P, C.c_int]),
        ]:
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        self.ws = lib.waltz_cpu_create()
        self.has_avx2 = bool(lib.waltz_cpu_has_avx2())

    def __del__(self):
        try:
            self.lib.waltz_cpu_destroy(self.ws)
        except Exception:
            pass

    def _cfg(self, shape, dtype, bound, backend):
        nz, ny, nx = shape
        return _RefConfig(nx, ny, nz, np.dtype(dtype).itemsize, bound, backend)

    def compress(self, volume, abs_bound, backend=SCALAR):
        a = np.ascontiguousarray(volume)
        cfg = self._cfg(a.shape, a.dtype, abs_bound, backend)
        cap = self.lib.waltz_cpu_max_compressed_bytes(C.byref(cfg))
        out = np.zeros(cap, np.uint8)
        written = _SZ()
        if self.lib.waltz_cpu_compress(self.ws, C.byref(cfg), a.ctypes.data, out.ctypes.data, cap, C.byref(written)):
            raise OracleError(self.lib.waltz_cpu_last_error(self.ws).decode())
        return out[:written.value].tobytes()

    def decompress(self, archive, shape, dtype, backend=SCALAR):
        src = np.frombuffer(bytes(archive), np.uint8)
        out = np.empty(shape, dtype)
        cfg = self._cfg(shape, dtype, 0.0, backend)
        if self.lib.waltz_cpu_decompress(self.ws, C.byref(cfg), src.ctypes.data, src.size, out.ctypes.data, out.nbytes):
            raise OracleError(self.lib.waltz_cpu_last_error(self.ws).decode())
        return out

    def transform(self, volume, lxy, lz, dyadic, inverse, avx2=False):
        a = np.ascontiguousarray(volume)
        nz, ny, nx = a.shape
        n = self.lib.ref_workspace_elements(nx, ny, nz)
        data = np.zeros(n, a.dtype)
        data[:a.size] = a.ravel()
        scratch = np.zeros(n, a.dtype)
        if self.lib.ref_transform(a.dtype.itemsize, int(inverse), data.ctypes.data, scratch.ctypes.data,
                                  nx, ny, nz, lxy, lz, int(dyadic), int(avx2)):
            raise OracleError("reference transform failed")
        return data[:a.size].reshape(a.shape)

    def wzp_encode(self, mag, sign, avx2=False):
        mag = np.ascontiguousarray(mag, np.uint16)
        sign = np.ascontiguousarray(sign, np.uint32)
        cap = self.lib.ref_wzp_max_bytes(mag.size)
        out = np.zeros(max(cap, 1), np.uint8)
        r = self.lib.ref_wzp_encode(mag.ctypes.data, sign.ctypes.data, mag.size, out.ctypes.data, cap, int(avx2))
        if r == 2**64 - 1:
            raise OracleError("reference wzp_encode failed")
        return out[:r].tobytes()


def reference_or_none():
    try:
        return Reference()
    except OracleError:
        return None
