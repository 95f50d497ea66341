"""Device-tensor codec API (mirrors the reference's pywaltz.compress /
decompress, python/pywaltz/codec.py:19-98, for torch tensors on the GPU).

Shapes are NumPy order: (z, y, x) or (y, x); x varies fastest. Archives are
the canonical bytes of the reference CPU codec, as uint8 device tensors.
All work runs on the current torch stream of the tensor's device.
"""
import ctypes as C
import math

import torch

from ._lib import WaltzError, check, lib

_DTYPES = {torch.float32: 4, torch.float64: 8}


def geometry(shape, dtype):
    """(nx, ny, nz), width for a (z,y,x) or (y,x) shape (python/pywaltz/codec.py:24-33)."""
    shape = tuple(int(v) for v in shape)
    if dtype not in _DTYPES:
        raise TypeError("Waltz requires float32 or float64")
    if len(shape) not in (2, 3) or min(shape) <= 0:
        raise ValueError("expected positive (y,x) or (z,y,x) dimensions")
    if math.prod(shape) * _DTYPES[dtype] > 0xFFFFFFFF:
        raise ValueError("field exceeds the native Waltz archive limit")
    return (*reversed(shape), 1)[:3], _DTYPES[dtype]


def _bound(value):
    value = float(value)
    if not math.isfinite(value) or value <= 0:
        raise ValueError("abs_error must be finite and positive")
    return value


def _stream(device):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def max_compressed_bytes(shape, dtype):
    (nx, ny, nz), width = geometry(shape, dtype)
    cap = lib().waltz_amd_max_compressed_bytes(nx, ny, nz, width)
    if not cap:
        raise ValueError("unsupported Waltz configuration")
    return cap


def compress(field, *, abs_error, out=None):
    """Return the canonical archive of a device tensor as a uint8 device tensor.

    ``out`` (optional) is a preallocated uint8 device buffer of at least
    max_compressed_bytes(); the result is then a view of it.
    """
    if not field.is_cuda:
        raise ValueError("compress expects a GPU tensor")
    (nx, ny, nz), width = geometry(field.shape, field.dtype)
    field = field.contiguous()
    cap = lib().waltz_amd_max_compressed_bytes(nx, ny, nz, width)
    if out is None:
        out = torch.empty(cap, dtype=torch.uint8, device=field.device)
    elif out.numel() < cap or out.dtype != torch.uint8 or out.device != field.device:
        raise ValueError("out must be a uint8 tensor on the field's device of at least max_compressed_bytes()")
    written = C.c_size_t()
    check(lib().waltz_amd_compress(width, field.data_ptr(), nx, ny, nz, _bound(abs_error), out.data_ptr(),
                                   out.numel(), C.byref(written), _stream(field.device)), "compress")
    return out[:written.value]


def decompress(archive, *, shape, dtype, out=None):
    """Decode a device archive (uint8 tensor) into a new (or given) tensor."""
    if not archive.is_cuda or archive.dtype != torch.uint8:
        raise ValueError("decompress expects a uint8 GPU tensor")
    (nx, ny, nz), width = geometry(shape, dtype)
    archive = archive.contiguous()
    if archive.numel() < 32:
        raise ValueError("truncated Waltz header")
    if out is None:
        out = torch.empty(tuple(shape), dtype=dtype, device=archive.device)
    elif tuple(out.shape) != tuple(shape) or out.dtype != dtype or not out.is_contiguous():
        raise ValueError("out must be a contiguous tensor of the given shape and dtype")
    check(lib().waltz_amd_decompress(width, archive.data_ptr(), archive.numel(), nx, ny, nz, out.data_ptr(),
                                     _stream(archive.device)), "decompress")
    return out


__all__ = ["compress", "decompress", "max_compressed_bytes", "geometry", "WaltzError"]
