"""Single pipeline stages on device tensors (for callers that drive one stage,
and for stage-level parity tests). Shapes are (z, y, x)."""
import ctypes as C

import torch

from ._lib import check, lib
from .codec import _stream

_W = {torch.float32: 4, torch.float64: 8}


def _dims(t):
    nz, ny, nx = t.shape
    return nx, ny, nz


def _scratch(volume):
    # two lifting scratch volumes from torch's stream-ordered allocator: freed
    # here, they are reused only by work queued after this call's kernels
    return torch.empty_like(volume), torch.empty_like(volume)


def transform(volume, levels_xy, levels_z, dyadic, inverse, out=None):
    """CDF 9/7 lifting (forward: samples -> coefficients; inverse: back),
    enqueued on the current stream without synchronising."""
    volume = volume.contiguous()
    out = torch.empty_like(volume) if out is None else out
    nx, ny, nz = _dims(volume)
    t0, t1 = _scratch(volume)
    check(lib().waltz_amd_transform_async(_W[volume.dtype], int(inverse), volume.data_ptr(), out.data_ptr(),
                                          t0.data_ptr(), t1.data_ptr(), nx, ny, nz, levels_xy, levels_z, int(dyadic),
                                          _stream(volume.device)), "transform")
    return out


def transform_adjoint(volume, levels_xy, levels_z, dyadic, of_inverse, out=None):
    """Transpose of transform() with the same geometry: of_inverse=False maps a
    coefficient gradient to samples (DWT^T), True maps a sample gradient to
    coefficients (IDWT^T). Enqueued on the current stream without synchronising."""
    volume = volume.contiguous()
    out = torch.empty_like(volume) if out is None else out
    nx, ny, nz = _dims(volume)
    t0, t1 = _scratch(volume)
    check(lib().waltz_amd_transform_adjoint_async(_W[volume.dtype], int(of_inverse), volume.data_ptr(),
                                                  out.data_ptr(), t0.data_ptr(), t1.data_ptr(), nx, ny, nz, levels_xy,
                                                  levels_z, int(dyadic), _stream(volume.device)),
          "transform_adjoint")
    return out


def quantize(coeff, bz, quantum):
    """Returns (dequantized, mag u16 as int16 view, sign words as int32, q indices, q values).

    q-outliers come back in device order (unsorted); the archive path orders them.
    """
    c = coeff.contiguous().clone()
    n = c.numel()
    nx, ny, nz = _dims(c)
    dev = c.device
    mag = torch.empty(n, dtype=torch.int16, device=dev)
    sign = torch.empty((n + 31) // 32, dtype=torch.int32, device=dev)
    qidx = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    qval = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    nq = C.c_uint32()
    check(lib().waltz_amd_quantize(_W[c.dtype], c.data_ptr(), nx, ny, nz, bz, float(quantum), mag.data_ptr(),
                                   sign.data_ptr(), qidx.data_ptr(), qval.data_ptr(), C.byref(nq), _stream(dev)),
          "quantize")
    return c, mag, sign, qidx[:nq.value], qval[:nq.value]


def wzp_encode(mag, sign):
    """WZP stream of packed u16 magnitudes (int16 tensor) and sign words (int32)."""
    n = mag.numel()
    chunks = (n + 8191) // 8192
    cap = ((chunks + 15) // 16) * 4 + chunks * 3 + 2 * n + (n + 7) // 8
    out = torch.zeros(max(cap, 1), dtype=torch.uint8, device=mag.device)
    written = C.c_size_t()
    check(lib().waltz_amd_wzp_encode(mag.data_ptr(), sign.data_ptr(), n, out.data_ptr(), cap, C.byref(written),
                                     _stream(mag.device)), "wzp_encode")
    return out[:written.value]


def wzp_decode(stream, n):
    dev = stream.device
    mag = torch.empty(max(n, 1), dtype=torch.int16, device=dev)
    sign = torch.zeros(max((n + 31) // 32, 1), dtype=torch.int32, device=dev)
    check(lib().waltz_amd_wzp_decode(stream.data_ptr(), stream.numel(), n, mag.data_ptr(), sign.data_ptr(),
                                     _stream(dev)), "wzp_decode")
    return mag[:n], sign[:(n + 31) // 32]


def autotune(shape, dtype, device=None):
    """Time the fused lifting kernels' bit-neutral launch choices for this
    (z, y, x) shape and dtype on the device and keep the fastest."""
    nz, ny, nx = shape
    dev = torch.device(device or "cuda")
    check(lib().waltz_amd_autotune(_W[dtype], nx, ny, nz, _stream(dev)), "autotune")


def set_tuning(shape, dtype, inverse, fused_levels, z_segments):
    """Fix the fused level count and the five z segments for one direction."""
    nz, ny, nx = shape
    seg = (C.c_uint32 * 5)(*[int(v) for v in z_segments])
    check(lib().waltz_amd_set_tuning(_W[dtype], nx, ny, nz, int(inverse), int(fused_levels), seg), "set_tuning")


def set_pass_tuning(shape, dtype, inverse, tiles, lds_pair_max):
    """Fix the per-axis passes' X/Y/Z line tiles and LDS pair limit for one direction."""
    nz, ny, nx = shape
    t = (C.c_uint32 * 3)(*[int(v) for v in tiles])
    check(lib().waltz_amd_set_pass_tuning(_W[dtype], nx, ny, nz, int(inverse), t, int(lds_pair_max)),
          "set_pass_tuning")


def clear_tuning():
    lib().waltz_amd_clear_tuning()


def tuning_report():
    import json
    cap = 1 << 16
    buf = C.create_string_buffer(cap)
    lib().waltz_amd_tuning_report(buf, cap)
    return json.loads(buf.value.decode())
