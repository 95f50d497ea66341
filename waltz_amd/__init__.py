"""Waltz for MI355X: error-bounded lossy compression of 3-D float32/float64
fields with hand-written gfx950 HIP kernels (CDF 9/7 lifting, Morton-brick
quantization, WZP magnitude/sign coding, native repairs). Archives are the
canonical bytes of the reference CPU codec.

    import waltz_amd
    archive = waltz_amd.compress(field_on_gpu, abs_error=1e-3)
    field2 = waltz_amd.decompress(archive, shape=field_on_gpu.shape, dtype=field_on_gpu.dtype)
"""
from ._lib import WaltzError, lib  # noqa: F401
from .codec import compress, decompress, max_compressed_bytes  # noqa: F401

__all__ = ["compress", "decompress", "max_compressed_bytes", "WaltzError"]
