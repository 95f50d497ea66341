"""waltz_cuda_decode_chunk_noalloc (fixed scratch inside the library, no
device allocation) against waltz_cuda_decode_chunk on the same archives: two
calls in a row, the second with a smaller chunk, both equal bit for bit (and
to the oracle); a chunk past the fixed scratch's limit is rejected with a
diagnostic."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from fields import field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

LIB = Path(__file__).resolve().parents[1] / "waltz_amd" / "libwaltz.so"
MAX_SAMPLES = 1048576


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    h = C.CDLL(str(LIB))
    p, sz = C.c_void_p, C.c_size_t
    for name in ("waltz_cuda_decode_chunk", "waltz_cuda_decode_chunk_noalloc"):
        fn = getattr(h, name)
        fn.restype, fn.argtypes = C.c_char_p, [C.c_uint32] * 3 + [C.c_int, p, sz, p, sz, p]
    return h


def _decode(fn, archive, shape, dtype, stream):
    nz, ny, nx = shape
    width = np.dtype(dtype).itemsize
    out = torch.full(shape, float("nan"), dtype=torch.float32 if width == 4 else torch.float64, device="cuda")
    stream.wait_stream(torch.cuda.current_stream())
    err = fn(nx, ny, nz, width, archive.data_ptr(), archive.numel(), out.data_ptr(), out.numel() * width,
             stream.cuda_stream)
    assert err is None, err
    stream.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_noalloc_matches_chunk_decoder(lib, oracle, dtype):
    stream = torch.cuda.Stream()
    # the largest chunk the fixed scratch takes (paired-brick geometry), then a smaller generic one
    for shape, pattern, bound in [((64, 128, 128), "noise", 1e-3), ((24, 40, 56), "outliers", 1e-2)]:
        a = oracle.compress(field(shape, pattern, dtype), bound)
        arc = torch.from_numpy(np.frombuffer(a, np.uint8).copy()).cuda()
        print(f"noalloc decode {shape} {np.dtype(dtype).name}", flush=True)
        got = _decode(lib.waltz_cuda_decode_chunk_noalloc, arc, shape, dtype, stream)
        ref = _decode(lib.waltz_cuda_decode_chunk, arc, shape, dtype, stream)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), shape
        assert np.array_equal(got.view(np.uint8), oracle.decompress(a, shape, dtype).view(np.uint8)), shape


def test_noalloc_rejects_large_and_bad_chunks(lib, oracle):
    shape = (65, 128, 128)  # one plane past the limit
    assert shape[0] * shape[1] * shape[2] > MAX_SAMPLES
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    arc = torch.zeros(64, dtype=torch.uint8, device="cuda")
    err = lib.waltz_cuda_decode_chunk_noalloc(128, 128, 65, 4, arc.data_ptr(), 64, out.data_ptr(), out.numel() * 4,
                                              None)
    assert err is not None and b"fixed decode scratch" in err
    # the chunk decoder's own argument rule applies too
    a = oracle.compress(field((8, 8, 8), "smooth", np.float32), 1e-3)
    arc = torch.from_numpy(np.frombuffer(a, np.uint8).copy()).cuda()
    small = torch.empty((8, 8, 8), dtype=torch.float32, device="cuda")
    err = lib.waltz_cuda_decode_chunk_noalloc(8, 8, 8, 4, arc.data_ptr(), len(a), small.data_ptr(),
                                              small.numel() * 4 + 4, None)
    assert err is not None and b"invalid chunk dimensions" in err
    err = lib.waltz_cuda_decode_chunk_noalloc(8, 8, 8, 4, arc.data_ptr(), len(a), small.data_ptr(),
                                              small.numel() * 4, None)
    assert err is None, err
    want = oracle.decompress(a, (8, 8, 8), np.float32)
    assert np.array_equal(small.cpu().numpy().view(np.uint32), want.view(np.uint32))
