"""Archive buffers too small for the archive: compress must fail with the
capacity error and write nothing at or past the stated capacity (the WZP
payloads go through per-chunk staging slots and the group pass checks each
group against the capacity; the section layout checks the rest). Then a
buffer of exactly the archive size must give the same bytes as the oracle."""
import ctypes as C

import numpy as np
import pytest

from fields import field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SENTINEL = 0xA5


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    from waltz_amd import _lib
    waltz_amd.lib()
    return waltz_amd, _lib


def compress_into(_lib, x, bound, capacity, extra=4096):
    nz, ny, nx = x.shape
    buf = torch.full((capacity + extra,), SENTINEL, dtype=torch.uint8, device=x.device)
    written = C.c_size_t()
    rc = _lib.lib().waltz_amd_compress(4 if x.dtype == torch.float32 else 8, x.data_ptr(), nx, ny, nz, bound,
                                       buf.data_ptr(), capacity, C.byref(written), None)
    torch.cuda.synchronize()
    err = _lib.lib().waltz_amd_last_error().decode() if rc else ""
    return rc, err, buf.cpu().numpy(), written.value


@pytest.mark.parametrize("shape", [(64, 64, 128), (40, 72, 100)], ids=["pairs", "bricks"])
def test_too_small_archive_buffer(wz, oracle, shape):
    waltz_amd, _lib = wz
    v = field(shape, "noise", np.float32, seed=11) * 4 + field(shape, "smooth", np.float32)
    x = torch.from_numpy(v).cuda()
    ref = oracle.compress(v, 1e-3)
    need = len(ref)
    header = 32 + 4 * ((len(range(0, v.size, 8192 * 16)))) + 3 * ((v.size + 8191) // 8192)
    for cap in (need - 1, (need + header) // 2, header + 16):
        rc, err, out, _ = compress_into(_lib, x, 1e-3, cap)
        assert rc != 0 and "capacity" in err, (cap, rc, err)
        assert (out[cap:] == SENTINEL).all(), cap
    rc, err, out, written = compress_into(_lib, x, 1e-3, need)
    assert rc == 0, err
    assert written == need and out[:need].tobytes() == ref
    assert (out[need:] == SENTINEL).all()
