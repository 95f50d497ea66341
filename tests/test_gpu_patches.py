"""Side lists on decode: dense native repairs (whole 64-sample row blocks),
side lists out of order (valid: decoded again with duplicate checks) and
invalid entries. Archives are rewritten with other side lists and decoded
against the oracle, which restates cpu/src/codec.cpp:260-330. Each case runs
with the launch choice these small fields get by default (per-axis passes,
repairs patched by a separate pass) and with the finest inverse level forced
to its fused kernel (repairs applied as it stores the samples)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
TDT = {np.float32: torch.float32, np.float64: torch.float64}


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    waltz_amd.lib()  # fails loudly if the HIP library is missing
    return waltz_amd


def pad8(n):
    return (n + 7) & ~7


def sections(a, width):
    """Header fields and side lists of an archive (include/utils/Blob.hpp:38-79)."""
    wzp = int.from_bytes(a[16:24], "little")
    np_, nq = int.from_bytes(a[24:28], "little"), int.from_bytes(a[28:32], "little")
    pi = 32 + pad8(wzp)
    pv = pi + pad8(4 * np_)
    qi = pv + pad8(width * np_)
    qv = qi + pad8(4 * nq)
    ft = np.float32 if width == 4 else np.float64
    return dict(wzp=wzp, body=bytes(a[:32 + wzp]),
                pidx=np.frombuffer(a, np.uint32, np_, pi).copy(), perr=np.frombuffer(a, ft, np_, pv).copy(),
                qidx=np.frombuffer(a, np.uint32, nq, qi).copy(), qval=np.frombuffer(a, np.uint32, nq, qv).copy())


def rebuild(s, pidx, perr, qidx=None, qval=None):
    """The archive with the given side lists (header counts and padding updated)."""
    qidx = s["qidx"] if qidx is None else qidx
    qval = s["qval"] if qval is None else qval
    h = bytearray(s["body"])
    width = perr.dtype.itemsize
    h[4:8] = int(len(pidx) * width).to_bytes(4, "little")
    h[24:28] = int(len(pidx)).to_bytes(4, "little")
    h[28:32] = int(len(qidx)).to_bytes(4, "little")
    out = bytearray(h)

    def put(b):
        out.extend(b)
        out.extend(b"\0" * (pad8(len(b)) - len(b)))

    out.extend(b"\0" * (pad8(len(out) - 32) - (len(out) - 32)))
    put(np.asarray(pidx, np.uint32).tobytes())
    put(np.asarray(perr, perr.dtype).tobytes())
    put(np.asarray(qidx, np.uint32).tobytes())
    put(np.asarray(qval, np.uint32).tobytes())
    return bytes(out)


@pytest.fixture(params=["passes", "fused"])
def finest(request, wz):
    """How the finest inverse level runs for the shapes of the test."""
    from waltz_amd import stages
    forced = []

    def use(shape, dtype):
        if request.param == "fused":
            stages.set_tuning(shape, TDT[dtype], True, 1, (16, 16, 16, 16, 16))
            forced.append(shape)

    yield use
    if forced:
        stages.clear_tuning()


def decode(wz, raw, shape, dtype):
    a = torch.from_numpy(np.frombuffer(raw, np.uint8).copy()).cuda()
    return wz.decompress(a, shape=shape, dtype=TDT[dtype]).cpu().numpy()


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def base_archive(oracle, shape, dtype, spikes=False):
    rng = np.random.default_rng(5)
    z, y, x = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    v = np.sin(0.09 * x) * np.cos(0.05 * y + 0.02 * z) + 0.03 * rng.standard_normal(shape)
    if spikes:  # coefficients beyond 65535 quanta: q-outliers
        v.reshape(-1)[rng.integers(0, v.size, 40)] = rng.choice([-1.0, 1.0], 40) * 5e3
    return oracle.compress(v.astype(dtype), 1e-3)


def dense_list(s, n, nx, rng, dtype):
    """The archive's repairs plus a whole row, most of a row block, the
    field's last samples and scattered singles; sorted, unique."""
    extra = [np.arange(5 * nx, 6 * nx),                     # every sample of one row
             7 * nx + np.arange(0, min(nx, 64), 3),         # 22 of the first block of another
             np.arange(n - 20, n),                          # the field's tail
             rng.choice(n, n // 50, replace=False)]
    idx = np.unique(np.concatenate([s["pidx"].astype(np.int64)] + [e.astype(np.int64) for e in extra]))
    err = (rng.standard_normal(idx.size) * 1e-4).astype(dtype)
    have = np.isin(idx, s["pidx"])
    err[have] = s["perr"][np.searchsorted(s["pidx"], idx[have])]
    return idx.astype(np.uint32), err


SHAPES = [(64, 64, 64), (36, 40, 100), (48, 72, 136)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_dense_repairs(wz, oracle, finest, shape, dtype):
    finest(shape, dtype)
    ref = base_archive(oracle, shape, dtype)
    s = sections(ref, np.dtype(dtype).itemsize)
    n = int(np.prod(shape))
    idx, err = dense_list(s, n, shape[2], np.random.default_rng(1), dtype)
    raw = rebuild(s, idx, err)
    want = oracle.decompress(raw, shape, dtype)
    assert same_bits(decode(wz, raw, shape, dtype), want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_unordered_lists_decode(wz, oracle, finest, dtype):
    """Out-of-order side lists without duplicates are valid archives."""
    shape = (64, 64, 64)
    finest(shape, dtype)
    ref = base_archive(oracle, shape, dtype, spikes=True)
    s = sections(ref, np.dtype(dtype).itemsize)
    assert len(s["qidx"]) >= 2 and len(s["pidx"]) >= 2
    want = oracle.decompress(ref, shape, dtype)
    rng = np.random.default_rng(2)
    p = rng.permutation(len(s["pidx"]))
    raw = rebuild(s, s["pidx"][p], s["perr"][p])
    assert same_bits(decode(wz, raw, shape, dtype), want)
    q = rng.permutation(len(s["qidx"]))
    raw = rebuild(s, s["pidx"], s["perr"], s["qidx"][q], s["qval"][q])
    assert same_bits(decode(wz, raw, shape, dtype), want)
    # and canonical archives decode as before afterwards
    assert same_bits(decode(wz, ref, shape, dtype), want)


@pytest.mark.parametrize("which", ["repair", "qoutlier"])
def test_duplicate_index_rejected(wz, oracle, which):
    shape = (64, 64, 64)
    ref = base_archive(oracle, shape, np.float32, spikes=True)
    s = sections(ref, 4)
    if which == "repair":
        idx = np.concatenate([s["pidx"][:1], s["pidx"]])
        raw = rebuild(s, idx, np.concatenate([s["perr"][:1], s["perr"]]))
    else:
        raw = rebuild(s, s["pidx"], s["perr"], np.concatenate([s["qidx"][:1], s["qidx"]]),
                      np.concatenate([s["qval"][:1], s["qval"]]))
    with pytest.raises(RuntimeError):
        oracle.decompress(raw, shape, np.float32)
    with pytest.raises(RuntimeError, match="corrupt"):
        decode(wz, raw, shape, np.float32)


@pytest.mark.parametrize("where", ["early", "late", "index"])
def test_bad_repair_rejected(wz, oracle, finest, where):
    """A non-finite correction (early or late in a dense row block) or an
    index beyond the field fails the decode."""
    shape = (64, 64, 64)
    finest(shape, np.float32)
    ref = base_archive(oracle, shape, np.float32)
    s = sections(ref, 4)
    idx, err = dense_list(s, int(np.prod(shape)), shape[2], np.random.default_rng(3), np.float32)
    row = np.flatnonzero(idx >= 5 * 64)[0]  # the full row: entries row .. row+63
    if where == "early":
        err[row + 2] = np.inf
    elif where == "late":
        err[row + 40] = np.nan
    else:
        idx[-1] = np.prod(shape)
    raw = rebuild(s, idx, err)
    with pytest.raises(RuntimeError):
        oracle.decompress(raw, shape, np.float32)
    with pytest.raises(RuntimeError, match="corrupt"):
        decode(wz, raw, shape, np.float32)
