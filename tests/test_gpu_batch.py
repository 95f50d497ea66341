"""Batched decode parity: every tile of a batch equals the oracle's decode of
the same archive bit for bit, through the mixed-shape batch ABI (eager and
replayed from a captured graph) and the reference's frame ABI
waltz_gpu_decode_batch (clipping at the frame edge, holes filled), following
python/tests/test_gpu_loader.py:104-125 of the reference (batch == per-chunk
CPU decode placed into the frame)."""
import ctypes as C

import numpy as np
import pytest

from fields import field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPES = [(32, 32, 32), (32, 64, 64), (48, 64, 96), (20, 17, 23), (8, 8, 8), (3, 40, 50), (33, 31, 129)]


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    from waltz_amd import batch
    waltz_amd.lib()
    return waltz_amd, batch


def _archives(waltz_amd, shapes, dtype, bound):
    pats = ["smooth", "noise", "outliers", "smooth"]
    fields, archives = [], []
    for k, s in enumerate(shapes):
        v = torch.from_numpy(field(s, pats[k % len(pats)], dtype)).cuda()
        fields.append(v)
        archives.append(waltz_amd.compress(v, abs_error=bound).clone())
    return fields, archives


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_mixed_batch_matches_oracle(wz, oracle, dtype):
    waltz_amd, batch = wz
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    _, archives = _archives(waltz_amd, SHAPES, dtype, 1e-3)
    packed, spans = batch.pack(archives)
    dec = batch.BatchDecoder(SHAPES, tdt)
    flat = dec.decode(packed, spans)
    for s, a, tile in zip(SHAPES, archives, dec.views(flat)):
        want = oracle.decompress(a.cpu().numpy().tobytes(), s, dtype)
        assert np.array_equal(tile.cpu().numpy().view(np.uint8), want.view(np.uint8)), s
    again = dec.decode(packed, spans)
    assert torch.equal(again, flat)


def test_graph_replay_matches_eager(wz):
    waltz_amd, batch = wz
    shapes = SHAPES[:4] * 3
    _, archives = _archives(waltz_amd, shapes, np.float32, 1e-3)
    packed, spans = batch.pack(archives)
    dec = batch.BatchDecoder(shapes, torch.float32)
    eager = dec.decode(packed, spans)
    out = torch.empty_like(eager)
    dec.decode(packed, spans, out=out)  # warm the plan before capture
    g = torch.cuda.CUDAGraph()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        with torch.cuda.graph(g, stream=stream):
            dec.decode(packed, spans, out=out, check=False)
    torch.cuda.current_stream().wait_stream(stream)
    for _ in range(3):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        dec.check()
        assert torch.equal(out, eager)
        _ = eager.sum()  # eager work between replays


def test_corrupt_archive_in_batch_is_reported(wz, oracle):
    waltz_amd, batch = wz
    shapes = SHAPES[:3]
    _, archives = _archives(waltz_amd, shapes, np.float32, 1e-3)
    bad = archives[1].clone()
    bad[1] = 7  # impossible level count
    packed, spans = batch.pack([archives[0], bad, archives[2]])
    dec = batch.BatchDecoder(shapes, torch.float32)
    with pytest.raises(Exception, match="corrupt"):
        dec.decode(packed, spans)
    packed, spans = batch.pack(archives)
    flat = dec.decode(packed, spans)  # usable afterwards
    want = oracle.decompress(archives[0].cpu().numpy().tobytes(), shapes[0], np.float32)
    assert np.array_equal(dec.views(flat)[0].cpu().numpy().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("width", [4, 8])
def test_reference_frame_abi(wz, oracle, width):
    """waltz_gpu_decode_batch: 16x16x8 tiles in 40x30x20 frames (clipped at the
    frame edge), two samples, two sides, a few holes filled with -7."""
    waltz_amd, batch = wz
    from waltz_amd._lib import BatchShape, lib
    dtype = np.float32 if width == 4 else np.float64
    tdt = torch.float32 if width == 4 else torch.float64
    tx, ty, tz, fx, fy, fz, samples, sides = 16, 16, 8, 40, 30, 20, 2, 2
    gx, gy, gz = -(-fx // tx), -(-fy // ty), -(-fz // tz)
    jobs = gx * gy * gz * samples * sides
    rng = np.random.default_rng(4)
    holes = set(rng.choice(jobs, 5, replace=False).tolist())
    archives, blobs = [], []
    for j in range(jobs):
        if j in holes:
            archives.append(None)
            continue
        v = torch.from_numpy(field((tz, ty, tx), ["smooth", "noise"][j % 2], dtype, seed=j)).cuda()
        a = waltz_amd.compress(v, abs_error=1e-3)
        archives.append(a)
    # packed layout: 16-byte aligned archives, then (offset, size) pairs
    body, table = bytearray(), []
    for a in archives:
        if a is None:
            table += [0, 0]
            continue
        raw = a.cpu().numpy().tobytes()
        table += [len(body), len(raw)]
        body += raw + bytes((-len(raw)) % 16)
    host = bytes(body) + np.array(table, dtype="<u8").tobytes()
    host_arr = np.frombuffer(host, np.uint8).copy()
    packed = torch.from_numpy(host_arr).cuda()
    shape = BatchShape(tx, ty, tz, fx, fy, fz, samples, sides, width)
    scratch_bytes = lib().waltz_gpu_batch_scratch_bytes(C.byref(shape))
    ref_formula = 3 * ((tx * ty * tz + 3) & ~3) * jobs * width + 8 * 256 * 32 * 4 + 16
    assert 0 < scratch_bytes <= ref_formula  # the reference loader's own scratch size suffices
    scratch = torch.empty(ref_formula, dtype=torch.uint8, device="cuda")
    out = torch.empty((sides, samples, fz, fy, fx), dtype=tdt, device="cuda")
    err = lib().waltz_gpu_decode_batch(C.byref(shape), host_arr.ctypes.data, packed.data_ptr(), packed.numel(),
                                       scratch.data_ptr(), scratch.numel(), out.data_ptr(), out.numel() * width,
                                       -7.0, torch.cuda.current_stream().cuda_stream)
    assert err is None, err
    want = np.empty((sides, samples, fz, fy, fx), dtype)
    for j, a in enumerate(archives):
        r = j
        ox = (r % gx) * tx
        r //= gx
        oy = (r % gy) * ty
        r //= gy
        oz = (r % gz) * tz
        r //= gz
        side, sample = r % sides, r // sides
        tile = (np.full((tz, ty, tx), -7.0, dtype) if a is None else
                oracle.decompress(a.cpu().numpy().tobytes(), (tz, ty, tx), dtype))
        ez, ey, ex = min(tz, fz - oz), min(ty, fy - oy), min(tx, fx - ox)
        want[side, sample, oz:oz + ez, oy:oy + ey, ox:ox + ex] = tile[:ez, :ey, :ex]
    assert np.array_equal(out.cpu().numpy().view(np.uint8), want.view(np.uint8))
