"""Batched compress (one launch per stage for every job of a mixed-shape
batch): every archive equals the oracle's canonical archive of its field byte
for byte (the single-field rule of cpu/src/codec.cpp:177-258 applied per job),
the spans table feeds the batched decoder directly, a captured graph replays
to the same bytes, and a failing job is reported by index while the others
stay valid."""
import numpy as np
import pytest

from fields import field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPES = [(32, 32, 32), (32, 64, 64), (48, 64, 96), (20, 17, 23), (8, 8, 8), (3, 40, 50), (33, 31, 129),
          (1, 1, 1), (16, 16, 16), (64, 128, 128)]
PATTERNS = ["smooth", "noise", "outliers", "ramp", "checker", "zero", "constant"]


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    from waltz_amd import batch
    waltz_amd.lib()
    return waltz_amd, batch


def _fields(shapes, dtype):
    return [field(s, PATTERNS[k % len(PATTERNS)], dtype, seed=k + 1) for k, s in enumerate(shapes)]


def _archive(archives, spans, j):
    off, size = int(spans[2 * j]), int(spans[2 * j + 1])
    return archives[off:off + size].cpu().numpy().tobytes()


@pytest.mark.parametrize("dtype,bound", [(np.float32, 1e-3), (np.float64, 1e-5), (np.float32, 0.05)],
                         ids=["f32-1e-3", "f64-1e-5", "f32-5e-2"])
def test_batch_compress_matches_oracle(wz, oracle, dtype, bound):
    _, batch = wz
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    hosts = _fields(SHAPES, dtype)
    plan = batch.BatchDecoder(SHAPES, tdt)
    flat = plan.flatten([torch.from_numpy(h).cuda() for h in hosts])
    archives, spans = plan.compress(flat, abs_error=bound)
    sp = spans.cpu().numpy()
    for j, (s, h) in enumerate(zip(SHAPES, hosts)):
        want = oracle.compress(h, bound)
        assert _archive(archives, sp, j) == want, (j, s)
    # the spans table is the batched decoder's input
    out = plan.decode(archives, spans)
    for j, (s, tile) in enumerate(zip(SHAPES, plan.views(out))):
        want = oracle.decompress(_archive(archives, sp, j), s, dtype)
        assert np.array_equal(tile.cpu().numpy().view(np.uint8), want.view(np.uint8)), (j, s)


def test_batch_compress_graph_replay(wz, oracle):
    _, batch = wz
    shapes = SHAPES[:6]
    hosts = _fields(shapes, np.float32)
    plan = batch.BatchDecoder(shapes, torch.float32)
    flat = plan.flatten([torch.from_numpy(h).cuda() for h in hosts])
    eager, spans = plan.compress(flat, abs_error=1e-3)
    eager = eager.clone()
    want_spans = spans.clone()
    archives = torch.zeros_like(eager)
    gspans = torch.zeros_like(spans)
    graph = torch.cuda.CUDAGraph()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        with torch.cuda.graph(graph, stream=stream):
            plan.compress(flat, abs_error=1e-3, archives=archives, spans=gspans, check=False)
    torch.cuda.current_stream().wait_stream(stream)
    for _ in range(3):
        archives.fill_(0xAB)
        gspans.fill_(-1)
        graph.replay()
        torch.cuda.synchronize()
        plan.compress_check()
        assert torch.equal(gspans, want_spans)
        sp = gspans.cpu().numpy()
        for j in range(len(shapes)):
            assert _archive(archives, sp, j) == _archive(eager, sp, j)
        _ = flat * 2.0  # eager work between the replays


def test_batch_compress_reports_failed_job(wz, oracle):
    waltz_amd, batch = wz
    shapes = [(16, 16, 16), (8, 12, 20), (10, 10, 10)]
    hosts = _fields(shapes, np.float32)
    hosts[1] = hosts[1].copy()
    hosts[1][3, 4, 5] = np.nan
    plan = batch.BatchDecoder(shapes, torch.float32)
    flat = plan.flatten([torch.from_numpy(h).cuda() for h in hosts])
    archives, spans = plan.compress(flat, abs_error=1e-3, check=False)
    with pytest.raises(waltz_amd.WaltzError, match="job 1: nonfinite input"):
        plan.compress_check()
    sp = spans.cpu().numpy()
    assert sp[3] == 0  # the failed job has no archive
    for j in (0, 2):
        assert _archive(archives, sp, j) == oracle.compress(hosts[j], 1e-3)
