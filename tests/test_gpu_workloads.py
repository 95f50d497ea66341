"""The benchmark workloads, checked in full against the oracle on the GPU.

- 512^3 float32 k^-5/3 Gaussian random field (+3.0): at REL 1e-3 and ABS 1e-4
  the archive equals the oracle's byte for byte, the decoded samples equal the
  oracle's decode bit for bit, and max |x - x^| <= the bound in FP64
  (cpu/tests/interop_gpu.cu:216-246 of the reference: the CPU codec is canonical).
- 256 float32 chunks of uneven shape (seeded mix of 32^3, 64x64x32, 96x64x48,
  128x128x64; each its own k^-5/3 field), ABS 1e-3: every archive equals the
  oracle's (one call per chunk and one batched compress of all of them), and every tile of the batched decode (eager, and replayed from a
  captured graph between eager work) equals the oracle's decode of its archive
  (python/tests/test_gpu_loader.py:104-125: batch == per-chunk CPU decode).
- 8 x 256^3 float32 fields, REL 1e-3: the transpose identity of the inverse
  transform and its adjoint at full size in FP64, and FP64 central finite
  differences of MSE(IDWT(c), target) on a 16^3 crop at one level per axis.

Fields come from waltz_amd.synthetic with the seeds bench.py uses.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    from waltz_amd import synthetic
    waltz_amd.lib()
    return waltz_amd, synthetic


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def grf512(wz):
    _, synthetic = wz
    x = synthetic.grf((512, 512, 512), slope=-5.0 / 3.0, seed=20261203, offset=3.0)
    return x, x.cpu().numpy()


@pytest.mark.parametrize("mode", ["rel1e-3", "abs1e-4"])
def test_grf_512_f32_matches_oracle(wz, oracle, grf512, mode):
    waltz_amd, _ = wz
    x, host = grf512
    bound = 1e-3 * (float(host.max()) - float(host.min())) if mode == "rel1e-3" else 1e-4
    _say(f"512^3 f32 GRF {mode}: abs bound {bound:.6g}")
    archive = waltz_amd.compress(x, abs_error=bound)
    got = archive.cpu().numpy().tobytes()
    want = oracle.compress(host, bound)
    _say(f"archive {len(got)} B (oracle {len(want)} B)")
    assert got == want
    out = waltz_amd.decompress(archive, shape=x.shape, dtype=torch.float32)
    err = float((out.double() - x.double()).abs().max())
    assert err <= bound, (err, bound)
    dec = oracle.decompress(want, host.shape, np.float32)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), dec.view(np.uint32))


def _batch_shapes():
    rng = np.random.default_rng(20261202 + 4)
    kinds = [(32, 32, 32), (32, 64, 64), (48, 64, 96), (64, 128, 128)]  # (z, y, x)
    return [kinds[k] for k in rng.integers(0, len(kinds), 256)]


def test_batched_256_chunks_match_oracle(wz, oracle):
    waltz_amd, synthetic = wz
    from waltz_amd import batch
    shapes = _batch_shapes()
    assert {s for s in shapes} == {(32, 32, 32), (32, 64, 64), (48, 64, 96), (64, 128, 128)}
    archives, hosts, fields = [], [], []
    for j, s in enumerate(shapes):
        f = synthetic.grf(s, slope=-5.0 / 3.0, seed=20261202 + 4 + 1000 * (j + 1))
        a = waltz_amd.compress(f, abs_error=1e-3).clone()
        h = a.cpu().numpy().tobytes()
        assert h == oracle.compress(f.cpu().numpy(), 1e-3), (j, s)
        archives.append(a)
        hosts.append(h)
        fields.append(f)
    _say(f"{len(shapes)} archives byte-identical to the oracle")
    # the same 256 fields through the batched compress: the same archives
    plan = batch.BatchDecoder(shapes, torch.float32)
    barch, bspans = plan.compress(plan.flatten(fields), abs_error=1e-3)
    sp = bspans.cpu().numpy()
    for j, h in enumerate(hosts):
        got = barch[int(sp[2 * j]):int(sp[2 * j]) + int(sp[2 * j + 1])].cpu().numpy().tobytes()
        assert got == h, (j, shapes[j])
    _say("batched compress: 256 archives byte-identical to the oracle")
    del barch
    wants = [oracle.decompress(h, s, np.float32) for h, s in zip(hosts, shapes)]
    packed, spans = batch.pack(archives)
    dec = batch.BatchDecoder(shapes, torch.float32)
    flat = dec.decode(packed, spans)

    def check(flat, what):
        for j, (tile, want) in enumerate(zip(dec.views(flat), wants)):
            assert np.array_equal(tile.cpu().numpy().view(np.uint32), want.view(np.uint32)), (what, j, shapes[j])

    check(flat, "eager")
    out = torch.full((dec.numel,), float("nan"), device="cuda")
    graph = torch.cuda.CUDAGraph()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        with torch.cuda.graph(graph, stream=stream):
            dec.decode(packed, spans, out=out, check=False)
    torch.cuda.current_stream().wait_stream(stream)
    for r in range(3):
        out.fill_(float("nan"))
        graph.replay()
        torch.cuda.synchronize()
        dec.check()
        check(out, f"replay {r}")
        dec.decode(packed, spans)  # eager work between the replays
    _say("batched decode: eager and 3 graph replays equal the oracle")


@pytest.fixture(scope="module")
def grf256x8(wz):
    _, synthetic = wz
    return [synthetic.grf((256, 256, 256), slope=-5.0 / 3.0, seed=20261202 + 5 + 1000 * (b + 1)) for b in range(8)]


def test_autograd_transpose_identity_full_size(wz, grf256x8):
    from waltz_amd import autograd, stages
    lxy, lz, dyadic = autograd.default_levels((256, 256, 256))
    g = torch.Generator(device="cuda").manual_seed(20261207)
    for b, f in enumerate(grf256x8):
        c = stages.transform(f.double(), lxy, lz, dyadic, inverse=False)  # coefficients of the field
        y = torch.randn(f.shape, dtype=torch.float64, device="cuda", generator=g)
        lhs = float((stages.transform(c, lxy, lz, dyadic, inverse=True) * y).sum())
        rhs = float((c * stages.transform_adjoint(y, lxy, lz, dyadic, of_inverse=True)).sum())
        scale = float(c.norm() * y.norm())
        assert abs(lhs - rhs) <= 1e-12 * scale, (b, lhs, rhs)


def test_autograd_finite_differences_on_crop(wz, grf256x8):
    from waltz_amd import autograd
    f = grf256x8[0]
    crop = f[100:116, 40:56, 7:23].contiguous().double()  # 16^3, one level per axis
    bound = 1e-3 * (float(f.max()) - float(f.min()))
    g = torch.Generator(device="cuda").manual_seed(20261208)
    target = crop + 0.05 * torch.randn(crop.shape, dtype=torch.float64, device="cuda", generator=g)
    levels = (1, 1, True)
    quantum = 1.5 * bound
    q = torch.round(autograd.dwt(crop, levels) / quantum)  # integer-valued magnitudes with signs

    def loss(qv):
        return torch.nn.functional.mse_loss(autograd.reconstruct(qv, quantum, levels), target)

    qv = q.clone().requires_grad_()
    loss(qv).backward()
    grad = qv.grad
    rng = np.random.default_rng(3)
    eps = 1e-3
    for _ in range(24):
        i = tuple(int(v) for v in rng.integers(0, 16, 3))
        up, dn = q.clone(), q.clone()
        up[i] += eps
        dn[i] -= eps
        fd = (float(loss(up)) - float(loss(dn))) / (2 * eps)
        # the loss is quadratic in q: central differences are exact up to
        # rounding (about 1e-16 of the loss over 2 eps, far below 1e-13)
        assert abs(fd - float(grad[i])) <= 1e-6 * abs(float(grad[i])) + 1e-13, (i, fd, float(grad[i]))


def test_wzp_lookback_beyond_one_round(wz, oracle):
    """A field of more than 1024 WZP groups (1024 x 16 x 8192 = 2^27 samples):
    the group pass's decoupled lookback needs a second 1024-descriptor round.
    The archive still equals the oracle's and decodes to its bits."""
    waltz_amd, synthetic = wz
    shape = (513, 512, 512)
    n = shape[0] * shape[1] * shape[2]
    groups = -(-(-(-n // 8192)) // 16)
    assert groups > 1024
    x = synthetic.grf(shape, slope=-5.0 / 3.0, seed=20261209, offset=1.0)
    host = x.cpu().numpy()
    _say(f"{shape} f32: {groups} WZP groups")
    archive = waltz_amd.compress(x, abs_error=2e-3)
    want = oracle.compress(host, 2e-3)
    assert archive.cpu().numpy().tobytes() == want
    out = waltz_amd.decompress(archive, shape=shape, dtype=torch.float32)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), oracle.decompress(want, shape, np.float32).view(np.uint32))
