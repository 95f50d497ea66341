"""Parity of the gfx950 HIP path with the oracle (bitwise: coefficients,
magnitudes, sign words, WZP bytes, archives, decoded samples) and with the
reference-written golden fixtures. Every case is its own test so the name is
printed before its GPU work starts."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from fields import field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
GOLDEN = json.loads((Path(__file__).parent / "golden" / "codec_golden.json").read_text())
DT = {"float32": np.float32, "float64": np.float64}
TDT = {np.float32: torch.float32, np.float64: torch.float64}


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    from waltz_amd import stages
    waltz_amd.lib()  # fails loudly if the HIP library is missing
    return waltz_amd, stages


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def archive_diff(got, want, n):
    """Where two archives of an n-sample field first differ (header, WZP
    directory / metas / payload, or side sections), for assertion messages."""
    if got == want:
        return "identical"
    m = min(len(got), len(want))
    d = next((i for i in range(m) if got[i] != want[i]), m)
    chunks = (n + 8191) // 8192
    groups = (chunks + 15) // 16
    wzp = int.from_bytes(want[16:24], "little")
    region = ("header" if d < 32 else "directory" if d < 32 + 4 * groups else
              "metas" if d < 32 + 4 * groups + 3 * chunks else "payload" if d < 32 + wzp else "sections")
    return (f"lengths {len(got)} / {len(want)}, first difference at byte {d} ({region}), "
            f"{sum(1 for i in range(m) if got[i] != want[i])} differing bytes")


# ------------------------------------------------------------------ transform
TRANSFORM_CASES = [
    ((1, 1, 1), 0, 0, True), ((4, 3, 2), 1, 1, True), ((5, 9, 7), 1, 0, False), ((13, 11, 9), 1, 1, True),
    ((19, 17, 15), 2, 2, True), ((8, 8, 16), 1, 0, False), ((33, 5, 17), 2, 0, False), ((17, 9, 57), 1, 2, False),
    ((19, 17, 65), 2, 1, False), ((64, 64, 64), 3, 3, True), ((8, 128, 128), 4, 0, False),
    ((40, 48, 56), 3, 3, True), ((96, 64, 32), 3, 4, False), ((70, 66, 67), 5, 5, True),
]


@pytest.mark.parametrize("case", TRANSFORM_CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_l{c[1]}{c[2]}")
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_transform_bitwise(wz, oracle, case, dtype):
    _, stages = wz
    shape, lxy, lz, dy = case
    x = field(shape, "noise", dtype) * 3 + field(shape, "smooth", dtype)
    f_gpu = host(stages.transform(dev(x), lxy, lz, dy, False))
    f_ref = oracle.transform(x, lxy, lz, dy, False)
    assert same_bits(f_gpu, f_ref)
    i_gpu = host(stages.transform(dev(f_ref), lxy, lz, dy, True))
    i_ref = oracle.transform(f_ref, lxy, lz, dy, True)
    assert same_bits(i_gpu, i_ref)


# ------------------------------------------------------------------ quantize
@pytest.mark.parametrize("shape", [(16, 16, 16), (17, 19, 15), (48, 64, 96), (2, 130, 18), (33, 5, 17)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_quantize_bitwise(wz, oracle, shape, dtype):
    _, stages = wz
    nx, ny, nz = shape[2], shape[1], shape[0]
    lx, lz, dyadic, bz = oracle.geometry(nx, ny, nz)
    c = oracle.transform(field(shape, "outliers", dtype), lx, lz, dyadic, False)
    quantum = 1.5e-3
    deq, mag, sign, qi, qv = oracle.quantize(c, bz, quantum)
    gd, gm, gs, gqi, gqv = stages.quantize(dev(c), bz, quantum)
    assert same_bits(host(gd), deq)
    assert np.array_equal(host(gm).view(np.uint16), mag)
    assert np.array_equal(host(gs).view(np.uint32), sign)
    order = np.argsort(host(gqi).view(np.uint32))
    assert np.array_equal(host(gqi).view(np.uint32)[order], qi)
    assert np.array_equal(host(gqv).view(np.uint32)[order], qv)


# ------------------------------------------------------------------ WZP
def _wzp_input(n, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "sparse":
        m = rng.integers(0, 4, n) * (rng.random(n) < 0.1)
    elif kind == "dense":
        m = rng.integers(0, 65536, n)
    elif kind == "geometric":
        m = rng.geometric(0.3, n) - 1
    elif kind == "zero":
        m = np.zeros(n)
    else:
        m = np.full(n, 3)
    s = rng.integers(0, 2**32, (n + 31) // 32, dtype=np.uint64).astype(np.uint32)
    return m.astype(np.uint16), s


@pytest.mark.parametrize("n", [1, 15, 17, 8191, 8192, 8193, 16 * 8192 + 37, 40 * 8192 + 5])
@pytest.mark.parametrize("kind", ["sparse", "dense", "geometric", "zero", "const"])
def test_wzp_bytes(wz, oracle, n, kind):
    _, stages = wz
    m, s = _wzp_input(n, kind, n)
    ref = oracle.wzp_encode(m, s)
    got = host(stages.wzp_encode(dev(m.view(np.int16)), dev(s.view(np.int32)))).tobytes()
    assert got == ref
    gm, gs = stages.wzp_decode(dev(np.frombuffer(ref, np.uint8)), n)
    om, os_ = oracle.wzp_decode(ref, n)
    assert np.array_equal(host(gm).view(np.uint16), om)
    assert np.array_equal(host(gs).view(np.uint32), os_)


# ------------------------------------------------------------------ whole codec vs golden
GOLDEN_GPU = [r for r in GOLDEN if np.prod(r["shape"]) <= 300_000]


@pytest.mark.parametrize("rec", GOLDEN_GPU, ids=lambda r: "%s_%s_%s_%g" % ("x".join(map(str, r["shape"])),
                                                                            r["pattern"], r["dtype"], r["bound"]))
def test_codec_golden(wz, rec):
    waltz_amd, _ = wz
    shape, dtype = tuple(rec["shape"]), DT[rec["dtype"]]
    v = field(shape, rec["pattern"], dtype)
    a = waltz_amd.compress(dev(v), abs_error=rec["bound"])
    ab = host(a).tobytes()
    assert len(ab) == rec["archive_bytes"]
    assert hashlib.sha256(ab).hexdigest() == rec["archive_sha256"]
    d = waltz_amd.decompress(a, shape=shape, dtype=TDT[dtype])
    assert hashlib.sha256(host(d).tobytes()).hexdigest() == rec["decoded_sha256"]


@pytest.mark.parametrize("rec", [r for r in GOLDEN if r.get("archive_file")], ids=lambda r: r["archive_file"])
def test_decode_reference_archives(wz, rec):
    """Archives written by the reference CPU codec decode to its exact bits."""
    waltz_amd, _ = wz
    raw = (Path(__file__).parent / "golden" / rec["archive_file"]).read_bytes()
    d = waltz_amd.decompress(dev(np.frombuffer(raw, np.uint8)), shape=tuple(rec["shape"]), dtype=TDT[DT[rec["dtype"]]])
    assert hashlib.sha256(host(d).tobytes()).hexdigest() == rec["decoded_sha256"]


@pytest.mark.parametrize("shape", [(64, 64, 64), (96, 64, 48), (24, 160, 200), (128, 128, 128)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_codec_matches_oracle(wz, oracle, shape, dtype):
    waltz_amd, _ = wz
    rng = np.random.default_rng(7)
    z, y, x = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    v = (np.sin(0.11 * x) * np.cos(0.07 * y) + 0.02 * z + 0.05 * rng.standard_normal(shape)).astype(dtype)
    for bound in (1e-3, 1e-5):
        ref = oracle.compress(v, bound)
        a = waltz_amd.compress(dev(v), abs_error=bound)
        got = host(a).tobytes()
        assert got == ref, archive_diff(got, ref, v.size)
        d = host(waltz_amd.decompress(a, shape=shape, dtype=TDT[dtype]))
        assert same_bits(d, oracle.decompress(ref, shape, dtype))
        assert np.max(np.abs(d.astype(np.float64) - v.astype(np.float64))) <= bound


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_codec_paired_bricks(wz, oracle, dtype):
    """Full 16^3 bricks with nx % 32 == 0 (quantizer fused into the WZP
    encoder, dequantizer into the decoder): a zero slab gives empty chunks,
    sparse spikes give q-outliers."""
    waltz_amd, _ = wz
    shape = (496, 144, 160)
    rng = np.random.default_rng(11)
    z, y, x = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    v = np.sin(0.05 * x) * np.cos(0.03 * y + 0.01 * z) + 0.01 * rng.standard_normal(shape)
    v[:200] = 0.0
    spikes = rng.integers(0, v.size, 300)
    v.reshape(-1)[spikes] = rng.choice([-1.0, 1.0], 300) * 5e3
    v = v.astype(dtype)
    ref = oracle.compress(v, 1e-3)
    a = waltz_amd.compress(dev(v), abs_error=1e-3)
    got = host(a).tobytes()
    assert got == ref, archive_diff(got, ref, v.size)
    d = host(waltz_amd.decompress(a, shape=shape, dtype=TDT[dtype]))
    assert same_bits(d, oracle.decompress(ref, shape, dtype))
    assert np.max(np.abs(d.astype(np.float64) - v.astype(np.float64))) <= 1e-3


def test_rejects_nonfinite(wz):
    waltz_amd, _ = wz
    v = np.ones((16, 16, 16), np.float32)
    v[3, 4, 5] = np.nan
    with pytest.raises(RuntimeError, match="nonfinite"):
        waltz_amd.compress(dev(v), abs_error=1e-3)
    # the library stays usable after a failed call
    a = waltz_amd.compress(dev(np.ones((16, 16, 16), np.float32)), abs_error=1e-3)
    assert a.numel() > 32


def test_rejects_corrupt_archive(wz, oracle):
    waltz_amd, _ = wz
    v = field((16, 16, 16), "smooth", np.float32)
    a = bytearray(oracle.compress(v, 1e-3))
    bad = bytes(a[:-8])  # truncated
    with pytest.raises(RuntimeError):
        waltz_amd.decompress(dev(np.frombuffer(bad, np.uint8)), shape=(16, 16, 16), dtype=torch.float32)
    a[1] = 7  # impossible level count
    with pytest.raises(RuntimeError):
        waltz_amd.decompress(dev(np.frombuffer(bytes(a), np.uint8)), shape=(16, 16, 16), dtype=torch.float32)
