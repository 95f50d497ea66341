"""The fused per-level lifting kernels (dyadic schedule) against the oracle,
bitwise: at a shape where the built-in heuristic fuses the finest level
(144 x 256 x 512, full 64 x 32 tiles, aligned rows), and at a small shape with
ragged tiles (80 x 100 x 120) under forced launch choices that fuse 0 to all 4
levels with different z segments. The codec round trip covers the fused
repair emitter (row-word and atomic bitmap paths) and the decoder's screen.
The autotuner's choice must not change a bit either."""
import numpy as np
import pytest

from fields import field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TDT = {np.float32: torch.float32, np.float64: torch.float64}


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    from waltz_amd import stages
    waltz_amd.lib()
    stages.clear_tuning()
    yield waltz_amd, stages
    stages.clear_tuning()


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def check_transform(stages, oracle, v, lxy, lz):
    c = stages.transform(torch.from_numpy(v).cuda(), lxy, lz, True, inverse=False).cpu().numpy()
    assert same_bits(c, oracle.transform(v, lxy, lz, True, False))
    r = stages.transform(torch.from_numpy(v).cuda(), lxy, lz, True, inverse=True).cpu().numpy()
    assert same_bits(r, oracle.transform(v, lxy, lz, True, True))


def check_codec(waltz_amd, oracle, v, bound):
    ref = oracle.compress(v, bound)
    a = waltz_amd.compress(torch.from_numpy(v).cuda(), abs_error=bound)
    assert a.cpu().numpy().tobytes() == ref
    d = waltz_amd.decompress(a, shape=v.shape, dtype=TDT[v.dtype.type]).cpu().numpy()
    assert same_bits(d, oracle.decompress(ref, v.shape, v.dtype.type))


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_heuristic_fused_level0(wz, oracle, dtype):
    waltz_amd, stages = wz
    stages.clear_tuning()
    shape = (144, 256, 512)
    v = field(shape, "noise", dtype) * 3 + field(shape, "smooth", dtype)
    check_transform(stages, oracle, v, 5, 5)
    check_codec(waltz_amd, oracle, v, 1e-3)


TUNINGS = [(0, (16, 16, 16, 16, 16)), (1, (16, 16, 16, 16, 16)), (2, (32, 16, 16, 16, 16)),
           (4, (64, 32, 16, 16, 16)), (4, (16, 16, 16, 16, 16)), (3, (128, 64, 16, 16, 16))]


@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: f"f{t[0]}_sz{t[1][0]}")
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_forced_tunings_ragged(wz, oracle, dtype, tuning):
    waltz_amd, stages = wz
    shape = (80, 100, 120)  # 4 dyadic levels; tiles ragged in x and y
    fused, sz = tuning
    for inverse in (False, True):
        stages.set_tuning(shape, TDT[dtype], inverse, fused, sz)
    try:
        v = field(shape, "smooth", dtype) + field(shape, "outliers", dtype, seed=3) / 64
        check_transform(stages, oracle, v, 4, 4)
        check_codec(waltz_amd, oracle, v, 1e-3)
        check_codec(waltz_amd, oracle, v, 1e-5)
    finally:
        stages.clear_tuning()


def test_autotuned_choice_is_bit_neutral(wz, oracle):
    waltz_amd, stages = wz
    shape = (144, 256, 512)
    stages.autotune(shape, torch.float32)
    rep = stages.tuning_report()
    assert any(r["shape_xyz"] == [512, 256, 144] and r["inverse"] for r in rep)
    v = field(shape, "smooth", np.float32)
    try:
        check_codec(waltz_amd, oracle, v, 1e-3)
    finally:
        stages.clear_tuning()


# brick-pair shapes with both leading inverse levels fused: compress runs the
# coarse inverse levels beside the encoder on requantized loads
SPLIT = [((64, 64, 64), 2, (16, 16, 16, 16, 16)), ((48, 80, 96), 3, (16, 16, 16, 16, 16)),
         ((96, 48, 160), 4, (32, 16, 16, 16, 16))]


@pytest.mark.parametrize("case", SPLIT, ids=lambda c: "x".join(map(str, c[0])) + f"_f{c[1]}")
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_split_compress_inverse(wz, oracle, dtype, case):
    waltz_amd, stages = wz
    shape, fused, sz = case
    stages.set_tuning(shape, TDT[dtype], True, fused, sz)
    try:
        v = field(shape, "smooth", dtype) + field(shape, "outliers", dtype, seed=3) / 64
        check_codec(waltz_amd, oracle, v, 1e-3)
        check_codec(waltz_amd, oracle, v, 1e-5)
    finally:
        stages.clear_tuning()


# per-axis pass choices: forward X/Y/Z line tiles, inverse X tile, LDS pair limit
PASS_TUNINGS = [((8, 8, 8), 0), ((16, 32, 16), 1024), ((32, 16, 32), 4096), ((32, 32, 8), 0), ((16, 8, 32), 4096)]


@pytest.mark.parametrize("pt", PASS_TUNINGS, ids=lambda p: "t" + "-".join(map(str, p[0])) + f"_pair{p[1]}")
@pytest.mark.parametrize("case", [((24, 100, 120), "plane", np.float64), ((40, 68, 72), "plane", np.float32),
                                  ((80, 100, 120), "dyadic-passes", np.float32)], ids=lambda c: c[1] + str(c[0][0]))
def test_forced_pass_tunings(wz, oracle, case, pt):
    """Line tiles and the LDS pair limit of the per-axis passes change no bit:
    the plane schedule (all axis passes) and a dyadic field with every level
    left to the passes (no fused level)."""
    waltz_amd, stages = wz
    shape, kind, dtype = case
    tiles, pair = pt
    if kind == "dyadic-passes":
        for inverse in (False, True):
            stages.set_tuning(shape, TDT[dtype], inverse, 0, (16, 16, 16, 16, 16))
    for inverse in (False, True):
        stages.set_pass_tuning(shape, TDT[dtype], inverse, tiles, pair)
    try:
        from waltz_amd.autograd import default_levels
        lxy, lz, dyadic = default_levels(shape)
        v = field(shape, "smooth", dtype) + field(shape, "noise", dtype, seed=5) / 16
        c = stages.transform(torch.from_numpy(v).cuda(), lxy, lz, dyadic, inverse=False).cpu().numpy()
        assert same_bits(c, oracle.transform(v, lxy, lz, dyadic, False))
        r = stages.transform(torch.from_numpy(v).cuda(), lxy, lz, dyadic, inverse=True).cpu().numpy()
        assert same_bits(r, oracle.transform(v, lxy, lz, dyadic, True))
        check_codec(waltz_amd, oracle, v, 1e-3)
    finally:
        stages.clear_tuning()


def test_autotuned_plane_schedule_is_bit_neutral(wz, oracle):
    waltz_amd, stages = wz
    shape = (24, 100, 120)
    stages.autotune(shape, torch.float64)
    rep = stages.tuning_report()
    assert any(r["shape_xyz"] == [120, 100, 24] and "pass_tiles_xyz" in r for r in rep)
    v = field(shape, "smooth", np.float64)
    try:
        check_codec(waltz_amd, oracle, v, 1e-6)
    finally:
        stages.clear_tuning()
