"""The reference workloads that are not the throughput line, as parity cases:
the 64^3 sinusoid field at REL 1e-3, and the 384 x 384 x 96 float64 field
(tanh fronts + k^-3 noise, plane schedule with 5 XY and 4 Z levels) at
REL 1e-6. Archives must equal the oracle's bytes, decoded samples its bits,
and the error bound must hold in FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    waltz_amd.lib()
    return waltz_amd


def sinusoids(shape, seed):
    """Sum of 8 random 3-D sinusoids (integer frequencies 1..8, amplitudes
    U[0.1, 1], phases U[0, 2pi)) plus N(0, 0.01^2) noise."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(s) / s for s in shape], indexing="ij")
    f = np.zeros(shape)
    for _ in range(8):
        fz, fy, fx = rng.integers(1, 9, 3)
        f += rng.uniform(0.1, 1.0) * np.sin(2 * np.pi * (fz * z + fy * y + fx * x) + rng.uniform(0, 2 * np.pi))
    return (f + 0.01 * rng.standard_normal(shape)).astype(np.float32)


def fronts(shape, seed):
    """Three tanh fronts of amplitude 40 plus 5 x a unit k^-3 random field."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    f = np.zeros(shape)
    for _ in range(3):
        nrm = rng.standard_normal(3)
        nrm /= np.linalg.norm(nrm)
        c = sum(nrm[k] * shape[k] * rng.uniform(0.3, 0.7) for k in range(3))
        w = rng.uniform(4.0, 12.0)
        f += 40.0 * np.tanh((nrm[0] * grids[0] + nrm[1] * grids[1] + nrm[2] * grids[2] - c) / w) / 3.0
    spec = np.fft.rfftn(rng.standard_normal(shape))
    k2 = sum(np.meshgrid(*[np.fft.fftfreq(s) ** 2 for s in shape[:2]] + [np.fft.rfftfreq(shape[2]) ** 2],
                         indexing="ij"))
    amp = np.where(k2 > 0, np.maximum(k2, 1e-30) ** (-3.0 / 4.0), 0.0)
    g = np.fft.irfftn(spec * amp, s=shape)
    g = (g - g.mean()) / g.std()
    return f + 5.0 * g


def roundtrip(wz, oracle, v, rel):
    shape = v.shape
    bound = rel * (float(v.max()) - float(v.min()))
    ref = oracle.compress(v, bound)
    tdt = torch.float32 if v.dtype == np.float32 else torch.float64
    a = wz.compress(torch.from_numpy(v).cuda(), abs_error=bound)
    assert a.cpu().numpy().tobytes() == ref
    d = wz.decompress(a, shape=shape, dtype=tdt).cpu().numpy()
    assert np.array_equal(d.view(np.uint8), oracle.decompress(ref, shape, v.dtype).view(np.uint8))
    assert np.max(np.abs(d.astype(np.float64) - v.astype(np.float64))) <= bound
    return ref


def test_sinusoids_64cube_rel_1e3(wz, oracle):
    roundtrip(wz, oracle, sinusoids((64, 64, 64), 20261202), 1e-3)


def test_fronts_384x384x96_f64_rel_1e6(wz, oracle):
    v = fronts((96, 384, 384), 20261202 + 2)
    assert oracle.geometry(384, 384, 96)[:2] == (5, 4)  # plane schedule
    ref = roundtrip(wz, oracle, v, 1e-6)
    assert ref[0] & 4 and not ref[0] & 1  # float64, plane schedule
