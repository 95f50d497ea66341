"""Pins the CPU oracle (oracle/waltz_oracle.c) against things other than itself:
the reference codec's golden digests, the reference's FNV-1a transform known
answers, oracle/_ref (the reference compiled from its sources) when built,
brute force on tiny inputs and closed forms."""
import hashlib

import numpy as np
import pytest

from fields import field, _lcg

DT = {"float32": np.float32, "float64": np.float64}


# --------------------------------------------------------------- whole codec
def test_golden_archives(oracle, golden):
    """Every reference-written archive digest (272 cases: dyadic and plane
    schedules, edge bricks, BZ 6, q-outliers, repairs, zero/constant fields)."""
    for rec in golden:
        shape, dtype = tuple(rec["shape"]), DT[rec["dtype"]]
        v = field(shape, rec["pattern"], dtype)
        assert hashlib.sha256(v.tobytes()).hexdigest() == rec["input_sha256"], rec
        a = oracle.compress(v, rec["bound"])
        assert len(a) == rec["archive_bytes"], rec
        assert hashlib.sha256(a).hexdigest() == rec["archive_sha256"], rec
        d = oracle.decompress(a, shape, dtype)
        assert hashlib.sha256(d.tobytes()).hexdigest() == rec["decoded_sha256"], rec
        assert np.all(np.abs(d.astype(np.float64) - v.astype(np.float64)) <= rec["bound"])


def test_golden_covers_side_channels(golden):
    assert sum(r["header"]["n_qo"] > 0 for r in golden) >= 10
    assert sum(r["header"]["n_pwe"] > 0 for r in golden) >= 50
    assert sum(not r["header"]["flags"] & 1 for r in golden) >= 50
    assert {r["header"]["bz"] for r in golden} >= {1, 2, 4, 6, 8, 16}


def test_matches_reference_random(oracle, reference):
    rng = np.random.default_rng(11)
    for shape in [(7, 5, 3), (31, 33, 9), (24, 70, 41), (66, 18, 20)]:
        for dtype in (np.float32, np.float64):
            v = (rng.standard_normal(shape) * 3 + 0.5).astype(dtype)
            for bound in (1e-4, 1e-2):
                a = oracle.compress(v, bound)
                assert a == reference.compress(v, bound, reference.SCALAR)
                d = oracle.decompress(a, shape, dtype)
                assert np.array_equal(d.view(np.uint8), reference.decompress(a, shape, dtype).view(np.uint8))


# --------------------------------------------------------------- transform
def _samples(shape, dtype):
    v = (_lcg(int(np.prod(shape)), 727).astype(dtype) / dtype(1048576)).reshape(shape)
    if v.size > 7:
        v.flat[0], v.flat[1], v.flat[2], v.flat[3] = 0, -0.0, 1, -1
    return v


# (nz, ny, nx), (lxy, lz): f32 forward, f32 inverse, f64 forward, f64 inverse.
# FNV-1a-64 known answers printed by the reference's transform bench
# (cpu/tests/test_transform.cpp:216-245); inverse rows transform the samples.
FNV = [
    ((64, 64, 64), (3, 3), "f0a15db4e9071e9f", "9415ed49e6df27fc", "78891ad96aa99727", "f8f874985b170e46"),
    ((8, 128, 128), (4, 0), "e9e4b23779fb765f", "8977c4c29e1b8484", "16fb3d9b2856ba7f", "026f6c05e31bdc45"),
    ((128, 128, 128), (4, 4), "392755e1c3817d28", "5faba50381c73cf9", "1e6091b507b6f516", "11df1b954e913138"),
    ((8, 512, 512), (5, 0), "0865348db4fe6421", "b3f5648dd3d90fac", "b83bd70c47c0fe43", "ab85502720a9b28f"),
]


@pytest.mark.parametrize("case", FNV, ids=lambda c: "x".join(map(str, c[0])))
def test_transform_known_answers(oracle, case):
    shape, (xy, z), *hashes = case
    got = []
    for dtype in (np.float32, np.float64):
        s = _samples(shape, dtype)
        got.append(oracle.fnv1a64(oracle.transform(s, xy, z, xy == z, False)))
        got.append(oracle.fnv1a64(oracle.transform(s, xy, z, xy == z, True)))
    assert got == list(hashes)


def test_transform_matches_reference(oracle, reference):
    shapes = [(1, 1, 1), (2, 2, 2), (4, 3, 2), (1, 2, 3), (5, 9, 7), (13, 11, 9), (19, 17, 15),
              (8, 8, 16), (9, 16, 8), (33, 5, 17), (17, 9, 57), (19, 17, 65), (257, 13, 11)]
    for shape in shapes:
        for dtype in (np.float32, np.float64):
            s = _samples(shape, dtype)
            for lxy, lz, dy in [(0, 0, True), (1, 1, True), (3, 3, True), (5, 5, True), (0, 3, False),
                                (2, 0, False), (1, 2, False), (3, 1, False)]:
                for inv in (False, True):
                    a = oracle.transform(s, lxy, lz, dy, inv)
                    b = reference.transform(s, lxy, lz, dy, inv)
                    assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (shape, lxy, lz, inv)


def test_transform_reconstruction_and_linearity(oracle):
    """Perfect reconstruction up to rounding, and linearity of the lifting
    (a dropped or sign-flipped lifting step breaks both)."""
    rng = np.random.default_rng(3)
    for shape, (lxy, lz) in [((32, 32, 32), (2, 2)), ((16, 40, 24), (2, 1)), ((9, 13, 21), (1, 1))]:
        x = rng.standard_normal(shape)
        c = oracle.transform(x, lxy, lz, lxy == lz, False)
        r = oracle.transform(c, lxy, lz, lxy == lz, True)
        assert np.max(np.abs(r - x)) < 1e-12
        y = rng.standard_normal(shape)
        cy = oracle.transform(y, lxy, lz, lxy == lz, False)
        cxy = oracle.transform(x + 2 * y, lxy, lz, lxy == lz, False)
        assert np.max(np.abs(cxy - (c + 2 * cy))) < 1e-11
        # energy is roughly preserved by the normalised 9/7 pair (not orthogonal)
        assert 0.5 < np.sum(c * c) / np.sum(x * x) < 2.0


def test_transform_lowpass_of_constant(oracle):
    """A constant field maps to low-pass gain k^3 per level in the LLL band and
    exact-ish zeros elsewhere (CDF 9/7 high-pass annihilates constants)."""
    x = np.full((16, 16, 16), 2.0)
    c = oracle.transform(x, 1, 1, True, False)
    k = 1.1496043988602418
    dc = 2.0 * (1 + 2 * -0.052980118572961 * (1 + 2 * -1.586134342059924)) * \
        (1 + 2 * 0.443506852043971 * (1 + 2 * 0.882911075530934 * 1))
    # the closed-form low-pass response to a constant: product of lifting steps
    hp = c.copy()
    hp[:8, :8, :8] = 0
    assert np.max(np.abs(hp)) < 1e-12
    assert np.allclose(c[:8, :8, :8], c[0, 0, 0])
    assert c[0, 0, 0] > 2.0 * k ** 3 * 0.9


# --------------------------------------------------------------- Morton
def test_morton_rank_bruteforce(oracle):
    def brute(x, y, z, nx, ny, nz):
        # explicit octree enumeration: z-lower half first, then y, then x
        order = []

        def visit(x0, y0, z0, sx, sy, sz):
            if sx * sy * sz == 0:
                return
            if sx == sy == sz == 1:
                order.append((x0, y0, z0))
                return
            hx, hy, hz = sx - sx // 2, sy - sy // 2, sz - sz // 2
            for cz in (0, 1):
                for cy in (0, 1):
                    for cx in (0, 1):
                        visit(x0 + cx * hx, y0 + cy * hy, z0 + cz * hz,
                              sx // 2 if cx else hx, sy // 2 if cy else hy, sz // 2 if cz else hz)
        visit(0, 0, 0, nx, ny, nz)
        return order.index((x, y, z))

    for dims in [(16, 16, 16), (5, 3, 2), (16, 7, 6), (3, 16, 4), (1, 1, 9)]:
        nx, ny, nz = dims
        ranks = sorted(oracle.morton_rank(x, y, z, nx, ny, nz)
                       for z in range(nz) for y in range(ny) for x in range(nx))
        assert ranks == list(range(nx * ny * nz))
        for x, y, z in [(0, 0, 0), (nx - 1, ny - 1, nz - 1), (nx // 2, 0, nz - 1), (min(1, nx - 1), ny // 2, nz // 3)]:
            assert oracle.morton_rank(x, y, z, nx, ny, nz) == brute(x, y, z, nx, ny, nz)
    # power-of-two brick = standard bit interleave z,y,x (MSB first)
    def interleave(x, y, z):
        r = 0
        for b in range(4):
            r |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
        return r
    for x, y, z in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (15, 3, 9), (7, 12, 5)]:
        assert oracle.morton_rank(x, y, z, 16, 16, 16) == interleave(x, y, z)


def test_levels_and_zband(oracle):
    assert [oracle.levels(n) for n in (1, 8, 9, 16, 17, 64, 96, 512, 10**6)] == [0, 0, 1, 1, 2, 3, 4, 5, 5]
    assert oracle.geometry(512, 512, 512) == (5, 5, True, 16)
    assert oracle.geometry(384, 384, 96) == (5, 4, False, 2)
    assert oracle.geometry(96, 64, 48) == (3, 3, True, 6)
    assert oracle.geometry(128, 128, 64) == (4, 3, False, 2)
    assert oracle.geometry(64, 64, 32) == (3, 2, False, 2)


# --------------------------------------------------------------- quantize
def test_quantize_closed_form(oracle):
    """llrint half-even of T(c*inv), INT32 saturation, +0 for zero codes."""
    q = 1.5e-3
    inv32 = np.float32(1.0 / q)
    vals = np.array([0.0, -0.0, 0.5 * q, -0.5 * q, 1.5 * q, -2.5 * q, 1e-3, 97.3, -98.4,
                     65535.4 * q, 65536 * q, -70000 * q, 3e9 * q, -4e12 * q], np.float32)
    c = np.zeros((1, 1, 16), np.float32)
    c.ravel()[:vals.size] = vals
    deq, mag, sign, qidx, qval = oracle.quantize(c, 1, q)
    scaled = c.ravel() * inv32
    code = np.rint(scaled.astype(np.float64))
    m = np.minimum(np.abs(code), 2**31 - 1)
    ranks = [oracle.morton_rank(i, 0, 0, 16, 1, 1) for i in range(16)]
    for i in range(16):
        assert mag[ranks[i]] == min(m[i], 65535)
        neg = code[i] < 0 and m[i] > 0
        assert bool(sign[0] >> ranks[i] & 1) == neg
        want = np.float32(0) if m[i] == 0 else np.float32(-m[i] if code[i] < 0 else m[i]) * np.float32(q)
        assert deq.ravel()[i].tobytes() == np.float32(want).tobytes()
    assert list(qidx) == [i for i in range(16) if m[i] > 65535]
    assert list(qval) == [int(m[i]) for i in range(16) if m[i] > 65535]


# --------------------------------------------------------------- repairs
def test_within_bound_twodiff(oracle):
    b = 0.1
    assert oracle.within_bound(1.0 + b / 2, 1.0, b)
    assert not oracle.within_bound(1.0 + 2 * b, 1.0, b)
    # |v - o| rounds to exactly b in double but the exact difference exceeds b
    o = 0.3
    v = o + b
    exact_gt = (np.longdouble(v) - np.longdouble(o)) > np.longdouble(b)
    assert oracle.within_bound(v, o, b) == (not exact_gt)
    o2, v2 = 1.0, 1.0 + 2.0**-52
    assert oracle.within_bound(v2, o2, 2.0**-52)
    assert not oracle.within_bound(v2, o2, 2.0**-53)


def test_repairs_closed_form(oracle):
    rng = np.random.default_rng(9)
    o = rng.standard_normal(4096).astype(np.float32)
    r = (o + rng.uniform(-2e-3, 2e-3, o.size)).astype(np.float32)
    idx, err = oracle.repairs(o, r, 1e-3)
    e = (r - o).astype(np.float32)
    want = np.nonzero(np.abs(e) > np.float32(1e-3))[0]
    assert np.array_equal(idx, want)
    assert np.array_equal(err, e[want])
    fixed = r.copy()
    fixed[idx] = fixed[idx] - err
    assert np.all(np.abs(fixed.astype(np.float64) - o) <= 1e-3)


# --------------------------------------------------------------- WZP
def test_bit2_permutation(oracle):
    """Element j bit k -> plane 15-k bit 15-j (each single-bit input)."""
    for j in (0, 5, 15):
        for k in (0, 7, 15):
            m = np.zeros(8192, np.uint16)
            m[j] = 1 << k
            m[100] = 0xFFFF  # keep the chunk nonzero elsewhere too
            s = np.zeros(256, np.uint32)
            raw = np.frombuffer(oracle.wzp_encode(m, s), np.uint8)
            mag, _ = oracle.wzp_decode(raw, 8192)
            assert np.array_equal(mag, m)


def _wzp_inputs(rng):
    for n in (1, 15, 16, 17, 8191, 8192, 8193, 16 * 8192, 16 * 8192 + 37, 33 * 8192 + 5):
        for kind in range(5):
            if kind == 0:
                m = np.zeros(n, np.uint16)
            elif kind == 1:
                m = rng.integers(0, 4, n).astype(np.uint16) * (rng.random(n) < 0.1)
            elif kind == 2:
                m = rng.integers(0, 65536, n).astype(np.uint16)
            elif kind == 3:
                m = np.ones(n, np.uint16) * 3
            else:
                m = (rng.geometric(0.3, n) - 1).astype(np.uint16)
            s = rng.integers(0, 2**32, (n + 31) // 32, dtype=np.uint64).astype(np.uint32)
            yield n, m.astype(np.uint16), s


def test_wzp_vs_reference(oracle, reference):
    rng = np.random.default_rng(1)
    for n, m, s in _wzp_inputs(rng):
        a = oracle.wzp_encode(m, s)
        assert a == reference.wzp_encode(m, s), (n,)


def test_wzp_roundtrip_and_sizes(oracle):
    rng = np.random.default_rng(2)
    for n, m, s in _wzp_inputs(rng):
        a = np.frombuffer(oracle.wzp_encode(m, s), np.uint8)
        chunks = (n + 8191) // 8192
        groups = (chunks + 15) // 16
        header = 4 * groups + 3 * chunks
        meta = a[4 * groups:header]
        total = 0
        offsets = []
        for g in range(groups):
            glen = 0
            for c in range(g * 16, min(chunks, g * 16 + 16)):
                v = int(meta[3 * c]) | int(meta[3 * c + 1]) << 8 | int(meta[3 * c + 2]) << 16
                code, sc = v & 0x3FFF, v >> 14
                el = min(8192, n - c * 8192)
                if code == 0:
                    continue
                glen += (2 * el if code == 0x3FFF else code) + (sc or 1024)
            offsets.append(int.from_bytes(a[4 * g:4 * g + 4].tobytes(), "little"))
            assert offsets[-1] == (total if glen else 0)
            total += glen
        assert a.size == header + total
        mag, sign = oracle.wzp_decode(a, n)
        assert np.array_equal(mag, m)
        keep = np.zeros((n + 31) // 32 * 32, bool)
        keep[:n] = m != 0
        bits = np.unpackbits(s.view(np.uint8), bitorder="little").astype(bool) & keep
        got = np.unpackbits(sign.view(np.uint8), bitorder="little").astype(bool)
        assert np.array_equal(got, bits)
