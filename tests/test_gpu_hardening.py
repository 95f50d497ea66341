"""Decoder hardening: the HIP decoder accepts or rejects a damaged archive
exactly when the oracle does, and decodes the same bits when both accept
(cpu/src/codec.cpp:276-336 and the WZP directory rule of
cpu/src/wzp_cpu.cpp:713-761; the malformed-archive cases follow
cpu/tests/test_codec.cpp:231-309 of the reference).

Cases: every truncation length of a small archive, every header byte with
single-bit and whole-byte flips, and on a two-group WZP stream mutated group
offsets (off by one, swapped, past the payload, nonzero for an empty group),
mutated chunk metas, and a valid reordering of the group payloads (accepted
by both: the directory only has to tile the payload).

Each case runs on both decode routes: the host reads the header first, or
(brick-pair fields whose finest inverse level is fused, forced here on small
fields) the header is checked on the device while the decode is already
launched, and the host checks the header it reads back at the end.
"""
import contextlib
import struct

import numpy as np
import pytest

from fields import field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    waltz_amd.lib()
    return waltz_amd


@contextlib.contextmanager
def route(planned, shape, dtype):
    """planned: force the fused finest inverse level for this shape, which
    makes a brick-pair field take the device-planned decode."""
    from waltz_amd import stages
    if planned:
        stages.set_tuning(shape, torch.float32 if dtype == np.float32 else torch.float64, True, 1,
                          (16, 16, 16, 16, 16))
    try:
        yield
    finally:
        if planned:
            stages.clear_tuning()


ROUTES = [False, True]
ROUTE_IDS = ["host-header", "planned"]


def _gpu_decode(wz, archive, shape, dtype):
    """Decoded array, or None when the HIP decoder rejects the archive."""
    from waltz_amd._lib import lib
    nz, ny, nx = shape
    width = np.dtype(dtype).itemsize
    buf = torch.zeros(max(len(archive), 8) + 8, dtype=torch.uint8, device="cuda")
    if archive:
        buf[:len(archive)] = torch.from_numpy(np.frombuffer(bytes(archive), np.uint8).copy()).cuda()
    out = torch.empty(shape, dtype=torch.float32 if width == 4 else torch.float64, device="cuda")
    rc = lib().waltz_amd_decompress(width, buf.data_ptr(), len(archive), nx, ny, nz, out.data_ptr(), None)
    return None if rc else out.cpu().numpy()


def _oracle_decode(oracle, archive, shape, dtype):
    from pyoracle import OracleError
    try:
        return oracle.decompress(bytes(archive), shape, dtype)
    except OracleError:
        return None


def _agree(wz, oracle, archive, shape, dtype, what):
    want = _oracle_decode(oracle, archive, shape, dtype)
    got = _gpu_decode(wz, archive, shape, dtype)
    assert (want is None) == (got is None), f"{what}: oracle {'rejects' if want is None else 'accepts'}"
    if want is not None:
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), what
    return want is not None


@pytest.mark.parametrize("planned", ROUTES, ids=ROUTE_IDS)
def test_every_truncation(wz, oracle, planned):
    shape = (16, 16, 32) if planned else (12, 14, 16)
    a = oracle.compress(field(shape, "smooth", np.float32), 1e-3)
    with route(planned, shape, np.float32):
        assert _agree(wz, oracle, a, shape, np.float32, "whole")
        for k in range(len(a)):
            assert not _agree(wz, oracle, a[:k], shape, np.float32, f"truncated to {k}")
        assert not _agree(wz, oracle, a + b"\0" * 8, shape, np.float32, "trailing bytes")


@pytest.mark.parametrize("planned", ROUTES, ids=ROUTE_IDS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_header_byte_flips(wz, oracle, dtype, planned):
    shape = (16, 16, 32) if planned else (16, 16, 16)
    a = bytearray(oracle.compress(field(shape, "outliers", dtype), 1e-2))
    accepted = 0
    with route(planned, shape, dtype):
        for pos in range(32):
            for mask in (0x01, 0x02, 0x10, 0x80, 0xFF):
                m = bytearray(a)
                m[pos] ^= mask
                accepted += _agree(wz, oracle, m, shape, dtype, f"header byte {pos} ^ {mask:#x}")
    assert accepted  # some flips (quantum low bits, the z band) still decode


@pytest.fixture(scope="module")
def two_groups(oracle):
    shape = (48, 64, 64)  # 196608 samples: 24 chunks, 2 groups
    v = field(shape, "noise", np.float32)
    return shape, bytearray(oracle.compress(v, 5e-2))


def _layout(a, n):
    chunks = (n + 8191) // 8192
    groups = (chunks + 15) // 16
    wzp = struct.unpack_from("<Q", a, 16)[0]
    header = 4 * groups + 3 * chunks
    return chunks, groups, wzp, header


@pytest.mark.parametrize("planned", ROUTES, ids=ROUTE_IDS)
def test_directory_mutations(wz, oracle, two_groups, planned):
    shape, a = two_groups
    with route(planned, shape, np.float32):
        _directory_mutations(wz, oracle, shape, a)


def _directory_mutations(wz, oracle, shape, a):
    n = int(np.prod(shape))
    chunks, groups, wzp, header = _layout(a, n)
    assert groups == 2
    payload = wzp - header
    off1 = struct.unpack_from("<I", a, 32 + 4)[0]
    assert 0 < off1 < payload
    cases = {
        "offset1+1": (1, off1 + 1), "offset1-1": (1, off1 - 1), "offset1=0": (1, 0), "offset1=payload": (1, payload),
        "offset1=max": (1, 0xFFFFFFFF), "offset0=1": (0, 1), "offset0=offset1": (0, off1),
    }
    for what, (g, value) in cases.items():
        m = bytearray(a)
        struct.pack_into("<I", m, 32 + 4 * g, value)
        assert not _agree(wz, oracle, m, shape, np.float32, what)
    # chunk metas: magnitude code and sign code of a few chunks
    for c in (0, 5, 16, chunks - 1):
        for delta in (1, -1, 0x4000):
            m = bytearray(a)
            p = 32 + 4 * groups + 3 * c
            meta = m[p] | (m[p + 1] << 8) | (m[p + 2] << 16)
            meta = (meta + delta) & 0xFFFFFF
            m[p:p + 3] = bytes([meta & 0xFF, (meta >> 8) & 0xFF, meta >> 16])
            _agree(wz, oracle, m, shape, np.float32, f"meta {c} {delta:+#x}")


@pytest.mark.parametrize("planned", ROUTES, ids=ROUTE_IDS)
def test_reordered_groups_accepted(wz, oracle, two_groups, planned):
    shape, a = two_groups
    with route(planned, shape, np.float32):
        _reordered_groups(wz, oracle, shape, a)


def _reordered_groups(wz, oracle, shape, a):
    n = int(np.prod(shape))
    chunks, groups, wzp, header = _layout(a, n)
    base = 32 + header
    off1 = struct.unpack_from("<I", a, 32 + 4)[0]
    g0, g1 = bytes(a[base:base + off1]), bytes(a[base + off1:32 + wzp])
    m = bytearray(a)
    m[base:32 + wzp] = g1 + g0  # group 1's payload first
    struct.pack_into("<I", m, 32, len(g1))
    struct.pack_into("<I", m, 36, 0)
    assert _agree(wz, oracle, m, shape, np.float32, "groups in reverse order")
    want = oracle.decompress(bytes(a), shape, np.float32)
    assert np.array_equal(_gpu_decode(wz, m, shape, np.float32).view(np.uint32), want.view(np.uint32))
    # the same payloads with overlapping spans (both at 0) are rejected
    struct.pack_into("<I", m, 32, 0)
    assert not _agree(wz, oracle, m, shape, np.float32, "overlapping groups")


@pytest.mark.parametrize("planned", ROUTES, ids=ROUTE_IDS)
def test_empty_group_offset(wz, oracle, planned):
    shape = (48, 64, 64)  # two groups, every chunk empty: payload 0, both offsets 0
    a = bytearray(oracle.compress(np.zeros(shape, np.float32), 1e-2))
    with route(planned, shape, np.float32):
        _empty_group(wz, oracle, shape, a)


def _empty_group(wz, oracle, shape, a):
    n = int(np.prod(shape))
    chunks, groups, wzp, header = _layout(a, n)
    assert wzp == header
    assert _agree(wz, oracle, a, shape, np.float32, "as written")
    for g in range(groups):
        m = bytearray(a)
        struct.pack_into("<I", m, 32 + 4 * g, 8)
        assert not _agree(wz, oracle, m, shape, np.float32, f"empty group {g} at offset 8")


@pytest.mark.parametrize("planned", ROUTES, ids=ROUTE_IDS)
def test_route_taken(wz, oracle, planned):
    """The forced launch choice does send the small brick-pair field down the
    planned decode (its header check is a stage of its own)."""
    from waltz_amd import _lib
    shape = (16, 16, 32)
    a = oracle.compress(field(shape, "smooth", np.float32), 1e-3)
    with route(planned, shape, np.float32):
        _lib.profile(True)
        try:
            _lib.profile_report()
            assert _agree(wz, oracle, a, shape, np.float32, "whole")
            stages = _lib.profile_report()
        finally:
            _lib.profile(False)
    assert ("decode_plan" in stages) == planned, sorted(stages)


# A plane-schedule field (lz = 1, lxy = 2): the planned decode with the
# inverse as fused XY levels and one Z phase, the repairs a pass of their own.
PLANE = (12, 32, 32)


def test_plane_route_taken(wz, oracle):
    from waltz_amd import _lib
    a = oracle.compress(field(PLANE, "outliers", np.float32), 1e-2)
    _lib.profile(True)
    try:
        _lib.profile_report()
        assert _agree(wz, oracle, a, PLANE, np.float32, "whole")
        stages = _lib.profile_report()
    finally:
        _lib.profile(False)
    assert "decode_plan" in stages and "patches" in stages, sorted(stages)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_plane_planned_mutations(wz, oracle, dtype):
    a = bytearray(oracle.compress(field(PLANE, "outliers", dtype), 1e-2))
    assert _agree(wz, oracle, a, PLANE, dtype, "whole")
    cuts = sorted(set(range(64)) | set(range(64, len(a), max(1, len(a) // 128))) | set(range(len(a) - 64, len(a))))
    for k in cuts:
        assert not _agree(wz, oracle, a[:k], PLANE, dtype, f"truncated to {k}")
    assert not _agree(wz, oracle, a + b"\0" * 8, PLANE, dtype, "trailing bytes")
    for pos in range(32):
        for mask in (0x01, 0x10, 0x80, 0xFF):
            m = bytearray(a)
            m[pos] ^= mask
            _agree(wz, oracle, m, PLANE, dtype, f"header byte {pos} ^ {mask:#x}")
