"""The drop-in boundary driven on the GPU, compared with the oracle.

- The `waltz` command-line tool (bin/waltz, C++ over include/waltz.hpp):
  REL resolved with WALTZ::d_extrema (examples/waltz.cu:50-55 of the
  reference), then WALTZ::d_compress / d_decompress; archive bytes and decoded
  bits equal the oracle's.
- The serial loader ABI exactly as python/pywaltz/gpu_dataloader.py:41-47
  binds it (waltz_hip_create / decode / error / destroy) and the CUDA-style
  chunk decoder waltz_cuda_decode_chunk (src/waltz.cu:836-865) on a torch
  stream, with the reference's argument rule (out_bytes must equal the
  field's bytes) and its canonical-archive rule.
- Overlapping input and archive buffers are rejected (cpu/src/codec.cpp:82-86,
  src/waltz.cu:346-349).
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from fields import field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "waltz_amd" / "bin" / "waltz"
LIB = ROOT / "waltz_amd" / "libwaltz.so"


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    waltz_amd.lib()
    return waltz_amd


def _cli(*args, timeout=120):
    r = subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


def _rel_abs(v, rel):
    """examples/waltz.cu:50-55: abs = rel * (double(max) - double(min))."""
    return rel * (float(v.max()) - float(v.min()))


@pytest.mark.parametrize("dtype,flag,shape,pattern", [
    (np.float32, "-f", (40, 48, 56), "smooth"),
    (np.float64, "-d", (24, 33, 20), "noise"),
    (np.float32, "-f", (64, 64, 64), "outliers"),
], ids=["f32-smooth", "f64-noise", "f32-outliers"])
def test_cli_rel_matches_oracle(wz, oracle, tmp_path, dtype, flag, shape, pattern):
    nz, ny, nx = shape
    v = field(shape, pattern, dtype)
    raw = tmp_path / "in.raw"
    raw.write_bytes(v.tobytes())
    arc, dec = tmp_path / "f.waltz", tmp_path / "f.out"
    rc, out, err = _cli(flag, "-i", raw, "-z", arc, "-3", nx, ny, nz, "-M", "REL", "1e-3")
    assert rc == 0, err
    bound = _rel_abs(v, 1e-3)
    assert f"ABS {bound:.17g}" in out, out  # d_extrema gave the exact extrema
    ref = oracle.compress(v, bound)
    assert arc.read_bytes() == ref
    rc, out, err = _cli(flag, "-z", arc, "-o", dec, "-3", nx, ny, nz)
    assert rc == 0, err
    got = np.frombuffer(dec.read_bytes(), dtype=dtype)
    want = oracle.decompress(ref, shape, dtype).reshape(-1)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    assert np.max(np.abs(got.astype(np.float64) - v.reshape(-1).astype(np.float64))) <= bound


def test_cli_compress_decode_verify(wz, oracle, tmp_path):
    shape = (32, 40, 48)
    v = field(shape, "ramp", np.float32)
    raw, dec = tmp_path / "in.f32", tmp_path / "out.f32"
    raw.write_bytes(v.tobytes())
    env = dict(os.environ, WALTZ_WARMUP="1", WALTZ_REPORT_CE2E="1")
    r = subprocess.run([str(CLI), "-f", "-i", str(raw), "-o", str(dec), "-3", "48", "40", "32", "-M", "ABS", "0.01"],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    assert "[WALTZ-CE2E] compress_ms=" in r.stdout and "[WALTZ-CE2E] decompress_ms=" in r.stdout
    max_e = float(r.stdout.split("Max_E  = ")[1].split(",")[0])
    assert max_e <= 0.01
    got = np.frombuffer(dec.read_bytes(), np.float32)
    want = oracle.decompress(oracle.compress(v, 0.01), shape, np.float32).reshape(-1)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_cli_argument_errors(wz, tmp_path):
    raw = tmp_path / "in.f32"
    raw.write_bytes(np.zeros(8 * 8 * 8, np.float32).tobytes())
    assert _cli()[0] != 0
    assert _cli("-h")[0] == 0
    rc, _, err = _cli("-f", "-i", raw, "-z", tmp_path / "a", "-3", 8, 8, 1, "-M", "ABS", "0.1")
    assert rc != 0 and "3D dimensions" in err  # a singleton axis collapses (Config's dimension rule)
    rc, _, err = _cli("-f", "-i", raw, "-z", tmp_path / "a", "-3", 8, 8, 9, "-M", "ABS", "0.1")
    assert rc != 0 and "File size" in err
    rc, _, err = _cli("-f", "-i", raw, "-z", tmp_path / "a", "-3", 8, 8, 8, "-M", "ABS", "-1")
    assert rc != 0


def _loader_binding():
    """The serial HIP binding of gpu_dataloader.py:41-47 on a fresh handle."""
    lib = C.CDLL(str(LIB))
    p, sz = C.c_void_p, C.c_size_t

    class HipConfig(C.Structure):
        _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32), ("width", C.c_int),
                    ("bound", C.c_double)]

    def bind(name, res, args):
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
        return fn

    bind("waltz_hip_create", p, [C.c_int])
    bind("waltz_hip_destroy", None, [p])
    bind("waltz_hip_error", C.c_char_p, [p])
    bind("waltz_hip_decode", C.c_int, [p, C.POINTER(HipConfig), p, sz, p, sz])
    bind("waltz_cuda_decode_chunk", C.c_char_p, [C.c_uint32] * 3 + [C.c_int, p, sz, p, sz, p])
    return lib, HipConfig


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_hip_session_abi_matches_oracle(wz, oracle, dtype):
    lib, HipConfig = _loader_binding()
    shapes = [(16, 24, 40), (32, 32, 32), (7, 9, 11)]
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    width = np.dtype(dtype).itemsize
    # archives packed back to back in one device buffer, as the loader's frames are
    arcs = [oracle.compress(field(s, "smooth", dtype, seed=k + 3), 1e-3) for k, s in enumerate(shapes)]
    offs = np.cumsum([0] + [(len(a) + 15) & ~15 for a in arcs])
    packed = np.zeros(int(offs[-1]), np.uint8)
    for o, a in zip(offs, arcs):
        packed[o:o + len(a)] = np.frombuffer(a, np.uint8)
    dev = torch.from_numpy(packed).cuda()
    session = lib.waltz_hip_create(torch.cuda.current_device())
    assert session
    try:
        for o, a, s in zip(offs, arcs, shapes):
            nz, ny, nx = s
            cfg = HipConfig(nx, ny, nz, width, 1e-3)
            out = torch.empty(s, dtype=tdt, device="cuda")
            rc = lib.waltz_hip_decode(session, C.byref(cfg), dev.data_ptr() + int(o), len(a), out.data_ptr(),
                                      out.numel() * width)
            assert rc == 0, lib.waltz_hip_error(session)
            want = oracle.decompress(a, s, dtype)
            assert np.array_equal(out.cpu().numpy().view(np.uint8), want.view(np.uint8))
        # a truncated archive fails with a message in the session
        nz, ny, nx = shapes[0]
        cfg = HipConfig(nx, ny, nz, width, 1e-3)
        out = torch.empty(shapes[0], dtype=tdt, device="cuda")
        rc = lib.waltz_hip_decode(session, C.byref(cfg), dev.data_ptr(), len(arcs[0]) - 8, out.data_ptr(),
                                  out.numel() * width)
        assert rc == -1 and b"trailing" in lib.waltz_hip_error(session)
    finally:
        lib.waltz_hip_destroy(session)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_cuda_decode_chunk_matches_oracle(wz, oracle, dtype):
    lib, _ = _loader_binding()
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    width = np.dtype(dtype).itemsize
    s = (48, 64, 96)
    nz, ny, nx = s
    v = field(s, "noise", dtype)
    a = oracle.compress(v, 2e-3)
    arc = torch.from_numpy(np.frombuffer(a, np.uint8).copy()).cuda()
    stream = torch.cuda.Stream()
    out = torch.full(s, float("nan"), dtype=tdt, device="cuda")
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        err = lib.waltz_cuda_decode_chunk(nx, ny, nz, width, arc.data_ptr(), len(a), out.data_ptr(),
                                          out.numel() * width, stream.cuda_stream)
    assert err is None, err
    stream.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint8), oracle.decompress(a, s, dtype).view(np.uint8))
    # the reference's argument rule: out_bytes must equal the field's bytes
    big = torch.empty(out.numel() + 4, dtype=tdt, device="cuda")
    err = lib.waltz_cuda_decode_chunk(nx, ny, nz, width, arc.data_ptr(), len(a), big.data_ptr(),
                                      big.numel() * width, stream.cuda_stream)
    assert err is not None and b"invalid chunk dimensions" in err
    for bad in [(0, ny, nz, width), (nx, ny, nz, 2)]:
        err = lib.waltz_cuda_decode_chunk(*bad, arc.data_ptr(), len(a), out.data_ptr(), out.numel() * width,
                                          stream.cuda_stream)
        assert err is not None and b"invalid chunk dimensions" in err
    # a canonical archive only: wrong length, or a header other than compress writes
    err = lib.waltz_cuda_decode_chunk(nx, ny, nz, width, arc.data_ptr(), len(a) + 8, out.data_ptr(),
                                      out.numel() * width, stream.cuda_stream)
    assert err is not None and b"canonical" in err
    mutated = bytearray(a)
    mutated[3] ^= 1  # the z band
    arc2 = torch.from_numpy(np.frombuffer(bytes(mutated), np.uint8).copy()).cuda()
    err = lib.waltz_cuda_decode_chunk(nx, ny, nz, width, arc2.data_ptr(), len(a), out.data_ptr(),
                                      out.numel() * width, stream.cuda_stream)
    assert err is not None and b"canonical" in err
    # the general decoder takes a non-canonical band exactly when the oracle does
    # (cpu/src/codec.cpp:281-283) and then decodes the same bits
    from pyoracle import OracleError
    try:
        want = oracle.decompress(bytes(mutated), s, dtype)
    except OracleError:
        want = None
    if want is None:
        with pytest.raises(wz.WaltzError):
            wz.decompress(arc2, shape=s, dtype=tdt)
    else:
        got = wz.decompress(arc2, shape=s, dtype=tdt).cpu().numpy()
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_overlapping_buffers_rejected(wz):
    from waltz_amd._lib import lib as L
    s = (16, 16, 16)
    n = 16 * 16 * 16
    cap = wz.max_compressed_bytes(s, torch.float32)
    buf = torch.zeros(cap + n * 4, dtype=torch.uint8, device="cuda")
    written = C.c_size_t()
    # the archive starts inside the input field
    rc = L().waltz_amd_compress(4, buf.data_ptr(), 16, 16, 16, 0.1, buf.data_ptr() + 1024, cap, C.byref(written),
                                None)
    assert rc == -1 and b"overlap" in L().waltz_amd_last_error()
    # decoding into the archive's own bytes
    v = torch.from_numpy(field(s, "smooth", np.float32)).cuda()
    arc = wz.compress(v, abs_error=1e-2)
    host = torch.zeros(arc.numel() + n * 4, dtype=torch.uint8, device="cuda")
    host[:arc.numel()] = arc
    rc = L().waltz_amd_decompress(4, host.data_ptr(), arc.numel(), 16, 16, 16, host.data_ptr() + 8, None)
    assert rc == -1 and b"overlap" in L().waltz_amd_last_error()
