"""Host-side logic of the package checked on the CPU (no GPU calls)."""
import pytest

torch = pytest.importorskip("torch")


def test_default_levels_match_oracle(oracle):
    from waltz_amd import autograd
    for n in list(range(1, 300)) + [511, 512, 513, 1000, 4096]:
        assert autograd.dwt_levels(n) == oracle.levels(n), n
    for shape in [(512, 512, 512), (96, 384, 384), (40, 36, 44), (1, 64, 64), (7, 130, 9)]:
        nz, ny, nx = shape
        lx, lz = oracle.levels(min(nx, ny)), oracle.levels(nz)
        assert autograd.default_levels(shape) == (lx, lz, lx == lz)
    assert autograd.default_levels((64, 48)) == (oracle.levels(48), 0, oracle.levels(48) == 0)


def test_bench_refuses_several_gpus():
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr
