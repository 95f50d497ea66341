"""CPU check of the device bit helpers (BIT_2 transpose, bit compress/expand,
RZE byte flags) against brute force, via a host build of bits.hpp."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.parametrize("cxx", ["g++", "/opt/rocm/llvm/bin/clang++"])
def test_bit_helpers(tmp_path, cxx):
    if shutil.which(cxx) is None and not Path(cxx).exists():
        pytest.skip(cxx + " not available")
    exe = tmp_path / "bits"
    subprocess.run([cxx, "-O2", "-std=c++17", str(HERE / "test_bits_host.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
