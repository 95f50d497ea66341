"""Build libwaltz.so (all HIP kernels + host pipelines + C/C++ ABI) for gfx950,
and the `waltz` command-line tool (bin/waltz) over its C++ API.

    python -m waltz_amd.build          # or waltz_amd.build.build()

Every translation unit is compiled with floating-point contraction off, no
fast-math and IEEE denormals (bitwise parity with the reference CPU codec),
and with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libwaltz.so"
ARCH = os.environ.get("WALTZ_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (Path(c).exists() or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = OBJ / (src.stem + ".o")
    cmd = [_hipcc(), *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"compile failed: {src.name}\n{r.stderr[-6000:]}")
    return obj


def build(verbose=False):
    """Compile every HIP source and link libwaltz.so (always rebuilds)."""
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("link failed\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_cli(verbose)
    return LIB


CLI_SRC = PKG / "cli" / "waltz.cpp"
CLI = PKG / "bin" / "waltz"


def build_cli(verbose=False):
    """The `waltz` command-line tool (host C++ over the C++ API, linked to libwaltz.so)."""
    CLI.parent.mkdir(exist_ok=True)
    rocm = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
    tmp = CLI.with_suffix(".tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}", f"-I{rocm / 'include'}",
           str(CLI_SRC), "-o", str(tmp), f"-L{PKG}", "-lwaltz", f"-L{rocm / 'lib'}", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{rocm / 'lib'}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("CLI build failed\n" + r.stderr[-4000:])
    os.replace(tmp, CLI)
    if verbose:
        print("built", CLI)
    return CLI


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
