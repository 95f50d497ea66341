"""Waltz on one MI355X: compress / decompress throughput of a 512^3 float32
field at REL 1e-3 (plus ABS 1e-4), byte identity against the reference CPU
codec, per-stage roofline, and the reference CPU codec timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 512] [--no-cpu]

One step = one compress + one decompress of the resident field. Prints ONE
JSON line (value = compress GB/s, decimal GB, field bytes / wall time of the
synchronous API call; inputs already in HBM).
"""
import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
METRIC = "compress/decompress GB/s at 512³ fp32 REL 1e-3 (1 GPU, % HBM peak); batched vs plain"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-only", action="store_true", help="few calls, for rocprofv3 runs")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_alg_bytes(name, n, w, wzp_bytes, sizes):
    """Algorithmic HBM bytes of one pipeline stage (minimum dataflow)."""
    if name.startswith(("forward_L", "inverse_L")):
        lvl = int(name.split("_L")[1][0])
        nb = sizes[lvl]
        return 2 * w * nb  # read + write the level's box
    table = {
        "quantize": 2 * w * n + 2 * n + n / 8,      # read coeff, write dequant, mags, signs
        "wzp_encode": 2 * n + n / 8 + wzp_bytes,    # read mags+signs, write stream
        "quantize_encode": 2 * w * n + wzp_bytes,   # read coeff, write dequant + stream
        "repairs": 2 * w * n,                        # read reconstruction + original
        "wzp_decode": wzp_bytes + 2 * n + n / 8,    # read stream, write mags+signs
        "dequantize": 2 * n + n / 8 + w * n,         # read mags+signs, write coefficients
        "decode_dequantize": wzp_bytes + w * n,      # read stream, write coefficients
    }
    return table.get(name)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (each MI355X "
              "compresses its own fields independently; see DESIGN.md for sharding)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import waltz_amd
    from waltz_amd import _lib, synthetic

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    S = args.size
    shape = (S, S, S)
    n = S ** 3
    w = 4
    seed = 20261202 + 1
    t0 = time.time()
    x = synthetic.grf(shape, slope=-5.0 / 3.0, seed=seed, offset=3.0, device=dev)
    lo, hi = float(x.min()), float(x.max())
    rel = 1e-3
    bound = rel * (float(np.float32(hi)) - float(np.float32(lo)))
    log(f"field {shape} f32 GRF k^-5/3 (+3.0) seed {seed}: range [{lo:.4f}, {hi:.4f}] abs bound {bound:.6g} "
        f"({time.time() - t0:.1f}s)")
    cap = waltz_amd.max_compressed_bytes(shape, torch.float32)
    arch = torch.empty(cap, dtype=torch.uint8, device=dev)
    out = torch.empty(shape, dtype=torch.float32, device=dev)

    def compress(b):
        return waltz_amd.compress(x, abs_error=b, out=arch)

    def decompress(a):
        return waltz_amd.decompress(a, shape=shape, dtype=torch.float32, out=out)

    steps, warm = (2, 1) if args.profile_only else (args.steps, args.warmup)
    for _ in range(warm):
        decompress(compress(bound))
    torch.cuda.synchronize()
    tc, td = [], []
    t_begin = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        archive = compress(bound)
        b = time.perf_counter()
        decompress(archive)
        c = time.perf_counter()
        tc.append(b - a)
        td.append(c - b)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_begin
    A = archive.numel()
    field_bytes = n * w
    c_gbps = field_bytes / statistics.median(tc) / 1e9
    d_gbps = field_bytes / statistics.median(td) / 1e9
    log(f"compress {statistics.median(tc) * 1e3:.3f} ms ({c_gbps:.1f} GB/s), decompress "
        f"{statistics.median(td) * 1e3:.3f} ms ({d_gbps:.1f} GB/s), archive {A} B (CR {field_bytes / A:.3f})")
    if args.profile_only:
        # calibration for the HBM byte counters: one plain elementwise kernel
        # that reads 256 MiB and writes 256 MiB
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"profile_only": True, "compress_ms": statistics.median(tc) * 1e3,
                          "decompress_ms": statistics.median(td) * 1e3, "calibration_bytes": 2 * src.numel() * 4}))
        return

    # correctness: strict bound in FP64
    max_err = float((out.double() - x.double()).abs().max())
    assert max_err <= bound, (max_err, bound)

    # per-stage live timing (HIP events on the pipeline's stream), one call each
    _lib.profile(True)
    _lib.profile_report()
    decompress(compress(bound))
    stages = _lib.profile_report()
    _lib.profile(False)
    sizes, side = [], S
    for lvl in range(5):
        sizes.append(side ** 3)
        side -= side // 2
    wzp_bytes = int.from_bytes(bytes(archive[16:24].cpu().numpy()), "little")
    rows = []
    for name, (ms, cnt) in stages.items():
        ab = stage_alg_bytes(name, n, w, wzp_bytes, sizes)
        rows.append((ms / max(cnt, 1), name, ab))
    rows.sort(reverse=True)
    dom_ms, dom_name, dom_bytes = next(r for r in rows if r[2])
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 2), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(achieved * 1e9 / HBM_PEAK, 4), "traffic": None,
                "alg_bytes_per_launch": int(dom_bytes), "avg_launch_ms": round(dom_ms, 4)}
    stage_table = {name: {"ms": round(ms, 4), "alg_GBps": round(ab / (ms * 1e-3) / 1e9, 1) if ab else None}
                   for ms, name, ab in rows}
    log("stages:", json.dumps(stage_table))

    # secondary bound: ABS 1e-4
    abs2 = 1e-4
    t2c, t2d = [], []
    for _ in range(3):
        a = time.perf_counter()
        arch2 = compress(abs2)
        b = time.perf_counter()
        decompress(arch2)
        c = time.perf_counter()
        t2c.append(b - a)
        t2d.append(c - b)
    arch2_host = arch2.cpu().numpy().tobytes()
    max_err2 = float((out.double() - x.double()).abs().max())
    assert max_err2 <= abs2

    # pre-quantization coefficients in the input's precision vs an FP64 transform
    from waltz_amd import autograd, stages as stage_api
    lxy, lz, dyadic = autograd.default_levels(shape)
    c32 = stage_api.transform(x, lxy, lz, dyadic, inverse=False)
    c64 = stage_api.transform(x.double(), lxy, lz, dyadic, inverse=False)
    coeff_err = float((c32.double() - c64).abs().max())
    coeff_max = float(c64.abs().max())
    del c32, c64
    torch.cuda.empty_cache()

    # restore the REL archive for the identity check
    archive = compress(bound)
    arch_host = archive.cpu().numpy().tobytes()

    # alg bytes of the whole pipelines (compress / decompress), for % of HBM peak
    C_terms = sum(sizes[1:]) * 4 * w
    alg_c = (4 * w + 4.25) * n + C_terms + A
    alg_d = 3 * w * n + sum(sizes[1:]) * 2 * w + A
    result = {
        "metric": METRIC, "value": round(c_gbps, 2), "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": round(total / steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{S}^3 float32 Gaussian random field, isotropic k^-5/3 spectrum, unit variance, "
                               f"+3.0 offset, seed {seed}; REL 1e-3 (abs {bound:.6g}); plain single-field codec",
                   "shape_zyx": list(shape), "rel_bound": rel, "abs_bound": bound},
        "compress_GBps": round(c_gbps, 2), "decompress_GBps": round(d_gbps, 2),
        "compress_ms": round(statistics.median(tc) * 1e3, 3), "decompress_ms": round(statistics.median(td) * 1e3, 3),
        "compress_pct_hbm_alg": round(100 * alg_c / statistics.median(tc) / HBM_PEAK, 2),
        "decompress_pct_hbm_alg": round(100 * alg_d / statistics.median(td) / HBM_PEAK, 2),
        "archive_bytes": A, "compression_ratio": round(field_bytes / A, 4), "max_abs_error_fp64": max_err,
        "coefficient_max_abs_error_vs_fp64_transform": coeff_err, "coefficient_max_abs": coeff_max,
        "abs_1e-4": {"compress_GBps": round(field_bytes / statistics.median(t2c) / 1e9, 2),
                     "decompress_GBps": round(field_bytes / statistics.median(t2d) / 1e9, 2),
                     "archive_bytes": len(arch2_host), "max_abs_error_fp64": max_err2},
        "stages_ms": stage_table, "roofline": roofline,
    }

    # CPU leg: byte identity with the reference CPU codec + its timing on this host
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle
        xh = x.cpu().numpy()
        ref = pyoracle.reference_or_none()
        try:
            os.sched_setaffinity(0, {min(os.sched_getaffinity(0))})
        except (AttributeError, OSError):
            pass
        if ref is not None:
            backend = ref.AVX2 if ref.has_avx2 else ref.SCALAR
            t = time.perf_counter()
            cpu_arch = ref.compress(xh, bound, backend)
            tcpu_c = time.perf_counter() - t
            result["byte_identical_vs_reference_cpu"] = cpu_arch == arch_host
            t = time.perf_counter()
            ref.decompress(cpu_arch, shape, np.float32, backend)
            tcpu_d = time.perf_counter() - t
            cpu2 = ref.compress(xh, abs2, backend)
            result["abs_1e-4"]["byte_identical_vs_reference_cpu"] = cpu2 == arch2_host
            reps_c, reps_d = [tcpu_c], [tcpu_d]
            for _ in range(2):
                t = time.perf_counter()
                ref.compress(xh, bound, backend)
                reps_c.append(time.perf_counter() - t)
                t = time.perf_counter()
                ref.decompress(cpu_arch, shape, np.float32, backend)
                reps_d.append(time.perf_counter() - t)
            cc = field_bytes / statistics.median(reps_c) / 1e9
            result["cpu_baseline"] = {
                "value": round(cc, 4), "unit": "GB/s", "cores": 1, "kind": "reference",
                "sample": f"whole {S}^3 f32 field, REL 1e-3, reference CPU codec "
                          f"({'AVX2/FMA' if backend == ref.AVX2 else 'scalar'} backend, single-threaded, "
                          f"pinned to 1 core), median of 3 compress calls",
                "decompress_GBps": round(field_bytes / statistics.median(reps_d) / 1e9, 4),
                "gpu_over_cpu_compress": round(c_gbps / cc, 1)}
        else:
            o = pyoracle.Oracle()
            sub = np.ascontiguousarray(xh[:128, :128, :128])
            t = time.perf_counter()
            o.compress(sub, bound)
            tc1 = time.perf_counter() - t
            result["cpu_baseline"] = {"value": round(sub.nbytes / tc1 / 1e9, 4), "unit": "GB/s", "cores": 1,
                                      "kind": "port", "sample": "128^3 corner of the field, oracle C port, 1 call"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
