"""Write tests/golden/codec_golden.json (and a few whole archives) from the
REFERENCE CPU codec built by ``make -C oracle ref`` from the reference's own
sources (cpu/src/*.cpp). Inputs come from tests/fields.py (integer arithmetic
only), so the fixture pins the reference's bytes, not this project's oracle.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))
from pyoracle import Reference  # noqa: E402
from fields import field  # noqa: E402

OUT = Path(__file__).resolve().parent

# (nz, ny, nx) shapes: dyadic, plane (lxy != lz), edge bricks, non-power-of-two
# z bands (96x64x48 -> BZ 6), thin and degenerate axes.
SHAPES = [
    (1, 1, 1), (4, 3, 2), (5, 9, 7), (9, 11, 13), (16, 16, 16), (17, 19, 15),
    (8, 64, 64), (33, 5, 17), (3, 40, 24), (32, 32, 32), (48, 64, 96), (19, 17, 65),
    (64, 24, 24), (2, 130, 18), (40, 40, 40),
]
CASES = []
for shape in SHAPES:
    for pattern in ("smooth", "noise", "checker", "outliers", "zero", "constant", "ramp"):
        for dtype in (np.float32, np.float64):
            bounds = (1e-3, 0.05) if pattern in ("smooth", "noise") else (1e-3,)
            for bound in bounds:
                CASES.append((shape, pattern, dtype, bound))
CASES.append(((8, 64, 64), "smooth", np.float64, 1e-7))
CASES.append(((16, 16, 16), "noise", np.float64, 1e-9))
# Five Z levels with nz >= 481 is the only way to reach BZ 16 (the 512^3 layout).
CASES.append(((496, 136, 130), "smooth", np.float32, 1e-3))

# Whole archives kept as binary fixtures (small ones) for device decode tests.
KEEP = {((16, 16, 16), "smooth", "float32", 1e-3), ((17, 19, 15), "noise", "float64", 1e-3),
        ((16, 16, 16), "outliers", "float32", 1e-3), ((48, 64, 96), "smooth", "float32", 0.05),
        ((33, 5, 17), "checker", "float64", 1e-3)}


def header(archive):
    b = archive[:32]
    return {
        "flags": b[0], "levels_xy": b[1], "levels_z": b[2], "bz": b[3],
        "wzp_bytes": int.from_bytes(b[16:24], "little"),
        "n_pwe": int.from_bytes(b[24:28], "little"), "n_qo": int.from_bytes(b[28:32], "little"),
        "hex": b.hex(),
    }


def main():
    ref = Reference()
    records = []
    arch_dir = OUT / "archives"
    arch_dir.mkdir(exist_ok=True)
    for shape, pattern, dtype, bound in CASES:
        v = field(shape, pattern, dtype)
        archive = ref.compress(v, bound, ref.SCALAR)
        if ref.has_avx2:
            assert archive == ref.compress(v, bound, ref.AVX2)
        dec = ref.decompress(archive, shape, dtype, ref.SCALAR)
        dname = np.dtype(dtype).name
        rec = {
            "shape": list(shape), "pattern": pattern, "dtype": dname, "bound": bound,
            "archive_bytes": len(archive), "archive_sha256": hashlib.sha256(archive).hexdigest(),
            "decoded_sha256": hashlib.sha256(dec.tobytes()).hexdigest(),
            "input_sha256": hashlib.sha256(v.tobytes()).hexdigest(),
            "header": header(archive),
        }
        if (shape, pattern, dname, bound) in KEEP:
            name = "%s_%s_%s_%g.bin" % ("x".join(map(str, shape)), pattern, dname, bound)
            (arch_dir / name).write_bytes(archive)
            rec["archive_file"] = "archives/" + name
        records.append(rec)
    (OUT / "codec_golden.json").write_text(json.dumps(records, indent=1) + "\n")
    print(len(records), "cases;",
          sum(r["header"]["n_qo"] > 0 for r in records), "with q-outliers;",
          sum(r["header"]["n_pwe"] > 0 for r in records), "with repairs;",
          sum(not (r["header"]["flags"] & 1) for r in records), "plane schedule")


if __name__ == "__main__":
    main()
