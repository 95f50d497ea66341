import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    """The reference CPU codec compiled by `make -C oracle ref`; skipped when absent."""
    from pyoracle import reference_or_none
    ref = reference_or_none()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return ref


@pytest.fixture(scope="session")
def golden():
    import json
    return json.loads((ROOT / "tests" / "golden" / "codec_golden.json").read_text())


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
