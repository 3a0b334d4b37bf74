"""Test configuration: the `gpu` marker and the in-tree libraries."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # build the in-tree libraries only if nothing is built yet
    if not (ROOT / "oracle" / "liboracle.so").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "liboracle.so"], check=True)
    if not (ROOT / "efem_amd" / "libefem_gpu.so").exists():
        subprocess.run([sys.executable, str(ROOT / "efem_amd" / "_build.py")], check=True)


@pytest.fixture(scope="session")
def rng():
    import numpy as np
    return np.random.default_rng(20240531)
