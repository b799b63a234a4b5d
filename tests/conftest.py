import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def lib():
    import pcaone_amd as pa
    return pa.lib()


@pytest.fixture(scope="session")
def gpu():
    import pcaone_amd as pa
    if pa.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests must run on the GPU host")
    g = pa.Gpu(0)
    yield g
    g.close()
    pa.shutdown()


def golden_cases():
    """The PCA fixtures: those holding the reference's eigenvalues and its CLI
    arguments (the LD and inbreeding fixtures have their own tests)."""
    out = []
    for p in sorted(GOLDEN.glob("*.npz")):
        with np.load(p, allow_pickle=False) as z:
            if "eigvals" in z.files and "args" in z.files:
                out.append(p)
    return out


def load_golden(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}
