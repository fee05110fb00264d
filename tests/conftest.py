import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU-side check (run with MW_RUN_SLOW=1)")


def pytest_collection_modifyitems(config, items):
    """slow tests (minutes of numpy on the full 8x8 array) run on request:
    MW_RUN_SLOW=1; the same golden vectors are checked on the GPU path."""
    if os.environ.get("MW_RUN_SLOW"):
        return
    skip = pytest.mark.skip(reason="slow CPU check; set MW_RUN_SLOW=1 to run it")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)


def golden(name):
    import numpy as np

    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


def rel_err(a, b):
    import numpy as np

    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="session")
def ref_accel():
    """The reference's compiled accelerator from oracle/_ref, if built."""
    import glob
    import importlib.util

    paths = glob.glob(os.path.join(ROOT, "oracle", "_ref", "_accelerators*.so"))
    if not paths:
        pytest.skip("oracle/_ref not built (make -C oracle ref)")
    spec = importlib.util.spec_from_file_location("_accelerators", paths[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod
