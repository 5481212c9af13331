import os
import subprocess
import sys

import numpy as np
import pytest

# Keep the brick path under test on small meshes: bricks for batches of any
# size (the library builds them from 4096 cells), and the plan-time tuner may
# not trade them for the per-cell kernel.
os.environ.setdefault("GMF_BRICK_MIN_CELLS", "0")
os.environ.setdefault("GMF_KEEP_BRICKS", "1")

# Parallel test workers (pytest -n): one BLAS thread each, or every worker's
# BLAS pool spins on all cores and the small dense solves of the tests crawl.
if "PYTEST_XDIST_WORKER" in os.environ:
    for _name in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS"):
        os.environ.setdefault(_name, "1")
    try:
        from threadpoolctl import threadpool_limits

        threadpool_limits(1)
    except ImportError:
        pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "grico.jl_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _ensure_built():
    lib = os.path.join(PKG, "lib")
    if not os.path.exists(os.path.join(lib, "libgrico_host.so")):
        subprocess.run([sys.executable, os.path.join(PKG, "build.py"), "host"], check=True)
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        import pyoracle
        pyoracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch
