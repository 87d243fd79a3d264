import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running (full-size workloads)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    pyoracle.load()
    threads = int(os.environ.get("DEMSO_ORACLE_THREADS", "0") or 0)
    if threads > 0:
        pyoracle.set_threads(threads)
    return pyoracle


@pytest.fixture(scope="session")
def gpu_ctx():
    """One HIP context for the GPU tests (fails loudly without the library)."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from demso_amd import api

    ctx = api.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(autouse=True)
def _default_step_variant(request):
    """Tests that change the shared context's kernel variants leave it as
    they found it (every variant gives the same results; this keeps each
    test's timing and coverage independent of the order)."""
    yield
    if "gpu_ctx" in request.fixturenames:
        request.getfixturevalue("gpu_ctx").set_step_variant()
