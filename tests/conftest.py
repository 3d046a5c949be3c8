import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and librt_gfx950.so")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import binding
    return binding.lib()


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test run without a visible GPU")
    from rustracer_amd import rt
    ctx = rt.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(autouse=True)
def _fresh_context_options(request):
    """Tests share one context: whatever options a test (or the autotuner it runs)
    set are put back to their defaults afterwards, so every test runs the settings
    it names."""
    yield
    if "gpu_ctx" in request.fixturenames:
        request.getfixturevalue("gpu_ctx").reset_options()
