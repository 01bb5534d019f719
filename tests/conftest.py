import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "pito-codecs_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.load()
    return O


@pytest.fixture(scope="session")
def pcme_lib():
    import pcme
    pcme.load()
    return pcme


@pytest.fixture(scope="session")
def gpu(pcme_lib):
    """The product library on a live gfx950 device; fails loudly otherwise."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    assert pcme_lib.device_count() >= 1, "no gfx950 device visible to libpcme"
    return pcme_lib
