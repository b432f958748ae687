import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) GPU")
    # the oracle and the library are compiled in-tree when missing (a machine
    # that only runs the tests receives them prebuilt)
    for d, out in (("oracle", "oracle/build/libktm_oracle.so"),
                   ("random_amd/csrc", "random_amd/libktm.so"),
                   ("random_amd/csrc", "random_amd/_build/eom_main")):
        if not os.path.exists(os.path.join(ROOT, out)):
            subprocess.run(["make", "-j8", "-C", os.path.join(ROOT, d)], check=True,
                           stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch.device("cuda:0")
