import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Compile-only run (no GPU needed): BITGEN_PREPARE_ONLY=1 pytest -m gpu
# compiles every code object the GPU tests load into the kernel cache and
# stops each test at its first device call, so that a later GPU run finds
# them cached.  Outcomes of such a run mean nothing.
PREPARE_ONLY = os.environ.get("BITGEN_PREPARE_ONLY") == "1"


class PrepareOnlyStop(Exception):
    """Raised at the first device call of a compile-only run."""


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs under -m gpu on the GPU box)")
    if PREPARE_ONLY:
        from bitgen_amd.backend.hip import runtime as rt

        def stop(*_a, **_k):
            raise PrepareOnlyStop("compile-only run")

        rt.init = stop


@pytest.fixture(scope="session")
def gpu():
    import torch

    if PREPARE_ONLY:
        return torch.device("cuda", 0)
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bitgen_amd.backend.hip import runtime as rt

    rt.init(0)
    return torch.device("cuda", 0)
