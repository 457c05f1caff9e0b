import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running GPU test")
    # One HIP runtime per process.  The PyTorch wheel ships its own
    # libamdhip64 with the same soname as ROCm's: when libffb_hip.so is loaded
    # first it brings ROCm's copy, and torch's GPU initialisation later in the
    # same process fails ("No HIP GPUs are available"); when torch is imported
    # first the library binds to torch's copy and both work.  The GPU tests
    # that pass torch tensors run in the same pytest process as the others, so
    # a session that may run GPU tests imports torch before any test loads the
    # library (the rule the torch bridge, fastfortranboltzmann_amd.adjoint,
    # follows by importing torch at import).
    expr = (config.getoption("markexpr") or "").replace(" ", "")
    if "notgpu" not in expr:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
