import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Some GPU tests hold device buffers in PyTorch.  libsfmm.so and PyTorch-ROCm
# must share one HIP runtime: PyTorch's copy is found by the dynamic loader
# only when torch is loaded before libsfmm.so (both carry the soname
# libamdhip64.so.7; loaded the other way round, the process holds two runtimes
# and PyTorch sees no GPU).
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libsfmm.so")
