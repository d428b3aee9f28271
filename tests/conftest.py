import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


# torch bundles its own HIP/HSA runtime (torch/lib/libamdhip64.so); libsegtaco
# binds to whichever libamdhip64.so.7 is already loaded. Import torch before
# anything loads libsegtaco so that one process holds one HIP runtime.
try:
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is plumbing for the GPU tests only
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (run with -m gpu)")
