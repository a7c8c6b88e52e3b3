import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "kdtree-rs_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

# Several GPU tests hand PyTorch device buffers to libkdtree_hip.so.  PyTorch
# ships its own libamdhip64.so.7; when it is loaded first, the library binds to
# that same runtime (one HIP runtime per process).  Loading the library first
# would pull /opt/rocm's copy in, and PyTorch could then not initialise.
try:
    import torch

    if torch.cuda.is_available():
        torch.cuda.init()
except Exception:  # no torch or no GPU: CPU tests only
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
