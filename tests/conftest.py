import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "stab-decomp-simulator-rust_amd")
for p in (ROOT, PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def pytest_terminal_summary(terminalreporter):
    """The largest relative FP64 error the GPU-vs-oracle checks met."""
    fp64 = sys.modules.get("fp64")
    if fp64 is not None and fp64.worst["where"]:
        terminalreporter.write_line(f"fp64 worst relative error {fp64.worst['rel']:.3e} ({fp64.worst['where']})")
