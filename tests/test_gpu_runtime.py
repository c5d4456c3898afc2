"""The library and PyTorch-ROCm share one HIP runtime whatever the import
order (PyTorch ships its own runtime; necstar imports torch, when present,
before it loads libnecstar_hip.so).  Run in a child process so that the
order is the first thing that process does."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_loaded_before_torch_use():
    code = (
        "import sys; sys.path[:0] = [%r, %r]\n"
        "from necstar import _lib, QuantumCircuit, QuantumState\n"
        "L = _lib.lib(); _lib.check(L.nx_init(0))\n"
        "qc = QuantumCircuit(2); qc.apply_h(0); qc.apply_t(0); qc.apply_cx(0, 1)\n"
        "st = QuantumState.from_circuit(qc)\n"
        "import torch\n"
        "x = torch.ones(4, device='cuda')\n"
        "print('ok', float(x.sum()), round(st.norm(), 12), flush=True)\n"
    ) % (ROOT, os.path.join(ROOT, "stab-decomp-simulator-rust_amd"))
    print("child: necstar first, then torch on cuda:0", flush=True)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "ok 4.0 1.0" in r.stdout
