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


def test_async_entry_points_match_sync():
    """nx_state_amplitudes_async / nx_state_norm_async (device outputs, no
    host wait, ordered against PyTorch's stream) give the synchronous
    entry points' values bit for bit; the inputs are produced by a PyTorch
    kernel on that stream just before the call, and the outputs are read by
    PyTorch after it, with no synchronization in between."""
    import ctypes as C
    sys.path[:0] = [ROOT, os.path.join(ROOT, "stab-decomp-simulator-rust_amd")]
    import numpy as np
    import torch
    import workloads as WL
    from necstar import QuantumState, _lib
    L = _lib.lib()
    _lib.check(L.nx_init(0))
    rng = np.random.default_rng(1601)
    n = 40
    st = QuantumState.from_circuit(WL.to_circuit(n, WL.random_circuit(rng, n, 400, 9)))
    bits = WL.random_bitstrings(rng, n, 3000)
    ref = st.amplitudes(bits)
    ref_norm = st.norm()
    dev = torch.device("cuda:0")
    for rep in range(6):
        # reps 3..5: the library enqueues on PyTorch's stream itself
        _lib.check(L.nx_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream), 1 if rep >= 3 else 0))
        host = torch.from_numpy(bits.reshape(-1).view(np.int64).copy())
        # a PyTorch kernel writes the input right before the call
        d_bits = torch.zeros(len(bits), dtype=torch.int64, device=dev)
        d_bits.add_(host.to(dev, non_blocking=False))
        d_out = torch.full((2 * len(bits),), float("nan"), dtype=torch.float64, device=dev)
        d_norm = torch.full((1,), float("nan"), dtype=torch.float64, device=dev)
        ts = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(L.nx_state_drop_cache(st._h))
        _lib.check(L.nx_state_amplitudes_async(st._h, len(bits), C.cast(d_bits.data_ptr(), C.POINTER(C.c_uint64)),
                                               1, C.cast(d_out.data_ptr(), C.POINTER(C.c_double)), ts))
        _lib.check(L.nx_state_norm_async(st._h, C.cast(d_norm.data_ptr(), C.POINTER(C.c_double)), ts))
        got = d_out.view(torch.complex128).cpu().numpy()  # PyTorch reads on its stream
        assert np.array_equal(got, ref), rep
        assert float(d_norm.item()) == ref_norm, rep
    _lib.check(L.nx_set_stream(None, 0))
    assert st.norm() == ref_norm
