"""nx_state_norm_noalloc: the allocation-free variant of nx_state_norm.

Called through the C-ABI on two states in a row (the second with fewer terms
and qubits), with exact Scalar and with complex128 coefficients: the result
equals nx_state_norm's bit for bit (same kernels, launch shape and summation
order), the library's allocation counter does not move during the calls,
and a state too large for the scratch array is refused with a clear error.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import workloads as WL

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    from necstar import _lib
    lib = _lib.lib()
    _lib.check(lib.nx_init(0))
    return lib


def allocations(L):
    from necstar import _lib
    ms, cnt = C.c_double(), C.c_int64()
    _lib.check(L.nx_profile_get(b"allocations", C.byref(ms), C.byref(cnt)))
    return cnt.value


def norms(L, st):
    from necstar import _lib
    a, b = C.c_double(), C.c_double()
    _lib.check(L.nx_state_norm(st._h, C.byref(a)))
    before = allocations(L)
    _lib.check(L.nx_state_norm_noalloc(st._h, C.byref(b)))
    assert allocations(L) == before, "nx_state_norm_noalloc allocated"
    return a.value, b.value


def test_norm_noalloc_equals_norm(L):
    import necstar as nx
    cfg = WL.config(1)
    big = nx.QuantumState.from_circuit(WL.to_circuit(cfg["n"], cfg["gates"]))
    rng = np.random.default_rng(7001)
    small = nx.QuantumState.from_circuit(WL.to_circuit(20, WL.random_circuit(rng, 20, 200, 6)))
    assert small.stabilizer_rank < big.stabilizer_rank
    say(f"noalloc: chi {big.stabilizer_rank} (n=64), then chi {small.stabilizer_rank} (n=20)")
    for st in (big, small):  # two calls in a row, the second smaller
        ref, got = norms(L, st)
        assert got == ref, (got, ref)
        assert abs(got - 1.0) <= 1e-12
    # complex128 coefficients (StabilizerDecomposedState<Complex64>)
    c = small.coefficients() * (1.0 + 1e-3 * (rng.standard_normal(small.stabilizer_rank)
                                               + 1j * rng.standard_normal(small.stabilizer_rank)))
    small.set_coefficients(c)
    ref, got = norms(L, small)
    assert got == ref and abs(got - 1.0) > 1e-9
    # two words per row
    wide = nx.QuantumState.from_circuit(WL.to_circuit(100, WL.random_circuit(rng, 100, 800, 8)))
    ref, got = norms(L, wide)
    assert got == ref
    # a state with no terms: norm 0
    z = nx.QuantumState.from_circuit(nx.QuantumCircuit(3))
    z.project_unnormalized(0, True)
    assert norms(L, z) == (0.0, 0.0)


def test_norm_noalloc_refuses_large_states(L):
    import necstar as nx
    from necstar import _lib
    rng = np.random.default_rng(7002)
    # 8,748 terms on two-word rows need ~73 MB of scratch (limit 64 MiB)
    st = nx.QuantumState.from_circuit(WL.to_circuit(100, WL.random_circuit(rng, 100, 3000, 32)))
    say(f"noalloc limit: chi {st.stabilizer_rank} at n=100")
    assert st.stabilizer_rank * 8392 > (64 << 20)
    out = C.c_double(-1.0)
    code = L.nx_state_norm_noalloc(st._h, C.byref(out))
    assert code == _lib.NX_E_VALUE
    assert "scratch" in L.nx_last_error().decode()
    assert out.value == -1.0
    # the library still works, and the scratch still serves a small state next
    small = nx.QuantumState.from_circuit(WL.to_circuit(12, WL.random_circuit(rng, 12, 100, 5)))
    ref, got = norms(L, small)
    assert got == ref
