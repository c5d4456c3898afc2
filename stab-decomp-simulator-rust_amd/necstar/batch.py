"""Batched norms and Pauli expectation values over many states (API
addition: the reference evaluates QuantumState.norm and exp_value one call
at a time, crates/necstar/src/state/stabilizer_decomposed_state/
{norm.rs:8-35, exp_value.rs:7-61}), and batched sampling.

    b = Batch([st0, st1, ...], [[P, Q, ...], [...], ...], autotune=True, graph=True)
    norms, values = b.run()      # norms[i], values[i][k] = st_i.exp_value(P_k)
    counts = sample_many([st0, st1, ...], [qargs0, qargs1, ...], [shots0, ...], [seed0, ...])
    # counts[i] == st_i.sample(qargs_i, shots_i, seed_i, exact=True)

Each state is copied when the batch is created.  All pairs of all states
and all Pauli strings of a state run in one persistent pair launch per
row-word width; with graph=True the launches are captured once into a HIP
graph and every run() replays it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, loaded
from .gate import _usize
from .state import parse_seed

NX_BATCH_AUTOTUNE = 1
NX_BATCH_GRAPH = 2


class Batch:
    def __init__(self, states, paulis, autotune: bool = True, graph: bool = True):
        states = list(states)
        paulis = [list(p) for p in paulis]
        if len(paulis) != len(states):
            raise ValueError("one Pauli list per state is required")
        rows = [st._dense_ops(pl)[: len(pl)] for st, pl in zip(states, paulis)]
        flat = np.concatenate([r.reshape(-1) for r in rows] + [np.zeros(1, np.int32)]).astype(np.int32)
        npauli = np.array([len(p) for p in paulis] + [0], np.int32)
        handles = (C.c_void_p * max(len(states), 1))(*[st._h for st in states])
        flags = (NX_BATCH_AUTOTUNE if autotune else 0) | (NX_BATCH_GRAPH if graph else 0)
        h = C.c_void_p()
        check(lib().nx_batch_create(len(states), handles, npauli.ctypes.data_as(C.POINTER(C.c_int32)),
                                    flat.ctypes.data_as(C.POINTER(C.c_int32)), flags, C.byref(h)))
        self._h = h
        self._counts = [len(p) for p in paulis]
        nout = C.c_int64()
        check(lib().nx_batch_num_outputs(self._h, C.byref(nout)))
        self._nout = nout.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                L = loaded()
            except Exception:  # interpreter shutdown: module state is gone
                return
            if L is not None:
                L.nx_batch_free(h)

    @property
    def shape(self):
        """(pair-chunk length, total waves, variant words) of the launch plan."""
        J, waves, vw = C.c_int32(), C.c_int64(), C.c_int32()
        check(lib().nx_batch_shape(self._h, C.byref(J), C.byref(waves), C.byref(vw)))
        return J.value, waves.value, vw.value

    def run_flat(self) -> np.ndarray:
        """Raw output: for each state, its norm then its expectation values."""
        out = np.zeros(max(self._nout, 1), np.float64)
        check(lib().nx_batch_run(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[: self._nout]

    def run(self):
        flat = self.run_flat()
        norms, values, pos = [], [], 0
        for k in self._counts:
            norms.append(float(flat[pos]))
            values.append(flat[pos + 1: pos + 1 + k].copy())
            pos += 1 + k
        return norms, values


def sample_many(states, qargs, shots, seeds=None, stats=None):
    """QuantumState.sample(qargs[i], shots[i], seeds[i], exact=True) for
    every state i in one call (sampling.rs:13-189 per state): each state's
    draws are those of its own call; the windows the states wait for at one
    time are evaluated in one launch plan.  Returns one outcome dict per
    state.  stats: a dict that receives the rounds, window passes and exact
    tails of the call, and the seconds spent evaluating rounds (round_s)
    and waiting for the states' host-side DFS steps (wait_s) and, of
    round_s, for the rounds' launches to finish (gpu_wait_s)."""
    states = list(states)
    ns = len(states)
    qargs = [[_usize(q) for q in qa] for qa in qargs]
    shots = [_usize(x) for x in shots]
    if len(qargs) != ns or len(shots) != ns:
        raise ValueError("one qubit list and one shot count per state are required")
    if seeds is None:
        seeds = [None] * ns
    if len(seeds) != ns:
        raise ValueError("one seed per state is required")
    import os
    seed_bytes = b"".join(parse_seed(sd) if sd is not None else os.urandom(32) for sd in seeds)
    for qa in qargs:
        if any(q >= 2 ** 31 for q in qa):
            raise OverflowError("qubit index out of range")
    nq = np.array([len(qa) for qa in qargs] + [0], np.int32)
    flat = np.array([q for qa in qargs for q in qa] + [0], np.int32)
    sh = np.array(shots + [0], np.uint64)
    sb = (C.c_uint8 * max(len(seed_bytes), 1)).from_buffer_copy(seed_bytes or bytes(1))
    handles = (C.c_void_p * max(ns, 1))(*[st._h for st in states])
    cap = sum(min(x, 1 << min(len(qa), 20)) for x, qa in zip(shots, qargs)) + 1
    nout = np.zeros(max(ns, 1), np.int64)
    st5 = np.zeros(6, np.int64)
    for _attempt in range(2):
        bits = np.zeros(2 * cap, np.uint64)
        counts = np.zeros(cap, np.uint64)
        nout[:] = 0
        code = lib().nx_batch_sample_exact(
            ns, handles, nq.ctypes.data_as(C.POINTER(C.c_int32)), flat.ctypes.data_as(C.POINTER(C.c_int32)),
            sh.ctypes.data_as(C.POINTER(C.c_uint64)), C.cast(sb, C.POINTER(C.c_uint8)), cap,
            bits.ctypes.data_as(C.POINTER(C.c_uint64)), counts.ctypes.data_as(C.POINTER(C.c_uint64)),
            nout.ctypes.data_as(C.POINTER(C.c_int64)), st5.ctypes.data_as(C.POINTER(C.c_int64)))
        need = int(nout[:ns].sum())
        if code == 0 or need <= cap:
            break
        cap = need
    check(code)
    if stats is not None:
        stats.update(rounds=int(st5[0]), windows=int(st5[1]), tails=int(st5[2]), round_s=st5[3] * 1e-6,
                     wait_s=st5[4] * 1e-6, gpu_wait_s=st5[5] * 1e-6)
    out, o = [], 0
    for i in range(ns):
        k = len(qargs[i])
        res = {}
        for m in range(int(nout[i])):
            word = int(bits[2 * (o + m)]) | (int(bits[2 * (o + m) + 1]) << 64)
            res["".join("1" if (word >> j) & 1 else "0" for j in range(k))] = int(counts[o + m])
        out.append(res)
        o += int(nout[i])
    return out
