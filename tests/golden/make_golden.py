"""Writes tests/golden/reference_kats.json.

Every value here is data the reference's own tests hold (transcribed from the
cited test file), a published standard test vector, or a closed form that the
mathematics fixes.  No reference code is imported or run.

Run: python tests/golden/make_golden.py
"""
import json
import math
import os

# gate op codes (include/necstar_hip.h)
H, X, Y, Z, S, SDG, SQRTX, SQRTXDG, CX, CZ, SWAP, T, TDG, CCX = range(14)

s2 = 1 / math.sqrt(2)


def cpx(z):
    return [z.real, z.imag]


def t_tensor(t):
    """(|0> + e^{i pi/4}|1>)/sqrt2 tensored t times; little-endian index
    (crates/necstar/src/state/magic_states/t_state.rs:82-121 builds the same
    vector with kron(v_{t-1}, v_1))."""
    one = [complex(s2, 0), complex(0.5, 0.5)]
    vec = [1.0 + 0j]
    for _ in range(t):
        vec = [a * b for a in vec for b in one]
    return vec


TOFFOLI = [(H, 0), (H, 1), (H, 2), (CX, 1, 2), (TDG, 2), (CX, 0, 2), (T, 2), (CX, 1, 2),
           (TDG, 2), (CX, 0, 2), (T, 1), (T, 2), (CX, 0, 1), (H, 2), (T, 0), (TDG, 1),
           (CX, 0, 1)]

kats = {
    # tests/test_state_initialization.py:5-22
    "h_t_cx": {
        "n": 2, "gates": [(H, 0), (T, 0), (CX, 0, 1)], "rank": 2,
        "statevector": [cpx(complex(s2, 0)), cpx(0j), cpx(0j), cpx(complex(0.5, 0.5))],
        "norm": 1.0, "str": "QuantumState(num_qubits=2, stabilizer_rank=2)",
    },
    # tests/test_state_initialization.py:25-31 (T on qubit 3 of a 2-qubit circuit)
    "invalid_t_qubit": {"n": 2, "gates": [(H, 0), (T, 3), (CX, 0, 1)]},
    # tests/test_state_calculation.py:6-60, tests/test_state_simulation.py
    "toffoli": {
        "n": 3, "gates": TOFFOLI, "exp_ZII": 0.5, "self_inner_product": 1.0, "norm": 1.0,
        "sample_outcomes": ["000", "100", "010", "111"],
        "sample_outcomes_reversed": ["000", "001", "010", "111"],
        # crates/necstar/src/state/stabilizer_decomposed_state/statevector.rs:33-49
        # (the same |Toffoli> state, little-endian)
        "statevector": [cpx(complex(v, 0)) for v in [0.5, 0.5, 0.5, 0, 0, 0, 0, 0.5]],
    },
    # crates/stabilizer-ch-form-rust/tests/gate_api.rs:7-35
    "hadamard": {"n": 1, "gates": [(H, 0)], "statevector": [cpx(complex(s2, 0))] * 2},
    "bell": {"n": 2, "gates": [(H, 0), (CX, 0, 1)],
             "statevector": [cpx(complex(s2, 0)), cpx(0j), cpx(0j), cpx(complex(s2, 0))]},
    # t_state.rs:82-121: |T>^{(x)t} for t = 1..11, reached here through the
    # compiler with H then T on every qubit
    "t_tensor": {str(t): [cpx(z) for z in t_tensor(t)] for t in range(1, 12)},
    # ChaCha block function known answers (key = 0, nonce = 0, counter = 0):
    # RFC 7539 section 2.3.2-style vector for 20 rounds (the ChaCha20 test
    # vector used by rand_chacha's own tests), and the 12-round vector of the
    # same inputs (Strombergson's ChaCha test vectors, TC1 256-bit key).
    "chacha20_zero_block": [0xade0b876, 0x903df1a0, 0xe56a5d40, 0x28bd8653, 0xb819d2bd,
                            0x1aed8da0, 0xccef36a8, 0xc70d778b, 0x7c5941da, 0x8d485751,
                            0x3fe02477, 0x374ad8b8, 0xf4b8436a, 0x1ca11815, 0x69b687c3,
                            0x8665eeb2],
    "chacha20_zero_block1": [0xbee7079f, 0x7a385155, 0x7c97ba98, 0x0d082d73, 0xa0290fcb,
                             0x6965e348, 0x3e53c612, 0xed7aee32, 0x7621b729, 0x434ee69c,
                             0xb03371d5, 0xd539d874, 0x281fed31, 0x45fb0a51, 0x1f0ae1ac,
                             0x6f4d794b],
    "chacha12_zero_bytes": "9bf49a6a0755f953811fce125f2683d50429c3bb49e074147e0089a52eae155f"
                           "0564f879d27ae3c02ce82834acfa8c793a629f2ca0de6919610be82f411326be",
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
    with open(out, "w") as f:
        json.dump(kats, f, indent=1)
    print("wrote", out)
