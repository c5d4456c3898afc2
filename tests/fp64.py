"""FP64 comparison of GPU sums against the oracle's.

Every value compared here is a sum of exact CH-form terms (coefficient x
zeta^k 2^{-r/2}); the GPU and the oracle add the same rounded terms in
different orders.  The bound is relative:

    |gpu - oracle| <= 1e-12 * max(|oracle|, 1e-3 * S)

with S the sum of the terms' magnitudes (for an amplitude,
|g| sum_i |c_i| |<x|phi_i>|; for a pair sum, |g_a g_b| sum_ij |c_i| |c_j|
|<phi_i|phi_j>|, the pair magnitudes taken from the exact pair tables).  The
floor only matters for sums that cancel below 1e-3 of their terms' scale: an
FP64 sum of those terms (the oracle's included) carries rounding errors of
order 1e-16 S, so it does not resolve such a value to 1e-12 relative.
`worst` keeps the largest relative error met, for the test log.
"""
import numpy as np

REL = 1e-12
FLOOR = 1e-3

worst = {"rel": 0.0, "where": ""}


def pair_mags(tab):
    """Magnitudes 2^{-r/2} of exact pair scalars (uint32 k | r << 8, 0xFFFFFFFF = zero)."""
    tab = np.asarray(tab, np.uint32)
    r = (tab >> np.uint32(8)).astype(np.float64)
    return np.where(tab != np.uint32(0xFFFFFFFF), 2.0 ** (-r / 2.0), 0.0)


def pair_scale(tab, ca, cb, ga=1.0, gb=1.0):
    """|g_a g_b| sum_ij |c_a,i| |c_b,j| |<a_i|b_j>|."""
    m = pair_mags(tab)
    return abs(ga) * abs(gb) * float(np.abs(ca) @ m @ np.abs(cb))


def state_scale(a, b):
    """Term scale of a.inner_product(b) (and of a.norm()**2 when b is a)."""
    return pair_scale(a.cross_pair_table(b), a.coefficients(), b.coefficients(), a.global_factor, b.global_factor)


def pauli_image(st, ops):
    """A copy of st with the dense Pauli ops (0 I, 1 X, 2 Y, 3 Z, little-endian)
    applied as gates."""
    from necstar import QuantumGate
    img = st.clone()
    gates = [(QuantumGate.x, QuantumGate.y, QuantumGate.z)[o - 1](q) for q, o in enumerate(ops) if o]
    if gates:
        img.apply_gates(gates)
    return img


def exp_scale(st, ops):
    """Term scale of st.exp_value(P): |g|^2 sum_ij |c_i| |c_j| |<phi_i|P phi_j>|."""
    return state_scale(st, pauli_image(st, ops)) if st.stabilizer_rank else 0.0


def amp_scale(o, per_term, g=1.0):
    """|g| sum_i |c_i| 2^{-r_i/2} over the nonzero terms of an oracle
    amplitude (per_term = (zero flags, k, r) from oracle State.amplitude)."""
    z, _, r = per_term
    c = np.abs(o.coeffs())
    return abs(g) * float(np.sum(np.where(z, 0.0, c * 2.0 ** (-np.asarray(r, np.float64) / 2.0))))


def check(got, ref, scale, where=""):
    """Asserts the bound above elementwise; returns the largest relative error."""
    got = np.atleast_1d(np.asarray(got))
    ref = np.atleast_1d(np.asarray(ref))
    scale = np.broadcast_to(np.asarray(scale, np.float64), ref.shape)
    err = np.abs(got - ref)
    bound = REL * np.maximum(np.abs(ref), FLOOR * scale)
    bad = np.nonzero(~(err <= bound))[0]
    assert bad.size == 0, (where, [(int(i), complex(got[i]), complex(ref[i]), float(scale[i])) for i in bad[:4]])
    # the largest relative error where the bound is purely relative
    strict = (np.abs(ref) > 0) & (np.abs(ref) >= FLOOR * scale)
    rel = float(np.max(err[strict] / np.abs(ref[strict]))) if np.any(strict) else 0.0
    if rel > worst["rel"]:
        worst["rel"], worst["where"] = rel, where
    return rel


def check_lazy(got, ref, scale_fn, where=""):
    """check() where the term scale is computed (scale_fn(i)) only for the
    entries that a purely relative 1e-12 bound does not already accept."""
    got = np.atleast_1d(np.asarray(got))
    ref = np.atleast_1d(np.asarray(ref))
    scale = np.zeros(ref.shape, np.float64)
    for i in np.nonzero(~(np.abs(got - ref) <= REL * np.abs(ref)))[0]:
        scale[i] = scale_fn(int(i))
    return check(got, ref, scale, where)
