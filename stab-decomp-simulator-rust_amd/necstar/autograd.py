"""torch.autograd.Function for Pauli-expectation losses of the decomposition
coefficients.

    L(c) = sum_k w_k <psi(c)| P_k |psi(c)>,  psi(c) = g sum_i c_i |phi_i>

(the unnormalised expectation of crates/necstar/src/state/stabilizer_decomposed_state/exp_value.rs:7-61,
with the coefficients as complex128 leaves).  The forward is the pair kernel's
Hermitian sum per Pauli; the backward is the conjugate Gram matvec
2 |g|^2 sum_k w_k G_k c on exact pair scalars (nx_state_grad_exp_values).
torch only carries the coefficient tensor; the state tables stay in the HIP
library.
"""
from __future__ import annotations

import numpy as np
import torch


class ExpValueLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coeffs, state, paulis, weights):
        c = coeffs.detach().to("cpu", torch.complex128).numpy()
        work = state.clone()
        work.set_coefficients(c)
        w = np.asarray(weights, np.float64)
        ev = work.exp_values(paulis)
        ctx.work, ctx.paulis, ctx.weights = work, list(paulis), w
        ctx.device = coeffs.device
        return torch.tensor(float(np.dot(w, ev)), dtype=torch.float64, device=coeffs.device)

    @staticmethod
    def backward(ctx, grad_out):
        _, grad = ctx.work.grad_exp_values(ctx.paulis, ctx.weights)
        g = torch.from_numpy(grad).to(ctx.device)
        return g * grad_out.to(torch.float64), None, None, None


def exp_value_loss(coeffs, state, paulis, weights):
    """L(c) with a backward through the HIP Gram kernel."""
    return ExpValueLoss.apply(coeffs, state, paulis, weights)
