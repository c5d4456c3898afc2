"""torch.autograd.Function for Pauli-expectation losses of the decomposition
coefficients.

    L(c) = sum_k w_k <psi(c)| P_k |psi(c)>,  psi(c) = g sum_i c_i |phi_i>

(the unnormalised expectation of crates/necstar/src/state/stabilizer_decomposed_state/exp_value.rs:7-61,
with the coefficients as complex128 leaves).  One device pass over the pairs
(nx_state_grad_exp_values_device) gives L and its gradient
2 |g|^2 sum_k w_k G_k c together; the forward keeps the gradient for the
backward, so the backward does no further device work.  CUDA tensors stay on
the device (the library enqueues on PyTorch's current stream); CPU tensors
go through the host entry point.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib


class ExpValueLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coeffs, state, paulis, weights):
        paulis = list(paulis)
        w = np.ascontiguousarray(np.asarray(weights, np.float64))
        if w.shape != (len(paulis),):
            raise ValueError("one weight per Pauli string is required")
        chi = state.stabilizer_rank
        if coeffs.shape != (chi,):
            raise ValueError("one coefficient per stabilizer term is required")
        if coeffs.is_cuda:
            c = coeffs.detach().to(torch.complex128).contiguous()
            loss = torch.zeros(1, dtype=torch.float64, device=coeffs.device)
            grad = torch.zeros(chi, dtype=torch.complex128, device=coeffs.device)
            rows = state._dense_ops(paulis)
            stream = C.c_void_p(torch.cuda.current_stream(coeffs.device).cuda_stream)
            P = C.POINTER(C.c_double)
            check(lib().nx_state_grad_exp_values_device(
                state._h, len(paulis), rows.ctypes.data_as(C.POINTER(C.c_int32)), w.ctypes.data_as(P),
                C.cast(c.data_ptr(), P), C.cast(loss.data_ptr(), P), C.cast(grad.data_ptr(), P), stream))
            loss = loss.reshape(())
        else:
            work = state.clone()
            work.set_coefficients(coeffs.detach().to(torch.complex128).numpy())
            lv, g = work.grad_exp_values(paulis, w)
            loss = torch.tensor(lv, dtype=torch.float64)
            grad = torch.from_numpy(g)
        ctx.save_for_backward(grad)
        return loss

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad * grad_out.to(torch.float64), None, None, None


def exp_value_loss(coeffs, state, paulis, weights):
    """L(c), with its gradient from the same device pass."""
    return ExpValueLoss.apply(coeffs, state, paulis, weights)
