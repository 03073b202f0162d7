"""torch.autograd binding of the adjoint gradient (C ABI fsa_adjoint_gradient).

    e = expectation(thetas, circuit, psi0, coeffs, paulis)
    e.backward()          # thetas.grad = d<H>/d theta from one reverse sweep

``thetas`` is a 1-D float64 (or float32) tensor with one angle per
pauliGadget / phaseGadget of ``circuit``, in recording order; the gadgets are
exp(+i theta P) (reference src/distributed_statevector.hpp:287-298), so the
gradient is d/dtheta <psi0| U(theta)^dag H U(theta) |psi0>.  The forward pass
runs the whole adjoint on the GPU (the state, the observable and the reverse
sweep stay in HBM); backward only scales the stored gradient.  The state's
precision (FP64 or FP32 amplitudes) selects the kernels; sums are FP64.
"""
from __future__ import annotations

import numpy as np
import torch

from . import adjoint_gradient


class AdjointExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, thetas, circuit, psi0, coeffs, paulis):
        angles = thetas.detach().to("cpu", torch.float64).numpy()
        if angles.ndim != 1 or angles.size != circuit.num_params:
            raise ValueError(f"thetas must hold {circuit.num_params} angles")
        circuit.set_params(np.ascontiguousarray(angles))
        expec, grads = adjoint_gradient(circuit, psi0, coeffs, paulis)
        ctx.save_for_backward(torch.as_tensor(grads, dtype=thetas.dtype, device=thetas.device))
        return torch.tensor(expec, dtype=thetas.dtype, device=thetas.device)

    @staticmethod
    def backward(ctx, grad_out):
        (grads,) = ctx.saved_tensors
        return grad_out * grads, None, None, None, None


def expectation(thetas, circuit, psi0, coeffs, paulis):
    """<psi0| U(thetas)^dag H U(thetas) |psi0> with H = sum_t coeffs[t] P_t
    (``paulis``: numTerms x numQubits codes, 0=I 1=X 2=Y 3=Z), differentiable
    in ``thetas`` through the adjoint method."""
    return AdjointExpectation.apply(thetas, circuit, psi0, coeffs, paulis)
