"""torch.autograd wrapper of the adjoint energy gradient (kq_energy_gradient over the C-ABI).

    energy = kettle_amd.autograd.energy(theta, circuit, hamiltonian)   # theta: 1-D float64 tensor
    energy.backward()                                                   # theta.grad = dE/dtheta

theta[i] is the value of the circuit's i-th parameter in creation order (or in ``params``
order).  The forward pass runs one forward + reverse adjoint sweep on the GPU and keeps the
gradient for backward, so backward costs nothing more.  The energy is real: H is Hermitian
(real string weights).
"""
from __future__ import annotations

from typing import Sequence

import torch

from . import _lib
from .api import PauliOperator, energy_and_gradient
from .circuit import QuantumCircuit


class AdjointEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, circuit: QuantumCircuit, hamiltonian: PauliOperator,
                params: Sequence[str] | None, precision: int) -> torch.Tensor:
        order = list(params) if params is not None else circuit.parameter_order()
        values = theta.detach().to("cpu", torch.float64)
        if values.ndim != 1 or values.numel() != len(order):
            raise ValueError(f"theta must be a 1-D tensor of {len(order)} parameter values")
        for pid, v in zip(order, values.tolist()):
            circuit.set_parameter_value(pid, v)
        r = energy_and_gradient(circuit, hamiltonian, order, precision=precision)
        ctx.save_for_backward(torch.from_numpy(r.gradient.copy()).to(theta.device, theta.dtype))
        return torch.tensor(r.energy, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None, None


def energy(theta: torch.Tensor, circuit: QuantumCircuit, hamiltonian: PauliOperator,
           params: Sequence[str] | None = None, precision: int = _lib.KQ_C64) -> torch.Tensor:
    """E(theta) = <psi(theta)|H|psi(theta)> as a differentiable scalar tensor."""
    return AdjointEnergy.apply(theta, circuit, hamiltonian, params, precision)
