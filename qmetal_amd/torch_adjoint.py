"""PyTorch (ROCm) autograd over the HIP adjoint sweep.

    import torch, qmetal_amd as qm
    from qmetal_amd.torch_adjoint import AdjointEnergy
    sim = qm.Simulator(n)
    theta = torch.rand(P, dtype=torch.float64, device="cuda", requires_grad=True)
    e = AdjointEnergy.apply(theta, sim, circuit, hamiltonian)
    e.backward()            # theta.grad from the same adjoint sweep

`circuit` is a template: its RX/RY/RZ gates, in order, take their angles from
`theta` (one entry each). The tensors cross the C ABI as raw device pointers
(qm_energy_and_gradient_device); the forward call runs the fused forward pass,
the co-state build and the backward kernels, and stores dE/dtheta for the
backward call. Tensors must be contiguous float64 on the simulator's device.
"""
from __future__ import annotations

from ctypes import POINTER, byref, c_double, c_int, c_size_t, c_void_p

import torch

from . import Circuit, Hamiltonian, QMetalError, Simulator, _check, _decl, _Hamiltonian

_eg_device = _decl("qm_energy_and_gradient_device", c_int, c_void_p, c_void_p,
                   POINTER(_Hamiltonian), c_void_p, c_size_t, c_void_p, c_void_p, c_void_p)


def rotation_count(circuit: Circuit) -> int:
    return sum(1 for kind, _, _ in circuit.gates if kind in (8, 9, 10))


def _stream_handle(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def energy_and_grad(sim: Simulator, circuit: Circuit, hamiltonian: Hamiltonian,
                    params: torch.Tensor):
    """(energy, grad) as device tensors, one gradient entry per rotation."""
    if not params.is_cuda or params.dtype != torch.float64:
        raise QMetalError("params must be a float64 tensor on the GPU")
    p = params.detach().contiguous()
    energy = torch.empty((), dtype=torch.float64, device=p.device)
    grad = torch.empty_like(p)
    h = hamiltonian._c()
    _check(_eg_device(sim._h, circuit._h, byref(h), c_void_p(p.data_ptr()), p.numel(),
                      c_void_p(energy.data_ptr()), c_void_p(grad.data_ptr()),
                      c_void_p(_stream_handle(p.device))))
    return energy, grad


class AdjointEnergy(torch.autograd.Function):
    """E(theta) = <psi(theta)|H|psi(theta)> with the adjoint gradient."""

    @staticmethod
    def forward(ctx, params, sim, circuit, hamiltonian):
        energy, grad = energy_and_grad(sim, circuit, hamiltonian, params)
        ctx.save_for_backward(grad)
        return energy

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None
