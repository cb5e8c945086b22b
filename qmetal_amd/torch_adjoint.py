"""PyTorch (ROCm) autograd over the HIP adjoint sweep.

    import torch, qmetal_amd as qm
    from qmetal_amd.torch_adjoint import AdjointEnergy
    sim = qm.Simulator(n)
    theta = torch.rand(P, dtype=torch.float64, device="cuda", requires_grad=True)
    e = AdjointEnergy.apply(theta, sim, circuit, hamiltonian)
    e.backward()            # theta.grad from the same adjoint sweep

`circuit` is a template: its RX/RY/RZ gates, in order, take their angles from
`theta` (one entry each). The tensors cross the C ABI as raw device pointers
(qm_energy_and_gradient_device); when theta requires a gradient, the forward
call runs the fused forward pass, the co-state build and the backward
kernels, and stores dE/dtheta for the backward call; otherwise (or through
`adjoint_energy` with gradients disabled) it runs the forward pass and the
energy reductions only. From 12 qubits the first call for a circuit template
compiles a device program and later calls only launch kernels on the current
stream (angles are written on the device from theta), so they can be captured
in a CUDA/HIP graph; below 12 qubits the angles are read back to the host and
the call waits for the stream. Tensors must be contiguous float64 on the
simulator's device.
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
                    params: torch.Tensor, need_grad: bool = True):
    """(energy, grad) as device tensors, one gradient entry per rotation;
    grad is None when need_grad is False (no co-state, no backward sweep)."""
    if not params.is_cuda or params.dtype != torch.float64:
        raise QMetalError("params must be a float64 tensor on the GPU")
    p = params.detach().contiguous()
    energy = torch.empty((), dtype=torch.float64, device=p.device)
    grad = torch.empty_like(p) if need_grad else None
    h = hamiltonian._c()
    _check(_eg_device(sim._h, circuit._h, byref(h), c_void_p(p.data_ptr()), p.numel(),
                      c_void_p(energy.data_ptr()),
                      c_void_p(grad.data_ptr() if grad is not None else None),
                      c_void_p(_stream_handle(p.device))))
    return energy, grad


class AdjointEnergy(torch.autograd.Function):
    """E(theta) = <psi(theta)|H|psi(theta)> with the adjoint gradient."""

    @staticmethod
    def forward(ctx, params, sim, circuit, hamiltonian):
        need = bool(ctx.needs_input_grad[0])
        energy, grad = energy_and_grad(sim, circuit, hamiltonian, params, need_grad=need)
        if need:
            ctx.save_for_backward(grad)
        return energy

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None


def adjoint_energy(params: torch.Tensor, sim: Simulator, circuit: Circuit,
                   hamiltonian: Hamiltonian) -> torch.Tensor:
    """AdjointEnergy.apply, or the energy alone (no backward sweep) when
    gradients are disabled or params does not require one."""
    if torch.is_grad_enabled() and params.requires_grad:
        return AdjointEnergy.apply(params, sim, circuit, hamiltonian)
    energy, _ = energy_and_grad(sim, circuit, hamiltonian, params, need_grad=False)
    return energy
