"""PyTorch autograd binding of the adjoint gate-fidelity gradient.

``gate_fidelity(context, controls, dt, target)`` returns F = |Tr(target^H U)|^2 / dim^2 for
the propagator U that ``context.equiprop(dt, *controls)`` computes, as a differentiable
0-d tensor: its backward is the library's adjoint kernel (Parament_gateFidelityGradientDevice),
which returns dF/dc for every control sample in one call.  PyTorch supplies only device
memory, the stream and the autograd graph; all arithmetic runs in libparament.

The forward pass of a tensor that requires grad runs the gradient kernel right away and
keeps the result for backward (the gradient costs about three forward passes; a separate
backward would have to redo the forward).  Without grad it runs equiprop only.
"""
from __future__ import annotations

import torch

__all__ = ["GateFidelity", "gate_fidelity"]


def _dtypes(context):
    if context._use_doubles:
        return torch.float64, torch.complex128
    return torch.float32, torch.complex64


class GateFidelity(torch.autograd.Function):
    """F(controls) with the adjoint gradient as backward.

    controls: real (amps, pts) tensor on the GPU (float64 for an 'fp64' context, float32 for
    'fp32'); target: (dim, dim) complex tensor on the same device.  The context must hold a
    Hamiltonian with at least ``amps`` controls.
    """

    @staticmethod
    def forward(ctx, controls, context, dt, target):
        real_t, cplx_t = _dtypes(context)
        if controls.dim() != 2 or controls.dtype != real_t or not controls.is_cuda:
            raise ValueError(f"controls must be an (amps, pts) {real_t} tensor on the GPU")
        if target.dtype != cplx_t or target.shape != (context.dim, context.dim) or target.device != controls.device:
            raise ValueError(f"target must be a ({context.dim}, {context.dim}) {cplx_t} tensor on the controls' device")
        amps, pts = controls.shape
        carr = torch.complex(controls.detach(), torch.zeros_like(controls)).contiguous()
        tgt = target.contiguous()
        context.set_stream(torch.cuda.current_stream(controls.device).cuda_stream)
        if ctx.needs_input_grad[0]:
            fid = torch.empty(1, dtype=real_t, device=controls.device)
            grad = torch.empty((amps, pts), dtype=real_t, device=controls.device)
            context.gate_fidelity_gradient_device(dt, carr.data_ptr(), pts, amps, tgt.data_ptr(), fid.data_ptr(),
                                                  grad.data_ptr())
            ctx.save_for_backward(grad)
            return fid[0].clone()
        u = torch.empty((context.dim, context.dim), dtype=cplx_t, device=controls.device)
        context.equiprop_device(dt, carr.data_ptr(), pts, amps, u.data_ptr())
        g = torch.sum(tgt.conj() * u)
        return (g.real ** 2 + g.imag ** 2) / float(context.dim) ** 2

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None


def gate_fidelity(context, controls, dt, target):
    """Differentiable gate fidelity |Tr(target^H U(controls))|^2 / dim^2 (see GateFidelity)."""
    return GateFidelity.apply(controls, context, dt, target)
