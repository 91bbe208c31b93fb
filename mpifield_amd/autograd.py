"""Differentiable gather and scatter: torch.autograd Functions whose backward passes are the
adjoint kernels.

The gather of a weighting is the transpose of its scatter (weighting.py:17-20), so

- gather(x, F): dL/dF = scatter(x, dL/dg); dL/dx_d = sum_c dL/dg_c * (d/dx_d gather(x, F))_c,
  the derivative-spline gather (mpf_gather_derivative);
- scatter(x, v): dL/dv = gather(x, dL/drho); dL/dx_d = sum_c v_c * (d/dx_d gather(x, dL/drho))_c.

Positions, values and fields are float64 or float32 GPU tensors; the field is float64.
The deposits in the backward use the current deposit mode (kernels.deposit_mode).
"""

from __future__ import annotations

import torch

from mpifield_amd import _lib
from mpifield_amd import kernels as K
from mpifield_amd.binning import kind_of


def gather_derivative(kind: int, degree: int, positions: torch.Tensor, field: torch.Tensor, geometry) -> torch.Tensor:
    """d/dx of the interpolant at the markers: (ndim, ncomp, N) in the positions' dtype."""
    ndim, lower, cell, num, code, _ = K._split(geometry)
    field = K._field(field)
    pos, ps = K.marker_rows(positions, ndim)
    ncomp = field.shape[-1]
    out = torch.empty((ndim, ncomp, pos.shape[1]), dtype=pos.dtype, device=field.device)
    with torch.cuda.device(field.device):
        _lib.call("mpf_gather_derivative", kind, degree, ndim, pos.data_ptr(), ps, K._dtype_code(pos), pos.shape[1],
                  field.data_ptr(), ncomp, out.data_ptr(), pos.shape[1], lower.ctypes.data, cell.ctypes.data,
                  num.ctypes.data, code.ctypes.data, None, K.stream_handle(field.device))
    return out


def _shape(geometry, ncomp: int) -> tuple[int, ...]:
    ndim = len(geometry) // 4
    return (*(int(v) for v in geometry[2 * ndim : 3 * ndim]), ncomp)


class GatherFunction(torch.autograd.Function):
    """g = gather(positions, field) with the adjoint backward."""

    @staticmethod
    def forward(ctx, positions, field, kind, degree, geometry):
        result = torch.empty((field.shape[-1], positions.shape[1]), dtype=positions.dtype, device=field.device)
        K.gather(kind, degree, positions, field.contiguous(), result, geometry)
        ctx.save_for_backward(positions, field)
        ctx.args = (kind, degree, geometry)
        return result

    @staticmethod
    def backward(ctx, grad):
        positions, field = ctx.saved_tensors
        kind, degree, geometry = ctx.args
        grad = grad.contiguous()
        grad_pos = grad_field = None
        if ctx.needs_input_grad[1]:
            grad_field = torch.zeros_like(field)
            K.scatter(kind, degree, positions.detach(), grad.to(positions.dtype), grad_field, geometry)
        if ctx.needs_input_grad[0]:
            d = gather_derivative(kind, degree, positions.detach(), field.detach().contiguous(), geometry)
            grad_pos = (d * grad.to(d.dtype).unsqueeze(0)).sum(dim=1)
        return grad_pos, grad_field, None, None, None


class ScatterFunction(torch.autograd.Function):
    """rho = scatter(positions, values) into a zero field, with the adjoint backward."""

    @staticmethod
    def forward(ctx, positions, values, kind, degree, geometry):
        field = torch.zeros(_shape(geometry, values.shape[0]), dtype=torch.float64, device=values.device)
        K.scatter(kind, degree, positions, values, field, geometry)
        ctx.save_for_backward(positions, values)
        ctx.args = (kind, degree, geometry)
        return field

    @staticmethod
    def backward(ctx, grad):
        positions, values = ctx.saved_tensors
        kind, degree, geometry = ctx.args
        grad = grad.contiguous().to(torch.float64)
        grad_pos = grad_val = None
        if ctx.needs_input_grad[1]:
            grad_val = torch.empty_like(values)
            K.gather(kind, degree, positions.detach(), grad, grad_val, geometry)
        if ctx.needs_input_grad[0]:
            d = gather_derivative(kind, degree, positions.detach(), grad, geometry)
            grad_pos = (d * values.detach().to(d.dtype).unsqueeze(0)).sum(dim=1)
            # a dead marker (all values zero) deposits nothing: no gradient through it
        return grad_pos, grad_val, None, None, None


def gather(positions: torch.Tensor, field: torch.Tensor, weighting, grid) -> torch.Tensor:
    """Differentiable ``weighting`` gather of ``field`` (num_nodes + (ncomp,)) at ``positions`` (ndim, N)."""
    kind, degree = kind_of(weighting)
    return GatherFunction.apply(positions, field, kind, degree, tuple(grid.geometry))


def scatter(positions: torch.Tensor, values: torch.Tensor, weighting, grid) -> torch.Tensor:
    """Differentiable ``weighting`` deposit of ``values`` (ncomp, N) at ``positions`` into a new zero field."""
    kind, degree = kind_of(weighting)
    return ScatterFunction.apply(positions, values, kind, degree, tuple(grid.geometry))

