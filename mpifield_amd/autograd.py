"""Differentiable gather and scatter: torch.autograd Functions whose backward passes are the
adjoint kernels.

The gather of a weighting is the transpose of its scatter (weighting.py:17-20), so

- gather(x, F): dL/dF = scatter(x, dL/dg); dL/dx_d = sum_c dL/dg_c * (d/dx_d gather(x, F))_c,
  the derivative-spline gather (mpf_gather_derivative);
- scatter(x, v): dL/dv = gather(x, dL/drho); dL/dx_d = sum_c v_c * (d/dx_d gather(x, dL/drho))_c.

Positions, values and fields are float64 or float32 GPU tensors; the field is float64.
The deposits in the backward use the current deposit mode (kernels.deposit_mode).

Large 3D batches on the tiled shapes (kernels.tiled_ok) bin their positions once: the bins
of the last few position tensors are kept (by tensor identity and version), so that a
gather of a scatter at the same positions, and both backward passes, sort the markers
once and use the tiled deposit, gather and gradient kernels.
"""

from __future__ import annotations

import weakref

import torch

from mpifield_amd import _lib
from mpifield_amd import binning
from mpifield_amd import kernels as K
from mpifield_amd.binning import kind_of

#: (weakref to the positions, their version, kind, degree, geometry, bins), newest last
_BINS: list = []
_BINS_KEPT = 2


def positions_bins(positions: torch.Tensor, kind: int, degree: int, geometry, ncomp: int):
    """Bins of ``positions`` (made from positions alone) for the tiled kernels, or None when
    the batch is small or the shape has no tiled kernel.  Reused while the same tensor is
    unchanged (same object, same version counter)."""
    ndim, _, _, num, code, _ = K._split(geometry)
    if ndim != 3 or positions.shape[1] < K.BINNED_MIN_MARKERS or not K.tiled_ok(kind, degree, ncomp, num, code):
        return None
    key = (kind, degree, tuple(geometry))
    for ref, version, k, bins in _BINS:
        if ref() is positions and version == positions._version and k == key:
            return bins
    bins = binning.bin_markers(positions.detach(), None, (kind, degree), geometry).prepare_items()
    _BINS.append((weakref.ref(positions), positions._version, key, bins))
    del _BINS[:-_BINS_KEPT]
    return bins


def _binned_deposit(bins, values: torch.Tensor, field: torch.Tensor) -> bool:
    """Deposit values with the tiled kernels when the mode is fixed point (their only mode)."""
    if bins is None or K.current_mode() != "fixed":
        return False
    binning.scatter_binned(binning.with_values(bins, values), field)
    return True


def _binned_gather(bins, field: torch.Tensor, dtype) -> torch.Tensor:
    out = torch.empty((field.shape[-1], bins.n), dtype=dtype, device=field.device)
    binning.gather_binned(bins, field, out)
    return bins.unsort(out)


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
        field_c = field.contiguous()
        bins = positions_bins(positions, kind, degree, geometry, field.shape[-1])
        if bins is not None:
            result = _binned_gather(bins, field_c, positions.dtype)
        else:
            result = torch.empty((field.shape[-1], positions.shape[1]), dtype=positions.dtype, device=field.device)
            K.gather(kind, degree, positions, field_c, result, geometry)
        ctx.save_for_backward(positions, field)
        ctx.args = (kind, degree, geometry)
        ctx.bins = bins
        return result

    @staticmethod
    def backward(ctx, grad):
        positions, field = ctx.saved_tensors
        kind, degree, geometry = ctx.args
        bins = ctx.bins
        grad = grad.contiguous()
        grad_pos = grad_field = None
        if ctx.needs_input_grad[1]:
            grad_field = torch.zeros_like(field)
            if not _binned_deposit(bins, grad.to(positions.dtype), grad_field):
                K.scatter(kind, degree, positions.detach(), grad.to(positions.dtype), grad_field, geometry)
        if ctx.needs_input_grad[0]:
            if bins is not None:
                grad_pos = binning.gather_derivative_binned(bins, field.detach(), grad.to(positions.dtype))
            else:
                d = gather_derivative(kind, degree, positions.detach(), field.detach().contiguous(), geometry)
                grad_pos = (d * grad.to(d.dtype).unsqueeze(0)).sum(dim=1)
        return grad_pos, grad_field, None, None, None


class ScatterFunction(torch.autograd.Function):
    """rho = scatter(positions, values) into a zero field, with the adjoint backward."""

    @staticmethod
    def forward(ctx, positions, values, kind, degree, geometry):
        field = torch.zeros(_shape(geometry, values.shape[0]), dtype=torch.float64, device=values.device)
        bins = positions_bins(positions, kind, degree, geometry, values.shape[0])
        if not _binned_deposit(bins, values, field):
            K.scatter(kind, degree, positions, values, field, geometry)
        ctx.save_for_backward(positions, values)
        ctx.args = (kind, degree, geometry)
        ctx.bins = bins
        return field

    @staticmethod
    def backward(ctx, grad):
        positions, values = ctx.saved_tensors
        kind, degree, geometry = ctx.args
        bins = ctx.bins
        grad = grad.contiguous().to(torch.float64)
        grad_pos = grad_val = None
        if ctx.needs_input_grad[1]:
            if bins is not None:
                grad_val = _binned_gather(bins, grad, positions.dtype).to(values.dtype)
            else:
                grad_val = torch.empty_like(values)
                K.gather(kind, degree, positions.detach(), grad, grad_val, geometry)
        if ctx.needs_input_grad[0]:
            # a dead marker (all values zero) deposits nothing: no gradient through it
            if bins is not None:
                grad_pos = binning.gather_derivative_binned(bins, grad, values.detach().to(positions.dtype))
            else:
                d = gather_derivative(kind, degree, positions.detach(), grad, geometry)
                grad_pos = (d * values.detach().to(d.dtype).unsqueeze(0)).sum(dim=1)
        return grad_pos, grad_val, None, None, None


def gather(positions: torch.Tensor, field: torch.Tensor, weighting, grid) -> torch.Tensor:
    """Differentiable ``weighting`` gather of ``field`` (num_nodes + (ncomp,)) at ``positions`` (ndim, N)."""
    kind, degree = kind_of(weighting)
    return GatherFunction.apply(positions, field, kind, degree, tuple(grid.geometry))


def scatter(positions: torch.Tensor, values: torch.Tensor, weighting, grid) -> torch.Tensor:
    """Differentiable ``weighting`` deposit of ``values`` (ncomp, N) at ``positions`` into a new zero field."""
    kind, degree = kind_of(weighting)
    return ScatterFunction.apply(positions, values, kind, degree, tuple(grid.geometry))

