"""Differentiable gather and scatter: torch.autograd Functions whose backward passes are the
adjoint kernels.

The gather of a weighting is the transpose of its scatter (weighting.py:17-20), so

- gather(x, F): dL/dF = scatter(x, dL/dg); dL/dx_d = sum_c dL/dg_c * (d/dx_d gather(x, F))_c,
  the derivative-spline gather (mpf_gather_derivative);
- scatter(x, v): dL/dv = gather(x, dL/drho); dL/dx_d = sum_c v_c * (d/dx_d gather(x, dL/drho))_c.

Positions, values and fields are float64 or float32 GPU tensors; the field is float64.
The deposits in the backward use the current deposit mode (kernels.deposit_mode).

Large 3D batches on the tiled shapes (kernels.binned_route) bin their positions and use the
tiled deposit, gather and gradient kernels; a Function keeps its bins from the forward pass
for its backward pass.  To share one binning between several Functions (a gather of a
scatter at the same positions), make the bins once with ``bin_positions`` and pass them as
``bins=``: the caller owns them and promises the positions are unchanged while they are in
use.  Nothing is reused implicitly, so a position tensor written in place by any route (an
in-place op, ``.data``, a pointer handed to a particle pusher) is binned afresh by the next
call that is not given bins.
"""

from __future__ import annotations

import torch

from mpifield_amd import _lib
from mpifield_amd import binning
from mpifield_amd import kernels as K
from mpifield_amd.binning import kind_of


def _positions_bins(positions: torch.Tensor, kind: int, degree: int, geometry, ncomp: int):
    """Fresh bins of ``positions`` (made from positions alone, with their work lists) for the
    tiled kernels, or None when the batch or the shape has no tiled kernel."""
    ndim, _, _, num, code, _ = K._split(geometry)
    if (ndim != 3 or not K.binned_route(kind, degree, ncomp, num, code, positions.shape[1])
            or not K.derivative_tiled_ok(kind, degree, ncomp, num, code)):
        return None
    return binning.bin_markers(positions.detach(), None, (kind, degree), geometry).prepare_items()


def bin_positions(positions: torch.Tensor, weighting, grid, ncomp: int = 1):
    """Bins of ``positions`` for ``weighting`` on ``grid``, to share between the differentiable
    gather and scatter at these positions (``bins=``), or None when the batch takes the
    direct kernels.  The bins hold a sorted copy of the positions: they stay valid only while
    the positions are unchanged."""
    kind, degree = kind_of(weighting)
    return _positions_bins(positions, kind, degree, tuple(grid.geometry), ncomp)


def _check_bins(bins, positions: torch.Tensor, kind: int, degree: int, geometry, ncomp: int):
    """The caller's bins where the tiled kernels take ``ncomp`` components, else None (the
    direct kernels); errors for bins of other markers, weightings or grids."""
    if bins is None:
        return None
    if not isinstance(bins, binning.MarkerBins) or bins.values is not None:
        raise TypeError("bins must be MarkerBins made from positions alone (bin_positions)")
    if (bins.kind, bins.degree, tuple(bins.geometry)) != (kind, degree, tuple(geometry)) or bins.n != positions.shape[1]:
        raise ValueError("these bins were made for another weighting, grid or number of markers")
    _, _, _, num, code, _ = K._split(geometry)
    return bins if K.derivative_tiled_ok(kind, degree, ncomp, num, code) else None


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


def _shape(geometry, kind: int, degree: int, ncomp: int) -> tuple[int, ...]:
    """The field a deposit of this weighting fills: num_nodes, plus degree - 1 on a clamped axis."""
    _, _, _, num, code, _ = K._split(geometry)
    deg = {_lib.KIND_LINEAR: 1, _lib.KIND_QUADRATIC: 2}.get(kind, degree)
    return (*K._expected_extent(num, code, deg, kind), ncomp)


class GatherFunction(torch.autograd.Function):
    """g = gather(positions, field) with the adjoint backward."""

    @staticmethod
    def forward(ctx, positions, field, kind, degree, geometry, bins=None):
        field_c = field.contiguous()
        bins = _check_bins(bins, positions, kind, degree, geometry, field.shape[-1])
        if bins is None:
            bins = _positions_bins(positions, kind, degree, geometry, field.shape[-1])
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
        return grad_pos, grad_field, None, None, None, None


class ScatterFunction(torch.autograd.Function):
    """rho = scatter(positions, values) into a zero field, with the adjoint backward."""

    @staticmethod
    def forward(ctx, positions, values, kind, degree, geometry, bins=None):
        field = torch.zeros(_shape(geometry, kind, degree, values.shape[0]), dtype=torch.float64,
                            device=values.device)
        bins = _check_bins(bins, positions, kind, degree, geometry, values.shape[0])
        if bins is None:
            bins = _positions_bins(positions, kind, degree, geometry, values.shape[0])
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
        return grad_pos, grad_val, None, None, None, None


def gather(positions: torch.Tensor, field: torch.Tensor, weighting, grid, bins=None) -> torch.Tensor:
    """Differentiable ``weighting`` gather of ``field`` (num_nodes + (ncomp,)) at ``positions`` (ndim, N);
    ``bins`` from bin_positions of the same positions, or None to bin them here."""
    kind, degree = kind_of(weighting)
    return GatherFunction.apply(positions, field, kind, degree, tuple(grid.geometry), bins)


def scatter(positions: torch.Tensor, values: torch.Tensor, weighting, grid, bins=None) -> torch.Tensor:
    """Differentiable ``weighting`` deposit of ``values`` (ncomp, N) at ``positions`` into a new zero field
    (num_nodes + degree - 1 on a clamped axis); ``bins`` as for gather."""
    kind, degree = kind_of(weighting)
    return ScatterFunction.apply(positions, values, kind, degree, tuple(grid.geometry), bins)

