"""Markers binned by stencil cell: the on-device counting sort and the tiled deposit and
gather that read the bins (mpf_bin_markers, mpf_scatter_binned, mpf_gather_binned).

A particle code that keeps its markers sorted by cell between steps (what the reference's
banded CPU buffers assume, scatter_linear_3d_pyccel.py:11-16) bins once and deposits and
gathers many times::

    bins = bin_markers(positions, values, kind_of(QUADRATIC), grid.geometry)
    scatter_binned(bins, field)              # adds into field, bit-reproducible
    gather_binned(bins, field, result)       # result (ncomp, N) in bins order

The forms deposit (faces, cells) reads bins of the spline nodes of its degree::

    bins = bin_markers(positions, values, (KIND_SPLINE, degree), grid.geometry)
    scatter_forms_binned(bins, faces, form=2)
    gather_forms_binned(bins, faces, result, form=2)   # result (ncomp, N) in bins order

Paths whose starts the caller keeps binned (the markers' positions at the start of a step)
deposit their current without a sort: Whitney paths read bins of the linear cell, the spline
edges of degree p bins of the spline nodes of p::

    bins = bin_markers(start, None, (KIND_SPLINE, 2), grid.geometry)
    deposit_paths_binned(bins, end_in_bins_order, charges_in_bins_order, edges, kind=1, degree=2)

``bins.order[i]`` is the input index of binned marker ``i``. `Weighting.scatter_block` /
`gather_block` take this path by themselves for large batches (binning inside the call).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from mpifield_amd import _lib
from mpifield_amd import kernels as K

#: bins of this many markers or more move rows between the input order and the bins' order
#: through a permutation plan (two passes of whole blocks) instead of scattered accesses
PLAN_MIN = 1 << 18

#: (kind, degree) of the built-in node weightings, by name and degree
_KINDS = {("linear", 1): (0, 1), ("quadratic", 2): (1, 2), ("cubic", 3): (2, 3)}


def kind_of(weighting) -> tuple[int, int]:
    """The kernel kind and degree of a built-in weighting (ValueError for others)."""
    # the clamping weightings run on the spline kernels (the only ones that take clamped
    # axes), whatever their name
    if weighting.degree is not None and weighting.clamps:
        return (2, int(weighting.degree))
    key = (weighting.name, weighting.degree)
    if key in _KINDS:
        return _KINDS[key]
    raise ValueError(f"the {weighting.name} weighting has no binned kernels")


@dataclass
class MarkerBins:
    """Markers sorted by stencil cell on the GPU, and what the tiled kernels need."""

    kind: int
    degree: int
    geometry: tuple
    positions: torch.Tensor  # (ndim, N) sorted, contiguous rows
    values: torch.Tensor | None  # (ncomp, N) sorted, or None for gather-only bins
    order: torch.Tensor  # (N,) int64: sorted index -> input index
    start: torch.Tensor  # (bins + 2,) int32 holding uint32 offsets (3D bins in tile order)
    scalars: torch.Tensor  # (3,) float64: [2^k, 2^-k, markers per register pass]
    tables: torch.Tensor | None = None  # (tiles, 4096) int16: the 3D tiles' work lists (prepare_items)
    table_counts: torch.Tensor | None = None  # (tiles,) int32
    plan: torch.Tensor | None = None  # uint8: the permutation plan (permutation_plan), made when it pays
    moves: int = 0  # single-row permutations so far without a plan

    @property
    def n(self) -> int:
        return int(self.positions.shape[1])

    @property
    def ncomp(self) -> int:
        return 0 if self.values is None else int(self.values.shape[0])

    def prepare_items(self) -> "MarkerBins":
        """Make the 3D tiles' work lists once (mpf_bin_items), for bins that are deposited more
        than once: each deposit launch then loads a tile's list instead of building it."""
        ndim, _, _, num, _ = _geometry_arrays(self.geometry)
        if ndim != 3 or self.tables is not None:
            return self
        tiles = (self.start.numel() - 2) // 1024
        dev = self.positions.device
        self.tables = torch.empty((max(tiles, 1), 4096), dtype=torch.int16, device=dev)
        self.table_counts = torch.empty(max(tiles, 1), dtype=torch.int32, device=dev)
        with torch.cuda.device(dev):
            _lib.call("mpf_bin_items", ndim, num.ctypes.data, self.start.data_ptr(), self.tables.data_ptr(),
                      self.table_counts.data_ptr(), None, K.stream_handle(dev))
        return self

    def _table_ptrs(self):
        if self.tables is None:
            return None, None
        return self.tables.data_ptr(), self.table_counts.data_ptr()

    def rebin(self, positions: torch.Tensor | None = None, values=None, scale_exp: int = _lib.SCALE_AUTO
              ) -> "MarkerBins":
        """New bins of these markers after they moved (mpf_rebin_markers): ``positions``
        (default: this object's sorted positions, updated in place by the caller) and
        ``values`` (default: the sorted values) are in this object's bins order.  The result
        holds its own arrays; its ``order`` still maps to the first input.  3D bins move only
        the markers that changed tile (to a neighbouring one); 1D/2D bins sort again."""
        ndim, lower, cell, num, code = _geometry_arrays(self.geometry)
        pos, ps = K.marker_rows(self.positions if positions is None else positions, ndim)
        n = pos.shape[1]
        if n != self.n:
            raise ValueError(f"these bins hold {self.n} markers, got {n} positions")
        if pos.dtype != self.positions.dtype or pos.device != self.positions.device:
            raise ValueError("positions must have the bins' dtype and device")
        dev = pos.device
        values = self.values if values is None else values
        if values is not None:
            val, vs = K.marker_rows(values, None, pos.dtype)
            if val.shape[1] != n:
                raise ValueError(f"{n} positions for {val.shape[1]} values")
            ncomp = val.shape[0]
            sval = torch.empty((ncomp, n), dtype=pos.dtype, device=dev)
            val_ptr, sval_ptr = val.data_ptr(), sval.data_ptr()
        else:
            vs, ncomp, sval, val_ptr, sval_ptr = 0, 0, None, None, None
        spos = torch.empty((ndim, n), dtype=pos.dtype, device=dev)
        order = torch.empty(n, dtype=torch.int64, device=dev)
        start = torch.empty_like(self.start)
        scalars = torch.empty(3, dtype=torch.float64, device=dev)
        with torch.cuda.device(dev):
            _lib.call("mpf_rebin_markers", self.kind, self.degree, ndim, pos.data_ptr(), ps, val_ptr, vs,
                      K._dtype_code(pos), n, ncomp, self.order.data_ptr(), self.start.data_ptr(), lower.ctypes.data,
                      cell.ctypes.data, num.ctypes.data, code.ctypes.data, spos.data_ptr(), sval_ptr,
                      order.data_ptr(), start.data_ptr(), scalars.data_ptr(), int(scale_exp), None,
                      K.stream_handle(dev))
        return MarkerBins(self.kind, self.degree, self.geometry, spos, sval, order, start, scalars)

    def permutation_plan(self) -> torch.Tensor | None:
        """The plan of this binning's permutation (mpf_perm_plan_build), made on first use and
        kept with the bins, so that rows move between the input order and the bins' order in
        two passes of whole blocks; None below PLAN_MIN markers (scattered accesses cost less
        there) and above 2^28."""
        if self.plan is None and self.n >= PLAN_MIN:
            size = ctypes.c_int64(0)
            _lib.call("mpf_perm_plan_bytes", self.n, ctypes.addressof(size))
            if size.value:
                dev = self.order.device
                plan = torch.empty(size.value, dtype=torch.uint8, device=dev)
                with torch.cuda.device(dev):
                    _lib.call("mpf_perm_plan_build", self.order.data_ptr(), self.n, plan.data_ptr(), None,
                              K.stream_handle(dev))
                self.plan = plan
        return self.plan

    def _plan_for(self, rows: int) -> torch.Tensor | None:
        """The plan for a permutation of ``rows`` rows: the existing one, or a new one when it
        pays -- several rows at once, or a second single-row permutation of these bins (one
        row of scattered accesses costs about what making the plan does)."""
        if self.plan is None and rows < 2 and self.moves < 1:
            self.moves += 1
            return None
        return self.permutation_plan()

    def unsort(self, sorted_rows: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Rows (k, N) in bins order back to the input order (mpf_perm_rows through the bins'
        plan, or mpf_unsort_rows)."""
        if out is None:
            out = torch.empty_like(sorted_rows)
        ok = (sorted_rows.dtype in (torch.float32, torch.float64) and out.dtype == sorted_rows.dtype
              and sorted_rows.is_contiguous() and out.is_contiguous() and sorted_rows.dim() == 2
              and tuple(out.shape) == tuple(sorted_rows.shape) and sorted_rows.shape[1] == self.n)
        if not ok:  # other dtypes or layouts: torch's own scatter
            out.index_copy_(1, self.order, sorted_rows)
            return out
        plan = self._plan_for(sorted_rows.shape[0])
        with torch.cuda.device(out.device):
            if plan is not None:
                _lib.call("mpf_perm_rows", plan.data_ptr(), self.n, 0, sorted_rows.data_ptr(), self.n, out.data_ptr(),
                          self.n, sorted_rows.shape[0], K._dtype_code(sorted_rows), None, K.stream_handle(out.device))
            else:
                _lib.call("mpf_unsort_rows", sorted_rows.data_ptr(), out.data_ptr(), self.order.data_ptr(),
                          sorted_rows.shape[0], self.n, K._dtype_code(sorted_rows), K.stream_handle(out.device))
        return out

    def to_bins_order(self, rows: torch.Tensor) -> torch.Tensor:
        """Rows (k, N) of these markers in the input order, in the bins' order (contiguous):
        mpf_perm_rows through the bins' plan, or torch.index_select."""
        plan = (self._plan_for(rows.shape[0]) if rows.dtype in (torch.float32, torch.float64) and rows.dim() == 2
                and rows.shape[1] == self.n and rows.is_cuda else None)
        if plan is None:
            return torch.index_select(rows, 1, self.order)
        if rows.shape[0] > 1 and (rows.stride(1) != 1 or rows.stride(0) < self.n):
            rows = rows.contiguous()
        stride = rows.stride(0) if rows.shape[0] > 1 else self.n
        out = torch.empty((rows.shape[0], self.n), dtype=rows.dtype, device=rows.device)
        with torch.cuda.device(rows.device):
            _lib.call("mpf_perm_rows", plan.data_ptr(), self.n, 1, rows.data_ptr(), stride, out.data_ptr(), self.n,
                      rows.shape[0], K._dtype_code(rows), None, K.stream_handle(rows.device))
        return out


def _geometry_arrays(geometry):
    ndim, lower, cell, num, code, _ = K._split(geometry)
    return ndim, lower, cell, num, code


def bin_markers(positions, values, kind_degree: tuple[int, int], geometry, scale_exp: int = _lib.SCALE_AUTO
                ) -> MarkerBins:
    """Sort markers by the stencil cell of the (kind, degree) weighting on ``geometry``."""
    kind, degree = kind_degree
    ndim, lower, cell, num, code = _geometry_arrays(geometry)
    pos, ps = K.marker_rows(positions, ndim)
    n = pos.shape[1]
    dev = pos.device
    if values is not None:
        val, vs = K.marker_rows(values, None, pos.dtype)
        if val.shape[1] != n:
            raise ValueError(f"{n} positions for {val.shape[1]} values")
        ncomp = val.shape[0]
        sval = torch.empty((ncomp, n), dtype=pos.dtype, device=dev)
        val_ptr, sval_ptr = val.data_ptr(), sval.data_ptr()
    else:
        vs, ncomp, sval, val_ptr, sval_ptr = 0, 0, None, None, None
    cells = ctypes_int64()
    _lib.call("mpf_bin_sizes", ndim, num.ctypes.data, cells.ptr)
    spos = torch.empty((ndim, n), dtype=pos.dtype, device=dev)
    order = torch.empty(n, dtype=torch.int64, device=dev)
    start = torch.empty(int(cells.value) + 2, dtype=torch.int32, device=dev)
    scalars = torch.empty(3, dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        _lib.call("mpf_bin_markers", kind, degree, ndim, pos.data_ptr(), ps, val_ptr, vs, K._dtype_code(pos), n,
                  ncomp, lower.ctypes.data, cell.ctypes.data, num.ctypes.data, code.ctypes.data, spos.data_ptr(),
                  sval_ptr, order.data_ptr(), start.data_ptr(), scalars.data_ptr(), int(scale_exp), None,
                  K.stream_handle(dev))
    return MarkerBins(kind, degree, tuple(geometry), spos, sval, order, start, scalars)


class ctypes_int64:
    """A host int64 slot for an out-parameter."""

    def __init__(self):
        self.array = np.zeros(1, dtype=np.int64)
        self.ptr = self.array.ctypes.data

    @property
    def value(self) -> int:
        return int(self.array[0])


def _bins_field(bins: MarkerBins, field, ncomp: int | None = None, form: int | None = None) -> torch.Tensor:
    """``field`` checked against the bins' grid and device before any launch: the tiled
    kernels index it by the grid's node counts, so a field of another grid would be read or
    written out of bounds.  ``ncomp`` None takes the field's own component count (a
    gather); ``form`` checks a k-form field of the bins' degree instead of a node field."""
    field = K._field(field)
    _, _, _, num, code = _geometry_arrays(bins.geometry)
    if form is not None:
        field = K._form_field(field, bins.degree, form, num, code)
    else:
        extent = K._expected_extent(num, code, bins.degree, bins.kind)  # num + degree - 1 on a clamped axis
        want = (*extent, field.shape[-1] if ncomp is None else ncomp)
        if tuple(field.shape) != want:
            raise ValueError(f"a field on this grid is {want}, got {tuple(field.shape)}")
    if field.device != bins.positions.device:
        raise ValueError(f"the field is on {field.device}, the bins on {bins.positions.device}")
    return field


def scatter_binned(bins: MarkerBins, field: torch.Tensor) -> None:
    """Add the binned markers' values into ``field`` (num_nodes + (ncomp,)) with the tiled kernels."""
    if bins.values is None:
        raise ValueError("these bins hold no values (binned for a gather)")
    field = _bins_field(bins, field, bins.ncomp)
    ndim, lower, cell, num, code = _geometry_arrays(bins.geometry)
    with torch.cuda.device(field.device):
        _lib.call("mpf_scatter_binned", bins.kind, bins.degree, ndim, bins.positions.data_ptr(),
                  bins.values.data_ptr(), K._dtype_code(bins.positions), bins.n, bins.ncomp, bins.start.data_ptr(),
                  *bins._table_ptrs(), bins.scalars.data_ptr(), field.data_ptr(), lower.ctypes.data,
                  cell.ctypes.data, num.ctypes.data, code.ctypes.data, None, K.stream_handle(field.device))


def gather_binned(bins: MarkerBins, field: torch.Tensor, result: torch.Tensor) -> None:
    """Interpolate ``field`` at the binned markers into ``result`` (ncomp, N) in bins order."""
    field = _bins_field(bins, field)
    ndim, lower, cell, num, code = _geometry_arrays(bins.geometry)
    ncomp = field.shape[-1]
    _check_result(bins, result, ncomp)
    with torch.cuda.device(field.device):
        # the gather takes one marker per work-item and reads no work lists
        _lib.call("mpf_gather_binned", bins.kind, bins.degree, ndim, bins.positions.data_ptr(),
                  K._dtype_code(bins.positions), bins.n, bins.start.data_ptr(), None, None,
                  field.data_ptr(), ncomp, result.data_ptr(), lower.ctypes.data, cell.ctypes.data, num.ctypes.data,
                  code.ctypes.data, None, K.stream_handle(field.device))


def with_values(bins: MarkerBins, values, scale_exp: int = _lib.SCALE_AUTO) -> MarkerBins:
    """The same bins carrying ``values`` (ncomp, N) in the input order: the values permuted to
    the bins' order and their fixed-point scale.  ``bins`` must hold every marker in a cell
    (binned from positions alone), since a marker in the dead bin deposits nothing."""
    if bins.values is not None:
        raise ValueError("these bins were made with values; bin positions alone to give them new values")
    val, vs = K.marker_rows(values, None, bins.positions.dtype)
    if val.shape[1] != bins.n:
        raise ValueError(f"{bins.n} binned markers for {val.shape[1]} values")
    sval = bins.to_bins_order(val)
    scalars = torch.empty(3, dtype=torch.float64, device=val.device)
    with torch.cuda.device(val.device):
        _lib.call("mpf_fixed_scale", val.data_ptr(), vs, K._dtype_code(val), bins.n, val.shape[0], int(scale_exp),
                  scalars.data_ptr(), None, K.stream_handle(val.device))
    return MarkerBins(bins.kind, bins.degree, bins.geometry, bins.positions, sval, bins.order, bins.start, scalars,
                      bins.tables, bins.table_counts, bins.plan, bins.moves)


def _check_result(bins: MarkerBins, result, ncomp: int) -> None:
    if (not isinstance(result, torch.Tensor) or tuple(result.shape) != (ncomp, bins.n) or not result.is_contiguous()
            or result.dtype != bins.positions.dtype or result.device != bins.positions.device):
        raise ValueError(f"result must be a contiguous ({ncomp}, {bins.n}) {bins.positions.dtype} tensor on "
                         f"{bins.positions.device}")


def gather_derivative_binned(bins: MarkerBins, field: torch.Tensor, coef: torch.Tensor) -> torch.Tensor:
    """sum_c coef[c] * grad F_c at the binned markers: (ndim, N) in the input order, with
    ``coef`` (ncomp, N) in the input order and ``field`` num_nodes + (ncomp,)."""
    field = _bins_field(bins, field)
    ndim, lower, cell, num, code = _geometry_arrays(bins.geometry)
    ncomp = field.shape[-1]
    k, ks = K.marker_rows(coef, ncomp, bins.positions.dtype)
    if k.shape[1] != bins.n:
        raise ValueError(f"{bins.n} binned markers for {k.shape[1]} coefficients")
    out = torch.empty((ndim, bins.n), dtype=bins.positions.dtype, device=field.device)
    with torch.cuda.device(field.device):
        _lib.call("mpf_gather_derivative_binned", bins.kind, bins.degree, ndim, bins.positions.data_ptr(),
                  K._dtype_code(bins.positions), bins.n, bins.start.data_ptr(), bins.order.data_ptr(),
                  field.data_ptr(), ncomp, k.data_ptr(), ks, out.data_ptr(), bins.n, lower.ctypes.data,
                  cell.ctypes.data, num.ctypes.data, code.ctypes.data, None, K.stream_handle(field.device))
    return out


def scatter_forms_binned(bins: MarkerBins, field: torch.Tensor, form: int) -> None:
    """Add the binned markers' values into the ``form``-form ``field`` with the tiled forms
    deposit (mpf_scatter_forms_binned): the fixed-point deposit of scatter_forms_spline_*d,
    bit for bit.  ``bins`` are of the spline nodes of the forms' degree (kind 2)."""
    if bins.values is None:
        raise ValueError("these bins hold no values (binned for a gather)")
    if bins.kind != _lib.KIND_SPLINE:
        raise ValueError("the forms deposit reads bins of the spline nodes (kind 2)")
    field = _bins_field(bins, field, form=form)
    if field.shape[-1] != bins.ncomp:
        raise ValueError(f"a {form}-form field has {field.shape[-1]} components, the bins {bins.ncomp} values")
    ndim, lower, cell, num, code = _geometry_arrays(bins.geometry)
    with torch.cuda.device(field.device):
        _lib.call("mpf_scatter_forms_binned", bins.degree, form, ndim, bins.positions.data_ptr(),
                  bins.values.data_ptr(), K._dtype_code(bins.positions), bins.n, bins.ncomp, bins.start.data_ptr(),
                  bins.scalars.data_ptr(), field.data_ptr(), lower.ctypes.data, cell.ctypes.data, num.ctypes.data,
                  code.ctypes.data, None, K.stream_handle(field.device))


def gather_forms_binned(bins: MarkerBins, field: torch.Tensor, result: torch.Tensor, form: int) -> None:
    """The ``form``-form ``field`` at the binned markers into ``result`` (ncomp, N) in bins order
    (mpf_gather_forms_binned): gather_forms_spline_*d's values bit for bit."""
    if bins.kind != _lib.KIND_SPLINE:
        raise ValueError("the forms gather reads bins of the spline nodes (kind 2)")
    field = _bins_field(bins, field, form=form)
    ndim, lower, cell, num, code = _geometry_arrays(bins.geometry)
    ncomp = field.shape[-1]
    _check_result(bins, result, ncomp)
    with torch.cuda.device(field.device):
        _lib.call("mpf_gather_forms_binned", bins.degree, form, ndim, bins.positions.data_ptr(),
                  K._dtype_code(bins.positions), bins.n, bins.start.data_ptr(), field.data_ptr(), ncomp,
                  result.data_ptr(), lower.ctypes.data, cell.ctypes.data, num.ctypes.data, code.ctypes.data, None,
                  K.stream_handle(field.device))


def deposit_paths_binned(bins: MarkerBins, end, charges, edges: torch.Tensor, kind: int, degree: int) -> None:
    """Add the current of paths from the binned start points ``bins.positions`` to ``end``
    (ndim, N) with ``charges`` (N,), both in the bins' order, into ``edges`` (mpf_deposit_paths_binned:
    mpf_deposit_paths's fixed-point deposit, bit for bit, without its sort).  ``kind`` 0 is
    WHITNEY (bins of the linear cell), 1 the spline edges of ``degree`` (bins of the spline
    nodes of that degree)."""
    want = (_lib.KIND_LINEAR, 1) if kind == 0 else (_lib.KIND_SPLINE, int(degree))
    if kind not in (0, 1):
        raise ValueError(f"path kind must be 0 (WHITNEY) or 1 (spline edges), got {kind}")
    if (bins.kind, bins.degree) != want:
        raise ValueError(f"these paths read bins of (kind, degree) {want}, got {(bins.kind, bins.degree)}")
    ndim, lower, cell, num, code = _geometry_arrays(bins.geometry)
    edges = K._field(edges, "edges")
    e, es = K.marker_rows(end, ndim, bins.positions.dtype)
    q = K.as_device(charges, bins.positions.dtype).reshape(-1).contiguous()
    if e.shape[1] != bins.n or q.numel() != bins.n:
        raise ValueError(f"{bins.n} binned starts for ends {tuple(e.shape)} and {q.numel()} charges")
    extent = K._expected_extent(num, code, want[1], _lib.KIND_LINEAR if kind == 0 else _lib.KIND_SPLINE)
    if tuple(edges.shape) != (*extent, ndim):
        raise ValueError(f"an edge field on this grid is {(*extent, ndim)}, got {tuple(edges.shape)}")
    if edges.device != bins.positions.device or e.device != edges.device or q.device != edges.device:
        raise ValueError("bins, ends, charges and edges must be on one device")
    with torch.cuda.device(edges.device):
        _lib.call("mpf_deposit_paths_binned", kind, want[1] if kind == 1 else 1, ndim, bins.positions.data_ptr(),
                  e.data_ptr(), es, q.data_ptr(), K._dtype_code(bins.positions), bins.n, bins.start.data_ptr(),
                  edges.data_ptr(), lower.ctypes.data, cell.ctypes.data, num.ctypes.data, code.ctypes.data, None,
                  K.stream_handle(edges.device))
