"""The kernel catalogue on the GPU: one function per kernel of the reference's catalogue
(src/mpifield/kernels/catalog.py:77-114), argument for argument, each a thin call into
libmpfhip.so on torch's current HIP stream.

Arrays are torch tensors on the GPU (any array convertible with ``torch.as_tensor`` is
moved there first); marker arrays are float32 or float64 rows with unit element stride,
fields and edge fields float64 and C-contiguous. The kernels add into ``field`` / ``edges``
and overwrite ``result``, in place, and return None.

The deposit mode (`deposit_mode`) selects float64 atomics ("atomic", the plain GPU form)
or int64 fixed point ("fixed", bit-reproducible under any marker order; the default).
"""

from __future__ import annotations

import contextlib
import threading

import numpy as np
import torch

from mpifield_amd import _lib

KERNEL_NAMES = (
    "gather_linear_1d", "gather_linear_2d", "gather_linear_3d",
    "scatter_linear_1d", "scatter_linear_2d", "scatter_linear_3d",
    "gather_quadratic_1d", "gather_quadratic_2d", "gather_quadratic_3d",
    "scatter_quadratic_1d", "scatter_quadratic_2d", "scatter_quadratic_3d",
    "gather_spline_1d", "gather_spline_2d", "gather_spline_3d",
    "scatter_spline_1d", "scatter_spline_2d", "scatter_spline_3d",
    "deposit_paths_1d", "deposit_paths_2d", "deposit_paths_3d",
    "deposit_paths_spline_1d", "deposit_paths_spline_2d", "deposit_paths_spline_3d",
    "gather_edges_1d", "gather_edges_2d", "gather_edges_3d",
    "gather_edges_spline_1d", "gather_edges_spline_2d", "gather_edges_spline_3d",
    "scatter_forms_spline_1d", "scatter_forms_spline_2d", "scatter_forms_spline_3d",
    "gather_forms_spline_1d", "gather_forms_spline_2d", "gather_forms_spline_3d",
)  # fmt: skip

_state = threading.local()


def current_mode() -> str:
    return getattr(_state, "mode", "fixed")


def set_deposit_mode(mode: str) -> None:
    """Select "fixed" (int64 fixed point, deterministic) or "atomic" (float64 atomics)."""
    if mode not in ("fixed", "atomic"):
        raise ValueError(f"deposit mode is 'fixed' or 'atomic', got {mode!r}")
    _state.mode = mode


@contextlib.contextmanager
def deposit_mode(mode: str):
    """Run the deposits of a ``with`` block in ``mode``."""
    old = current_mode()
    set_deposit_mode(mode)
    try:
        yield
    finally:
        set_deposit_mode(old)


def _mode_code() -> int:
    return _lib.MODE_FIXED if current_mode() == "fixed" else _lib.MODE_ATOMIC


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise _lib.MpfError("mpifield_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(dev: torch.device | None = None) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def as_device(a, dtype=None) -> torch.Tensor:
    """``a`` as a tensor on the current GPU (no copy if it already is one of ``dtype``)."""
    if isinstance(a, torch.Tensor):
        t = a if a.is_cuda else a.to(device())
    else:
        t = torch.as_tensor(np.asarray(a), device=device())
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t


def marker_rows(a, rows: int | None = None, dtype=None) -> tuple[torch.Tensor, int]:
    """A (rows, N) float32/float64 tensor with unit element stride and its row stride."""
    t = as_device(a)
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if t.dim() == 1:
        t = t.reshape(1, -1)
    if t.dim() != 2 or (rows is not None and t.shape[0] != rows):
        raise ValueError(f"expected markers of shape ({rows}, N), got {tuple(t.shape)}")
    if t.shape[1] > 1 and t.stride(1) != 1:
        t = t.contiguous()
    stride = t.stride(0) if t.shape[0] > 1 else t.shape[1]
    return t, stride


def _dtype_code(t: torch.Tensor) -> int:
    return _lib.F32 if t.dtype == torch.float32 else _lib.F64


def _field(field: torch.Tensor, what: str = "field") -> torch.Tensor:
    if not isinstance(field, torch.Tensor) or not field.is_cuda:
        raise TypeError(f"{what} must be a torch tensor on the GPU")
    if field.dtype != torch.float64 or not field.is_contiguous():
        raise ValueError(f"{what} must be a C-contiguous float64 tensor")
    return field


def _split(geometry, ndim_extra: int = 0):
    values = list(geometry)
    if ndim_extra:
        values, extra = values[:-ndim_extra], values[-ndim_extra:]
    else:
        extra = []
    if len(values) % 4:
        raise ValueError(f"geometry is (*lower, *cell_size, *num_nodes, *code), got {len(values)} values")
    ndim = len(values) // 4
    lower = np.asarray(values[:ndim], dtype=np.float64)
    cell = np.asarray(values[ndim : 2 * ndim], dtype=np.float64)
    num = np.asarray([int(v) for v in values[2 * ndim : 3 * ndim]], dtype=np.int64)
    code = np.asarray([int(v) for v in values[3 * ndim :]], dtype=np.int64)
    return ndim, lower, cell, num, code, extra


def _expected_extent(num, code, degree, kind: int = _lib.KIND_SPLINE):
    if kind != _lib.KIND_SPLINE and any(int(c) == 2 for c in code):
        raise ValueError("a clamped axis (code 2) needs the spline kernels")
    return tuple(int(n) + (degree - 1 if c == 2 else 0) for n, c in zip(num, code))


#: batches at least this large go through the binned (tiled) kernels where they apply (the
#: library's own threshold for mpf_scatter / mpf_gather, BINNED_MIN_MARKERS in mpf_internal.h)
BINNED_MIN_MARKERS = 1 << 18
#: bins hold 32-bit offsets: larger batches take the direct kernels
BINNED_MAX_MARKERS = (1 << 31) - 1
_TILE = (8, 8, 16)


def binned_min() -> int:
    """The smallest batch that mpf_scatter / mpf_gather bin inside the call (the library's
    current setting: BINNED_MIN_MARKERS, or the tuned value the library loaded)."""
    return int(_lib.load().mpf_set_binned_min(-1))


@contextlib.contextmanager
def binned_from(markers: int):
    """Within the block, mpf_scatter / mpf_gather bin batches of at least ``markers``
    (mpf_set_binned_min; process-wide, restored on exit)."""
    lib = _lib.load()
    old = lib.mpf_set_binned_min(int(markers))
    try:
        yield
    finally:
        lib.mpf_set_binned_min(old)


def direct_kernels():
    """Within the block, mpf_scatter / mpf_gather keep every batch on the direct kernels."""
    return binned_from(1 << 62)


def binned_route(kind: int, degree: int, ncomp: int, num, code, n: int) -> bool:
    """Whether a batch of ``n`` markers goes through bins and tiles (mpf_binned.hip's
    binned_route): at least binned_min(), below 2^31, and a shape tiled_ok takes."""
    return binned_min() <= n <= BINNED_MAX_MARKERS and tiled_ok(kind, degree, ncomp, num, code)


def tiled_ok(kind: int, degree: int, ncomp: int, num, code) -> bool:
    """Whether the tiled deposit and gather take this case (the rules of mpf_binned.hip's
    tiled_supported and tiled_2d_supported): 2D with 1-3 components on bounded or periodic
    axes; 3D with any degree and 1-64 components (groups of three where p <= 2), bounded axes
    and the splines' clamped axes of at least a tile, periodic axes of at least two tiles and
    a footprint (partial last tiles allowed)."""
    if len(num) not in (2, 3) or (any(int(c) == 2 for c in code) and (len(num) == 2 or kind != _lib.KIND_SPLINE)):
        return False
    if kind == _lib.KIND_SPLINE and not 0 <= degree <= 3:
        return False
    if len(num) == 2:
        return 1 <= ncomp <= 3
    if not 1 <= ncomp <= 64:
        return False
    p = {0: 1, 1: 2}.get(kind, degree)
    span = 2 if kind == _lib.KIND_LINEAR else p + 1
    for n, c, t in zip(num, code, _TILE):
        n = int(n)
        if n < t or (int(c) == 1 and ((n + t - 1) // t < 2 or n < t + span - 1)):
            return False
    return True


def derivative_tiled_ok(kind: int, degree: int, ncomp: int, num, code) -> bool:
    """Whether the tiled gradient (mpf_gather_derivative_binned) takes this 3D case: 1 or 3
    components up to degree 2, the cubic with one, whole periodic tiles."""
    if len(num) != 3 or not tiled_ok(kind, degree, ncomp, num, code):
        return False
    p = {0: 1, 1: 2}.get(kind, degree)
    if not ((ncomp in (1, 3) and p <= 2) or (p == 3 and ncomp == 1)):
        return False
    return all(int(c) == 0 or (int(c) == 1 and int(n) % t == 0) for n, c, t in zip(num, code, _TILE))


def scatter(kind: int, degree: int, positions, values, field: torch.Tensor, geometry, *, mode: int | None = None,
            scale_exp: int = _lib.SCALE_AUTO) -> None:
    """The generic node deposit (mpf_scatter): adds into ``field`` num_nodes + (ncomp,).

    Large batches in fixed-point mode are binned and deposited with the tiled kernels inside
    the library (binned_route)."""
    ndim, lower, cell, num, code, _ = _split(geometry)
    field = _field(field)
    pos, ps = marker_rows(positions, ndim)
    val, vs = marker_rows(values, None, pos.dtype)
    if val.shape[1] != pos.shape[1]:
        raise ValueError(f"{pos.shape[1]} positions for {val.shape[1]} values")
    ncomp = val.shape[0]
    deg = {0: 1, 1: 2}.get(kind, degree)
    extent = _expected_extent(num, code, deg, kind)
    if tuple(field.shape) != (*extent, ncomp):
        raise ValueError(f"a field on this grid is {(*extent, ncomp)}, got {tuple(field.shape)}")
    mode_code = _mode_code() if mode is None else mode
    with torch.cuda.device(field.device):
        _lib.call("mpf_scatter", kind, degree, ndim, pos.data_ptr(), ps, val.data_ptr(), vs, _dtype_code(pos),
                  pos.shape[1], ncomp, field.data_ptr(), lower.ctypes.data, cell.ctypes.data, num.ctypes.data,
                  code.ctypes.data, mode_code, scale_exp, None, stream_handle(field.device))


def gather(kind: int, degree: int, positions, field: torch.Tensor, result: torch.Tensor, geometry) -> None:
    """The generic node gather (mpf_gather): overwrites ``result`` (ncomp, N)."""
    ndim, lower, cell, num, code, _ = _split(geometry)
    field = _field(field)
    if not isinstance(result, torch.Tensor) or not result.is_cuda:
        raise TypeError("result must be a torch tensor on the GPU")
    pos, ps = marker_rows(positions, ndim, result.dtype if result.dtype in (torch.float32, torch.float64) else None)
    ncomp = field.shape[-1]
    deg = {0: 1, 1: 2}.get(kind, degree)
    extent = _expected_extent(num, code, deg, kind)
    if tuple(field.shape) != (*extent, ncomp):
        raise ValueError(f"a field on this grid is {(*extent, ncomp)}, got {tuple(field.shape)}")
    if result.dim() == 1:
        result2 = result.reshape(1, -1)
    else:
        result2 = result
    if tuple(result2.shape) != (ncomp, pos.shape[1]):
        raise ValueError(f"result must be ({ncomp}, {pos.shape[1]}), got {tuple(result.shape)}")
    direct = result2.dtype == pos.dtype and (result2.shape[1] <= 1 or result2.stride(1) == 1)
    out = result2 if direct else torch.empty((ncomp, pos.shape[1]), dtype=pos.dtype, device=field.device)
    rs = out.stride(0) if out.shape[0] > 1 else out.shape[1]
    with torch.cuda.device(field.device):
        _lib.call("mpf_gather", kind, degree, ndim, pos.data_ptr(), ps, _dtype_code(pos), pos.shape[1],
                  field.data_ptr(), ncomp, out.data_ptr(), rs, lower.ctypes.data, cell.ctypes.data,
                  num.ctypes.data, code.ctypes.data, None, stream_handle(field.device))
    if not direct:
        result2.copy_(out)


def deposit_paths(kind: int, degree: int, start, end, charges, edges: torch.Tensor, geometry, *,
                  mode: int | None = None) -> None:
    """The generic path deposit (mpf_deposit_paths): adds into ``edges``."""
    ndim, lower, cell, num, code, _ = _split(geometry)
    edges = _field(edges, "edges")
    s, ss = marker_rows(start, ndim)
    e, es = marker_rows(end, ndim, s.dtype)
    q = as_device(charges, s.dtype).reshape(-1).contiguous()
    if e.shape != s.shape or q.numel() != s.shape[1]:
        raise ValueError(f"paths from {tuple(s.shape)} to {tuple(e.shape)} for {q.numel()} charges")
    deg = 1 if kind == 0 else degree
    extent = _expected_extent(num, code, deg, _lib.KIND_LINEAR if kind == 0 else _lib.KIND_SPLINE)
    if tuple(edges.shape) != (*extent, ndim):
        raise ValueError(f"an edge field on this grid is {(*extent, ndim)}, got {tuple(edges.shape)}")
    with torch.cuda.device(edges.device):
        _lib.call("mpf_deposit_paths", kind, degree, ndim, s.data_ptr(), ss, e.data_ptr(), es, q.data_ptr(),
                  _dtype_code(s), s.shape[1], edges.data_ptr(), lower.ctypes.data, cell.ctypes.data,
                  num.ctypes.data, code.ctypes.data, _mode_code() if mode is None else mode, _lib.SCALE_AUTO,
                  None, stream_handle(edges.device))


def gather_edges(kind: int, degree: int, positions, edges: torch.Tensor, result: torch.Tensor, geometry) -> None:
    """The generic edge gather (mpf_gather_edges): overwrites ``result`` (ndim, N) with the field
    of the edge field ``edges`` at ``positions``; kind 0 WHITNEY, 1 the spline edges of ``degree``."""
    ndim, lower, cell, num, code, _ = _split(geometry)
    edges = _field(edges, "edges")
    if not isinstance(result, torch.Tensor) or not result.is_cuda:
        raise TypeError("result must be a torch tensor on the GPU")
    pos, ps = marker_rows(positions, ndim, result.dtype if result.dtype in (torch.float32, torch.float64) else None)
    deg = 1 if kind == 0 else degree
    extent = _expected_extent(num, code, deg, _lib.KIND_LINEAR if kind == 0 else _lib.KIND_SPLINE)
    if tuple(edges.shape) != (*extent, ndim):
        raise ValueError(f"an edge field on this grid is {(*extent, ndim)}, got {tuple(edges.shape)}")
    result2 = result.reshape(1, -1) if result.dim() == 1 else result
    if tuple(result2.shape) != (ndim, pos.shape[1]):
        raise ValueError(f"result must be ({ndim}, {pos.shape[1]}), got {tuple(result.shape)}")
    direct = result2.dtype == pos.dtype and (result2.shape[1] <= 1 or result2.stride(1) == 1)
    out = result2 if direct else torch.empty((ndim, pos.shape[1]), dtype=pos.dtype, device=edges.device)
    rs = out.stride(0) if out.shape[0] > 1 else out.shape[1]
    with torch.cuda.device(edges.device):
        _lib.call("mpf_gather_edges", kind, degree, ndim, pos.data_ptr(), ps, _dtype_code(pos), pos.shape[1],
                  edges.data_ptr(), out.data_ptr(), rs, lower.ctypes.data, cell.ctypes.data, num.ctypes.data,
                  code.ctypes.data, None, stream_handle(edges.device))
    if not direct:
        result2.copy_(out)


def form_components(form: int, ndim: int) -> int:
    """The components of a ``form``-form in ``ndim`` dimensions, ``ndim`` choose ``form``
    (spline_forms.py:39-43)."""
    if not 1 <= form <= ndim <= 3:
        raise ValueError(f"forms of {ndim} dimensions are 1 to {ndim}, got {form}")
    return 1 if form == ndim else ndim


def _form_field(field, degree: int, form: int, num, code) -> torch.Tensor:
    field = _field(field)
    ncomp = form_components(form, len(num))
    extent = _expected_extent(num, code, degree)
    if tuple(field.shape) != (*extent, ncomp):
        raise ValueError(f"a {form}-form field on this grid is {(*extent, ncomp)}, got {tuple(field.shape)}")
    return field


def scatter_forms(degree: int, form: int, positions, values, field: torch.Tensor, geometry, *,
                  mode: int | None = None) -> None:
    """The k-form deposit (mpf_scatter_forms): adds ``values`` (ncomp, N) into the
    ``form``-form ``field`` num_nodes + (ncomp,) (num + degree - 1 on a clamped axis)."""
    ndim, lower, cell, num, code, _ = _split(geometry)
    field = _form_field(field, degree, form, num, code)
    pos, ps = marker_rows(positions, ndim)
    val, vs = marker_rows(values, field.shape[-1], pos.dtype)
    if val.shape[1] != pos.shape[1]:
        raise ValueError(f"{pos.shape[1]} positions for {val.shape[1]} values")
    mode_code = _mode_code() if mode is None else mode
    if (mode_code == _lib.MODE_FIXED and binned_min() <= pos.shape[1] < 1 << 31
            and forms_tiled_ok(degree, num, code)):
        from mpifield_amd import binning

        bins = binning.bin_markers(pos, val, (_lib.KIND_SPLINE, degree), geometry)
        binning.scatter_forms_binned(bins, field, form)
        return
    with torch.cuda.device(field.device):
        _lib.call("mpf_scatter_forms", degree, form, ndim, pos.data_ptr(), ps, val.data_ptr(), vs, _dtype_code(pos),
                  pos.shape[1], val.shape[0], field.data_ptr(), lower.ctypes.data, cell.ctypes.data,
                  num.ctypes.data, code.ctypes.data, mode_code, _lib.SCALE_AUTO, None, stream_handle(field.device))


def forms_tiled_ok(degree: int, num, code) -> bool:
    """Whether the tiled forms deposit takes this grid (mpf_forms.hip's forms_tiled_supported):
    2D or 3D, no clamped axis, every periodic axis at least a tile and degree nodes long."""
    if len(num) not in (2, 3):
        return False
    tiles = (_TILE if len(num) == 3 else (32, 32))
    return all(int(c) != 2 and (int(c) != 1 or int(n) >= t + degree) for n, c, t in zip(num, code, tiles))


def gather_forms(degree: int, form: int, positions, field: torch.Tensor, result: torch.Tensor, geometry) -> None:
    """The k-form gather (mpf_gather_forms): overwrites ``result`` (ncomp, N) with sum W f."""
    ndim, lower, cell, num, code, _ = _split(geometry)
    field = _form_field(field, degree, form, num, code)
    if not isinstance(result, torch.Tensor) or not result.is_cuda:
        raise TypeError("result must be a torch tensor on the GPU")
    pos, ps = marker_rows(positions, ndim, result.dtype if result.dtype in (torch.float32, torch.float64) else None)
    ncomp = field.shape[-1]
    result2 = result.reshape(1, -1) if result.dim() == 1 else result
    if tuple(result2.shape) != (ncomp, pos.shape[1]):
        raise ValueError(f"result must be ({ncomp}, {pos.shape[1]}), got {tuple(result.shape)}")
    if binned_min() <= pos.shape[1] < 1 << 31 and forms_tiled_ok(degree, num, code):
        from mpifield_amd import binning

        bins = binning.bin_markers(pos, None, (_lib.KIND_SPLINE, degree), geometry)
        sorted_out = torch.empty((ncomp, pos.shape[1]), dtype=pos.dtype, device=field.device)
        binning.gather_forms_binned(bins, field, sorted_out, form)
        result2.copy_(bins.unsort(sorted_out))
        return
    direct = result2.dtype == pos.dtype and (result2.shape[1] <= 1 or result2.stride(1) == 1)
    out = result2 if direct else torch.empty((ncomp, pos.shape[1]), dtype=pos.dtype, device=field.device)
    rs = out.stride(0) if out.shape[0] > 1 else out.shape[1]
    with torch.cuda.device(field.device):
        _lib.call("mpf_gather_forms", degree, form, ndim, pos.data_ptr(), ps, _dtype_code(pos), pos.shape[1],
                  field.data_ptr(), ncomp, out.data_ptr(), rs, lower.ctypes.data, cell.ctypes.data,
                  num.ctypes.data, code.ctypes.data, None, stream_handle(field.device))
    if not direct:
        result2.copy_(out)


# ---- the catalogue ---------------------------------------------------------------------


def _node_kernels(kind_name: str, kind: int):
    def make_scatter(ndim):
        def kernel(positions, values, field, *geometry):
            scatter(kind, 0, positions, values, field, geometry)

        kernel.__name__ = f"scatter_{kind_name}_{ndim}d"
        kernel.__doc__ = f"kernels/scatter_{kind_name}_{ndim}d: add the deposit of values at positions into field."
        return kernel

    def make_gather(ndim):
        def kernel(positions, field, result, *geometry):
            gather(kind, 0, positions, field, result, geometry)

        kernel.__name__ = f"gather_{kind_name}_{ndim}d"
        kernel.__doc__ = f"kernels/gather_{kind_name}_{ndim}d: overwrite result with field at positions."
        return kernel

    return make_scatter, make_gather


def _spline_kernels():
    def make_scatter(ndim):
        def kernel(positions, values, field, *geometry_degree):
            *geometry, degree = geometry_degree
            scatter(_lib.KIND_SPLINE, int(degree), positions, values, field, geometry)

        kernel.__name__ = f"scatter_spline_{ndim}d"
        kernel.__doc__ = f"kernels/scatter_spline_{ndim}d: B-spline deposit of degree 0..3 (the last argument)."
        return kernel

    def make_gather(ndim):
        def kernel(positions, field, result, *geometry_degree):
            *geometry, degree = geometry_degree
            gather(_lib.KIND_SPLINE, int(degree), positions, field, result, geometry)

        kernel.__name__ = f"gather_spline_{ndim}d"
        kernel.__doc__ = f"kernels/gather_spline_{ndim}d: B-spline gather of degree 0..3 (the last argument)."
        return kernel

    return make_scatter, make_gather


_qs, _qg = _node_kernels("quadratic", _lib.KIND_QUADRATIC)
scatter_quadratic_1d, scatter_quadratic_2d, scatter_quadratic_3d = (_qs(d) for d in (1, 2, 3))
gather_quadratic_1d, gather_quadratic_2d, gather_quadratic_3d = (_qg(d) for d in (1, 2, 3))
_ls, _lg = _node_kernels("linear", _lib.KIND_LINEAR)
scatter_linear_2d = _ls(2)
gather_linear_1d, gather_linear_2d = _lg(1), _lg(2)
_ss, _sg = _spline_kernels()
scatter_spline_1d, scatter_spline_2d, scatter_spline_3d = (_ss(d) for d in (1, 2, 3))
gather_spline_1d, gather_spline_2d, gather_spline_3d = (_sg(d) for d in (1, 2, 3))


def scatter_linear_1d(positions, values, field, lower_x, cell_size, num_gridpoints, periodic_x) -> None:
    """kernels/scatter_linear_1d: ONE component, ``(N,)`` markers into an ``(nx,)`` grid column.

    As in the reference (weighting.py:259-295) the 1D weighting calls it once per component.
    """
    pos = as_device(positions).reshape(1, -1)
    val = as_device(values).reshape(1, -1)
    if not isinstance(field, torch.Tensor) or field.dim() != 1:
        raise ValueError("the 1D linear deposit takes one (nx,) column of the field")
    column = field if field.is_contiguous() else field.contiguous()
    scatter(_lib.KIND_LINEAR, 0, pos, val, column.reshape(-1, 1), (lower_x, cell_size, num_gridpoints, periodic_x))
    if column is not field:
        field.copy_(column)


def scatter_linear_3d(positions, values, rows, *geometry) -> None:
    """kernels/scatter_linear_3d: the grid as rows ``(nx * ny * nz, ncomp)``."""
    n = [int(v) for v in geometry[6:9]]
    scatter(_lib.KIND_LINEAR, 0, positions, values, rows.view(*n, rows.shape[-1]), geometry)


def gather_linear_3d(positions, rows, result, *geometry) -> None:
    """kernels/gather_linear_3d: the grid as rows ``(nx * ny * nz, ncomp)``."""
    n = [int(v) for v in geometry[6:9]]
    gather(_lib.KIND_LINEAR, 0, positions, rows.view(*n, rows.shape[-1]), result, geometry)


def _path_kernel(ndim: int, spline: bool):
    if spline:
        def kernel(start, end, charges, edges, *geometry_degree):
            *geometry, degree = geometry_degree
            deposit_paths(1, int(degree), start, end, charges, edges, geometry)

        kernel.__name__ = f"deposit_paths_spline_{ndim}d"
    else:
        def kernel(start, end, charges, edges, *geometry):
            deposit_paths(0, 1, start, end, charges, edges, geometry)

        kernel.__name__ = f"deposit_paths_{ndim}d"
    kernel.__doc__ = f"kernels/{kernel.__name__}: add the current of the paths into edges."
    return kernel


deposit_paths_1d, deposit_paths_2d, deposit_paths_3d = (_path_kernel(d, False) for d in (1, 2, 3))
deposit_paths_spline_1d, deposit_paths_spline_2d, deposit_paths_spline_3d = (_path_kernel(d, True) for d in (1, 2, 3))


def _edge_gather_kernel(ndim: int, spline: bool):
    if spline:
        def kernel(positions, edges, result, *geometry_degree):
            *geometry, degree = geometry_degree
            gather_edges(1, int(degree), positions, edges, result, geometry)

        kernel.__name__ = f"gather_edges_spline_{ndim}d"
    else:
        def kernel(positions, edges, result, *geometry):
            gather_edges(0, 1, positions, edges, result, geometry)

        kernel.__name__ = f"gather_edges_{ndim}d"
    kernel.__doc__ = f"kernels/{kernel.__name__}: overwrite result with the field of edges at positions."
    return kernel


gather_edges_1d, gather_edges_2d, gather_edges_3d = (_edge_gather_kernel(d, False) for d in (1, 2, 3))
gather_edges_spline_1d, gather_edges_spline_2d, gather_edges_spline_3d = (_edge_gather_kernel(d, True)
                                                                          for d in (1, 2, 3))


def _forms_kernel(ndim: int, scatter: bool):
    if scatter:
        def kernel(positions, values, field, *geometry_degree_form):
            *geometry, degree, form = geometry_degree_form
            scatter_forms(int(degree), int(form), positions, values, field, geometry)

        kernel.__name__ = f"scatter_forms_spline_{ndim}d"
        kernel.__doc__ = (f"kernels/{kernel.__name__}: add the deposit of the k-form values at positions into field "
                          "(degree and form the last two arguments).")
    else:
        def kernel(positions, field, result, *geometry_degree_form):
            *geometry, degree, form = geometry_degree_form
            gather_forms(int(degree), int(form), positions, field, result, geometry)

        kernel.__name__ = f"gather_forms_spline_{ndim}d"
        kernel.__doc__ = (f"kernels/{kernel.__name__}: overwrite result with the k-form field at positions "
                          "(degree and form the last two arguments).")
    return kernel


scatter_forms_spline_1d, scatter_forms_spline_2d, scatter_forms_spline_3d = (_forms_kernel(d, True) for d in (1, 2, 3))
gather_forms_spline_1d, gather_forms_spline_2d, gather_forms_spline_3d = (_forms_kernel(d, False) for d in (1, 2, 3))

#: every kernel by its catalogue name (catalog.py:51-88)
KERNELS = {name: globals()[name] for name in KERNEL_NAMES}


def halo(op: str, data: torch.Tensor, halo_width: int, periodic) -> None:
    """Fill (op "fill") or fold (op "fold") the halos of one block's storage in place."""
    data = _field(data, "storage")
    ndim = data.dim() - 1
    shape = np.asarray(data.shape[:ndim], dtype=np.int64)
    per = np.asarray([int(bool(p)) for p in periodic], dtype=np.int64)
    with torch.cuda.device(data.device):
        _lib.call("mpf_halo_fill" if op == "fill" else "mpf_halo_fold", ndim, shape.ctypes.data, data.shape[-1],
                  int(halo_width), per.ctypes.data, data.data_ptr(), stream_handle(data.device))


def divergence(edges: torch.Tensor, periodic) -> torch.Tensor:
    edges = _field(edges, "edges")
    ndim = edges.dim() - 1
    shape = np.asarray(edges.shape[:ndim], dtype=np.int64)
    per = np.asarray([int(bool(p)) for p in periodic], dtype=np.int64)
    out = torch.empty(edges.shape[:ndim], dtype=torch.float64, device=edges.device)
    with torch.cuda.device(edges.device):
        _lib.call("mpf_divergence", ndim, shape.ctypes.data, per.ctypes.data, edges.data_ptr(), out.data_ptr(),
                  stream_handle(edges.device))
    return out


def gradient(nodes: torch.Tensor, periodic) -> torch.Tensor:
    nodes = _field(nodes, "nodes")
    ndim = nodes.dim()
    shape = np.asarray(nodes.shape, dtype=np.int64)
    per = np.asarray([int(bool(p)) for p in periodic], dtype=np.int64)
    out = torch.empty((*nodes.shape, ndim), dtype=torch.float64, device=nodes.device)
    with torch.cuda.device(nodes.device):
        _lib.call("mpf_gradient", ndim, shape.ctypes.data, per.ctypes.data, nodes.data_ptr(), out.data_ptr(),
                  stream_handle(nodes.device))
    return out
