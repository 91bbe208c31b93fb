"""ctypes front of the CPU oracle (``mpf_oracle.c``) -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import
this module, always as the checker or the timed host path, never as the product. The
functions take NumPy float64 arrays in the reference's kernel convention
(``src/mpifield/weighting.py:22-38``, ``src/mpifield/edges.py:23-31``): positions
``(ndim, N)`` with any row stride, values/results ``(ncomp, N)``, fields
``num_nodes + (ncomp,)`` C-contiguous, geometry ``(*lower, *cell_size, *num_nodes,
*code)``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libmpforacle.so"
REF_PATH = HERE / "_ref" / "libmpfref.so"

KIND = {"linear": 0, "quadratic": 1, "spline": 2}

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_INT = ctypes.c_int
_DBL = ctypes.c_double


def build() -> None:
    """Compile the oracle (and oracle/_ref where the reference tree is mounted)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def _load() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.orc_scatter.argtypes = [_INT, _INT, _INT, _P, _I64, _P, _I64, _I64, _I64, _P, _P, _P, _P, _P, _INT]
    lib.orc_scatter_policy.argtypes = [_INT, _INT, _INT, _P, _I64, _P, _I64, _I64, _I64, _P, _P, _P, _P, _P, _INT,
                                       _INT]
    lib.orc_gather.argtypes = [_INT, _INT, _INT, _P, _I64, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _INT]
    lib.orc_deposit_whitney.argtypes = [_INT, _P, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _INT]
    lib.orc_deposit_spline.argtypes = [_INT, _INT, _P, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _P]
    lib.orc_gather_edges_whitney.argtypes = [_INT, _P, _I64, _I64, _P, _P, _I64, _P, _P, _P, _P]
    lib.orc_gather_edges_spline.argtypes = [_INT, _INT, _P, _I64, _I64, _P, _P, _I64, _P, _P, _P, _P]
    lib.orc_gather_forms.argtypes = [_INT, _INT, _INT, _P, _I64, _I64, _P, _P, _I64, _P, _P, _P, _P, _INT]
    lib.orc_scatter_forms.argtypes = [_INT, _INT, _INT, _P, _I64, _P, _I64, _I64, _P, _P, _P, _P, _P, _INT]
    lib.orc_form_components.argtypes = [_INT, _INT]
    lib.orc_form_components.restype = _INT
    for name in ("orc_halo_fill", "orc_halo_fold"):
        getattr(lib, name).argtypes = [_INT, _P, _I64, _I64, _P, _P]
    lib.orc_divergence.argtypes = [_INT, _P, _P, _P, _P]
    lib.orc_gradient.argtypes = [_INT, _P, _P, _P, _P]
    lib.orc_axis_stencil.argtypes = [_INT, _INT, _DBL, _I64, _I64, _P, _P]
    lib.orc_axis_stencil.restype = _INT
    lib.orc_along_stencil.argtypes = [_DBL, _I64, _I64, _INT, _P, _P, _P]
    lib.orc_path_cuts.argtypes = [_P, _P, _INT, _DBL, _P, _P, _P]
    lib.orc_path_cuts.restype = _INT
    return lib


_LIB: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _rows(a: np.ndarray) -> tuple[np.ndarray, int]:
    """A float64 (rows, N) array with contiguous rows, and its row stride in elements."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    if a.strides[1] != 8 or a.strides[0] % 8:
        a = np.ascontiguousarray(a)
    return a, (a.strides[0] // 8 if a.shape[0] > 1 else a.shape[1])


def _geometry(geometry, ndim: int):
    g = list(geometry)
    lower = np.array(g[:ndim], dtype=np.float64)
    cell = np.array(g[ndim : 2 * ndim], dtype=np.float64)
    num = np.array(g[2 * ndim : 3 * ndim], dtype=np.int64)
    code = np.array(g[3 * ndim : 4 * ndim], dtype=np.int64)
    return lower, cell, num, code


def threads_default() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", "1"))


def scatter(kind: str, positions, values, field: np.ndarray, *geometry, degree: int = 0, threads: int = 1,
            guard: bool = True) -> None:
    """Add the deposit of ``values`` at ``positions`` into ``field`` (in place); ``guard=False``
    lifts the reference's MAX_BUFFER_VALUES bound on the threads' private copies."""
    ndim = len(geometry) // 4
    pos, ps = _rows(positions)
    val, vs = _rows(values)
    assert field.dtype == np.float64 and field.flags.c_contiguous
    lower, cell, num, code = _geometry(geometry, ndim)
    lib().orc_scatter_policy(KIND[kind], degree, ndim, _ptr(pos), ps, _ptr(val), vs, pos.shape[1], val.shape[0],
                             _ptr(field), _ptr(lower), _ptr(cell), _ptr(num), _ptr(code), threads, 1 if guard else 0)


def gather(kind: str, positions, field: np.ndarray, result: np.ndarray, *geometry, degree: int = 0,
           threads: int = 1) -> None:
    """Overwrite ``result`` ``(ncomp, N)`` with the interpolation of ``field`` at ``positions``."""
    ndim = len(geometry) // 4
    pos, ps = _rows(positions)
    field = np.ascontiguousarray(field, dtype=np.float64)
    assert result.dtype == np.float64 and result.strides[1] == 8
    rs = result.strides[0] // 8 if result.shape[0] > 1 else result.shape[1]
    lower, cell, num, code = _geometry(geometry, ndim)
    lib().orc_gather(KIND[kind], degree, ndim, _ptr(pos), ps, pos.shape[1], _ptr(field), field.shape[-1],
                     _ptr(result), rs, _ptr(lower), _ptr(cell), _ptr(num), _ptr(code), threads)


def deposit_whitney(start, end, charges, edges: np.ndarray, *geometry, threads: int = 1) -> None:
    ndim = len(geometry) // 4
    s, ss = _rows(start)
    e, es = _rows(end)
    q = np.ascontiguousarray(charges, dtype=np.float64).reshape(-1)
    assert edges.dtype == np.float64 and edges.flags.c_contiguous
    lower, cell, num, code = _geometry(geometry, ndim)
    lib().orc_deposit_whitney(ndim, _ptr(s), ss, _ptr(e), es, _ptr(q), q.size, _ptr(edges), _ptr(lower),
                              _ptr(cell), _ptr(num), _ptr(code), threads)


def deposit_spline(start, end, charges, edges: np.ndarray, *geometry, degree: int) -> None:
    ndim = len(geometry) // 4
    s, ss = _rows(start)
    e, es = _rows(end)
    q = np.ascontiguousarray(charges, dtype=np.float64).reshape(-1)
    assert edges.dtype == np.float64 and edges.flags.c_contiguous
    lower, cell, num, code = _geometry(geometry, ndim)
    lib().orc_deposit_spline(ndim, degree, _ptr(s), ss, _ptr(e), es, _ptr(q), q.size, _ptr(edges), _ptr(lower),
                             _ptr(cell), _ptr(num), _ptr(code))


def gather_edges(positions, edges: np.ndarray, result: np.ndarray, *geometry, degree: int = 0) -> None:
    """Overwrite ``result`` ``(ndim, N)`` with the field of the edge field ``edges`` at
    ``positions``: Whitney for degree 0 (gather_edges_*d), spline edges for degree 1..3
    (gather_edges_spline_*d)."""
    ndim = len(geometry) // 4
    pos, ps = _rows(positions)
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    assert result.dtype == np.float64 and result.strides[1] == 8 and result.shape == (ndim, pos.shape[1])
    rs = result.strides[0] // 8 if ndim > 1 else result.shape[1]
    lower, cell, num, code = _geometry(geometry, ndim)
    if degree == 0:
        lib().orc_gather_edges_whitney(ndim, _ptr(pos), ps, pos.shape[1], _ptr(edges), _ptr(result), rs, _ptr(lower),
                                       _ptr(cell), _ptr(num), _ptr(code))
    else:
        lib().orc_gather_edges_spline(ndim, degree, _ptr(pos), ps, pos.shape[1], _ptr(edges), _ptr(result), rs,
                                      _ptr(lower), _ptr(cell), _ptr(num), _ptr(code))


def form_components(form: int, ndim: int) -> int:
    """The components of a ``form``-form in ``ndim`` dimensions (0 for no such form)."""
    return int(lib().orc_form_components(form, ndim))


def _form_field(field: np.ndarray, form: int, ndim: int) -> None:
    if not (1 <= form <= ndim <= 3):
        raise ValueError(f"forms of {ndim} dimensions are 1 to {ndim}, got {form}")
    if field.shape[-1] != form_components(form, ndim):
        raise ValueError(f"a {form}-form in {ndim}D has {form_components(form, ndim)} components")


def gather_forms(positions, field: np.ndarray, result: np.ndarray, *geometry, degree: int, form: int,
                 threads: int = 1) -> None:
    """Overwrite ``result`` ``(ncomp, N)`` with the ``form``-form ``field`` (*extent, ncomp) at
    ``positions`` (gather_forms_spline_*d)."""
    ndim = len(geometry) // 4
    pos, ps = _rows(positions)
    field = np.ascontiguousarray(field, dtype=np.float64)
    _form_field(field, form, ndim)
    assert result.dtype == np.float64 and result.strides[1] == 8 and result.shape == (field.shape[-1], pos.shape[1])
    rs = result.strides[0] // 8 if result.shape[0] > 1 else result.shape[1]
    lower, cell, num, code = _geometry(geometry, ndim)
    lib().orc_gather_forms(ndim, degree, form, _ptr(pos), ps, pos.shape[1], _ptr(field), _ptr(result), rs,
                           _ptr(lower), _ptr(cell), _ptr(num), _ptr(code), threads)


def scatter_forms(positions, values, field: np.ndarray, *geometry, degree: int, form: int, threads: int = 1) -> None:
    """Add the deposit of ``values`` (ncomp, N) at ``positions`` into the ``form``-form ``field``
    (scatter_forms_spline_*d)."""
    ndim = len(geometry) // 4
    pos, ps = _rows(positions)
    val, vs = _rows(values)
    assert field.dtype == np.float64 and field.flags.c_contiguous
    _form_field(field, form, ndim)
    assert val.shape == (field.shape[-1], pos.shape[1])
    lower, cell, num, code = _geometry(geometry, ndim)
    lib().orc_scatter_forms(ndim, degree, form, _ptr(pos), ps, _ptr(val), vs, pos.shape[1], _ptr(field),
                            _ptr(lower), _ptr(cell), _ptr(num), _ptr(code), threads)


def _halo(name: str, data: np.ndarray, halo: int, periodic) -> None:
    assert data.dtype == np.float64 and data.flags.c_contiguous
    ndim = data.ndim - 1
    shape = np.array(data.shape[:ndim], dtype=np.int64)
    per = np.array([int(bool(p)) for p in periodic], dtype=np.int64)
    getattr(lib(), name)(ndim, _ptr(shape), data.shape[-1], halo, _ptr(per), _ptr(data))


def halo_fill(data: np.ndarray, halo: int, periodic) -> None:
    _halo("orc_halo_fill", data, halo, periodic)


def halo_fold(data: np.ndarray, halo: int, periodic) -> None:
    _halo("orc_halo_fold", data, halo, periodic)


def divergence(edges: np.ndarray, periodic) -> np.ndarray:
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    ndim = edges.ndim - 1
    shape = np.array(edges.shape[:ndim], dtype=np.int64)
    per = np.array([int(bool(p)) for p in periodic], dtype=np.int64)
    out = np.zeros(edges.shape[:ndim])
    lib().orc_divergence(ndim, _ptr(shape), _ptr(per), _ptr(edges), _ptr(out))
    return out


def gradient(nodes: np.ndarray, periodic) -> np.ndarray:
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    ndim = nodes.ndim
    shape = np.array(nodes.shape, dtype=np.int64)
    per = np.array([int(bool(p)) for p in periodic], dtype=np.int64)
    out = np.zeros((*nodes.shape, ndim))
    lib().orc_gradient(ndim, _ptr(shape), _ptr(per), _ptr(nodes), _ptr(out))
    return out


def axis_stencil(kind: str, rel: float, num: int, code: int, degree: int = 0):
    node = np.zeros(4, dtype=np.int64)
    weight = np.zeros(4)
    size = lib().orc_axis_stencil(KIND[kind], degree, float(rel), num, code, _ptr(node), _ptr(weight))
    return node[:size].copy(), weight[:size].copy()


def along_stencil(rel: float, num: int, code: int, degree: int):
    edge = np.zeros(4, dtype=np.int64)
    weight = np.zeros(4)
    exists = np.zeros(4, dtype=np.int32)
    lib().orc_along_stencil(float(rel), num, code, degree, _ptr(edge), _ptr(weight), _ptr(exists))
    return edge[:degree].copy(), weight[:degree].copy(), exists[:degree].astype(bool)


def path_cuts(r0, r1, shift: float, num, code) -> np.ndarray:
    r0 = np.ascontiguousarray(r0, dtype=np.float64)
    r1 = np.ascontiguousarray(r1, dtype=np.float64)
    num = np.ascontiguousarray(num, dtype=np.int64)
    code = np.ascontiguousarray(code, dtype=np.int64)
    cuts = np.zeros(32)
    count = lib().orc_path_cuts(_ptr(r0), _ptr(r1), r0.size, float(shift), _ptr(num), _ptr(code), _ptr(cuts))
    return cuts[:count].copy()


# ----------------------------------------------------------------------------------------
# The reference's own C host kernels, built from its sources into oracle/_ref.

_REF: ctypes.CDLL | None = None


def ref_available() -> bool:
    return REF_PATH.exists()


def ref_lib() -> ctypes.CDLL:
    global _REF
    if _REF is None:
        _REF = ctypes.CDLL(str(REF_PATH))
        _REF.scatter_linear_1d.argtypes = [_P, _P, _P, _P, _P, _P, _DBL, _DBL, _I64, _I64]
        _REF.scatter_rows_3d.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P] + [_DBL] * 6 + [_I64] * 6
    return _REF


def ref_scatter_linear_1d(positions: np.ndarray, values: np.ndarray, field: np.ndarray, lower, cell, num, periodic):
    """``scatter_linear_1d_c.c``: one component, ``(N,)`` markers into an ``(n,)`` grid."""
    positions = np.ascontiguousarray(positions, dtype=np.float64)
    values = np.ascontiguousarray(values, dtype=np.float64)
    assert field.flags.c_contiguous and field.dtype == np.float64
    shp = np.array([positions.size], dtype=np.int64)
    vshp = np.array([values.size], dtype=np.int64)
    fshp = np.array([field.size], dtype=np.int64)
    ref_lib().scatter_linear_1d(_ptr(positions), _ptr(shp), _ptr(values), _ptr(vshp), _ptr(field), _ptr(fshp),
                                float(lower), float(cell), int(num), int(periodic))


def ref_scatter_rows_3d(positions: np.ndarray, values: np.ndarray, rows: np.ndarray, *geometry):
    """``scatter_rows_3d_c.c``: one component, positions rows px/py/pz, grid as ``(nx*ny*nz,)`` rows."""
    px, py, pz = (np.ascontiguousarray(r, dtype=np.float64) for r in positions)
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    assert rows.flags.c_contiguous and rows.dtype == np.float64
    shp = np.array([px.size], dtype=np.int64)
    fshp = np.array([rows.size], dtype=np.int64)
    ref_lib().scatter_rows_3d(_ptr(px), _ptr(shp), _ptr(py), _ptr(shp), _ptr(pz), _ptr(shp), _ptr(v), _ptr(shp),
                              _ptr(rows), _ptr(fshp), *[float(x) for x in geometry[:6]],
                              *[int(x) for x in geometry[6:]])
