"""Weightings: a deposit and its gather for each number of dimensions (the reference's
plugin API, src/mpifield/weighting.py:100-399), on the gfx950 kernels.

The kernel convention is the reference's (weighting.py:22-38)::

    scatter(positions, values, field, *geometry)   # adds the values into field
    gather(positions, field, result, *geometry)    # overwrites result

with ``geometry = (*lower, *cell_size, *num_nodes, *code)``. Any callable with those
arguments plugs into a `Weighting`; the built-in ones call libmpfhip.so. Fields passed
as NumPy arrays are computed on the GPU and copied back: there is no host kernel.
"""

from __future__ import annotations

from collections.abc import Callable, Mapping
from dataclasses import dataclass
from types import MappingProxyType
from typing import Any

import torch

from mpifield_amd import kernels as K
from mpifield_amd._types import BlockGrid

__all__ = ["CUBIC", "LINEAR", "QUADRATIC", "Weighting"]

KernelFunction = Callable[..., Any]


def _frozen(kernels: Mapping[int, KernelFunction], role: str) -> Mapping[int, Any]:
    for ndim in kernels:
        if ndim not in (1, 2, 3):
            raise ValueError(f"{role} kernels are for 1, 2 or 3 dimensions, got {ndim}")
    return MappingProxyType(dict(kernels))


def _check_clamps(weighting: Any, grid: BlockGrid) -> None:
    if any(grid.clamped) and not weighting.clamps:
        raise ValueError(
            f"the {weighting.name} weighting does not take clamped axes; use mpifield_amd.splines(degree, clamped=True)"
        )


class _DeviceArray:
    """A float64 C-contiguous GPU view of ``array`` for a kernel to write, copied back on exit
    when ``array`` is a NumPy array or a tensor the kernel cannot take as it is."""

    def __init__(self, array, dtype=torch.float64, contiguous: bool = True):
        self.array = array
        if isinstance(array, torch.Tensor) and array.is_cuda and array.dtype == dtype and (
            array.is_contiguous() or not contiguous
        ):
            self.tensor, self.copy_back = array, False
        else:
            src = torch.as_tensor(array) if not isinstance(array, torch.Tensor) else array
            self.tensor = src.to(device=K.device(), dtype=dtype).contiguous()
            self.copy_back = True

    def __enter__(self) -> torch.Tensor:
        return self.tensor

    def __exit__(self, *exc) -> None:
        if exc[0] is None and self.copy_back:
            if isinstance(self.array, torch.Tensor):
                self.array.copy_(self.tensor)
            else:
                self.array[...] = self.tensor.cpu().numpy()


def _is_bins(a) -> bool:
    from mpifield_amd.binning import MarkerBins

    return isinstance(a, MarkerBins)


def _markers(a) -> torch.Tensor:
    """Marker rows on the GPU, float32 kept as float32, anything else as float64."""
    t = K.as_device(a)
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    return t


@dataclass(frozen=True, eq=False, repr=False)
class Weighting:
    """A deposit and its gather, for each number of dimensions they cover.

    ``halo`` is the halo nodes a split field needs, ``(s + 1) // 2`` for a stencil of
    ``s`` nodes per axis; ``degree`` the B-spline degree (or None); ``clamps`` says that the
    kernels take clamped axes (boundary code 2).
    """

    name: str
    scatter: Mapping[int, KernelFunction]
    gather: Mapping[int, KernelFunction]
    halo: int = 1
    degree: int | None = None
    clamps: bool = False

    def __post_init__(self) -> None:
        if self.halo < 0:
            raise ValueError(f"halo must be >= 0, got {self.halo}")
        if self.degree is not None and self.degree < 0:
            raise ValueError(f"degree must be >= 0, got {self.degree}")
        object.__setattr__(self, "scatter", _frozen(self.scatter, "scatter"))
        object.__setattr__(self, "gather", _frozen(self.gather, "gather"))

    def _kernel(self, kernels: Mapping[int, KernelFunction], ndim: int, role: str):
        try:
            return kernels[ndim]
        except KeyError:
            have = ", ".join(f"{n}D" for n in sorted(kernels)) or "none"
            raise ValueError(f"the {self.name} weighting has no {ndim}D {role} kernel (has {have})") from None

    def scatter_kernel(self, ndim: int) -> KernelFunction:
        return self._kernel(self.scatter, ndim, "scatter")

    def gather_kernel(self, ndim: int) -> KernelFunction:
        return self._kernel(self.gather, ndim, "gather")

    def bin(self, positions, grid: BlockGrid):
        """The positions sorted by this weighting's stencil cell on ``grid`` (a `MarkerBins`),
        to pass as ``positions`` to `scatter_block` and `gather_block` while the markers stay
        where they are: those calls then skip the sort.  The built-in weightings only; the
        caller keeps the bins and bins again after the markers move."""
        from mpifield_amd import binning

        _check_clamps(self, grid)
        return binning.bin_markers(_markers(positions), None, binning.kind_of(self), tuple(grid.geometry))

    def _bins(self, bins, grid: BlockGrid):
        from mpifield_amd import binning

        if bins.values is not None:
            raise ValueError("bins passed as positions must be made from positions alone (Weighting.bin)")
        if (bins.kind, bins.degree) != binning.kind_of(self) or tuple(bins.geometry) != tuple(grid.geometry):
            raise ValueError(f"these bins were made for another weighting or grid than {self.name} on this block")
        return binning

    def scatter_block(self, positions, values, field, grid: BlockGrid) -> None:
        """Deposit ``values`` at ``positions`` into ``field``, a block on ``grid`` (weighting.py:195-213).

        ``field`` is ``num_nodes + (ncomp,)`` on the GPU (a torch tensor) or the host (a
        NumPy array, computed on the GPU and copied back); a 1-D ``values`` is one component.
        ``positions`` may be the `MarkerBins` of `bin` (values still in the markers' order).
        """
        _check_clamps(self, grid)
        if _is_bins(positions):
            binning = self._bins(positions, grid)
            val = _markers(values).to(positions.positions.dtype).reshape(-1, positions.n)
            with _DeviceArray(field) as buffer:
                if K.current_mode() == "fixed":
                    binning.scatter_binned(binning.with_values(positions, val), buffer)
                else:  # float64 atomics: the direct kernels on the markers in bins order
                    sval = positions.to_bins_order(val)
                    self.scatter_kernel(len(grid.periodic))(positions.positions, sval, buffer, *grid.geometry)
            return
        kernel = self.scatter_kernel(len(grid.periodic))
        pos = _markers(positions)
        val = _markers(values).reshape(-1, pos.shape[-1])
        with _DeviceArray(field) as buffer:
            kernel(pos, val, buffer, *grid.geometry)

    def gather_block(self, positions, field, result, grid: BlockGrid) -> None:
        """Interpolate ``field`` at ``positions`` into ``result`` ``(ncomp, N)`` (weighting.py:233-250);
        ``positions`` may be the `MarkerBins` of `bin` (the result is in the markers' order)."""
        _check_clamps(self, grid)
        if _is_bins(positions):
            binning = self._bins(positions, grid)
            values = K.as_device(field, torch.float64).contiguous()
            sorted_out = torch.empty((values.shape[-1], positions.n), dtype=positions.positions.dtype,
                                     device=values.device)
            binning.gather_binned(positions, values, sorted_out)
            with _DeviceArray(result, dtype=result.dtype if isinstance(result, torch.Tensor) and result.dtype in (
                    torch.float32, torch.float64) else torch.float64, contiguous=False) as out:
                out.copy_(positions.unsort(sorted_out).reshape(out.shape))
            return
        kernel = self.gather_kernel(len(grid.periodic))
        pos = _markers(positions)
        values = K.as_device(field, torch.float64).contiguous()
        out_dtype = result.dtype if isinstance(result, torch.Tensor) else torch.float64
        with _DeviceArray(result, dtype=out_dtype if out_dtype in (torch.float32, torch.float64) else torch.float64,
                          contiguous=False) as out:
            kernel(pos, values, out, *grid.geometry)

    def __repr__(self) -> str:
        dims = sorted(set(self.scatter) | set(self.gather))
        degree = "" if self.degree is None else f", degree={self.degree}"
        return f"Weighting({self.name!r}, ndim={tuple(dims)}, halo={self.halo}{degree})"


def _scatter_linear_1d(positions, values, field, lower_x, cell_size_x, num_x, periodic_x) -> None:
    """LINEAR's 1D deposit: the kernel one component at a time (weighting.py:259-295)."""
    for component in range(field.shape[1]):
        column = field[:, component]
        K.scatter_linear_1d(positions[0], values[component], column, lower_x, cell_size_x, num_x, periodic_x)


def _rows(field):
    if not field.is_contiguous():
        raise ValueError("the linear 3D kernels need a C-contiguous field")
    return field.reshape(-1, field.shape[-1])


def _scatter_linear_3d(positions, values, field, *geometry) -> None:
    """LINEAR's 3D deposit on the grid's rows (weighting.py:305-329)."""
    K.scatter_linear_3d(positions, values, _rows(field), *geometry)


def _gather_linear_3d(positions, field, result, *geometry) -> None:
    K.gather_linear_3d(positions, _rows(field), result, *geometry)


#: Linear (cloud-in-cell) weights, two nodes per axis.
LINEAR = Weighting(
    "linear",
    scatter={1: _scatter_linear_1d, 2: K.scatter_linear_2d, 3: _scatter_linear_3d},
    gather={1: K.gather_linear_1d, 2: K.gather_linear_2d, 3: _gather_linear_3d},
    halo=1,
    degree=1,
)

#: Quadratic B-spline ("triangular-shaped cloud") weights, three nodes per axis.
QUADRATIC = Weighting(
    "quadratic",
    scatter={1: K.scatter_quadratic_1d, 2: K.scatter_quadratic_2d, 3: K.scatter_quadratic_3d},
    gather={1: K.gather_quadratic_1d, 2: K.gather_quadratic_2d, 3: K.gather_quadratic_3d},
    halo=2,
    degree=2,
)


@dataclass(frozen=True)
class _WithDegree:
    """A spline kernel with its degree appended: a kernel of the weighting convention."""

    kernel: KernelFunction
    degree: int

    def __call__(self, *arguments: Any) -> None:
        self.kernel(*arguments, self.degree)


#: Cubic B-spline weights, four nodes per axis, on the spline kernels (degree 3).
CUBIC = Weighting(
    "cubic",
    scatter={n: _WithDegree(k, 3) for n, k in
             ((1, K.scatter_spline_1d), (2, K.scatter_spline_2d), (3, K.scatter_spline_3d))},
    gather={n: _WithDegree(k, 3) for n, k in
            ((1, K.gather_spline_1d), (2, K.gather_spline_2d), (3, K.gather_spline_3d))},
    halo=2,
    degree=3,
    clamps=True,
)


def spline_nodes(name: str, degree: int) -> Weighting:
    """The node weighting of a degree on the spline kernels (any boundary code)."""
    scatter = {1: K.scatter_spline_1d, 2: K.scatter_spline_2d, 3: K.scatter_spline_3d}
    gather = {1: K.gather_spline_1d, 2: K.gather_spline_2d, 3: K.gather_spline_3d}
    return Weighting(
        name,
        scatter={n: _WithDegree(scatter[n], degree) for n in (1, 2, 3)},
        gather={n: _WithDegree(gather[n], degree) for n in (1, 2, 3)},
        halo=2 if degree > 1 else 1,
        degree=degree,
        clamps=True,
    )
