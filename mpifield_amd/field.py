"""A field on the nodes of a uniform grid, stored on one GPU (the hot-path part of the
reference's Field, src/mpifield/field.py: block_grid 150-245, node_volumes 563-588,
fill_halos / exchange_halos 601-630, gather / scatter / deposit_density 1097-1218).

One process, one block: the whole grid is stored on the GPU as a float64 torch tensor
``num_gridpoints (+ 2 halo per axis) + (num_components,)``. With ``num_ghost_points`` > 0
the block carries halo nodes like a field split over ranks; on a periodic axis they hold
the wrapped nodes, and `exchange_halos` / `fill_halos` fold and fill them across the seam
on the device (the single-block case of mpiarray's halo exchange).

Grid spacing follows simgeometry's rule: a bounded axis of n nodes over [lo, hi] has
h = (hi - lo) / (n - 1), a periodic one h = (hi - lo) / n (n nodes, n cells).
"""

from __future__ import annotations

from typing import Any

import numpy as np
import torch

from mpifield_amd import kernels as K
from mpifield_amd._types import BlockGrid, Bounds
from mpifield_amd.weighting import LINEAR, Weighting, _is_bins

__all__ = ["Field"]


def _periodic_flags(periodic, ndim: int) -> tuple[bool, ...]:
    if periodic is None:
        return (False,) * ndim
    if isinstance(periodic, (bool, np.bool_)):
        return (bool(periodic),) * ndim
    flags = tuple(bool(p) for p in periodic)
    if len(flags) != ndim:
        raise ValueError(f"periodic needs {ndim} entries, got {len(flags)}")
    return flags


class Field:
    """A field on the nodes of a uniform grid, on one GPU.

    ``comm`` and ``domain_decomposition`` are accepted for the reference's signature; a
    field here is always one block (``distributed`` when it carries halo nodes).
    """

    def __init__(
        self,
        ndim: int,
        lower_bound: tuple,
        upper_bound: tuple,
        num_gridpoints: int | tuple[int, ...],
        num_components: int = 1,
        comm: Any = None,
        num_ghost_points: int = 0,
        domain_decomposition: Any = None,
        periodic: bool | tuple[bool, ...] | None = None,
        dtype: type = float,
        device: torch.device | str | None = None,
    ) -> None:
        if isinstance(num_gridpoints, (int, np.integer)):
            num_gridpoints = (int(num_gridpoints),) * ndim
        if len(num_gridpoints) != ndim:
            raise ValueError(f"num_gridpoints needs {ndim} entries")
        if num_components < 1:
            raise ValueError(f"num_components must be >= 1, got {num_components}")
        if np.dtype(dtype) != np.float64:
            raise ValueError("the kernels take float64 fields")
        self.ndim = ndim
        self.num_gridpoints = tuple(int(n) for n in num_gridpoints)
        self.periodic = _periodic_flags(periodic, ndim)
        lower = np.asarray(lower_bound, dtype=float).reshape(-1)
        upper = np.asarray(upper_bound, dtype=float).reshape(-1)
        if lower.size == 1 and ndim > 1:
            lower = np.full(ndim, lower[0])
        if upper.size == 1 and ndim > 1:
            upper = np.full(ndim, upper[0])
        self.bounds = Bounds(lower_bound=lower, upper_bound=upper)
        cells = np.array([n if p else n - 1 for n, p in zip(self.num_gridpoints, self.periodic)], dtype=float)
        self.grid_spacing = (upper - lower) / cells
        self.num_ghost_points = int(num_ghost_points)
        self._num_components = int(num_components)
        self.comm = comm
        self.device = torch.device(device) if device is not None else K.device()
        h = self.num_ghost_points
        shape = tuple(n + 2 * h for n in self.num_gridpoints) + (self._num_components,)
        self._storage = torch.zeros(shape, dtype=torch.float64, device=self.device)

    # ---- layout ---------------------------------------------------------------------

    @property
    def num_components(self) -> int:
        return self._num_components

    @property
    def dtype(self):
        return np.dtype(np.float64)

    @property
    def distributed(self) -> bool:
        """True when the block carries halo nodes (the halo path of a split field)."""
        return self.num_ghost_points > 0

    @property
    def local_with_halos(self) -> torch.Tensor:
        """The stored block, halo nodes included: the array the kernels read and write."""
        return self._storage

    @property
    def local(self) -> torch.Tensor:
        """The owned nodes, without halo nodes: a writable view."""
        h = self.num_ghost_points
        if h == 0:
            return self._storage
        return self._storage[tuple(slice(h, -h) for _ in range(self.ndim))]

    def block_grid(self, cpu: bool = False, inside_domain: bool = False) -> BlockGrid:
        """The grid of the stored block, halos included (field.py:150-245).

        A periodic axis is periodic for the kernels only without halo nodes. With
        ``inside_domain`` the halo nodes beyond a bounded edge are left out.
        """
        h = self.num_ghost_points
        first, last = [], []
        for axis in range(self.ndim):
            trim = inside_domain and not self.periodic[axis]
            low = h if trim else 0
            high = h if trim else 0
            first.append(low)
            last.append(self.num_gridpoints[axis] + 2 * h - high)
        spacing = np.asarray(self.grid_spacing, dtype=float)
        start = np.array([-h + low for low in first], dtype=float)
        nodes = np.array([stop - low for low, stop in zip(first, last)], dtype=int)
        lower = self.bounds.lower_bound + start * spacing
        upper = lower + (nodes - 1) * spacing
        periodic = tuple(bool(p) and h == 0 for p in self.periodic)
        return BlockGrid(
            bounds=Bounds(lower_bound=lower, upper_bound=upper),
            grid_spacing=spacing,
            num_gridpoints=nodes,
            periodic=periodic,
            geometry=(*(float(v) for v in lower), *(float(v) for v in spacing), *(int(n) for n in nodes),
                      *(int(p) for p in periodic)),
            index=tuple(slice(low, stop) for low, stop in zip(first, last)),
        )

    # ---- data -----------------------------------------------------------------------

    def set_field_data(self, data) -> None:
        """Set the owned nodes from ``data`` ``num_gridpoints`` (+ ``(num_components,)``)."""
        t = torch.as_tensor(np.asarray(data) if not isinstance(data, torch.Tensor) else data)
        t = t.to(device=self.device, dtype=torch.float64)
        if t.dim() == self.ndim:
            t = t.unsqueeze(-1)
        self.local[...] = t

    def to_numpy(self) -> np.ndarray:
        return self.local.cpu().numpy()

    def _like(self, num_components: int | None = None) -> "Field":
        f = Field.__new__(Field)
        f.__dict__.update({k: v for k, v in self.__dict__.items() if k != "_storage"})
        f._num_components = self._num_components if num_components is None else num_components
        shape = tuple(self._storage.shape[:-1]) + (f._num_components,)
        f._storage = torch.zeros(shape, dtype=torch.float64, device=self.device)
        return f

    def node_volumes(self) -> "Field":
        """The volume each node collects in a deposit (field.py:563-588): one cell, half of
        it at the first and last node of a bounded axis; halos filled."""
        volumes = self._like(num_components=1)
        local = torch.ones(self.num_gridpoints, dtype=torch.float64, device=self.device)
        for axis in range(self.ndim):
            n = self.num_gridpoints[axis]
            factor = torch.full((n,), float(self.grid_spacing[axis]), dtype=torch.float64, device=self.device)
            if not self.periodic[axis]:
                factor[0] *= 0.5
                factor[-1] *= 0.5
            shape = [1] * self.ndim
            shape[axis] = n
            local = local * factor.reshape(shape)
        volumes.local[..., 0] = local
        if volumes.distributed:
            volumes.fill_halos()
        return volumes

    # ---- halos ----------------------------------------------------------------------

    def fill_halos(self) -> None:
        """Fill the halo nodes across the periodic seams (update_halos, field.py:601-613)."""
        if self.num_ghost_points:
            K.halo("fill", self._storage, self.num_ghost_points, self.periodic)

    def exchange_halos(self) -> None:
        """Add the halo nodes into their owners across the seams and clear them
        (accumulate_halos, field.py:615-630)."""
        if self.num_ghost_points:
            K.halo("fold", self._storage, self.num_ghost_points, self.periodic)

    # ---- gather and scatter ---------------------------------------------------------

    def _check_halo(self, weighting: Weighting) -> None:
        if self.num_ghost_points < weighting.halo and self.distributed:
            raise ValueError(
                f"the {weighting.name} weighting needs num_ghost_points >= {weighting.halo} on a field split over "
                f"ranks, this one has {self.num_ghost_points}"
            )

    def _positions(self, positions) -> torch.Tensor:
        if isinstance(positions, float):
            if self.ndim != 1:
                raise ValueError(f"a float is one point of a 1D field, not {self.ndim}D")
            positions = [[positions]]
        t = K.as_device(positions)
        if t.dtype not in (torch.float32, torch.float64):
            t = t.to(torch.float64)
        if t.dim() == 1:
            if self.ndim == 1:
                return t.reshape(1, -1)
            if t.shape[0] != self.ndim:
                raise ValueError(f"Expected {self.ndim} coordinates, got {t.shape[0]}")
            return t.reshape(self.ndim, 1)
        if t.dim() == 2:
            if t.shape[0] != self.ndim:
                raise ValueError(f"Expected positions with shape ({self.ndim}, N), got {tuple(t.shape)}")
            return t
        raise ValueError(f"Unsupported shape: {tuple(t.shape)}")

    def _values(self, values, n: int, dtype) -> torch.Tensor:
        if isinstance(values, (float, int)):
            return torch.full((self.num_components, n), float(values), dtype=dtype, device=self.device)
        t = K.as_device(values)
        if t.dtype not in (torch.float32, torch.float64):
            t = t.to(torch.float64)
        if t.dim() == 1:
            if t.shape[0] == n:
                t = t.reshape(1, n)
            elif n == 1 and t.shape[0] == self.num_components:
                t = t.reshape(self.num_components, 1)
            else:
                raise ValueError(f"Expected {n} values (or {self.num_components} components of one marker), "
                                 f"got {tuple(t.shape)}")
        elif t.dim() != 2 or t.shape[1] != n:
            raise ValueError(f"Expected values with shape ({self.num_components}, {n}), got {tuple(t.shape)}")
        if t.shape[0] != self.num_components:
            raise ValueError(f"Expected {self.num_components} components, got {t.shape[0]}")
        return t

    def bin_markers(self, positions, weighting: Weighting = LINEAR):
        """The positions binned for ``weighting`` on this field's block (`Weighting.bin`): pass
        them as ``positions`` to `gather` and `scatter` to skip the sort while the markers
        stay where they are."""
        return weighting.bin(self._positions(positions), self.block_grid(inside_domain=True))

    def gather(self, positions, weighting: Weighting = LINEAR):
        """Interpolate the field at marker positions (field.py:1097-1126); ``(ncomp, N)``.

        The result is a GPU tensor of the positions' float dtype, or a NumPy array when the
        positions were given on the host.  ``positions`` may be the bins of `bin_markers`.
        """
        self._check_halo(weighting)
        host = not isinstance(positions, torch.Tensor) and not _is_bins(positions)
        pos = positions if _is_bins(positions) else self._positions(positions)
        grid = self.block_grid(inside_domain=True)
        stored = self._storage[grid.index]
        if not stored.is_contiguous():
            stored = stored.contiguous()
        n, dtype = (pos.n, pos.positions.dtype) if _is_bins(pos) else (pos.shape[1], pos.dtype)
        result = torch.zeros((self.num_components, n), dtype=dtype, device=self.device)
        weighting.gather_block(pos, stored, result, grid)
        return result.cpu().numpy().astype(np.float64) if host else result

    def scatter(self, positions, values, weighting: Weighting = LINEAR) -> None:
        """Add marker values to the stored nodes, halos included (field.py:1128-1180).

        On a field with halos call `exchange_halos` afterwards.  ``positions`` may be the
        bins of `bin_markers` (the values in the markers' order).
        """
        self._check_halo(weighting)
        if _is_bins(positions):
            pos = positions
            val = self._values(values, pos.n, pos.positions.dtype)
        else:
            pos = self._positions(positions)
            val = self._values(values, pos.shape[1], pos.dtype)
        grid = self.block_grid(inside_domain=True)
        block = self._storage[grid.index]
        weighting.scatter_block(pos, val, block, grid)

    def deposit_density(self, positions, values, weighting: Weighting = LINEAR) -> None:
        """Add the density of the marker values: a deposit over the node volumes (field.py:1182-1218)."""
        deposit = self._like()
        deposit.scatter(positions, values, weighting)
        if deposit.distributed:
            deposit.exchange_halos()
        self.local.add_(deposit.local / self.node_volumes().local)
