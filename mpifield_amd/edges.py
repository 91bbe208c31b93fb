"""Edge fields on a block: the charge-conserving path deposit and the discrete gradient /
divergence that check it (the reference's EdgeWeighting, src/mpifield/edges.py:74-309).

An edge field is ``num_nodes + (ndim,)`` float64 (``num_nodes + degree - 1`` on a clamped
axis): ``edges[..., a]`` is the line integral along the edge from a node to its
neighbour along axis ``a``. ``deposit_paths_block`` adds the current of charges moving
along straight paths (not wrapped on a periodic axis, clamped on a bounded one), so that
``divergence_block`` of it is the change of the node weighting's charge of the same degree.
``gather_block`` writes the field of an edge field at markers (the transpose of the deposit:
the current times an edge field is the work of the field along the paths).
"""

from __future__ import annotations

from collections.abc import Callable, Mapping
from dataclasses import dataclass
from types import MappingProxyType
from typing import Any

import torch

from mpifield_amd import kernels as K
from mpifield_amd._types import BlockGrid
from mpifield_amd.weighting import _DeviceArray, _WithDegree, _check_clamps, _markers

__all__ = ["CUBIC_EDGES", "QUADRATIC_EDGES", "WHITNEY", "EdgeWeighting"]

KernelFunction = Callable[..., Any]


def _checked(kernels: Mapping[int, KernelFunction], role: str) -> Mapping[int, Any]:
    for ndim in kernels:
        if ndim not in (1, 2, 3):
            raise ValueError(f"{role} kernels are for 1, 2 or 3 dimensions, got {ndim}")
    return MappingProxyType(dict(kernels))


@dataclass(frozen=True, eq=False, repr=False)
class EdgeWeighting:
    """The path deposit (and, where mapped, the edge gather) for each number of dimensions."""

    name: str
    gather: Mapping[int, KernelFunction]
    deposit_paths: Mapping[int, KernelFunction]
    halo: int = 1
    degree: int | None = None
    clamps: bool = False

    def __post_init__(self) -> None:
        if self.halo < 0:
            raise ValueError(f"halo must be >= 0, got {self.halo}")
        if self.degree is not None and self.degree < 1:
            raise ValueError(f"degree must be >= 1, got {self.degree}")
        object.__setattr__(self, "gather", _checked(self.gather, "gather"))
        object.__setattr__(self, "deposit_paths", _checked(self.deposit_paths, "deposit_paths"))

    def _kernel(self, kernels: Mapping[int, KernelFunction], ndim: int, role: str):
        try:
            return kernels[ndim]
        except KeyError:
            have = ", ".join(f"{n}D" for n in sorted(kernels)) or "none"
            raise ValueError(f"the {self.name} edge weighting has no {ndim}D {role} kernel (has {have})") from None

    def gather_kernel(self, ndim: int) -> KernelFunction:
        return self._kernel(self.gather, ndim, "gather")

    def deposit_paths_kernel(self, ndim: int) -> KernelFunction:
        return self._kernel(self.deposit_paths, ndim, "deposit_paths")

    def _check(self, edges, grid: BlockGrid) -> int:
        """The number of dimensions of ``grid``, checking the shape of ``edges`` (edges.py:145-167)."""
        _check_clamps(self, grid)
        ndim = len(grid.periodic)
        ends = grid.clamped or (False,) * ndim
        extra = 0 if self.degree is None else self.degree - 1
        expected = (*(int(n) + (extra if end else 0) for n, end in zip(grid.num_gridpoints, ends)), ndim)
        if tuple(edges.shape) != expected:
            raise ValueError(f"an edge field on this grid is {expected}, got {tuple(edges.shape)}")
        return ndim

    def gather_block(self, positions, edges, result, grid: BlockGrid) -> None:
        kernel = self.gather_kernel(self._check(edges, grid))
        pos = _markers(positions)
        with _DeviceArray(result, contiguous=False) as out:
            kernel(pos, K.as_device(edges, torch.float64).contiguous(), out, *grid.geometry)

    def deposit_paths_block(self, start, end, charges, edges, grid: BlockGrid) -> None:
        """Add the current of ``charges`` moving from ``start`` to ``end`` to ``edges`` (edges.py:187-215)."""
        kernel = self.deposit_paths_kernel(self._check(edges, grid))
        s = _markers(start)
        e = _markers(end).to(s.dtype)
        q = _markers(charges).to(s.dtype).reshape(-1)
        if tuple(s.shape) != tuple(e.shape) or s.shape[-1] != q.shape[0]:
            raise ValueError(f"paths from {tuple(s.shape)} to {tuple(e.shape)} for {q.shape[0]} charges")
        with _DeviceArray(edges) as buffer:
            kernel(s, e, q, buffer, *grid.geometry)

    @staticmethod
    def gradient_block(nodes, grid: BlockGrid):
        """``G phi``: a node field's difference along every edge (edges.py:218-237)."""
        host = not isinstance(nodes, torch.Tensor)
        out = K.gradient(K.as_device(nodes, torch.float64).contiguous(), grid.periodic)
        return out.cpu().numpy() if host else out

    @staticmethod
    def divergence_block(edges, grid: BlockGrid):
        """``G^T j``: the change of the node charge a deposited current carries (edges.py:240-259)."""
        host = not isinstance(edges, torch.Tensor)
        out = K.divergence(K.as_device(edges, torch.float64).contiguous(), grid.periodic)
        return out.cpu().numpy() if host else out

    def __repr__(self) -> str:
        dims = sorted(set(self.gather) | set(self.deposit_paths))
        degree = "" if self.degree is None else f", degree={self.degree}"
        return f"EdgeWeighting({self.name!r}, ndim={tuple(dims)}, halo={self.halo}{degree})"


#: The lowest-order (Whitney) edge weighting: its deposit conserves LINEAR's charge.
WHITNEY = EdgeWeighting(
    "whitney",
    gather={1: K.gather_edges_1d, 2: K.gather_edges_2d, 3: K.gather_edges_3d},
    deposit_paths={1: K.deposit_paths_1d, 2: K.deposit_paths_2d, 3: K.deposit_paths_3d},
    halo=1,
    degree=1,
)


def spline_edges(name: str, degree: int) -> EdgeWeighting:
    """The edge weighting of B-spline degree ``degree`` on the spline path kernels."""
    return EdgeWeighting(
        name,
        gather={n: _WithDegree(k, degree) for n, k in
                ((1, K.gather_edges_spline_1d), (2, K.gather_edges_spline_2d), (3, K.gather_edges_spline_3d))},
        deposit_paths={n: _WithDegree(k, degree) for n, k in
                       ((1, K.deposit_paths_spline_1d), (2, K.deposit_paths_spline_2d),
                        (3, K.deposit_paths_spline_3d))},
        halo=2,
        degree=degree,
        clamps=True,
    )


#: The edge weighting that goes with QUADRATIC: its deposit conserves QUADRATIC's charge.
QUADRATIC_EDGES = spline_edges("quadratic", 2)
#: The edge weighting that goes with CUBIC.
CUBIC_EDGES = spline_edges("cubic", 3)

