"""The small value types of the kernel convention: ``Bounds`` and ``BlockGrid``.

``BlockGrid`` keeps the reference's fields and its ``from_geometry`` constructor
(src/mpifield/_types.py:20-96): ``geometry`` is the tuple the kernels take after their
arrays, ``(*lower, *cell_size, *num_nodes, *code)`` with code 0 bounded, 1 periodic,
2 clamped spline knots.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np


class Bounds(NamedTuple):
    """The lower and upper corner of a grid block, one entry per axis."""

    lower_bound: np.ndarray
    upper_bound: np.ndarray


class BlockGrid(NamedTuple):
    """The grid of one stored block as the kernels see it.

    ``geometry`` is ``(*lower, *cell_size, *num_nodes, *code)``; ``index`` selects the
    block's nodes in a field's storage with halos; ``clamped`` marks axes with clamped
    spline knots, where an array of degree p holds ``num_nodes + p - 1`` coefficients.
    """

    bounds: Bounds
    grid_spacing: np.ndarray
    num_gridpoints: np.ndarray
    periodic: tuple[bool, ...]
    geometry: tuple[float | int, ...]
    index: tuple[slice, ...]
    clamped: tuple[bool, ...] = ()

    @classmethod
    def from_geometry(
        cls,
        lower,
        cell_size,
        num_nodes,
        periodic=None,
        clamped=None,
    ) -> "BlockGrid":
        """The grid of a whole array of ``num_nodes`` nodes from ``lower`` on.

        ``periodic`` defaults to bounded axes, ``clamped`` to none; an axis cannot be both.
        """
        lower = tuple(float(v) for v in np.asarray(lower, dtype=float).reshape(-1))
        cell_size = tuple(float(v) for v in np.asarray(cell_size, dtype=float).reshape(-1))
        num_nodes = tuple(int(n) for n in np.asarray(num_nodes).reshape(-1))
        ndim = len(num_nodes)
        flags = (False,) * ndim if periodic is None else tuple(bool(p) for p in periodic)
        ends = (False,) * ndim if clamped is None else tuple(bool(c) for c in clamped)
        if not len(lower) == len(cell_size) == len(flags) == len(ends) == ndim:
            raise ValueError("lower, cell_size, num_nodes, periodic and clamped differ in length")
        if any(f and e for f, e in zip(flags, ends)):
            raise ValueError("an axis is either periodic or clamped")
        lower_bound = np.asarray(lower, dtype=float)
        spacing = np.asarray(cell_size, dtype=float)
        nodes = np.asarray(num_nodes, dtype=int)
        return cls(
            bounds=Bounds(lower_bound=lower_bound, upper_bound=lower_bound + (nodes - 1) * spacing),
            grid_spacing=spacing,
            num_gridpoints=nodes,
            periodic=flags,
            geometry=(*lower, *cell_size, *num_nodes, *(2 if e else int(f) for f, e in zip(flags, ends))),
            index=(slice(None),) * ndim,
            clamped=ends,
        )
