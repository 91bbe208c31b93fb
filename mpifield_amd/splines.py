"""A B-spline degree's weightings of the spline sequence (the reference's Splines,
src/mpifield/splines.py:100-272): the nodes (0-forms), the edges (1-forms) that match
them, the faces (2-forms, 3D: B_p along the normal, B_{p-1} across) and the cells (the top
form, B_{p-1} along every axis).  The face and cell values are integrals over the faces
or cells; their deposit and gather are each other's transpose (sum W f, no measures).
"""

from __future__ import annotations

from dataclasses import dataclass

from typing import Any

from mpifield_amd import kernels as K
from mpifield_amd.edges import CUBIC_EDGES, QUADRATIC_EDGES, WHITNEY, EdgeWeighting, spline_edges
from mpifield_amd.weighting import CUBIC, LINEAR, QUADRATIC, Weighting, spline_nodes

__all__ = ["CLAMPED_SPLINES", "SPLINES", "Splines", "splines"]


@dataclass(frozen=True, eq=False, repr=False)
class Splines:
    """The node weighting of B-spline degree ``degree`` and its edge weighting."""

    degree: int
    nodes: Weighting
    edges: EdgeWeighting | None = None
    faces: Weighting | None = None
    cells: Weighting | None = None
    clamped: bool = False

    def __post_init__(self) -> None:
        if self.nodes.degree != self.degree:
            raise ValueError(
                f"the node weighting {self.nodes.name!r} has degree {self.nodes.degree}, not {self.degree}"
            )
        if self.edges is not None and self.edges.degree != self.degree:
            raise ValueError(
                f"the edge weighting {self.edges.name!r} has degree {self.edges.degree}, not {self.degree}"
            )
        if self.clamped:
            for weighting in (self.nodes, self.edges, self.faces, self.cells):
                if weighting is not None and not weighting.clamps:
                    raise ValueError(f"the {weighting.name} weighting does not take clamped axes")

    @property
    def halo(self) -> int:
        edges = 0 if self.edges is None else self.edges.halo
        return max(self.nodes.halo, edges)

    def require_edges(self) -> EdgeWeighting:
        if self.edges is None:
            have = ", ".join(str(s.degree) for s in SPLINES.values() if s.edges)
            raise ValueError(f"degree {self.degree} has no edge weighting yet (degrees with one: {have})")
        return self.edges

    def __repr__(self) -> str:
        edges = None if self.edges is None else self.edges.name
        clamped = ", clamped=True" if self.clamped else ""
        return f"Splines(degree={self.degree}, nodes={self.nodes.name!r}, edges={edges!r}{clamped})"


@dataclass(frozen=True)
class _WithForm:
    """A forms kernel with its degree and form appended: a kernel of the weighting convention."""

    kernel: Any
    degree: int
    form: int

    def __call__(self, *arguments: Any) -> None:
        self.kernel(*arguments, self.degree, self.form)


_SCATTER_FORMS = {1: K.scatter_forms_spline_1d, 2: K.scatter_forms_spline_2d, 3: K.scatter_forms_spline_3d}
_GATHER_FORMS = {1: K.gather_forms_spline_1d, 2: K.gather_forms_spline_2d, 3: K.gather_forms_spline_3d}


def _faces(name: str, degree: int) -> Weighting:
    """The 2-forms of a degree in 3D: one component per face normal (splines.py:185-193)."""
    return Weighting(f"{name} faces", scatter={3: _WithForm(_SCATTER_FORMS[3], degree, 2)},
                     gather={3: _WithForm(_GATHER_FORMS[3], degree, 2)}, halo=2 if degree > 1 else 1, clamps=True)


def _cells(name: str, degree: int) -> Weighting:
    """The top forms of a degree in 1D, 2D and 3D: one component (splines.py:196-204)."""
    return Weighting(f"{name} cells", scatter={n: _WithForm(_SCATTER_FORMS[n], degree, n) for n in (1, 2, 3)},
                     gather={n: _WithForm(_GATHER_FORMS[n], degree, n) for n in (1, 2, 3)},
                     halo=2 if degree > 1 else 1, clamps=True)


_NAMES = {1: "linear", 2: "quadratic", 3: "cubic"}
_FACES = {p: _faces(_NAMES[p], p) for p in (1, 2, 3)}
_CELLS = {p: _cells(_NAMES[p], p) for p in (1, 2, 3)}

#: The built-in shapes, by degree.
SPLINES: dict[int, Splines] = {
    1: Splines(1, LINEAR, WHITNEY, _FACES[1], _CELLS[1]),
    2: Splines(2, QUADRATIC, QUADRATIC_EDGES, _FACES[2], _CELLS[2]),
    3: Splines(3, CUBIC, CUBIC_EDGES, _FACES[3], _CELLS[3]),
}

#: The built-in shapes whose weightings all take clamped axes (the spline kernels).
CLAMPED_SPLINES: dict[int, Splines] = {
    1: Splines(1, spline_nodes("linear", 1), spline_edges("linear", 1), _FACES[1], _CELLS[1], clamped=True),
    2: Splines(2, spline_nodes("quadratic", 2), QUADRATIC_EDGES, _FACES[2], _CELLS[2], clamped=True),
    3: Splines(3, CUBIC, CUBIC_EDGES, _FACES[3], _CELLS[3], clamped=True),
}


def splines(degree: int | Splines, clamped: bool = False) -> Splines:
    """The built-in `Splines` of a degree (a `Splines` is returned as is)."""
    if isinstance(degree, Splines):
        if clamped and not degree.clamped:
            raise ValueError(f"{degree!r} does not take clamped axes")
        return degree
    if isinstance(degree, bool) or int(degree) != degree:
        raise ValueError(f"a spline degree is an integer, got {degree!r}")
    try:
        return (CLAMPED_SPLINES if clamped else SPLINES)[int(degree)]
    except KeyError:
        raise ValueError(f"no built-in splines of degree {degree}; degrees {tuple(SPLINES)}") from None
