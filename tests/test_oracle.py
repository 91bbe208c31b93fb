"""Pins of the CPU oracle (oracle/mpf_oracle.c) before it is trusted as the checker.

Every oracle function is checked against something the oracle's own formulas do not
restate: the reference tests' known answers (tests/golden/known_answers.json, each with its
source file:line), closed forms, scipy's BSpline for clamped knots, brute-force line
integrals, invariants (charge, adjoint, continuity, halo equivalence), and the
reference's own C kernels compiled from its sources into oracle/_ref.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc

GOLDEN = json.loads((Path(__file__).parent / "golden" / "known_answers.json").read_text())


def geom(lower, cell, num, code):
    return (*map(float, lower), *map(float, cell), *map(int, num), *map(int, code))


def spacing(bounds, shape, periodic):
    return [(hi - lo) / (n if p else n - 1) for (lo, hi), n, p in zip(bounds, shape, periodic)]


# ----------------------------------------------------------------------------------------
# Known answers of the reference tests


def test_quadratic_node_and_midway():
    case = GOLDEN["quadratic_node_and_midway"]
    lo, h, n, code = case["geometry"]
    field = np.zeros((n, 1))
    for x, v in zip(case["positions"], case["values"]):
        orc.scatter("quadratic", np.array([[x]]), np.array([[v]]), field, lo, h, n, code)
    np.testing.assert_allclose(field[:, 0], case["expected"], atol=1e-15)


def test_quadratic_fold_and_wrap():
    case = GOLDEN["quadratic_fold_at_bounded_edge"]
    lo, h, n, code = case["geometry"]
    field = np.zeros((n, 1))
    orc.scatter("quadratic", np.array([case["positions"]]), np.array([case["values"]]), field, lo, h, n, code)
    np.testing.assert_allclose(field[:3, 0], case["expected_first3"])
    case = GOLDEN["quadratic_wrap_periodic"]
    lo, h, n, code = case["geometry"]
    field = np.zeros((n, 1))
    orc.scatter("quadratic", np.array([case["positions"]]), np.array([case["values"]]), field, lo, 1.0 / 11, n, code)
    np.testing.assert_allclose(field[[-1, 0, 1], 0], case["expected_last_first_second"])


def test_linear_1d_known_answers():
    case = GOLDEN["linear_1d_past_bounded_ends"]
    lo, h, n, code = case["geometry"]
    field = np.zeros((n, 1))
    orc.scatter("linear", np.array([case["positions"]]), np.array([case["values"]]), field, lo, h, n, code)
    assert field[-1, 0] == case["expected_last"] and field[0, 0] == case["expected_first"]
    assert field.sum() == case["expected_sum"]
    for key in ("linear_1d_periodic_gather", "linear_1d_gather_edge_cell"):
        case = GOLDEN[key]
        lo, h, n, code = case["geometry"]
        result = np.zeros((1, len(case["positions"])))
        orc.gather("linear", np.array([case["positions"]]), np.array(case["field"])[:, None], result, lo, h, n, code)
        np.testing.assert_allclose(result[0], case["expected"])


def test_block_grid_quadratic_2d():
    case = GOLDEN["block_grid_quadratic_2d"]
    array = np.zeros((3, 3, 1))
    orc.scatter("quadratic", np.array(case["positions"]), np.array(case["values"]), array, *case["geometry"])
    np.testing.assert_allclose(array[:, :, 0].sum(axis=1), case["expected_row_sums"])


def test_degree_zero_is_the_nearest_node():
    case = GOLDEN["degree_zero_nearest_node"]
    lo, h, n, code = case["geometry"]
    field = np.zeros((n, 1))
    pts = np.array([case["positions"]])
    orc.scatter("spline", pts, np.ones_like(pts), field, lo, h, n, code, degree=0)
    np.testing.assert_array_equal(field[:, 0], case["expected"])


def test_cubic_weights_closed_form_and_moments():
    for periodic in (0, 1):
        for d in np.linspace(0.0, 0.99, 12):
            node, w = orc.axis_stencil("spline", 3.0 + d, 10, periodic, degree=3)
            assert list(node) == [2, 3, 4, 5]
            np.testing.assert_allclose(w[0], (1 - d) ** 3 / 6, atol=1e-15)
            np.testing.assert_allclose(w[3], d**3 / 6, atol=1e-15)
            np.testing.assert_allclose(w.sum(), 1.0, atol=1e-15)
            np.testing.assert_allclose((w * node).sum(), 3.0 + d, atol=1e-14)
            np.testing.assert_allclose((w * (node - 3.0 - d) ** 2).sum(), 1.0 / 3.0, atol=1e-14)


def test_cubic_fold_and_wrap():
    case = GOLDEN["cubic_fold_and_wrap"]
    b = case["bounded"]
    node, _ = orc.axis_stencil("spline", b["rel"][0], b["num"], 0, degree=3)
    np.testing.assert_array_equal(node, b["first_nodes_of_rel0"])
    _, w = orc.axis_stencil("spline", b["rel"][1], b["num"], 0, degree=3)
    np.testing.assert_allclose(w, b["weights_of_rel_minus2"], atol=1e-15)
    p = case["periodic"]
    nodes = np.array([orc.axis_stencil("spline", r, p["num"], 1, degree=3)[0] for r in p["rel"]]).T
    np.testing.assert_array_equal(nodes, p["nodes"])


def test_edges_past_a_bounded_wall_do_not_exist():
    case = GOLDEN["edges_past_a_bounded_wall"]
    exists = []
    for rel in case["rel"]:
        _, w, ex = orc.along_stencil(rel, case["num"], 0, case["degree"])
        np.testing.assert_allclose(w.sum(), 1.0)
        exists.append(list(ex))
    assert [list(e) for e in np.array(exists).T] == case["exists_by_edge"]


def _seam_field(case):
    shape, periodic = case["shape"], case["periodic"]
    h = spacing([(0.0, 1.0)] * len(shape), shape, periodic)
    return shape, periodic, h


def test_seam_known_answers():
    case = GOLDEN["seam_cell_centres_2d"]
    shape, periodic, h = _seam_field(case)
    xs = (np.arange(6) + 0.5) * h[0]
    ys = (np.arange(6) + 0.5) * h[1]
    pts = np.array([(x, y) for x in xs for y in ys]).T
    data = np.zeros((*shape, 1))
    orc.scatter("linear", pts, np.ones((1, pts.shape[1])), data, *geom([0, 0], h, shape, periodic))
    assert np.allclose(data[1:-1, :, 0], case["interior"])
    assert np.allclose(data[0, :, 0], case["bounded_edges"]) and np.allclose(data[-1, :, 0], case["bounded_edges"])

    for key in ("seam_marker_past_last_node", "seam_3d", "bounded_clamp_2d"):
        case = GOLDEN[key]
        shape, periodic, h = _seam_field(case)
        point = np.array([[c * s] for c, s in zip(case["cells"], h)])
        data = np.zeros((*shape, 1))
        g = geom([0] * len(shape), h, shape, periodic)
        orc.scatter("linear", point, np.array([[1.0]]), data, *g)
        for where, value in case["expected"].items():
            idx = tuple(int(i) for i in where.split(","))
            assert data[idx + (0,)] == pytest.approx(value)
        if "sum" in case:
            assert data.sum() == pytest.approx(case["sum"])
        if "gathered" in case:
            out = np.zeros((1, 1))
            orc.gather("linear", point, data, out, *g)
            assert out[0, 0] == pytest.approx(case["gathered"])

    case = GOLDEN["seam_gather_cos"]
    shape, periodic, h = _seam_field(case)
    y = np.arange(shape[1]) * h[1]
    data = np.broadcast_to(np.cos(2 * np.pi * y)[None, :, None], (*shape, 1)).copy()
    point = np.array([[c * s] for c, s in zip(case["cells"], h)])
    out = np.zeros((1, 1))
    orc.gather("linear", point, data, out, *geom([0, 0], h, shape, periodic))
    assert out[0, 0] == pytest.approx(0.5 * (np.cos(2 * np.pi * 7 * h[1]) + 1.0))


# ----------------------------------------------------------------------------------------
# Closed forms and library routines


def _knots(num, degree):
    return np.concatenate([np.zeros(degree), np.arange(num, dtype=float), np.full(degree, num - 1.0)])


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_clamped_stencils_are_scipy_b_splines(degree):
    """test_clamped.py:40-65: Cox-de Boor on clamped knots = scipy's design matrix; staggered derivative."""
    from scipy.interpolate import BSpline

    num = 7
    knots = _knots(num, degree)
    xs = np.linspace(0.0, num - 1.0, 400)[1:-1] + 0.0013
    xs = xs[xs < num - 1]
    size = num + degree - 1
    nodes = np.zeros((xs.size, size))
    for r, x in enumerate(xs):
        node, w = orc.axis_stencil("spline", x, num, 2, degree=degree)
        np.add.at(nodes[r], node, w)
    np.testing.assert_allclose(nodes, BSpline.design_matrix(xs, knots, degree).toarray(), atol=1e-14)
    staggered = np.zeros((xs.size, size - 1))
    for r, x in enumerate(xs):
        edge, w, ex = orc.along_stencil(x, num, 2, degree)
        assert ex.all()
        np.add.at(staggered[r], edge, w)
    derivative = np.stack(
        [BSpline(knots, np.eye(size)[i], degree).derivative()(xs) for i in range(size)], axis=1
    )
    padded = np.pad(staggered, ((0, 0), (1, 1)))
    np.testing.assert_allclose(derivative, padded[:, :-1] - padded[:, 1:], atol=1e-12)


def _cardinal(p, x):
    """The centred cardinal B-spline of degree p (support |x| < (p + 1) / 2)."""
    a = np.abs(x)
    if p == 0:
        return ((x >= -0.5) & (x < 0.5)).astype(float)
    if p == 1:
        return np.maximum(0.0, 1.0 - a)
    if p == 2:
        return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))
    return np.where(a < 1.0, (4 - 6 * a * a + 3 * a**3) / 6, np.where(a < 2.0, (2 - a) ** 3 / 6, 0.0))


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_periodic_weights_are_cardinal_b_splines(degree):
    """Brute force: the weight of node k at rel is B_p(rel - k) summed over the periodic images."""
    rng = np.random.default_rng(degree)
    num = 9
    for rel in rng.uniform(-20, 20, 50):
        node, w = orc.axis_stencil("spline", rel, num, 1, degree=degree)
        dense = np.zeros(num)
        np.add.at(dense, node, w)
        k = np.arange(num)
        brute = sum(_cardinal(degree, rel - k - m * num) for m in range(-4, 5))
        np.testing.assert_allclose(dense, brute, atol=1e-12)


def test_linear_and_quadratic_equal_spline_one_and_two():
    """test_splines.py:62-77 (atol 1e-13)."""
    for ndim, periodic in [(1, (0,)), (1, (1,)), (2, (0, 1)), (2, (1, 1)), (3, (0, 0, 0)), (3, (1, 0, 1))]:
        g, pts, vals, field = _case(ndim, periodic)
        for kind, degree in (("linear", 1), ("quadratic", 2)):
            a = np.zeros_like(field)
            orc.scatter(kind, pts, vals, a, *g)
            b = np.zeros_like(field)
            orc.scatter("spline", pts, vals, b, *g, degree=degree)
            np.testing.assert_allclose(a, b, atol=1e-13, rtol=0)
            ra = np.zeros_like(vals)
            orc.gather(kind, pts, field, ra, *g)
            rb = np.zeros_like(vals)
            orc.gather("spline", pts, field, rb, *g, degree=degree)
            np.testing.assert_allclose(ra, rb, atol=1e-13, rtol=0)


def _case(ndim, periodic, seed=0, n=400):
    """test_splines.py:37-49: a block grid, markers partly outside, two components."""
    rng = np.random.default_rng(seed)
    shape = (7, 6, 5)[:ndim]
    lower = np.array([0.0, -1.0, 2.0])[:ndim]
    step = np.array([0.25, 0.5, 0.1])[:ndim]
    upper = lower + step * np.array(shape)
    pts = rng.uniform(lower - 0.3, upper + 0.3, (n, ndim)).T.copy()
    vals = rng.normal(size=(2, n))
    field = rng.normal(size=(*shape, 2))
    return geom(lower, step, shape, periodic), pts, vals, field


# ----------------------------------------------------------------------------------------
# Invariants


@pytest.mark.parametrize("kind,degree", [("linear", 1), ("quadratic", 2), ("spline", 0), ("spline", 1),
                                         ("spline", 2), ("spline", 3)])
@pytest.mark.parametrize("ndim,periodic", [(1, (0,)), (1, (1,)), (2, (0, 1)), (2, (1, 1)), (3, (0, 0, 0)),
                                           (3, (1, 0, 1))])
def test_charge_total_and_adjoint(kind, degree, ndim, periodic):
    """test_splines.py:80-105: no charge lost (also beyond bounded edges), sum g q = sum G rho."""
    g, pts, vals, field = _case(ndim, periodic, seed=degree)
    vals[:, ::9] = 0.0
    pts[:, ::9] = np.nan  # dead markers may sit anywhere
    deposit = np.zeros_like(field)
    orc.scatter(kind, pts, vals, deposit, *g, degree=degree)
    np.testing.assert_allclose(deposit.reshape(-1, 2).sum(axis=0), vals.sum(axis=1), rtol=1e-12)
    alive = ~np.isnan(pts[0])
    gathered = np.zeros((2, int(alive.sum())))
    orc.gather(kind, pts[:, alive].copy(), field, gathered, *g, degree=degree)
    np.testing.assert_allclose((gathered * vals[:, alive]).sum(axis=1),
                               (field * deposit).reshape(-1, 2).sum(axis=0), rtol=1e-12)


@pytest.mark.parametrize("ndim,code", [(1, 2), (2, 2), (3, 2)])
def test_clamped_charge_and_adjoint(ndim, code):
    rng = np.random.default_rng(11)
    for degree in (1, 2, 3):
        shape = (7, 6, 5)[:ndim]
        extent = tuple(s + degree - 1 for s in shape)
        g = geom([0.0] * ndim, [0.5] * ndim, shape, [code] * ndim)
        pts = rng.uniform(-0.5, 3.5, (ndim, 300))
        vals = rng.normal(size=(1, 300))
        dep = np.zeros((*extent, 1))
        orc.scatter("spline", pts, vals, dep, *g, degree=degree)
        np.testing.assert_allclose(dep.sum(), vals.sum(), rtol=1e-12)
        field = rng.normal(size=(*extent, 1))
        out = np.zeros((1, 300))
        orc.gather("spline", pts, field, out, *g, degree=degree)
        np.testing.assert_allclose((out * vals).sum(), (field * dep).sum(), rtol=1e-12)


def test_threaded_scatter_matches_serial():
    """test_parallel_scatter.py:89-115: the per-thread copies summed in thread order."""
    rng = np.random.default_rng(3)
    n = 60000
    g = geom([0, 0, 0], [0.1, 0.1, 0.1], [9, 8, 7], [1, 0, 1])
    pts = rng.uniform(-0.2, 1.0, (3, n))
    vals = rng.normal(size=(2, n))
    for kind, degree in (("linear", 1), ("quadratic", 2), ("spline", 3)):
        one = np.zeros((9, 8, 7, 2))
        orc.scatter(kind, pts, vals, one, *g, degree=degree, threads=1)
        many = np.zeros_like(one)
        orc.scatter(kind, pts, vals, many, *g, degree=degree, threads=4)
        np.testing.assert_allclose(many, one, atol=1e-10)
        again = np.zeros_like(one)
        orc.scatter(kind, pts, vals, again, *g, degree=degree, threads=4)
        np.testing.assert_array_equal(again, many)  # reproducible for a thread count


# ----------------------------------------------------------------------------------------
# The reference's own C kernels (oracle/_ref, built from the reference sources)


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_reference_c_kernels_equal_the_oracle():
    rng = np.random.default_rng(7)
    for periodic in (0, 1):
        pts = rng.uniform(-0.3, 1.3, 5000)
        vals = rng.normal(size=5000)
        vals[::13] = 0.0
        ref = np.zeros(33)
        orc.ref_scatter_linear_1d(pts, vals, ref, 0.0, 1.0 / 32, 33, periodic)
        mine = np.zeros((33, 1))
        orc.scatter("linear", pts[None], vals[None], mine, 0.0, 1.0 / 32, 33, periodic)
        np.testing.assert_array_equal(mine[:, 0], ref)  # same operations, bit for bit
        g = geom([0.0, -1.0, 2.0], [0.1, 0.2, 0.05], [11, 9, 7], [periodic, 0, 1])
        p3 = rng.uniform([-0.2, -1.3, 1.9], [1.3, 0.9, 2.4], (5000, 3)).T.copy()
        ref3 = np.zeros(11 * 9 * 7)
        orc.ref_scatter_rows_3d(p3, vals, ref3, *g)
        mine3 = np.zeros((11, 9, 7, 1))
        orc.scatter("linear", p3, vals[None], mine3, *g)
        np.testing.assert_array_equal(mine3.reshape(-1), ref3)


# ----------------------------------------------------------------------------------------
# Path deposits


def _paths(ndim, periodic, seed, n=200, reach=1.0):
    rng = np.random.default_rng(seed)
    shape = (9, 8, 7)[:ndim]
    lower = np.array([0.0, -1.0, 2.0])[:ndim]
    step = np.array([0.25, 0.5, 0.1])[:ndim]
    upper = lower + step * (np.array(shape) - 1)
    start = rng.uniform(lower - 3 * step, upper + 3 * step, (n, ndim)).T.copy()
    end = start + rng.normal(scale=reach * step[:, None], size=start.shape)
    q = rng.normal(size=n)
    return geom(lower, step, shape, [periodic] * ndim), shape, start, end, q


def _node_charge(kind, degree, pts, q, shape, g):
    out = np.zeros((*shape, 1))
    orc.scatter(kind, pts, q[None], out, *g, degree=degree)
    return out[..., 0]


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("periodic", [0, 1])
def test_whitney_continuity_and_spline_degree_one(ndim, periodic):
    """test_edges.py:58-66 and test_spline_edges.py:46-64."""
    g, shape, start, end, q = _paths(ndim, periodic, seed=1, reach=3.0)
    current = np.zeros((*shape, ndim))
    orc.deposit_whitney(start, end, q, current, *g)
    change = _node_charge("linear", 1, end, q, shape, g) - _node_charge("linear", 1, start, q, shape, g)
    np.testing.assert_allclose(orc.divergence(current, [periodic] * ndim), change, atol=1e-12)
    spline = np.zeros_like(current)
    orc.deposit_spline(start, end, q, spline, *g, degree=1)
    np.testing.assert_allclose(spline, current, atol=1e-13, rtol=0)
    for axis in range(ndim):  # test_edges.py:69-78 bounded axes leave the last edges empty
        last = current.take(-1, axis=axis)[..., axis]
        assert (np.abs(last).max() > 0) if periodic else np.all(last == 0.0)


@pytest.mark.parametrize("degree", [2, 3])
@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("periodic", [0, 1])
def test_spline_path_continuity(degree, ndim, periodic):
    """test_spline_edges.py:67-87: G^T j = rho(end) - rho(start), atol 1e-12."""
    g, shape, start, end, q = _paths(ndim, periodic, seed=degree, reach=3.0)
    current = np.zeros((*shape, ndim))
    orc.deposit_spline(start, end, q, current, *g, degree=degree)
    kind = "quadratic" if degree == 2 else "spline"
    change = _node_charge(kind, degree, end, q, shape, g) - _node_charge(kind, degree, start, q, shape, g)
    np.testing.assert_allclose(orc.divergence(current, [periodic] * ndim), change, atol=1e-12)


def clamped_paths(ndim, degree, seed=0, n=300):
    """The reference's clamped case (tests/unit/test_clamped.py:78-90): 8 x 7 x 6 nodes with
    spacings 0.5, 0.4, 0.3, the first and third axes clamped and the second periodic, starts
    0.2 beyond the walls, ends two cells away on average."""
    rng = np.random.default_rng(seed)
    num = np.array((8, 7, 6)[:ndim])
    h = np.array((0.5, 0.4, 0.3)[:ndim])
    codes = [2, 1, 2][:ndim]
    extent = tuple(int(k) + (degree - 1 if c == 2 else 0) for k, c in zip(num, codes))
    upper = h * (num - 1)
    start = np.ascontiguousarray(rng.uniform(-0.2, upper + 0.2, (n, ndim)).T)
    q = rng.normal(size=n)
    end = np.ascontiguousarray(start + rng.normal(scale=2 * h[:, None], size=start.shape))
    return geom(np.zeros(ndim), h, num, codes), extent, codes, start, end, q


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_clamped_axes_keep_the_charge_and_conserve_it_along_paths(ndim, degree):
    """tests/unit/test_clamped.py:103-131: on clamped axes (the node arrays num + degree - 1
    long) the clamped spline deposit keeps the whole charge, and the spline-edge current of
    the paths has G^T j = rho(end) - rho(start) at 1e-12, the walls and clamped knots cut."""
    g, extent, codes, start, end, q = clamped_paths(ndim, degree, seed=degree)
    before = np.zeros((*extent, 1))
    orc.scatter("spline", start, q[None], before, *g, degree=degree)
    after = np.zeros((*extent, 1))
    orc.scatter("spline", end, q[None], after, *g, degree=degree)
    assert before.sum() == pytest.approx(q.sum(), abs=1e-12)
    current = np.zeros((*extent, ndim))
    orc.deposit_spline(start, end, q, current, *g, degree=degree)
    div = orc.divergence(current, [c == 1 for c in codes])
    np.testing.assert_allclose(div, (after - before)[..., 0], atol=1e-12)
    # the edges of a clamped axis stop one short of its nodes (num + degree - 2 of them)
    for axis, c in enumerate(codes):
        if c == 2:
            assert np.all(current.take(-1, axis=axis)[..., axis] == 0.0)


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_clamped_continuity_fails_for_a_planted_mistake(degree):
    """The clamped continuity check above is sharp: a current deposited with the walls'
    cuts missing (the path taken as periodic along the clamped axis) breaks it."""
    g, extent, codes, start, end, q = clamped_paths(2, degree, seed=7)
    before = np.zeros((*extent, 1))
    orc.scatter("spline", start, q[None], before, *g, degree=degree)
    after = np.zeros((*extent, 1))
    orc.scatter("spline", end, q[None], after, *g, degree=degree)
    wrong = list(g)
    wrong[6] = 1  # the first axis (clamped) read as periodic
    current = np.zeros((extent[0] - (degree - 1), extent[1], 2))
    orc.deposit_spline(start, end, q, current, *wrong, degree=degree)
    padded = np.zeros((*extent, 2))
    padded[: current.shape[0]] = current
    div = orc.divergence(padded, [c == 1 for c in codes])
    assert np.abs(div - (after - before)[..., 0]).max() > 1e-3
    """Brute force on a periodic 2D grid: j_(k,a) = q int B_(p-1)(r_a - k_a - 1/2) B_p(r_b - k_b) dr_a."""
    rng = np.random.default_rng(degree)
    shape, h = (8, 6), (0.5, 0.25)
    g = geom([0.0, 0.0], h, shape, [1, 1])
    start = rng.uniform(0, 1, (2, 5)) * np.array([[4.0], [1.5]])
    end = start + rng.normal(scale=0.6, size=start.shape) * np.array([[0.5], [0.25]])
    q = rng.normal(size=5)
    current = np.zeros((*shape, 2))
    orc.deposit_spline(start, end, q, current, *g, degree=degree)
    samples = 40000
    t = (np.arange(samples) + 0.5) / samples
    brute = np.zeros_like(current)
    for p in range(5):
        r = (start[:, p, None] + t * (end - start)[:, p, None]) / np.array(h)[:, None]  # (2, S) in cells
        dr = (end - start)[:, p] / np.array(h) / samples
        w = []
        for axis in range(2):
            k = np.arange(shape[axis])[:, None]
            node = sum(_cardinal(degree, r[axis] - k - m * shape[axis]) for m in range(-3, 4))
            edge = sum(_cardinal(degree - 1, r[axis] - k - 0.5 - m * shape[axis]) for m in range(-3, 4))
            w.append((node, edge))
        brute[..., 0] += q[p] * dr[0] * w[0][1] @ w[1][0].T
        brute[..., 1] += q[p] * dr[1] * w[0][0] @ w[1][1].T
    np.testing.assert_allclose(current, brute, atol=5e-4 if degree == 1 else 1e-6)


def test_dead_and_non_finite_paths_deposit_nothing():
    """test_edges.py:102-107."""
    g = geom([0, 0], [0.3, 0.4], [6, 5], [0, 1])
    start = np.array([[0.1, np.nan, 0.2], [0.3, 0.1, np.inf]])
    end = start + 0.4
    for deposit in ("w", 2, 3):
        current = np.zeros((6, 5, 2))
        if deposit == "w":
            orc.deposit_whitney(start, end, np.array([0.0, 1.0, 2.0]), current, *g)
        else:
            orc.deposit_spline(start, end, np.array([0.0, 1.0, 2.0]), current, *g, degree=deposit)
        assert np.all(current == 0.0)


def test_path_cuts_are_sorted_and_bounded():
    cuts = orc.path_cuts([0.2, 3.9], [3.7, 0.1], 0.5, [8, 8], [0, 1])
    assert cuts[0] == 0.0 and cuts[-1] == 1.0 and np.all(np.diff(cuts) >= 0)
    # planes at k + 1/2 crossed: x from 0.2 to 3.7 crosses 0.5,1.5,2.5,3.5; y 3.9 -> 0.1 crosses 3.5..0.5
    assert len(cuts) == 2 + 4 + 4


# ----------------------------------------------------------------------------------------
# Halo fold and fill


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("kind,halo", [("linear", 1), ("quadratic", 2)])
def test_halo_fold_and_fill_equal_the_bare_periodic_grid(ndim, kind, halo):
    """test_field.py:541-577: a halo'd block with a periodic first axis deposits and gathers as the bare grid."""
    shape = (7, 5, 4)[:ndim]
    periodic = (1,) + (0,) * (ndim - 1)
    h = spacing([(0.0, 1.0)] * ndim, shape, periodic)
    rng = np.random.default_rng(5)
    pts = rng.uniform(0.0, 1.0, (ndim, 400))
    pts[1:] *= 0.999
    vals = rng.uniform(0.5, 1.5, (1, 400))
    bare = np.zeros((*shape, 1))
    orc.scatter(kind, pts, vals, bare, *geom([0.0] * ndim, h, shape, periodic))
    # the stored block: halo nodes on every axis; the kernels see the periodic axis bounded with its halos,
    # the bounded axes without theirs (block_grid(inside_domain=True), field.py:150-202)
    storage = np.zeros((*(n + 2 * halo for n in shape), 1))
    inner = tuple(slice(halo, -halo) if d else slice(None) for d in range(ndim))
    sub_shape = (shape[0] + 2 * halo, *shape[1:])
    sub = np.zeros((*sub_shape, 1))
    lower = [-halo * h[0]] + [0.0] * (ndim - 1)
    sub_geom = geom(lower, h, sub_shape, [0] * ndim)
    orc.scatter(kind, pts, vals, sub, *sub_geom)
    storage[inner] = sub
    orc.halo_fold(storage, halo, periodic)
    local = storage[tuple(slice(halo, -halo) for _ in range(ndim))]
    np.testing.assert_allclose(local, bare, atol=1e-12)
    # halos were cleared
    assert np.abs(storage).sum() == pytest.approx(np.abs(local).sum())
    orc.halo_fill(storage, halo, periodic)
    bare_out = np.zeros((1, 400))
    orc.gather(kind, pts, bare, bare_out, *geom([0.0] * ndim, h, shape, periodic))
    padded_out = np.zeros((1, 400))
    orc.gather(kind, pts, np.ascontiguousarray(storage[inner]), padded_out, *sub_geom)
    np.testing.assert_allclose(padded_out, bare_out, atol=1e-12)


def test_gradient_and_divergence_are_transposes():
    rng = np.random.default_rng(2)
    for periodic in ([0, 1, 1], [1, 0, 0]):
        phi = rng.normal(size=(5, 4, 3))
        j = rng.normal(size=(5, 4, 3, 3))
        grad = orc.gradient(phi, periodic)
        div = orc.divergence(j, periodic)
        np.testing.assert_allclose((grad * j).sum(), (phi * div).sum(), rtol=1e-12)
        # edges.py:218-237: periodic roll, bounded drop
        np.testing.assert_allclose(grad[..., 1], (np.roll(phi, -1, axis=1) - phi) if periodic[1] else
                                   np.concatenate([np.diff(phi, axis=1), np.zeros((5, 1, 3))], axis=1))


# ----------------------------------------------------------------------------------------
# Edge gathers


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("periodic", [0, 1])
def test_whitney_field_of_a_gradient_is_the_gradient_of_linear(ndim, periodic):
    """test_edges.py:124-150: inside a cell, away from its faces, the Whitney field of the
    edge differences of a node potential is the derivative of the linear interpolant (a
    central difference of it is exact there)."""
    rng = np.random.default_rng(9)
    shape = (9, 8, 7)[:ndim]
    lower, h = np.array([0.0, -1.0, 2.0])[:ndim], np.array([0.25, 0.5, 0.1])[:ndim]
    g = geom(lower, h, shape, [periodic] * ndim)
    potential = rng.normal(size=shape)
    edges = orc.gradient(potential, [periodic] * ndim)
    cells = np.array([rng.integers(0, n - 1, 50) for n in shape])
    points = lower[:, None] + (cells + rng.uniform(0.2, 0.8, cells.shape)) * h[:, None]
    field = np.zeros_like(points)
    orc.gather_edges(points, edges, field, *g, degree=0)
    step = 0.05
    for axis in range(ndim):
        shift = np.zeros((ndim, 1))
        shift[axis] = step * h[axis]
        values = []
        for sign in (1.0, -1.0):
            value = np.zeros((1, points.shape[1]))
            orc.gather("linear", points + sign * shift, potential[..., None], value, *g)
            values.append(value[0])
        np.testing.assert_allclose(field[axis], (values[0] - values[1]) / (2.0 * shift[axis]), rtol=1e-10,
                                   atol=1e-10)


def _work_along(start, end, q, edges, g, degree):
    """sum q * integral of E . dl along the paths (clamped to the grid on bounded and clamped
    axes), exact: each path cut where the field's pieces change (the knot planes: integers
    for WHITNEY, odd degrees and clamped axes, half-integers for even degrees elsewhere; and
    the walls), 5-point Gauss-Legendre on every piece (E . dl is a polynomial of degree at
    most 8 in t there)."""
    ndim = start.shape[0]
    lower = np.asarray(g[:ndim])
    h = np.asarray(g[ndim:2 * ndim])
    num = np.asarray(g[2 * ndim:3 * ndim])
    codes = [int(c) for c in g[3 * ndim:4 * ndim]]
    even = 0.5 if degree % 2 == 0 and degree > 0 else 0.0
    nodes, weights = np.polynomial.legendre.leggauss(5)
    walled = np.array([c != 1 for c in codes])[:, None]

    def clamp(points):
        inside = np.clip(points, lower[:, None], (lower + h * (num - 1))[:, None])
        return np.where(walled, inside, points)

    work = 0.0
    for p in range(q.size):
        r0, r1 = (start[:, p] - lower) / h, (end[:, p] - lower) / h
        cuts = [0.0, 1.0]
        for d in range(ndim):
            if r1[d] == r0[d]:
                continue
            lo, hi = min(r0[d], r1[d]), max(r0[d], r1[d])
            shift = 0.0 if codes[d] == 2 else even
            planes = np.arange(np.floor(lo - shift) + shift, hi + 1.0)
            if codes[d] != 1:
                planes = np.concatenate([planes, [0.0, num[d] - 1.0]])
            t = (planes - r0[d]) / (r1[d] - r0[d])
            cuts.extend(t[(t > 0.0) & (t < 1.0)])
        cuts = np.unique(cuts)
        for t0, t1 in zip(cuts[:-1], cuts[1:]):
            a = clamp(start[:, p:p + 1] + t0 * (end - start)[:, p:p + 1])
            b = clamp(start[:, p:p + 1] + t1 * (end - start)[:, p:p + 1])
            points = np.ascontiguousarray(a + 0.5 * (nodes + 1.0) * (b - a))
            field = np.zeros_like(points)
            orc.gather_edges(points, edges, field, *g, degree=degree)
            work += q[p] * 0.5 * np.sum(weights * np.sum(field * (b - a), axis=0))
    return work


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
@pytest.mark.parametrize("ndim,periodic", [(1, 0), (1, 1), (2, 1), (2, 0), (3, 0), (3, 1)])
def test_the_current_is_the_transpose_of_the_edge_gather(degree, ndim, periodic):
    """test_spline_edges.py:88-115 and test_edges.py:153-172 (degree 0 here is WHITNEY):
    j . e = sum q * integral of E . dl along the clamped paths, integrated exactly piece by
    piece; paths partly outside the grid and some several cells long."""
    g, shape, start, end, q = _paths(ndim, periodic, seed=5 + degree, n=12, reach=3.0)
    edges = np.random.default_rng(degree).normal(size=(*shape, ndim))
    current = np.zeros_like(edges)
    if degree == 0:
        orc.deposit_whitney(start, end, q, current, *g)
    else:
        orc.deposit_spline(start, end, q, current, *g, degree=degree)
    work = _work_along(start, end, q, edges, g, degree)
    assert float(np.sum(current * edges)) == pytest.approx(work, rel=1e-11, abs=1e-12)


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_clamped_current_is_the_transpose_of_the_edge_gather(degree, ndim):
    """The work identity j . e = sum q * integral of E . dl (test_spline_edges.py:88-115) on
    the clamped case of tests/unit/test_clamped.py:78-90: the walls of the clamped axes
    (code 2) clamp the paths and cut them, the clamped knots sit at the nodes."""
    g, extent, codes, start, end, q = clamped_paths(ndim, degree, seed=20 + degree, n=16)
    edges = np.random.default_rng(degree).normal(size=(*extent, ndim))
    current = np.zeros_like(edges)
    orc.deposit_spline(start, end, q, current, *g, degree=degree)
    work = _work_along(start, end, q, edges, g, degree)
    assert float(np.sum(current * edges)) == pytest.approx(work, rel=1e-11, abs=1e-12)


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("periodic", [0, 1])
def test_spline_edge_gather_of_degree_one_is_whitney(ndim, periodic):
    """The degree-1 spline edge functions are the Whitney ones (test_spline_edges.py:46-64)."""
    g, shape, start, _, _ = _paths(ndim, periodic, seed=3, n=500)
    edges = np.random.default_rng(4).normal(size=(*shape, ndim))
    whitney, spline = np.zeros_like(start), np.zeros_like(start)
    orc.gather_edges(start, edges, whitney, *g, degree=0)
    orc.gather_edges(start, edges, spline, *g, degree=1)
    np.testing.assert_allclose(spline, whitney, rtol=1e-13, atol=1e-13)


# ----------------------------------------------------------------------------------------
# Forms of the spline sequence (faces, cells): gather_forms_spline_*d, scatter_forms_spline_*d

FORM_CODES = [(1, (1,)), (1, (0,)), (1, (2,)), (2, (1, 0)), (2, (2, 1)), (3, (0, 1, 2)), (3, (1, 1, 1)),
              (3, (0, 0, 0)), (3, (2, 2, 2))]


def _form_grid(ndim, codes, degree):
    num = np.array([7, 6, 8])[:ndim]
    h = np.array([0.5, 0.4, 0.3])[:ndim]
    lower = np.array([-1.0, 0.5, 2.0])[:ndim]
    extent = tuple(int(n) + (degree - 1 if c == 2 else 0) for n, c in zip(num, codes))
    return geom(lower, h, num, codes), lower, h, num, extent


def _forward(values, axis, code):
    """values[k + 1] - values[k] along ``axis``; a bounded or clamped axis has no last difference."""
    out = np.roll(values, -1, axis=axis) - values
    if code != 1:
        index = [slice(None)] * values.ndim
        index[axis] = -1
        out[tuple(index)] = 0.0
    return out


def _exterior(cochain, form, codes):
    """The discrete exterior derivative of a ``form``-cochain (*extent, ncomp): gradient,
    curl (the circulation around each face) and divergence."""
    ndim = len(codes)
    if form == 0:
        return np.stack([_forward(cochain[..., 0], a, codes[a]) for a in range(ndim)], axis=-1)
    if form == 1 and ndim == 2:
        return (_forward(cochain[..., 1], 0, codes[0]) - _forward(cochain[..., 0], 1, codes[1]))[..., None]
    if form == 1:
        return np.stack([_forward(cochain[..., (c + 2) % 3], (c + 1) % 3, codes[(c + 1) % 3])
                         - _forward(cochain[..., (c + 1) % 3], (c + 2) % 3, codes[(c + 2) % 3]) for c in range(3)],
                        axis=-1)
    return sum(_forward(cochain[..., a], a, codes[a]) for a in range(3))[..., None]


def _form_field(points, cochain, form, g, degree):
    out = np.zeros((cochain.shape[-1], points.shape[1]))
    if form == 0:
        orc.gather("spline", points, cochain, out, *g, degree=degree)
    else:
        orc.gather_forms(points, cochain, out, *g, degree=degree, form=form)
    return out


def _inside_pieces(rng, ndim, codes, num, degree, n):
    """Positions in cells, each at least 0.15 cells from the planes where the interpolants change
    piece (integers for odd degrees and clamped axes, half-integers for even degrees)."""
    rel = np.empty((ndim, n))
    for d in range(ndim):
        half = degree % 2 == 0 and codes[d] != 2
        top = int(num[d]) - (1 if codes[d] == 1 else (3 if half else 2))
        rel[d] = rng.integers(0, top + 1, n) + (0.5 if half else 0.0) + rng.uniform(0.15, 0.85, n)
    return rel


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("ndim,codes", FORM_CODES)
def test_form_fields_of_the_exterior_derivative_are_the_derivatives(degree, ndim, codes):
    """test_spline_forms.py:72-100/103-148 and test_clamped.py:152-195: the (k+1)-form field of
    d c is the exterior derivative (gradient, curl, divergence) of the k-form field of c,
    for every k, on periodic, bounded and clamped axes.  The derivatives in cells by a
    five-point difference, exact for the cubic pieces, at points inside one piece."""
    rng = np.random.default_rng(degree + 10 * ndim)
    g, lower, h, num, extent = _form_grid(ndim, codes, degree)
    rel = _inside_pieces(rng, ndim, codes, num, degree, 80)
    points = lower[:, None] + rel * h[:, None]
    step = 0.02
    for form in range(ndim):
        ncomp = 1 if form == 0 else orc.form_components(form, ndim)
        cochain = rng.normal(size=(*extent, ncomp))
        found = _form_field(points, _exterior(cochain, form, codes), form + 1, g, degree)
        partial = []
        for axis in range(ndim):
            shift = np.zeros((ndim, 1))
            shift[axis] = step * h[axis]
            f = [_form_field(points + s * shift, cochain, form, g, degree) for s in (-2, -1, 1, 2)]
            partial.append((f[0] - 8.0 * f[1] + 8.0 * f[2] - f[3]) / (12.0 * step))
        if form == 0:
            expected = np.array([partial[a][0] for a in range(ndim)])
        elif form == 1 and ndim == 2:
            expected = (partial[0][1] - partial[1][0])[None]
        elif form == 1:
            expected = np.array([partial[(c + 1) % 3][(c + 2) % 3] - partial[(c + 2) % 3][(c + 1) % 3]
                                 for c in range(3)])
        else:
            expected = sum(partial[a][a] for a in range(3))[None]
        np.testing.assert_allclose(found, expected, rtol=0, atol=1e-9)


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("ndim,codes", FORM_CODES)
def test_one_forms_are_the_edge_field_times_the_spacing(degree, ndim, codes):
    """The 1-form kernels are the spline edge gather without its division by the cell size."""
    rng = np.random.default_rng(7 * degree + ndim)
    g, lower, h, num, extent = _form_grid(ndim, codes, degree)
    points = lower[:, None] + rng.uniform(-0.5, 1.0, (ndim, 300)) * (h * (num - 1))[:, None] * 1.2
    edges = rng.normal(size=(*extent, ndim))
    field, forms = np.zeros((ndim, 300)), np.zeros((ndim, 300))
    orc.gather_edges(points, edges, field, *g, degree=degree)
    orc.gather_forms(points, edges, forms, *g, degree=degree, form=1)
    np.testing.assert_allclose(forms, field * h[:, None], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("ndim,codes", FORM_CODES)
def test_form_scatter_is_the_transpose_of_the_gather(degree, ndim, codes):
    """test_spline_forms.py:151-181: sum(gathered * values) == sum(field * deposit), markers
    partly outside the grid, every 7th dead."""
    rng = np.random.default_rng(3 * degree + ndim)
    g, lower, h, num, extent = _form_grid(ndim, codes, degree)
    points = lower[:, None] + rng.uniform(-0.3, 1.2, (ndim, 400)) * (h * (num - 1))[:, None]
    for form in range(1, ndim + 1):
        ncomp = orc.form_components(form, ndim)
        values = rng.normal(size=(ncomp, 400))
        values[:, ::7] = 0.0
        field = rng.normal(size=(*extent, ncomp))
        deposit = np.zeros_like(field)
        orc.scatter_forms(points, values, deposit, *g, degree=degree, form=form)
        gathered = np.zeros_like(values)
        orc.gather_forms(points, field, gathered, *g, degree=degree, form=form)
        assert np.sum(gathered * values) == pytest.approx(np.sum(field * deposit), rel=1e-12)


def test_cells_sum_to_one_inside_and_hold_the_whole_deposit_when_periodic():
    """test_spline_forms.py:184-195."""
    g = geom([0.0, 0.0], [0.5, 0.4], [9, 8], [1, 1])
    points = np.random.default_rng(1).uniform(0.0, 3.0, (2, 200))
    for degree in (1, 2, 3):
        deposit = np.zeros((9, 8, 1))
        orc.scatter_forms(points, np.ones((1, 200)), deposit, *g, degree=degree, form=2)
        assert deposit.sum() == pytest.approx(200.0)
        gathered = np.zeros((1, 200))
        orc.gather_forms(points, np.ones((9, 8, 1)), gathered, *g, degree=degree, form=2)
        np.testing.assert_allclose(gathered, 1.0, rtol=1e-14)


def test_forms_skip_dead_markers_and_thread_like_the_serial_deposit():
    """A marker whose values are all zero is not read (NaN positions deposit nothing); the
    per-thread copies summed in thread order equal the serial deposit to rounding."""
    rng = np.random.default_rng(12)
    g = geom([0.0] * 3, [0.1] * 3, [8, 8, 8], [1, 0, 2])
    points = rng.uniform(0.0, 0.7, (3, 30000))
    values = rng.normal(size=(3, 30000))
    points[:, ::5] = np.nan
    values[:, ::5] = 0.0
    serial, threaded = np.zeros((8, 8, 9, 3)), np.zeros((8, 8, 9, 3))
    orc.scatter_forms(points, values, serial, *g, degree=2, form=2)
    orc.scatter_forms(points, values, threaded, *g, degree=2, form=2, threads=4)
    assert np.isfinite(serial).all()
    gathered = [np.zeros((3, 30000)), np.zeros((3, 30000))]
    for out, threads in zip(gathered, (1, 4)):
        orc.gather_forms(points, serial, out, *g, degree=2, form=2, threads=threads)
    assert np.array_equal(gathered[0], gathered[1], equal_nan=True)
    np.testing.assert_allclose(threaded, serial, rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError, match="has 3 components"):
        orc.scatter_forms(points, values[:1], np.zeros((8, 8, 9, 1)), *g, degree=2, form=2)
