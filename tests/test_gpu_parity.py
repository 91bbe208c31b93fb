"""GPU parity of libmpfhip.so against the CPU oracle, through the C ABI (mpifield_amd.kernels).

Bars: float64 markers agree with the oracle within 1e-12 relative to the field's largest
value (scatter, MODE_ATOMIC) or the gathered values (gather); MODE_FIXED deposits round every
contribution to a quantum of 2^-k (k from the batch's size and largest value) and are held
to 1e-10, far inside the 1e-6 that the fixed-point form promises; float32
markers run the same float64 cell selection, so the oracle fed the same float32 values
agrees as tightly, and the float32 gather result within float32 rounding (1e-6).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from mpifield_amd import _lib
    from mpifield_amd import kernels as K

KINDS = [("linear", 0, 1), ("quadratic", 1, 2), ("spline", 2, 0), ("spline", 2, 1), ("spline", 2, 2),
         ("spline", 2, 3)]
CASES = [(1, (0,)), (1, (1,)), (2, (0, 1)), (2, (1, 1)), (3, (0, 0, 0)), (3, (1, 0, 1)), (3, (1, 1, 1))]
MODES = ["atomic", "fixed"]
BAR = {"atomic": 1e-12, "fixed": 1e-10}


def geom(lower, cell, num, code):
    return (*map(float, lower), *map(float, cell), *map(int, num), *map(int, code))


def case(ndim, periodic, seed=0, n=2000, ncomp=2, clamped=False, degree=1):
    rng = np.random.default_rng(seed)
    shape = (7, 6, 5)[:ndim]
    lower = np.array([0.0, -1.0, 2.0])[:ndim]
    step = np.array([0.25, 0.5, 0.1])[:ndim]
    upper = lower + step * np.array(shape)
    pts = rng.uniform(lower - 0.3, upper + 0.3, (n, ndim)).T.copy()
    vals = rng.normal(size=(ncomp, n))
    code = [2] * ndim if clamped else list(periodic)
    extent = tuple(s + (degree - 1 if clamped else 0) for s in shape)
    field = rng.normal(size=(*extent, ncomp))
    return geom(lower, step, shape, code), pts, vals, field


def cuda(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def rel_err(got, want):
    scale = max(np.abs(want).max(), 1e-300)
    return np.abs(got - want).max() / scale


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("ndim,periodic", CASES)
@pytest.mark.parametrize("name,kind,degree", KINDS)
def test_scatter_matches_oracle(name, kind, degree, ndim, periodic, mode):
    print(f"scatter {name} p={degree} {ndim}D periodic={periodic} mode={mode}", flush=True)
    g, pts, vals, field = case(ndim, periodic, seed=kind * 10 + degree)
    vals[:, ::9] = 0.0
    pts[:, ::9] = np.nan  # dead markers may sit anywhere
    want = field.copy()
    orc.scatter(name, pts, vals, want, *g, degree=degree)
    got = cuda(field)
    with K.deposit_mode(mode):
        K.scatter(kind, degree, cuda(pts), cuda(vals), got, g)
    assert rel_err(got.cpu().numpy(), want) < BAR[mode]


@pytest.mark.parametrize("ndim,periodic", CASES)
@pytest.mark.parametrize("name,kind,degree", KINDS)
def test_gather_matches_oracle(name, kind, degree, ndim, periodic):
    print(f"gather {name} p={degree} {ndim}D periodic={periodic}", flush=True)
    g, pts, _, field = case(ndim, periodic, seed=kind * 10 + degree + 1)
    want = np.zeros((field.shape[-1], pts.shape[1]))
    orc.gather(name, pts, field, want, *g, degree=degree)
    got = torch.zeros((field.shape[-1], pts.shape[1]), dtype=torch.float64, device="cuda")
    K.gather(kind, degree, cuda(pts), cuda(field), got, g)
    assert rel_err(got.cpu().numpy(), want) < 1e-12


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("degree", [1, 2, 3])
def test_clamped_axes_match_oracle(degree, ndim, mode):
    print(f"clamped p={degree} {ndim}D mode={mode}", flush=True)
    g, pts, vals, field = case(ndim, (0,) * ndim, seed=degree, clamped=True, degree=degree)
    want = field.copy()
    orc.scatter("spline", pts, vals, want, *g, degree=degree)
    got = cuda(field)
    with K.deposit_mode(mode):
        K.scatter(2, degree, cuda(pts), cuda(vals), got, g)
    assert rel_err(got.cpu().numpy(), want) < BAR[mode]
    want_g = np.zeros((field.shape[-1], pts.shape[1]))
    orc.gather("spline", pts, field, want_g, *g, degree=degree)
    got_g = torch.zeros_like(cuda(want_g))
    K.gather(2, degree, cuda(pts), cuda(field), got_g, g)
    assert rel_err(got_g.cpu().numpy(), want_g) < 1e-12


@pytest.mark.parametrize("mode", MODES)
def test_2d_linear_cloud_in_cell_on_101_by_51_nodes(mode):
    """The reference's plain CPU case: LINEAR (cloud-in-cell) scatter then gather on a 101 x 51
    node grid over [0, 0.1] x [0, 0.05], 10,000 seeded uniform float64 markers, unit weights
    (1, N), one bounded block -- through the catalogue kernels, against the oracle."""
    rng = np.random.default_rng(0)
    n = 10_000
    pos = np.stack([rng.uniform(0.0, 0.1, n), rng.uniform(0.0, 0.05, n)])
    g = geom([0.0, 0.0], [0.1 / 100, 0.05 / 50], [101, 51], [0, 0])
    weights = np.ones((1, n))
    want = np.zeros((101, 51, 1))
    orc.scatter("linear", pos, weights, want, *g)
    field = torch.zeros((101, 51, 1), dtype=torch.float64, device="cuda")
    with K.deposit_mode(mode):
        K.scatter_linear_2d(cuda(pos), cuda(weights), field, *g)
    assert rel_err(field.cpu().numpy(), want) < BAR[mode]
    assert float(field.sum()) == pytest.approx(n, rel=1e-12)
    back = np.zeros((1, n))
    orc.gather("linear", pos, want, back, *g)
    got = torch.empty((1, n), dtype=torch.float64, device="cuda")
    K.gather_linear_2d(cuda(pos), cuda(want), got, *g)
    assert rel_err(got.cpu().numpy(), back) < 1e-12


def test_strided_rows_and_float32_markers():
    """Rows with gaps (storage[:, :n], weighting.py:33-34) and float32 markers."""
    g, pts, vals, field = case(3, (1, 0, 1), seed=3, n=5000, ncomp=3)
    storage_p = torch.zeros((3, 8000), dtype=torch.float32, device="cuda")
    storage_v = torch.zeros((3, 8000), dtype=torch.float32, device="cuda")
    storage_p[:, :5000] = cuda(pts.astype(np.float32))
    storage_v[:, :5000] = cuda(vals.astype(np.float32))
    p32 = pts.astype(np.float32).astype(np.float64)
    v32 = vals.astype(np.float32).astype(np.float64)
    want = field.copy()
    orc.scatter("quadratic", p32, v32, want, *g)
    for mode in MODES:
        got = cuda(field)
        with K.deposit_mode(mode):
            K.scatter(1, 0, storage_p[:, :5000], storage_v[:, :5000], got, g)
        assert rel_err(got.cpu().numpy(), want) < BAR[mode], mode
    want_g = np.zeros((3, 5000))
    orc.gather("quadratic", p32, field, want_g, *g)
    got_g = torch.zeros((3, 5000), dtype=torch.float32, device="cuda")
    K.gather(1, 0, storage_p[:, :5000], cuda(field), got_g, g)
    np.testing.assert_allclose(got_g.cpu().numpy(), want_g, rtol=1e-6, atol=1e-6 * np.abs(want_g).max())


def test_fixed_point_is_bit_reproducible_under_any_marker_order():
    g, pts, vals, field = case(3, (1, 1, 1), seed=8, n=200000, ncomp=3)
    first = torch.zeros(field.shape, dtype=torch.float64, device="cuda")
    with K.deposit_mode("fixed"):
        K.scatter(1, 0, cuda(pts), cuda(vals), first, g)
        for seed in range(3):
            order = np.random.default_rng(seed).permutation(pts.shape[1])
            again = torch.zeros_like(first)
            K.scatter(1, 0, cuda(pts[:, order]), cuda(vals[:, order]), again, g)
            assert torch.equal(again, first)


# ---- path deposits ---------------------------------------------------------------------


def paths(ndim, periodic, seed, n=400, reach=3.0):
    rng = np.random.default_rng(seed)
    shape = (9, 8, 7)[:ndim]
    lower = np.array([0.0, -1.0, 2.0])[:ndim]
    step = np.array([0.25, 0.5, 0.1])[:ndim]
    upper = lower + step * (np.array(shape) - 1)
    start = rng.uniform(lower - 3 * step, upper + 3 * step, (n, ndim)).T.copy()
    end = start + rng.normal(scale=reach * step[:, None], size=start.shape)
    q = rng.normal(size=n)
    q[::11] = 0.0
    end[:, ::13] = np.nan
    return geom(lower, step, shape, [periodic] * ndim), shape, start, end, q


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_path_deposits_match_oracle(degree, ndim, periodic, mode):
    """degree 0 here means WHITNEY (deposit_paths_*d); 1..3 the spline edges."""
    print(f"paths p={degree} {ndim}D periodic={periodic} mode={mode}", flush=True)
    g, shape, start, end, q = paths(ndim, periodic, seed=degree)
    want = np.zeros((*shape, ndim))
    if degree == 0:
        orc.deposit_whitney(start, end, q, want, *g)
    else:
        orc.deposit_spline(start, end, q, want, *g, degree=degree)
    got = torch.zeros((*shape, ndim), dtype=torch.float64, device="cuda")
    with K.deposit_mode(mode):
        K.deposit_paths(0 if degree == 0 else 1, max(degree, 1), cuda(start), cuda(end), cuda(q), got, g)
    assert rel_err(got.cpu().numpy(), want) < BAR[mode]


def many_paths(ndim, periodic, seed, n=80000):
    """Enough paths for the binned, tiled deposit: short ones (a third of a cell), every 50th
    long (about 6 cells, past the tile's halo), zero charges and NaN ends among them; one 3D
    axis has fewer nodes than a tile's footprint."""
    rng = np.random.default_rng(seed)
    shape = (70, 48) if ndim == 2 else (20, 12, 36)
    lower = np.array([0.0, -1.0, 2.0])[:ndim]
    step = np.array([0.05, 0.1, 0.025])[:ndim]
    upper = lower + step * (np.array(shape) - 1)
    start = rng.uniform(lower - 2 * step, upper + 2 * step, (n, ndim)).T.copy()
    reach = np.full(n, 0.3)
    reach[::50] = 6.0
    end = start + rng.normal(size=start.shape) * reach * step[:, None]
    q = rng.normal(size=n)
    q[::11] = 0.0
    end[:, ::97] = np.nan
    return geom(lower, step, shape, [periodic] * ndim), shape, start, end, q


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_tiled_path_deposit_matches_oracle_and_is_reproducible(degree, ndim, periodic):
    """80000 paths take the binned, tiled fixed-point deposit (degree 0 here means WHITNEY,
    1..3 the spline edges): it equals the oracle, the float64-atomic deposit, and itself bit
    for bit under any path order."""
    print(f"tiled paths p={degree} {ndim}D periodic={periodic}", flush=True)
    g, shape, start, end, q = many_paths(ndim, periodic, seed=30 + degree)
    want = np.zeros((*shape, ndim))
    if degree == 0:
        orc.deposit_whitney(start, end, q, want, *g)
    else:
        orc.deposit_spline(start, end, q, want, *g, degree=degree)
    kind, deg = (0, 1) if degree == 0 else (1, degree)
    fixed = torch.zeros((*shape, ndim), dtype=torch.float64, device="cuda")
    K.deposit_paths(kind, deg, cuda(start), cuda(end), cuda(q), fixed, g, mode=_lib.MODE_FIXED)
    # the fixed-point quantum is 2^-k with n * max|q| (length + 2) * 2^k < 2^61: about 2^-36
    # here (long paths), and an edge sums up to ~10^4 rounded contributions
    assert rel_err(fixed.cpu().numpy(), want) < 2e-9
    atomic = torch.zeros_like(fixed)
    K.deposit_paths(kind, deg, cuda(start), cuda(end), cuda(q), atomic, g, mode=_lib.MODE_ATOMIC)
    assert rel_err(atomic.cpu().numpy(), want) < BAR["atomic"]
    perm = np.random.default_rng(degree).permutation(start.shape[1])
    again = torch.zeros_like(fixed)
    K.deposit_paths(kind, deg, cuda(start[:, perm]), cuda(end[:, perm]), cuda(q[perm]), again, g,
                    mode=_lib.MODE_FIXED)
    assert torch.equal(again, fixed)


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("degree", [2, 3])
def test_path_deposit_continuity_on_the_device(degree, ndim, periodic):
    """G^T j = rho(end) - rho(start) with the node weighting of the same degree, all on the GPU."""
    g, shape, start, end, q = paths(ndim, periodic, seed=10 + degree, n=3000)
    ok = np.isfinite(end).all(axis=0)
    start, end, q = start[:, ok], end[:, ok], q[ok]
    with K.deposit_mode("fixed"):
        j = torch.zeros((*shape, ndim), dtype=torch.float64, device="cuda")
        K.deposit_paths(1, degree, cuda(start), cuda(end), cuda(q), j, g)
        rho0 = torch.zeros((*shape, 1), dtype=torch.float64, device="cuda")
        rho1 = torch.zeros_like(rho0)
        K.scatter(2, degree, cuda(start), cuda(q[None]), rho0, g)
        K.scatter(2, degree, cuda(end), cuda(q[None]), rho1, g)
    div = K.divergence(j, [periodic] * ndim)
    np.testing.assert_allclose(div.cpu().numpy(), (rho1 - rho0)[..., 0].cpu().numpy(), atol=1e-11)


# ---- edge gathers ----------------------------------------------------------------------


@pytest.mark.parametrize("code", [0, 1, 2])
@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_edge_gathers_match_oracle(degree, ndim, code):
    """degree 0 here means WHITNEY (gather_edges_*d); 1..3 the spline edges; markers partly
    outside the grid; float64 markers at 1e-12, float32 ones at float32 rounding."""
    if code == 2 and degree == 0:
        pytest.skip("WHITNEY does not take clamped axes")
    print(f"edge gather p={degree} {ndim}D code={code}", flush=True)
    g, shape, start, _, _ = paths(ndim, code, seed=20 + degree, n=3000)
    deg = max(degree, 1)
    extent = tuple(n + (deg - 1 if code == 2 else 0) for n in shape)
    edges = np.random.default_rng(degree).normal(size=(*extent, ndim))
    want = np.zeros((ndim, start.shape[1]))
    orc.gather_edges(start, edges, want, *g, degree=degree)
    kind = 0 if degree == 0 else 1
    got = torch.empty((ndim, start.shape[1]), dtype=torch.float64, device="cuda")
    K.gather_edges(kind, deg, cuda(start), cuda(edges), got, g)
    assert rel_err(got.cpu().numpy(), want) < 1e-12
    p32 = start.astype(np.float32)
    want32 = np.zeros_like(want)
    orc.gather_edges(p32.astype(np.float64), edges, want32, *g, degree=degree)
    got32 = torch.empty((ndim, start.shape[1]), dtype=torch.float32, device="cuda")
    K.gather_edges(kind, deg, cuda(p32), cuda(edges), got32, g)
    np.testing.assert_allclose(got32.cpu().numpy(), want32, rtol=1e-6, atol=1e-6 * np.abs(want32).max())


def test_edge_weightings_gather_through_the_api():
    """EdgeWeighting.gather_block (edges.py:169-185) on device tensors, and the current times
    an edge field is the work of the field along the paths (test_spline_edges.py:88-115)."""
    import mpifield_amd as m

    for weighting, degree in ((m.WHITNEY, 0), (m.QUADRATIC_EDGES, 2), (m.CUBIC_EDGES, 3)):
        g, shape, start, end, q = paths(2, 1, seed=40 + degree, n=2000)
        ok = np.isfinite(end).all(axis=0)
        start, end, q = start[:, ok], end[:, ok], q[ok]
        grid = m.BlockGrid.from_geometry(g[0:2], g[2:4], g[4:6], (True, True))
        edges = np.random.default_rng(degree).normal(size=(*shape, 2))
        field = torch.empty((2, start.shape[1]), dtype=torch.float64, device="cuda")
        weighting.gather_block(cuda(start), cuda(edges), field, grid)
        want = np.zeros((2, start.shape[1]))
        orc.gather_edges(start, edges, want, *g, degree=degree)
        assert rel_err(field.cpu().numpy(), want) < 1e-12
        current = torch.zeros((*shape, 2), dtype=torch.float64, device="cuda")
        with K.deposit_mode("atomic"):
            weighting.deposit_paths_block(cuda(start), cuda(end), cuda(q), current, grid)
        oracle_current = np.zeros((*shape, 2))
        if degree == 0:
            orc.deposit_whitney(start, end, q, oracle_current, *g)
        else:
            orc.deposit_spline(start, end, q, oracle_current, *g, degree=degree)
        work = float((current * cuda(edges)).sum())
        assert work == pytest.approx(float((oracle_current * edges).sum()), rel=1e-11)


# ---- halos, gradient, divergence --------------------------------------------------------


@pytest.mark.parametrize("halo", [1, 2])
@pytest.mark.parametrize("periodic", [(1, 0, 1), (1, 1, 1), (0, 1, 0)])
def test_halo_fold_and_fill_match_oracle(periodic, halo):
    rng = np.random.default_rng(halo)
    data = rng.normal(size=(9, 8, 7, 3))
    for op, fn in (("fold", orc.halo_fold), ("fill", orc.halo_fill)):
        want = data.copy()
        fn(want, halo, periodic)
        got = cuda(data)
        K.halo(op, got, halo, periodic)
        np.testing.assert_array_equal(got.cpu().numpy(), want)


def test_gradient_and_divergence_match_oracle():
    rng = np.random.default_rng(0)
    for periodic in ([0, 1, 1], [1, 0, 0], [1, 1]):
        shape = (6, 5, 4)[: len(periodic)]
        phi = rng.normal(size=shape)
        j = rng.normal(size=(*shape, len(periodic)))
        np.testing.assert_allclose(K.gradient(cuda(phi), periodic).cpu().numpy(), orc.gradient(phi, periodic),
                                   atol=1e-14)
        np.testing.assert_allclose(K.divergence(cuda(j), periodic).cpu().numpy(), orc.divergence(j, periodic),
                                   atol=1e-14)


# ---- errors come back through the ABI ---------------------------------------------------


def test_errors_are_raised_with_the_library_message():
    f = torch.zeros((5, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="spline degree"):
        K.scatter(2, 7, cuda(np.zeros((1, 3))), cuda(np.ones((1, 3))), f, (0.0, 1.0, 5, 0))
    with pytest.raises(ValueError, match="clamped"):
        K.scatter(1, 0, cuda(np.zeros((1, 3))), cuda(np.ones((1, 3))), f, (0.0, 1.0, 5, 2))
    with pytest.raises(ValueError, match="a field on this grid"):
        K.scatter(1, 0, cuda(np.zeros((1, 3))), cuda(np.ones((1, 3))), f, (0.0, 1.0, 6, 0))
    with pytest.raises(ValueError, match="spline edge degrees"):
        K.deposit_paths(1, 4, cuda(np.zeros((1, 1))), cuda(np.ones((1, 1))), cuda(np.ones(1)),
                        torch.zeros((4, 1), dtype=torch.float64, device="cuda"), (0.0, 1.0, 4, 0))
    assert _lib.load().mpf_abi_version() == 1
