"""GPU parity of the path deposits on clamped axes and with float32 paths, through the C ABI.

Clamped axes (code 2) follow the reference's clamped case (tests/unit/test_clamped.py:78-131):
the GPU current equals the oracle's (float64 atomics at 1e-12, fixed point at 1e-10), and its
divergence is the change of the clamped spline node charge. Float32 paths run the same
float64 segment split as float64 ones, so the oracle fed the same float32 values (widened
to float64) is the checker: the per-path kernels below the tiled threshold, and the binned,
tiled fixed-point deposit at and above it on the shape of the 2D quadratic-edge workload
(periodic grid, uniform starts, Maxwellian displacements clipped to half a cell) scaled
down.
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

BAR = {"atomic": 1e-12, "fixed": 1e-10}


def geom(lower, cell, num, code):
    return (*map(float, lower), *map(float, cell), *map(int, num), *map(int, code))


def cuda(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def rel_err(got, want):
    scale = max(np.abs(want).max(), 1e-300)
    return np.abs(got - want).max() / scale


def oracle_current(degree, start, end, q, shape, g):
    """degree 0 is WHITNEY, 1..3 the spline edges; float64 paths."""
    want = np.zeros(shape)
    if degree == 0:
        orc.deposit_whitney(start, end, q, want, *g)
    else:
        orc.deposit_spline(start, end, q, want, *g, degree=degree)
    return want


def clamped_paths(ndim, degree, seed=0, n=300, dtype=np.float64):
    """tests/unit/test_clamped.py:78-90: 8 x 7 x 6 nodes, spacings 0.5 / 0.4 / 0.3, the first
    and third axes clamped, the second periodic; starts up to 0.2 beyond the walls, ends two
    cells away on average; every 11th charge zero."""
    rng = np.random.default_rng(seed)
    num = np.array((8, 7, 6)[:ndim])
    h = np.array((0.5, 0.4, 0.3)[:ndim])
    codes = [2, 1, 2][:ndim]
    extent = tuple(int(k) + (degree - 1 if c == 2 else 0) for k, c in zip(num, codes))
    upper = h * (num - 1)
    start = rng.uniform(-0.2, upper + 0.2, (n, ndim)).T.astype(dtype)
    q = rng.normal(size=n).astype(dtype)
    q[::11] = 0.0
    end = (start + rng.normal(scale=2 * h[:, None], size=start.shape)).astype(dtype)
    return geom(np.zeros(ndim), h, num, codes), extent, codes, start, end, q


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["atomic", "fixed"])
@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("degree", [1, 2, 3])
def test_clamped_path_deposits_match_oracle(degree, ndim, mode, dtype):
    """deposit_paths_spline_{1,2,3}d on clamped axes equal the oracle's current."""
    print(f"clamped paths p={degree} {ndim}D mode={mode} {np.dtype(dtype).name}", flush=True)
    g, extent, codes, start, end, q = clamped_paths(ndim, degree, seed=degree + 3 * ndim, n=2000, dtype=dtype)
    want = oracle_current(degree, start.astype(np.float64), end.astype(np.float64), q.astype(np.float64),
                          (*extent, ndim), g)
    got = torch.zeros((*extent, ndim), dtype=torch.float64, device="cuda")
    with K.deposit_mode(mode):
        K.deposit_paths(1, degree, cuda(start), cuda(end), cuda(q), got, g)
    assert rel_err(got.cpu().numpy(), want) < BAR[mode]


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("degree", [1, 2, 3])
def test_clamped_splines_conserve_the_charge_through_the_api(degree, ndim):
    """splines(degree, clamped=True): the edges' deposit_paths_block and the nodes'
    scatter_block on a clamped BlockGrid, all on the device; G^T j = rho(end) - rho(start)
    (test_clamped.py:103-131) and no charge lost at the walls."""
    import mpifield_amd as m

    g, extent, codes, start, end, q = clamped_paths(ndim, degree, seed=40 + degree, n=3000)
    grid = m.BlockGrid.from_geometry(g[:ndim], g[ndim:2 * ndim], g[2 * ndim:3 * ndim],
                                     [c == 1 for c in codes], [c == 2 for c in codes])
    shape = m.splines(degree, clamped=True)
    before = torch.zeros((*extent, 1), dtype=torch.float64, device="cuda")
    after = torch.zeros_like(before)
    current = torch.zeros((*extent, ndim), dtype=torch.float64, device="cuda")
    shape.nodes.scatter_block(cuda(start), cuda(q[None]), before, grid)
    shape.nodes.scatter_block(cuda(end), cuda(q[None]), after, grid)
    shape.edges.deposit_paths_block(cuda(start), cuda(end), cuda(q), current, grid)
    assert float(before.sum()) == pytest.approx(float(q.sum()), abs=1e-10)
    div = shape.edges.divergence_block(current, grid)
    np.testing.assert_allclose(div.cpu().numpy(), (after - before)[..., 0].cpu().numpy(), atol=1e-10)
    want = oracle_current(degree, start, end, q, (*extent, ndim), g)
    assert rel_err(current.cpu().numpy(), want) < BAR["fixed"]


def short_paths(ndim, periodic, n, seed, nodes):
    """The 2D quadratic-edge workload's paths at a smaller size: a periodic (or bounded) unit
    box of ``nodes`` per axis, float32 starts uniform in it, per-axis displacements
    N(0, h/4) clipped to half a cell, ends = starts + displacement (float32, not wrapped),
    unit charges; every 11th charge zero, every 97th end NaN."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (nodes if periodic else nodes - 1)
    start = rng.random((ndim, n), dtype=np.float32)
    step = np.clip(rng.normal(0.0, 0.25 * h, (ndim, n)), -0.5 * h, 0.5 * h).astype(np.float32)
    end = start + step
    q = np.ones(n, dtype=np.float32)
    q[::11] = 0.0
    end[:, ::97] = np.nan
    g = geom([0.0] * ndim, [h] * ndim, [nodes] * ndim, [periodic] * ndim)
    return g, (nodes,) * ndim, start, end, q


@pytest.mark.parametrize("mode", ["atomic", "fixed"])
@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_float32_paths_match_oracle(degree, ndim, periodic, mode):
    """Float32 paths below the tiled threshold (the per-path kernels); degree 0 is WHITNEY."""
    print(f"float32 paths p={degree} {ndim}D periodic={periodic} mode={mode}", flush=True)
    nodes = {1: 64, 2: 24, 3: 10}[ndim]
    g, shape, start, end, q = short_paths(ndim, periodic, 20000, seed=50 + degree, nodes=nodes)
    want = oracle_current(degree, start.astype(np.float64), end.astype(np.float64), q.astype(np.float64),
                          (*shape, ndim), g)
    got = torch.zeros((*shape, ndim), dtype=torch.float64, device="cuda")
    with K.deposit_mode(mode):
        K.deposit_paths(0 if degree == 0 else 1, max(degree, 1), cuda(start), cuda(end), cuda(q), got, g)
    assert rel_err(got.cpu().numpy(), want) < BAR[mode]


@pytest.mark.parametrize("periodic", [1, 0])
@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_float32_tiled_paths_match_oracle_and_conserve_charge(degree, ndim, periodic):
    """2^17 float32 paths take the binned, tiled fixed-point deposit (float32 records with
    the charge and the end point as value rows): equal to the oracle fed the same float32
    values, bit-reproducible under any path order, and G^T j = rho(end) - rho(start) with
    the node weighting of the same degree, computed on the device."""
    print(f"float32 tiled paths p={degree} {ndim}D periodic={periodic}", flush=True)
    nodes = 256 if ndim == 2 else 48
    n = 1 << 17
    g, shape, start, end, q = short_paths(ndim, periodic, n, seed=60 + degree, nodes=nodes)
    want = oracle_current(degree, start.astype(np.float64), end.astype(np.float64), q.astype(np.float64),
                          (*shape, ndim), g)
    kind, deg = (0, 1) if degree == 0 else (1, degree)
    got = torch.zeros((*shape, ndim), dtype=torch.float64, device="cuda")
    K.deposit_paths(kind, deg, cuda(start), cuda(end), cuda(q), got, g, mode=_lib.MODE_FIXED)
    # the quantum is 2^-k with n * max|q| (length + 2) * 2^k < 2^61; an edge sums a few
    # hundred rounded contributions here
    assert rel_err(got.cpu().numpy(), want) < 1e-10
    perm = np.random.default_rng(degree).permutation(n)
    again = torch.zeros_like(got)
    K.deposit_paths(kind, deg, cuda(start[:, perm]), cuda(end[:, perm]), cuda(q[perm]), again, g,
                    mode=_lib.MODE_FIXED)
    assert torch.equal(again, got)
    ok = np.isfinite(end).all(axis=0)
    node_kind = (0, 1) if degree in (0, 1) else ((1, 2) if degree == 2 else (2, 3))
    with K.deposit_mode("fixed"):
        rho0 = torch.zeros((*shape, 1), dtype=torch.float64, device="cuda")
        rho1 = torch.zeros_like(rho0)
        K.scatter(*node_kind, cuda(start[:, ok]), cuda(q[None, ok]), rho0, g)
        K.scatter(*node_kind, cuda(end[:, ok]), cuda(q[None, ok]), rho1, g)
    div = K.divergence(got, [periodic] * ndim)
    change = (rho1 - rho0)[..., 0]
    assert float((div - change).abs().max()) < 1e-9 * max(float(change.abs().max()), 1.0)


@pytest.mark.parametrize("periodic", [1, 0])
@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_binned_path_deposit_equals_the_sorting_deposit(degree, ndim, periodic):
    """Paths kept binned by their start (binning.deposit_paths_binned) deposit the same edges,
    bit for bit, as mpf_deposit_paths, which bins them inside the call; dead paths (zero
    charge, NaN end) included; wrong bins are refused."""
    from mpifield_amd import binning

    print(f"binned paths p={degree} {ndim}D periodic={periodic}", flush=True)
    nodes = 256 if ndim == 2 else 48
    n = 1 << 17
    g, shape, start, end, q = short_paths(ndim, periodic, n, seed=70 + degree, nodes=nodes)
    kind, deg = (0, 1) if degree == 0 else (1, degree)
    want = torch.zeros((*shape, ndim), dtype=torch.float64, device="cuda")
    K.deposit_paths(kind, deg, cuda(start), cuda(end), cuda(q), want, g, mode=_lib.MODE_FIXED)
    bins = binning.bin_markers(cuda(start), None, (0, 1) if degree == 0 else (2, degree), g)
    order = bins.order
    got = torch.zeros_like(want)
    binning.deposit_paths_binned(bins, cuda(end)[:, order], cuda(q)[order], got, kind, deg)
    assert torch.equal(got, want)
    with pytest.raises(ValueError):
        binning.deposit_paths_binned(bins, cuda(end)[:, order], cuda(q)[order], got, 1, deg + 1 if kind else 1)
