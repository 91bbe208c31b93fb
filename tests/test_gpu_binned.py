"""GPU parity of the binned (tiled) hot path: on-device counting sort, tiled deposit
(int64 fixed point through exact float64 register sums) and tiled gather, against the CPU
oracle; bit-reproducibility under marker permutations; full-size invariants.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from mpifield_amd import binning
    from mpifield_amd import kernels as K

NAMES = {(0, 1): "linear", (1, 2): "quadratic", (2, 0): "spline", (2, 1): "spline", (2, 2): "spline", (2, 3): "spline"}
TILED = [(0, 1, 1), (0, 1, 3), (1, 2, 1), (1, 2, 3), (2, 3, 1),  # (kind, degree, ncomp)
         (2, 0, 1), (2, 0, 3), (2, 1, 1), (2, 1, 3), (2, 2, 1), (2, 2, 3)]
GRIDS = [
    ((32, 16, 32), (1, 1, 1)),
    ((24, 16, 32), (0, 1, 0)),   # bounded x and z: folds at the walls
    ((21, 13, 37), (0, 0, 0)),   # bounded, partial tiles on every axis
]


def geom(num, code, lower=(0.0, -1.0, 0.5), cell=(0.0625, 0.125, 0.03125)):
    return (*lower, *cell, *num, *code)


def markers(num, code, n, ncomp, seed, dtype=np.float64, lower=(0.0, -1.0, 0.5), cell=(0.0625, 0.125, 0.03125)):
    rng = np.random.default_rng(seed)
    lo = np.array(lower)
    span = np.array(cell) * np.array(num)
    margin = np.where(np.array(code) == 1, 1.0, 0.1) * span  # periodic axes wrap, bounded ones clamp
    pos = rng.uniform(lo - margin, lo + span + margin, (n, 3)).T.astype(dtype)
    val = rng.normal(1.0, 0.3, (ncomp, n)).astype(dtype)
    val[:, ::97] = 0.0  # dead markers
    pos[:, ::97] = np.nan
    return np.ascontiguousarray(pos), np.ascontiguousarray(val)


def cuda(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


@pytest.mark.parametrize("num,code", GRIDS)
@pytest.mark.parametrize("kind,degree,ncomp", TILED)
def test_binned_scatter_and_gather_match_oracle(kind, degree, ncomp, num, code):
    print(f"binned kind={kind} p={degree} C={ncomp} num={num} code={code}", flush=True)
    g = geom(num, code)
    pos, val = markers(num, code, 60000, ncomp, seed=kind + 3 * degree)
    assert K.tiled_ok(kind, degree, ncomp, num, code)
    bins = binning.bin_markers(cuda(pos), cuda(val), (kind, degree), g)
    # the sort is a permutation and the offsets are the cells' counts
    order = bins.order.cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(pos.shape[1]))
    start = bins.start.cpu().numpy().astype(np.int64)
    assert start[0] == 0 and start[-1] == pos.shape[1] and np.all(np.diff(start) >= 0)
    dead = (val == 0).all(axis=0)
    assert start[-1] - start[-2] == dead.sum()
    np.testing.assert_array_equal(bins.positions.cpu().numpy(), pos[:, order])

    field0 = np.random.default_rng(1).normal(size=(*num, ncomp))
    want = field0.copy()
    orc.scatter(NAMES[(kind, degree)], pos, val, want, *g, degree=degree)
    got = cuda(field0)
    binning.scatter_binned(bins, got)
    err = np.abs(got.cpu().numpy() - want).max() / np.abs(want).max()
    assert err < 1e-12, err
    # the tiles' work lists made once give the same deposit and gather, bit for bit
    tabled = binning.bin_markers(cuda(pos), cuda(val), (kind, degree), g)
    tabled.values.copy_(bins.values)
    tabled.positions.copy_(bins.positions)
    tabled.start.copy_(bins.start)
    tabled.prepare_items()
    again = cuda(field0)
    binning.scatter_binned(tabled, again)
    assert torch.equal(again, got)
    # and the float32 gather kernel through the entry point with the lists
    from mpifield_amd import _lib

    f32_bins = binning.bin_markers(cuda(pos.astype(np.float32)), None, (kind, degree), g)
    r1 = torch.empty((ncomp, pos.shape[1]), dtype=torch.float32, device="cuda")
    r2 = torch.empty_like(r1)
    binning.gather_binned(f32_bins, got, r1)
    f32_bins.prepare_items()
    ndim_, lw, cs, nm, cd, _ = K._split(g)
    _lib.call("mpf_gather_binned", kind, degree, 3, f32_bins.positions.data_ptr(), _lib.F32, f32_bins.n,
              f32_bins.start.data_ptr(), f32_bins.tables.data_ptr(), f32_bins.table_counts.data_ptr(), got.data_ptr(),
              ncomp, r2.data_ptr(), lw.ctypes.data, cs.ctypes.data, nm.ctypes.data, cd.ctypes.data, None,
              K.stream_handle(got.device))
    torch.testing.assert_close(r1, r2, rtol=0, atol=0, equal_nan=True)  # NaN positions give NaN

    live = ~dead
    want_g = np.zeros((ncomp, int(live.sum())))
    orc.gather(NAMES[(kind, degree)], pos[:, live], field0, want_g, *g, degree=degree)
    res = torch.empty((ncomp, pos.shape[1]), dtype=torch.float64, device="cuda")
    binning.gather_binned(bins, cuda(field0), res)
    back = bins.unsort(res).cpu().numpy()
    gerr = np.abs(back[:, live] - want_g).max() / np.abs(want_g).max()
    assert gerr < 1e-12, gerr


def anchor_cell(kind, degree, rel, num, code):
    """The stencil cell of one marker on one axis: the node at the stencil's offset (0 for
    linear, 1 for quadratic, degree // 2 for a spline) in the oracle's node list."""
    nodes, _ = orc.axis_stencil(NAMES[(kind, degree)], rel, num, code, degree)
    off = 0 if kind == 0 else (1 if kind == 1 else degree // 2)
    return int(nodes[off])


def anchor_cells(kind, degree, rel, num, code):
    """anchor_cell for many markers at once (numpy); pinned to anchor_cell on a sample."""
    r = rel.astype(np.float64).copy()
    if code == 1:
        bad = ~((r >= 0.0) & (r < num))
        r[bad] = r[bad] - num * np.floor(r[bad] / num)
    if kind == 0:
        i = np.floor(np.clip(np.nan_to_num(r, nan=0.0), 0.0, num)).astype(np.int64)
        return np.minimum(i, num - 1) if code == 1 else np.minimum(i, num - 2)
    if code != 1:
        r = np.clip(r, 0.0, num - 1.0)
    p = 2 if kind == 1 else degree
    mid = np.floor(r + 0.5) if p % 2 == 0 else np.floor(r)
    mid = mid.astype(np.int64)
    return np.where(mid >= num, mid - num, mid) if code == 1 else np.clip(mid, 0, num - 1)


def decode_bin(k, num):
    """(cell per axis) of a 3D bin in tile order (tiles of 8 x 8 x 16, x-major, z fastest)."""
    t, c = divmod(int(k), 8 * 8 * 16)
    nty, ntz = -(-num[1] // 8), -(-num[2] // 16)
    tx, rest = divmod(t, nty * ntz)
    ty, tz = divmod(rest, ntz)
    cx, rest = divmod(c, 8 * 16)
    cy, cz = divmod(rest, 16)
    return (tx * 8 + cx, ty * 8 + cy, tz * 16 + cz)


@pytest.mark.parametrize("case", ["3d-quadratic", "3d-linear-bounded", "3d-cubic-gather", "2d-large", "1d"])
def test_bins_hold_each_marker_in_its_stencil_cell(case):
    """Every binned marker sits in the bin of the cell the oracle's stencil anchors it to (3D
    bins in tiles of 8 x 8 x 16 cells, 2D in tiles of 32 x 32); the 2D case has enough cells
    (2^25) and few enough markers that the sort works on groups of 4096 keys."""
    print(f"bins {case}", flush=True)
    rng = np.random.default_rng(7)
    if case.startswith("3d"):
        num, code = ((40, 24, 48), (1, 0, 1)) if case != "3d-linear-bounded" else ((21, 13, 37), (0, 0, 0))
        kind, degree = {"3d-quadratic": (1, 2), "3d-linear-bounded": (0, 1), "3d-cubic-gather": (2, 3)}[case]
        lower, cell = np.array([0.0, -1.0, 0.5]), np.array([0.0625, 0.125, 0.03125])
    elif case == "2d-large":
        num, code, kind, degree = (8192, 4096), (1, 1), 1, 2
        lower, cell = np.array([0.0, 0.0]), np.array([1 / 8192, 1 / 4096])
    else:
        num, code, kind, degree = (5000,), (0,), 2, 3
        lower, cell = np.array([-2.0]), np.array([0.001])
    ndim = len(num)
    span = cell * np.array(num)
    n = 400000
    pos = rng.uniform(lower - 0.2 * span, lower + 1.2 * span, (n, ndim)).T.copy()
    gather_only = case == "3d-cubic-gather"
    val = None if gather_only else rng.normal(size=(1, n))
    if val is not None:
        val[:, ::50] = 0.0
    g = (*lower, *cell, *num, *code)
    bins = binning.bin_markers(cuda(pos), None if val is None else cuda(val), (kind, degree), g)
    start = bins.start.cpu().numpy().astype(np.int64)
    order = bins.order.cpu().numpy()
    cells = start.size - 2
    assert start[0] == 0 and start[-1] == n and np.all(np.diff(start) >= 0)
    assert np.array_equal(np.sort(order), np.arange(n))
    np.testing.assert_array_equal(bins.positions.cpu().numpy(), pos[:, order])
    if val is not None:
        np.testing.assert_array_equal(bins.values.cpu().numpy(), val[:, order])
        dead = (val == 0).all(axis=0)
        assert start[-1] - start[-2] == dead.sum() and dead[order[start[-2]:]].all()
    else:
        assert start[-1] == start[-2]
    # every binned marker sits in the bin of its anchor cell; the vectorised anchors are
    # checked against the oracle's stencils on a sample first
    spos = bins.positions.cpu().numpy()
    live = int(start[-2])
    rel = [(spos[d, :live] - lower[d]) * (1.0 / cell[d]) for d in range(ndim)]
    anchors = [anchor_cells(kind, degree, rel[d], num[d], code[d]) for d in range(ndim)]
    for i in rng.choice(live, size=min(500, live), replace=False):
        want = tuple(anchor_cell(kind, degree, rel[d][i], num[d], code[d]) for d in range(ndim))
        assert tuple(int(a[i]) for a in anchors) == want, (i, want)
    bin_of = np.searchsorted(start, np.arange(live), side="right") - 1
    if ndim == 3:
        t, c = np.divmod(bin_of, 8 * 8 * 16)
        nty, ntz = -(-num[1] // 8), -(-num[2] // 16)
        tx, rest = np.divmod(t, nty * ntz)
        ty, tz = np.divmod(rest, ntz)
        cx, rest = np.divmod(c, 8 * 16)
        cy, cz = np.divmod(rest, 16)
        got = (tx * 8 + cx, ty * 8 + cy, tz * 16 + cz)
        assert decode_bin(bin_of[live // 2], num) == tuple(int(x[live // 2]) for x in got)
    elif ndim == 2:
        t, c = np.divmod(bin_of, 32 * 32)  # tiles of 32 x 32 cells, x-major, y fastest
        tx, ty = np.divmod(t, -(-num[1] // 32))
        cx, cy = np.divmod(c, 32)
        got = (tx * 32 + cx, ty * 32 + cy)
    else:
        got = (bin_of,)
    for d in range(ndim):
        bad = np.nonzero(got[d] != anchors[d])[0]
        assert bad.size == 0, (d, bad[:5], got[d][bad[:5]], anchors[d][bad[:5]])


def test_binned_deposit_is_bit_reproducible_under_any_marker_order():
    num, code = (32, 32, 32), (1, 1, 1)
    g = geom(num, code)
    pos, val = markers(num, code, 300000, 3, seed=4, dtype=np.float32)
    first = torch.zeros((*num, 3), dtype=torch.float64, device="cuda")
    binning.scatter_binned(binning.bin_markers(cuda(pos), cuda(val), (1, 2), g), first)
    for seed in range(3):
        perm = np.random.default_rng(seed).permutation(pos.shape[1])
        again = torch.zeros_like(first)
        binning.scatter_binned(binning.bin_markers(cuda(pos[:, perm]), cuda(val[:, perm]), (1, 2), g), again)
        assert torch.equal(again, first)


def test_weighting_routes_large_batches_through_the_bins():
    """Weighting/kernels calls above BINNED_MIN_MARKERS bin by themselves; same answers."""
    num, code = (32, 16, 32), (1, 0, 1)
    g = geom(num, code)
    n = K.BINNED_MIN_MARKERS + 1000
    pos, val = markers(num, code, n, 3, seed=9, dtype=np.float32)
    want = np.zeros((*num, 3))
    orc.scatter("quadratic", pos.astype(np.float64), val.astype(np.float64), want, *g)
    got = torch.zeros((*num, 3), dtype=torch.float64, device="cuda")
    with K.deposit_mode("fixed"):
        K.scatter_quadratic_3d(cuda(pos), cuda(val), got, *g)
    assert np.abs(got.cpu().numpy() - want).max() / np.abs(want).max() < 1e-12
    res = torch.empty((3, n), dtype=torch.float32, device="cuda")
    K.gather_quadratic_3d(cuda(pos), cuda(want), res, *g)
    live = ~(val == 0).all(axis=0)
    want_g = np.zeros((3, int(live.sum())))
    orc.gather("quadratic", pos[:, live].astype(np.float64), want, want_g, *g)
    np.testing.assert_allclose(res.cpu().numpy()[:, live], want_g, rtol=2e-6, atol=2e-6 * np.abs(want_g).max())


def test_full_size_invariants():
    """256^3 periodic, 3 components, 2^26 float32 markers: the fixed-point deposit keeps the total
    charge to its quantum and stays within a few quanta of the float64-atomic deposit per node,
    and the adjoint identity <scatter(v), F> = <v, gather(F)> holds to float64 rounding."""
    from mpifield_amd import _lib

    nodes, n = 256, 1 << 26
    g = (0.0, 0.0, 0.0, 1 / nodes, 1 / nodes, 1 / nodes, nodes, nodes, nodes, 1, 1, 1)
    gen = torch.Generator(device="cuda").manual_seed(2)
    pos = torch.rand((3, n), generator=gen, device="cuda", dtype=torch.float32)
    val = (1.0 + 0.1 * torch.randn((3, n), generator=gen, device="cuda")).to(torch.float32)
    bins = binning.bin_markers(pos, val, (1, 2), g)
    field = torch.zeros((nodes, nodes, nodes, 3), dtype=torch.float64, device="cuda")
    binning.scatter_binned(bins, field)
    quantum = float(bins.scalars[1].item())
    total = field.sum(dim=(0, 1, 2))
    want = val.double().sum(dim=1)
    # each of the 27 contributions of a marker is rounded to the nearest quantum
    assert torch.all((total - want).abs() <= 27 * 0.5 * quantum * n + 1e-12 * want.abs())
    atomic = torch.zeros_like(field)
    K.scatter(1, 0, bins.positions, bins.values, atomic, g, mode=_lib.MODE_ATOMIC)
    per_node = (field - atomic).abs().max().item()
    assert per_node <= 200 * quantum, per_node / quantum
    rng = torch.Generator(device="cuda").manual_seed(5)
    probe = torch.randn((nodes, nodes, nodes, 3), generator=rng, device="cuda", dtype=torch.float64)
    res = torch.empty((3, n), dtype=torch.float64, device="cuda")
    bins64 = binning.bin_markers(pos.double(), val.double(), (1, 2), g)
    binning.gather_binned(bins64, probe, res)
    lhs = (atomic * probe).sum(dim=(0, 1, 2))
    rhs = (res * bins64.values).sum(dim=1)
    size = (atomic.abs() * probe.abs()).sum(dim=(0, 1, 2))  # the sum's natural scale
    assert torch.all((lhs - rhs).abs() <= 1e-12 * size), ((lhs - rhs).abs() / size)


KINDS_2D = [(0, 1), (1, 2), (2, 0), (2, 1), (2, 2), (2, 3)]  # linear, quadratic, spline 0..3


@pytest.mark.parametrize("code", [(1, 1), (0, 1), (0, 0)])
@pytest.mark.parametrize("ncomp", [1, 2, 3])
@pytest.mark.parametrize("kind,degree", KINDS_2D)
def test_2d_tiles_match_the_oracle_and_the_direct_fixed_deposit(kind, degree, ncomp, code):
    """2D bins (tiles of 32 x 32 cells): the tiled deposit equals the one-marker-per-work-item
    fixed-point deposit bit for bit and the oracle to the quantum; the tiled gather equals the
    oracle to float64 rounding (dead markers through the dead bin)."""
    from mpifield_amd import _lib

    print(f"2d tiles kind={kind} p={degree} C={ncomp} code={code}", flush=True)
    num = (70, 45)
    lower, cell = (0.5, -1.0), (0.03125, 0.0625)
    g = (*lower, *cell, *num, *code)
    rng = np.random.default_rng(kind * 7 + degree + 3 * ncomp)
    n = 120000
    lo, span = np.array(lower), np.array(cell) * np.array(num)
    margin = np.where(np.array(code) == 1, 1.0, 0.1) * span
    pos = rng.uniform(lo - margin, lo + span + margin, (n, 2)).T.copy()
    val = rng.normal(1.0, 0.3, (ncomp, n))
    val[:, ::97] = 0.0
    pos[:, ::97] = np.nan
    name = {(0, 1): "linear", (1, 2): "quadratic"}.get((kind, degree), "spline")
    want = np.zeros((*num, ncomp))
    orc.scatter(name, pos, val, want, *g, degree=degree)
    bins = binning.bin_markers(cuda(pos), cuda(val), (kind, degree), g)
    tiled = torch.zeros((*num, ncomp), dtype=torch.float64, device="cuda")
    binning.scatter_binned(bins, tiled)
    direct = torch.zeros_like(tiled)
    ndim, lw, cs, nm, cd, _ = K._split(g)
    p_d, v_d = cuda(pos), cuda(val)
    _lib.call("mpf_scatter", kind, degree, 2, p_d.data_ptr(), n, v_d.data_ptr(), n, _lib.F64, n, ncomp,
              direct.data_ptr(), lw.ctypes.data, cs.ctypes.data, nm.ctypes.data, cd.ctypes.data, _lib.MODE_FIXED,
              _lib.SCALE_AUTO, None, K.stream_handle(direct.device))
    assert torch.equal(tiled, direct)
    assert np.abs(tiled.cpu().numpy() - want).max() <= 1e-10 * np.abs(want).max()
    field = rng.normal(size=(*num, ncomp))
    live = ~(val == 0).all(axis=0)
    want_g = np.zeros((ncomp, int(live.sum())))
    orc.gather(name, pos[:, live], field, want_g, *g, degree=degree)
    res = torch.empty((ncomp, n), dtype=torch.float64, device="cuda")
    binning.gather_binned(bins, cuda(field), res)
    back = bins.unsort(res).cpu().numpy()
    np.testing.assert_allclose(back[:, live], want_g, rtol=1e-12, atol=1e-12 * np.abs(want_g).max())


@pytest.mark.parametrize("n", [7, 300, 5000])
def test_2d_tiles_with_few_markers_per_tile(n):
    """Sparse 2D tiles (fewer markers than a lane's run of 8, empty tiles, partial runs): the
    tiled node and forms deposits equal the one-marker-per-work-item fixed-point deposits bit
    for bit."""
    from mpifield_amd import _lib

    print(f"sparse 2d tiles n={n}", flush=True)
    num = (70, 45)
    g = (0.5, -1.0, 0.03125, 0.0625, *num, 0, 1)
    rng = np.random.default_rng(n)
    lo, span = np.array([0.5, -1.0]), np.array([0.03125, 0.0625]) * np.array(num)
    pos = rng.uniform(lo, lo + span, (n, 2)).T.copy()
    for kind, degree, ncomp in ((1, 2, 1), (1, 2, 3), (2, 3, 1)):
        val = rng.normal(1.0, 0.3, (ncomp, n))
        bins = binning.bin_markers(cuda(pos), cuda(val), (kind, degree), g)
        tiled = torch.zeros((*num, ncomp), dtype=torch.float64, device="cuda")
        binning.scatter_binned(bins, tiled)
        direct = torch.zeros_like(tiled)
        with K.deposit_mode("fixed"):
            K.scatter(kind, degree, cuda(pos), cuda(val), direct, g)
        assert torch.equal(tiled, direct)
    for form, ncomp in ((1, 2), (2, 1)):
        val = rng.normal(1.0, 0.3, (ncomp, n))
        bins = binning.bin_markers(cuda(pos), cuda(val), (_lib.KIND_SPLINE, 2), g)
        tiled = torch.zeros((*num, ncomp), dtype=torch.float64, device="cuda")
        binning.scatter_forms_binned(bins, tiled, form)
        direct = torch.zeros_like(tiled)
        with K.deposit_mode("fixed"):
            K.scatter_forms(2, form, cuda(pos), cuda(val), direct, g)
        assert torch.equal(tiled, direct)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bins_are_reused_across_steps_with_new_values(dtype):
    """One sort of the positions serves many deposits: ``with_values`` re-orders each step's
    values through the bins' permutation, and the deposit equals the one from bins made with
    those values, bit for bit.  Bins made with values refuse new ones."""
    num, code = (32, 16, 32), (1, 0, 1)
    g = geom(num, code)
    pos, _ = markers(num, code, 200000, 3, seed=12, dtype=dtype)
    pos[:, ::97] = np.array([[0.3], [-0.2], [1.1]], dtype=dtype)  # finite everywhere
    bins = binning.bin_markers(cuda(pos), None, (1, 2), g)
    rng = np.random.default_rng(13)
    for step in range(2):
        values = rng.normal(0.0, 2.0 + step, (3, pos.shape[1])).astype(dtype)
        reused = torch.zeros((*num, 3), dtype=torch.float64, device="cuda")
        binning.scatter_binned(binning.with_values(bins, cuda(values)), reused)
        fresh = torch.zeros_like(reused)
        binning.scatter_binned(binning.bin_markers(cuda(pos), cuda(values), (1, 2), g), fresh)
        assert torch.equal(reused, fresh), step
        want = np.zeros((*num, 3))
        orc.scatter("quadratic", pos.astype(np.float64), values.astype(np.float64), want, *g)
        # the fixed-point quantum of the automatic scale, a few hundred terms per node
        assert np.abs(reused.cpu().numpy() - want).max() / np.abs(want).max() < 1e-10
    with pytest.raises(ValueError, match="made with values"):
        binning.with_values(binning.bin_markers(cuda(pos), cuda(values), (1, 2), g), cuda(values))


def test_no_markers_leave_the_field_and_raise_nothing():
    """Empty inputs (the reference's kernels loop over zero markers): every entry point takes
    n = 0, adds nothing to the field and writes no results; empty bins hold zero offsets."""
    num, code = (32, 16, 32), (1, 0, 1)
    g = geom(num, code)
    for dtype in (torch.float32, torch.float64):
        pos = torch.empty((3, 0), dtype=dtype, device="cuda")
        val = torch.empty((3, 0), dtype=dtype, device="cuda")
        field = torch.full((*num, 3), 0.5, dtype=torch.float64, device="cuda")
        res = torch.empty((3, 0), dtype=dtype, device="cuda")
        for mode in ("atomic", "fixed"):
            with K.deposit_mode(mode):
                K.scatter(1, 2, pos, val, field, g)
                K.scatter_forms(2, 2, pos, val, field, g)
        K.gather(1, 2, pos, field, res, g)
        K.gather_forms(2, 2, pos, field, res, g)
        bins = binning.bin_markers(pos, val, (1, 2), g)
        assert bins.n == 0 and not bool(bins.start.any())
        binning.scatter_binned(bins, field)
        binning.gather_binned(bins, field, res)
        spline_bins = binning.bin_markers(pos, val, (2, 2), g)  # the forms read bins of the spline nodes
        binning.scatter_forms_binned(spline_bins, field, 2)
        binning.gather_forms_binned(spline_bins, field, res, 2)
        assert bins.unsort(res).shape == (3, 0)
        bare = binning.bin_markers(pos, None, (1, 2), g)
        binning.scatter_binned(binning.with_values(bare, val), field)
        binning.gather_binned(bare.prepare_items(), field, res)
        grid2 = (0.0, 0.0, 0.1, 0.1, 16, 16, 1, 1)
        edges = torch.zeros((16, 16, 2), dtype=torch.float64, device="cuda")
        for kind, degree in ((0, 1), (1, 2)):
            K.deposit_paths(kind, degree, pos[:2], pos[:2], torch.empty(0, dtype=dtype, device="cuda"), edges,
                            grid2)
            K.gather_edges(kind, degree, pos[:2], edges, res[:2], grid2)
        torch.cuda.synchronize()
        assert bool((field == 0.5).all()) and not bool(edges.any())


@pytest.mark.parametrize("dtype,ncomp", [(torch.float32, 3), (torch.float32, 1), (torch.float64, 3)])
def test_binning_of_a_batch_agrees_with_binning_its_halves(dtype, ncomp):
    """5.2M markers on 128^3 (2049 groups of 1024 keys) and each half binned on its own (1025
    groups): every marker lands in the same bin either way, the order is a permutation and
    the sorted rows are the input rows in that order."""
    print(f"batch vs halves dtype={dtype} ncomp={ncomp}", flush=True)
    nodes, n = 128, (1 << 22) + (1 << 20) + 4321
    g = (0.0, 0.0, 0.0, 1 / nodes, 1 / nodes, 1 / nodes, nodes, nodes, nodes, 0, 0, 0)
    gen = torch.Generator(device="cuda").manual_seed(11)
    pos = torch.rand((3, n), generator=gen, device="cuda", dtype=torch.float32).to(dtype)
    val = torch.randn((ncomp, n), generator=gen, device="cuda", dtype=torch.float32).to(dtype)

    def bin_of_input(bins):
        count = bins.order.numel()
        start = bins.start.long() & 0xFFFFFFFF
        idx = torch.arange(count, device="cuda")
        sorted_bin = torch.searchsorted(start, idx, right=True) - 1
        per_input = torch.empty_like(sorted_bin)
        per_input[bins.order] = sorted_bin
        return per_input

    full = binning.bin_markers(pos, val, (1, 2), g)
    assert int(full.start[-1].item()) == n
    assert torch.equal(torch.sort(full.order).values, torch.arange(n, device="cuda"))
    assert torch.equal(full.positions, pos[:, full.order])
    assert torch.equal(full.values, val[:, full.order])
    whole = bin_of_input(full)
    half = n // 2
    for lo, hi in ((0, half), (half, n)):
        part = binning.bin_markers(pos[:, lo:hi].contiguous(), val[:, lo:hi].contiguous(), (1, 2), g)
        assert torch.equal(bin_of_input(part), whole[lo:hi]), (lo, hi)
