"""The 3D tiles on the shapes they took last: periodic extents that are not whole tiles
(a partial last tile next to the seam), component counts other than 1 and 3 (groups of
three, then single components), and the cubic stencil with several components (one at a
time).  Each through mpf_scatter / mpf_gather, which bin inside the call above 2^18 markers,
against the oracle fed the same values, and bit-reproducible under any marker order.
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

N = (1 << 18) + 5000

#: (kind, degree, ncomp, nodes, codes): kind 0 linear, 1 quadratic, 2 spline
CASES = [
    (1, 2, 2, (100, 36, 52), (1, 1, 1)),   # partial last tiles on every periodic axis, two components
    (1, 2, 5, (40, 24, 50), (1, 0, 1)),    # five components: a group of three, then two single ones
    (2, 3, 3, (48, 40, 34), (1, 1, 1)),    # cubic, three components one at a time, partial tiles
    (0, 1, 4, (17, 9, 33), (1, 0, 1)),     # linear on short periodic axes (two tiles, the second partial)
    (2, 0, 2, (24, 16, 20), (0, 1, 0)),    # nearest node
    (2, 1, 3, (10, 16, 19), (1, 1, 1)),    # periodic axes just a footprint long
]
KIND_NAME = {0: "linear", 1: "quadratic", 2: "spline"}


def markers(nodes, codes, ncomp, dtype, seed):
    rng = np.random.default_rng(seed)
    h = np.array([1.0 / n if c == 1 else 1.0 / (n - 1) for n, c in zip(nodes, codes)])
    pos = rng.uniform(-0.03, 1.03, (3, N))
    pos[:, ::9] = rng.uniform(0.0, 1.0, (3, pos[:, ::9].shape[1]))
    val = rng.normal(1.0, 0.5, (ncomp, N))
    geo = (0.0, 0.0, 0.0, *h, *nodes, *codes)
    return geo, pos.astype(dtype), val.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_wide_tiles_match_oracle(case, dtype):
    kind, degree, ncomp, nodes, codes = CASES[case]
    print(f"wide tiles case {case} {np.dtype(dtype).name}", flush=True)
    assert K.binned_route(kind, degree, ncomp, nodes, codes, N)
    geo, pos, val = markers(nodes, codes, ncomp, dtype, seed=case)
    want = np.zeros((*nodes, ncomp))
    orc.scatter(KIND_NAME[kind], pos.astype(np.float64), val.astype(np.float64), want, *geo, degree=degree)
    field = torch.zeros((*nodes, ncomp), dtype=torch.float64, device="cuda")
    K.scatter(kind, degree, torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda"), field, geo,
              mode=_lib.MODE_FIXED)
    got = field.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    # the gather of a smooth random field, against the oracle on the same field
    f = np.random.default_rng(case + 100).normal(size=(*nodes, ncomp))
    want_g = np.zeros((ncomp, N))
    orc.gather(KIND_NAME[kind], pos.astype(np.float64), f, want_g, *geo, degree=degree)
    res = torch.empty((ncomp, N), dtype=torch.float32 if dtype == np.float32 else torch.float64, device="cuda")
    K.gather(kind, degree, torch.as_tensor(pos, device="cuda"), torch.as_tensor(f, device="cuda"), res, geo)
    tol = 2e-6 if dtype == np.float32 else 1e-12
    np.testing.assert_allclose(res.cpu().numpy(), want_g, rtol=tol, atol=tol * np.abs(want_g).max())


@pytest.mark.parametrize("case", [0, 2])
def test_wide_tiles_are_bit_reproducible_under_any_marker_order(case):
    kind, degree, ncomp, nodes, codes = CASES[case]
    geo, pos, val = markers(nodes, codes, ncomp, np.float32, seed=case + 7)
    first = torch.zeros((*nodes, ncomp), dtype=torch.float64, device="cuda")
    K.scatter(kind, degree, torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda"), first, geo,
              mode=_lib.MODE_FIXED)
    perm = np.random.default_rng(case).permutation(N)
    again = torch.zeros_like(first)
    K.scatter(kind, degree, torch.as_tensor(pos[:, perm], device="cuda"), torch.as_tensor(val[:, perm], device="cuda"),
              again, geo, mode=_lib.MODE_FIXED)
    assert torch.equal(again, first)


#: clamped spline axes (code 2: extent num + degree - 1, cells 0 .. num - 2, Cox-de Boor
#: weights near the walls): (degree, ncomp, nodes, codes)
CLAMPED = [
    (2, 1, (40, 36, 44), (2, 1, 2)),   # quadratic, clamped x and z around a periodic y
    (3, 1, (24, 20, 33), (2, 2, 2)),   # cubic, every axis clamped, partial last tiles
    (1, 3, (16, 24, 40), (0, 2, 1)),   # linear spline, three components: bounded, clamped, periodic
    (0, 2, (20, 16, 18), (2, 0, 2)),   # nearest node: num - 1 nodes on a clamped axis
    (2, 5, (8, 8, 16), (2, 2, 2)),     # one tile per axis; a group of three components, then two
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", range(len(CLAMPED)))
def test_clamped_tiles_match_oracle_and_the_direct_kernels(case, dtype):
    """Clamped axes (the reference's clamped splines, tests/unit/test_clamped.py) take the
    tiles: mpf_scatter above 2^18 markers bins inside the call and deposits tile by tile;
    the field equals the oracle's clamped deposit and the direct fixed-point deposit of the
    same markers to the fixed-point quantum, the charge stays on the grid, and the tiled
    gather equals the oracle's and the direct kernels' gather."""
    degree, ncomp, nodes, codes = CLAMPED[case]
    print(f"clamped tiles case {case} p={degree} C={ncomp} {np.dtype(dtype).name}", flush=True)
    assert K.binned_route(2, degree, ncomp, nodes, codes, N)
    geo, pos, val = markers(nodes, codes, ncomp, dtype, seed=50 + case)
    extent = tuple(n + (degree - 1 if c == 2 else 0) for n, c in zip(nodes, codes))
    want = np.zeros((*extent, ncomp))
    orc.scatter("spline", pos.astype(np.float64), val.astype(np.float64), want, *geo, degree=degree)
    p_d, v_d = torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda")
    tiled = torch.zeros((*extent, ncomp), dtype=torch.float64, device="cuda")
    K.scatter(2, degree, p_d, v_d, tiled, geo, mode=_lib.MODE_FIXED)
    got = tiled.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    if all(c != 0 for c in codes):  # no bounded wall folds: every marker's charge is on the grid
        np.testing.assert_allclose(got.sum(axis=(0, 1, 2)), val.astype(np.float64).sum(axis=1), rtol=1e-10)
    # the direct fixed-point deposit rounds each contribution's float64 product to quanta,
    # the tiles round the exact product: the two differ by at most a quantum per marker
    direct = torch.zeros_like(tiled)
    with K.direct_kernels():
        K.scatter(2, degree, p_d, v_d, direct, geo, mode=_lib.MODE_FIXED)
    assert float((tiled - direct).abs().max()) <= 1e-11 * np.abs(want).max()
    if dtype == np.float32:  # any marker order gives the same bits
        perm = torch.as_tensor(np.random.default_rng(case).permutation(N), device="cuda")
        again = torch.zeros_like(tiled)
        K.scatter(2, degree, p_d[:, perm].contiguous(), v_d[:, perm].contiguous(), again, geo, mode=_lib.MODE_FIXED)
        assert torch.equal(again, tiled)

    f = np.random.default_rng(case + 200).normal(size=(*extent, ncomp))
    want_g = np.zeros((ncomp, N))
    orc.gather("spline", pos.astype(np.float64), f, want_g, *geo, degree=degree)
    rdt = torch.float32 if dtype == np.float32 else torch.float64
    res = torch.empty((ncomp, N), dtype=rdt, device="cuda")
    K.gather(2, degree, p_d, torch.as_tensor(f, device="cuda"), res, geo)
    tol = 2e-6 if dtype == np.float32 else 1e-12
    np.testing.assert_allclose(res.cpu().numpy(), want_g, rtol=tol, atol=tol * np.abs(want_g).max())
    if dtype == np.float64:  # the tiles' float64 gather sums in the direct kernels' order
        res_direct = torch.empty_like(res)
        with K.direct_kernels():
            K.gather(2, degree, p_d, torch.as_tensor(f, device="cuda"), res_direct, geo)
        torch.testing.assert_close(res, res_direct, rtol=1e-14, atol=1e-14 * float(res.abs().max()))


def test_clamped_bins_are_kept_and_binned_again():
    """Bins on a clamped grid (binning.bin_markers with the spline weighting) serve the
    deposit and gather of new values, and rebin (a fresh sort on clamped axes) gives bins
    whose deposit equals the deposit of the moved markers."""
    from mpifield_amd import binning

    degree, ncomp, nodes, codes = CLAMPED[0]
    geo, pos, val = markers(nodes, codes, 1, np.float32, seed=77)
    extent = tuple(n + (degree - 1 if c == 2 else 0) for n, c in zip(nodes, codes))
    bins = binning.bin_markers(torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda"),
                               (2, degree), geo)
    got = torch.zeros((*extent, 1), dtype=torch.float64, device="cuda")
    binning.scatter_binned(bins, got)
    want = torch.zeros_like(got)
    K.scatter(2, degree, torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda"), want, geo,
              mode=_lib.MODE_FIXED)
    assert torch.equal(got, want)
    moved = bins.positions.clone()
    moved[0] += 0.02  # x moves: some markers cross cells and clamp at the upper wall
    again = bins.rebin(moved)
    order = again.order.cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(N))
    first = bins.order.cpu().numpy()
    moved_in = np.empty_like(pos)
    moved_in[:, first] = moved.cpu().numpy()  # the moved markers in the input order
    np.testing.assert_array_equal(again.positions.cpu().numpy(), moved_in[:, order])
    np.testing.assert_array_equal(again.values.cpu().numpy(), val[:, order])
    field = torch.zeros_like(got)
    binning.scatter_binned(again, field)
    want2 = torch.zeros_like(got)
    K.scatter(2, degree, torch.as_tensor(moved_in, device="cuda"), torch.as_tensor(val, device="cuda"), want2, geo,
              mode=_lib.MODE_FIXED)
    assert torch.equal(field, want2)
