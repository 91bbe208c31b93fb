"""Re-binning markers kept in bins order (mpf_rebin_markers, MarkerBins.rebin) against a
fresh binning of the same moved markers: the same bin offsets and scale, the same markers in
every bin, the deposit bit for bit and the gather, on periodic grids with partial tiles and a
periodic axis of two tiles, bounded axes, markers whose values become zero, markers that
jump further than a tile (the call sorts from scratch then), gather-only bins, float64, and
2D bins (sorted again).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from mpifield_amd import binning
    from mpifield_amd import kernels as K

N = 300000

#: (kind, degree, ncomp, nodes, codes); kind 0 linear, 1 quadratic, 2 spline
CASES = [
    (1, 2, 3, (64, 48, 40), (1, 1, 1)),   # a partial last tile along z
    (1, 2, 1, (24, 16, 36), (1, 1, 1)),   # a periodic axis of two tiles (y), three along x
    (0, 1, 3, (40, 24, 33), (0, 1, 0)),   # bounded x and z
    (2, 3, 1, (17, 9, 16), (1, 0, 1)),    # one tile along z, partial tiles along x and y
]


def geometry(nodes, codes):
    h = [1.0 / n if c == 1 else 1.0 / (n - 1) for n, c in zip(nodes, codes)]
    return (0.0,) * len(nodes) + tuple(h) + tuple(nodes) + tuple(codes)


def moved(bins, geo, nodes, codes, rng, far=0):
    """The bins' sorted positions moved by up to half a cell per axis (wrapped on periodic
    axes, clipped on bounded ones), and ``far`` markers moved anywhere."""
    ndim = len(nodes)
    pos = bins.positions.double().cpu().numpy()
    h = np.array(geo[ndim:2 * ndim])
    pos = pos + rng.uniform(-0.5, 0.5, pos.shape) * h[:, None]
    if far:
        idx = rng.choice(pos.shape[1], far, replace=False)
        pos[:, idx] = rng.uniform(0.0, 1.0, (ndim, far))
    for d, c in enumerate(codes):
        pos[d] = np.mod(pos[d], 1.0) if c == 1 else np.clip(pos[d], 0.0, 1.0)
    return torch.as_tensor(pos, dtype=bins.positions.dtype, device="cuda")


def cell_of_sorted(bins):
    counts = (bins.start[1:].long() - bins.start[:-1].long())
    return torch.repeat_interleave(torch.arange(counts.numel(), device="cuda"), counts)


def same_markers_per_bin(a_cells, a_ids, b_cells, b_ids):
    ka = a_cells.cpu().numpy() * (1 << 32) + a_ids.cpu().numpy()
    kb = b_cells.cpu().numpy() * (1 << 32) + b_ids.cpu().numpy()
    return np.array_equal(np.sort(ka), np.sort(kb))


def check_against_fresh(bins0, new, pos1, val1, nodes):
    """new (rebinned from bins0 with pos1/val1 in bins0's order) against a fresh binning of
    pos1/val1."""
    fresh = binning.bin_markers(pos1, val1, (bins0.kind, bins0.degree), bins0.geometry)
    assert torch.equal(new.start, fresh.start)
    assert torch.equal(new.scalars, fresh.scalars)
    # the same markers (by first-input index) in every bin
    fresh_first = bins0.order[fresh.order]
    assert same_markers_per_bin(cell_of_sorted(new), new.order, cell_of_sorted(fresh), fresh_first)
    # each sorted row holds its marker's moved position (and value)
    inv0 = torch.empty_like(bins0.order)
    inv0[bins0.order] = torch.arange(bins0.n, device="cuda")
    at = inv0[new.order]
    assert torch.equal(new.positions, pos1[:, at])
    if val1 is not None:
        assert torch.equal(new.values, val1[:, at])
    _, _, _, num, code = binning._geometry_arrays(bins0.geometry)
    if val1 is not None and K.tiled_ok(bins0.kind, bins0.degree, new.ncomp, num, code):
        f_new = torch.zeros((*nodes, new.ncomp), dtype=torch.float64, device="cuda")
        f_fresh = torch.zeros_like(f_new)
        binning.scatter_binned(new, f_new)
        binning.scatter_binned(fresh, f_fresh)
        assert torch.equal(f_new, f_fresh)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_rebin_equals_a_fresh_binning(case, dtype):
    kind, degree, ncomp, nodes, codes = CASES[case]
    print(f"rebin case {case} {dtype}", flush=True)
    rng = np.random.default_rng(case)
    geo = geometry(nodes, codes)
    pos = torch.as_tensor(rng.uniform(0.0, 1.0, (3, N)), dtype=dtype, device="cuda")
    val = torch.as_tensor(rng.normal(1.0, 0.5, (ncomp, N)), dtype=dtype, device="cuda")
    val[:, ::97] = 0.0  # dead markers
    bins0 = binning.bin_markers(pos, val, (kind, degree), geo)
    pos1 = moved(bins0, geo, nodes, codes, rng)
    val1 = bins0.values.clone()
    val1[:, ::89] = 0.0  # markers that become dead
    val1[:, 5::193] = 2.0  # dead ones that come back to life (far: sorted again)
    new = bins0.rebin(pos1, val1)
    check_against_fresh(bins0, new, pos1, val1, nodes)
    # markers that stay live: the incremental pass alone
    val2 = bins0.values.clone()
    val2[:, ::89] = 0.0
    new2 = bins0.rebin(pos1, val2)
    check_against_fresh(bins0, new2, pos1, val2, nodes)
    # twice in a row: bins of bins
    pos2 = moved(new2, geo, nodes, codes, rng)
    again = new2.rebin(pos2)
    inv0 = torch.empty_like(bins0.order)
    inv0[bins0.order] = torch.arange(N, device="cuda")
    fresh2 = binning.bin_markers(pos2, new2.values, (kind, degree), geo)
    assert torch.equal(again.start, fresh2.start)
    assert same_markers_per_bin(cell_of_sorted(again), again.order, cell_of_sorted(fresh2), new2.order[fresh2.order])


def test_rebin_far_markers_sort_from_scratch():
    kind, degree, ncomp, nodes, codes = CASES[0]
    rng = np.random.default_rng(11)
    geo = geometry(nodes, codes)
    pos = torch.as_tensor(rng.uniform(0.0, 1.0, (3, N)), dtype=torch.float32, device="cuda")
    val = torch.as_tensor(rng.normal(size=(ncomp, N)), dtype=torch.float32, device="cuda")
    bins0 = binning.bin_markers(pos, val, (kind, degree), geo)
    pos1 = moved(bins0, geo, nodes, codes, rng, far=1000)
    new = bins0.rebin(pos1)
    check_against_fresh(bins0, new, pos1, bins0.values, nodes)


def test_rebin_gather_only_bins_and_the_gather():
    kind, degree, _, nodes, codes = CASES[1]
    rng = np.random.default_rng(12)
    geo = geometry(nodes, codes)
    pos = torch.as_tensor(rng.uniform(0.0, 1.0, (3, N)), dtype=torch.float32, device="cuda")
    bins0 = binning.bin_markers(pos, None, (kind, degree), geo)
    pos1 = moved(bins0, geo, nodes, codes, rng)
    new = bins0.rebin(pos1)
    check_against_fresh(bins0, new, pos1, None, nodes)
    f = torch.as_tensor(rng.normal(size=(*nodes, 2)), device="cuda")
    out = torch.empty((2, N), dtype=torch.float32, device="cuda")
    binning.gather_binned(new, f, out)
    fresh = binning.bin_markers(pos1, None, (kind, degree), geo)
    out_f = torch.empty_like(out)
    binning.gather_binned(fresh, f, out_f)
    # both back to the first input's order
    a = new.unsort(out)
    b = bins0.unsort(fresh.unsort(out_f))
    assert torch.equal(a, b)


def test_rebin_2d_sorts_again():
    nodes, codes = (96, 80), (1, 0)
    rng = np.random.default_rng(13)
    geo = geometry(nodes, codes)
    pos = torch.as_tensor(rng.uniform(0.0, 1.0, (2, N)), dtype=torch.float32, device="cuda")
    val = torch.as_tensor(rng.normal(size=(1, N)), dtype=torch.float32, device="cuda")
    bins0 = binning.bin_markers(pos, val, (2, 2), geo)
    pos1 = moved(bins0, geo, nodes, codes, rng)
    new = bins0.rebin(pos1)
    check_against_fresh(bins0, new, pos1, bins0.values, nodes)


def test_rebin_checks_its_arrays():
    """A marker count other than the bins', and new bins written over the old ones (the call
    reads the old arrays while it writes the new)."""
    from mpifield_amd import _lib

    nodes, codes = (32, 32, 32), (1, 1, 1)
    geo = geometry(nodes, codes)
    pos = torch.rand((3, 1000), device="cuda")
    bins0 = binning.bin_markers(pos, None, (1, 2), geo)
    with pytest.raises(ValueError, match="hold 1000 markers"):
        bins0.rebin(pos[:, :144])
    ndim, lower, cell, num, code = binning._geometry_arrays(geo)
    p = bins0.positions
    with pytest.raises(ValueError, match="arrays of their own"):
        _lib.call("mpf_rebin_markers", 1, 2, 3, p.data_ptr(), 1000, None, 0, _lib.F32, 1000, 0,
                  bins0.order.data_ptr(), bins0.start.data_ptr(), lower.ctypes.data, cell.ctypes.data,
                  num.ctypes.data, code.ctypes.data, p.data_ptr(), None, torch.empty_like(bins0.order).data_ptr(),
                  torch.empty_like(bins0.start).data_ptr(), torch.empty(3, dtype=torch.float64, device="cuda").data_ptr(),
                  _lib.SCALE_AUTO, None, K.stream_handle(p.device))
