"""The permutation plan: marker rows between the input order and the bins' order in two
passes of whole blocks (mpf_perm_plan_build, mpf_perm_rows), against torch's index_copy_ /
index_select on the same permutation, bit for bit (a permutation moves values unchanged).
Also the routes that use it: mpf_unsort_rows from 2^20 markers (a plan made in the stream's
default context), the in-call unsort of mpf_gather, MarkerBins.unsort / to_bins_order and
with_values."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from mpifield_amd import _lib, binning
    from mpifield_amd import kernels as K


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def build_plan(order: torch.Tensor) -> torch.Tensor:
    n = order.numel()
    size = ctypes.c_int64(0)
    _lib.call("mpf_perm_plan_bytes", n, ctypes.addressof(size))
    assert size.value > 0
    plan = torch.empty(size.value, dtype=torch.uint8, device="cuda")
    _lib.call("mpf_perm_plan_build", order.data_ptr(), n, plan.data_ptr(), None, stream())
    return plan


@pytest.mark.parametrize("n", [1, 999, 1 << 14, (1 << 14) + 1, 1 << 15, (1 << 15) + 1, 100_003, (1 << 20) + 12_345])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("rows", [1, 3])
def test_plan_moves_rows_like_index_ops(n, dtype, rows):
    """Both directions, on rows with gaps between them (row strides n + 7 and n + 64),
    against index_copy_ (bins order -> input order) and index_select (input -> bins)."""
    print(f"perm plan n={n} {dtype} rows={rows}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(n)
    order = torch.randperm(n, device="cuda", generator=g)
    plan = build_plan(order)
    src_store = torch.randn((rows, n + 7), dtype=torch.float64, device="cuda", generator=g).to(dtype)
    src = src_store[:, :n]
    dst_store = torch.full((rows, n + 64), -1.0, dtype=dtype, device="cuda")
    code = _lib.F32 if dtype == torch.float32 else _lib.F64
    _lib.call("mpf_perm_rows", plan.data_ptr(), n, 0, src.data_ptr(), n + 7, dst_store.data_ptr(), n + 64, rows,
              code, None, stream())
    want = torch.empty((rows, n), dtype=dtype, device="cuda")
    want.index_copy_(1, order, src)
    assert torch.equal(dst_store[:, :n], want)
    assert torch.all(dst_store[:, n:] == -1.0)  # the gaps between rows untouched
    back = torch.full((rows, n + 64), -1.0, dtype=dtype, device="cuda")
    _lib.call("mpf_perm_rows", plan.data_ptr(), n, 1, dst_store.data_ptr(), n + 64, back.data_ptr(), n + 64, rows,
              code, None, stream())
    assert torch.equal(back[:, :n], src)
    assert torch.equal(back[:, :n], torch.index_select(dst_store[:, :n], 1, order))
    assert torch.all(back[:, n:] == -1.0)


def test_plan_reused_across_calls_and_rows():
    """One plan serves many calls: 1, 2 and 5 rows in a row, each equal to index_copy_."""
    n = 300_001
    order = torch.randperm(n, device="cuda")
    plan = build_plan(order)
    for rows in (1, 2, 5):
        src = torch.randn((rows, n), device="cuda")
        out = torch.empty_like(src)
        _lib.call("mpf_perm_rows", plan.data_ptr(), n, 0, src.data_ptr(), n, out.data_ptr(), n, rows, _lib.F32, None,
                  stream())
        want = torch.empty_like(src)
        want.index_copy_(1, order, src)
        assert torch.equal(out, want), rows


@pytest.mark.parametrize("n", [(1 << 20) - 1, (1 << 20) + 777])
@pytest.mark.parametrize("rows", [1, 3])
def test_unsort_rows_below_and_above_the_planned_size(n, rows):
    """mpf_unsort_rows: scattered writes below 2^20 markers or for one row, a plan made in the
    default context otherwise; the same values either way."""
    order = torch.randperm(n, device="cuda")
    src = torch.randn((rows, n), device="cuda")
    out = torch.empty_like(src)
    _lib.call("mpf_unsort_rows", src.data_ptr(), out.data_ptr(), order.data_ptr(), rows, n, _lib.F32, stream())
    want = torch.empty_like(src)
    want.index_copy_(1, order, src)
    assert torch.equal(out, want)


def test_single_row_moves_plan_on_the_second_use():
    """One row of fresh bins moves by scattered accesses, the second single-row move makes
    the plan, and both equal the index ops."""
    n, nodes = 1 << 19, 64
    rng = np.random.default_rng(9)
    pos = torch.as_tensor(rng.random((3, n), dtype=np.float32), device="cuda")
    geometry = (0.0, 0.0, 0.0, 1 / nodes, 1 / nodes, 1 / nodes, nodes, nodes, nodes, 1, 1, 1)
    bins = binning.bin_markers(pos, None, (2, 3), geometry)
    row = torch.randn((1, n), device="cuda")
    assert torch.equal(bins.to_bins_order(row), torch.index_select(row, 1, bins.order))
    assert bins.plan is None
    want = torch.empty_like(row)
    want.index_copy_(1, bins.order, row)
    assert torch.equal(bins.unsort(row), want)
    assert bins.plan is not None


def test_bins_unsort_to_bins_order_and_with_values():
    """MarkerBins of a real binning (2^19 markers on 64^3): unsort and to_bins_order through
    the bins' plan equal the index ops, with_values keeps the plan and sorts values as
    index_select does, and the binned gather in the input order matches the direct gather."""
    n, nodes = 1 << 19, 64
    rng = np.random.default_rng(5)
    pos = torch.as_tensor(rng.random((3, n), dtype=np.float32), device="cuda")
    val = torch.as_tensor(rng.normal(1.0, 0.1, (3, n)).astype(np.float32), device="cuda")
    geometry = (0.0, 0.0, 0.0, 1 / nodes, 1 / nodes, 1 / nodes, nodes, nodes, nodes, 1, 1, 1)
    bins = binning.bin_markers(pos, None, (1, 2), geometry)
    assert bins.permutation_plan() is not None
    rows = torch.randn((3, n), device="cuda")
    want = torch.empty_like(rows)
    want.index_copy_(1, bins.order, rows)
    assert torch.equal(bins.unsort(rows), want)
    assert torch.equal(bins.to_bins_order(rows), torch.index_select(rows, 1, bins.order))
    # a strided view of the values (rows n + 5 apart) and a single row
    store = torch.zeros((3, n + 5), device="cuda")
    store[:, :n] = val
    assert torch.equal(bins.to_bins_order(store[:, :n]), torch.index_select(val, 1, bins.order))
    assert torch.equal(bins.to_bins_order(val[1:2]), torch.index_select(val[1:2], 1, bins.order))
    withv = binning.with_values(bins, val)
    assert withv.plan is bins.plan
    assert torch.equal(withv.values, torch.index_select(val, 1, bins.order))
    field = torch.zeros((nodes, nodes, nodes, 3), dtype=torch.float64, device="cuda")
    binning.scatter_binned(withv, field)
    sorted_res = torch.empty((3, n), dtype=torch.float32, device="cuda")
    binning.gather_binned(withv, field, sorted_res)
    direct = torch.empty((3, n), dtype=torch.float32, device="cuda")
    with K.direct_kernels():  # the one-marker-per-work-item gather in the input order
        K.gather_quadratic_3d(pos, field, direct, *geometry)
    assert torch.allclose(withv.unsort(sorted_res), direct, rtol=1e-6, atol=1e-6)
    # the in-call route of mpf_gather (bins, tiles and a planned unsort inside the call)
    routed = torch.empty_like(direct)
    with K.binned_from(1 << 16):
        K.gather_quadratic_3d(pos, field, routed, *geometry)
    assert torch.equal(routed, withv.unsort(sorted_res))
