"""The catalogue's per-name C symbols called through raw ctypes, the tiled route inside the
library, explicit bin reuse in the plugin API, and one default context per stream.

mpf_scatter_quadratic_3d / mpf_gather_quadratic_3d (and every per-name symbol, since they
forward to mpf_scatter / mpf_gather) bin large batches inside the call and run the tiled
kernels; a C or FFI caller gets that route with plain device pointers, as a maintainer
binding the library from the reference's catalogue would (catalog.py:77-114).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    import mpifield_amd as m
    from mpifield_amd import _lib, binning
    from mpifield_amd import kernels as K


def raw_lib():
    """A fresh ctypes handle on the built library with the header's argument types (no use
    of the package's wrappers)."""
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    P, I64, INT, DBL = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    geom3 = [DBL] * 6 + [I64] * 6
    lib.mpf_scatter_quadratic_3d.argtypes = [P, I64, P, I64, INT, I64, I64, P] + geom3 + [INT, P, P]
    lib.mpf_gather_quadratic_3d.argtypes = [P, I64, INT, I64, P, I64, P, I64] + geom3 + [P, P]
    lib.mpf_scatter_spline_2d.argtypes = [P, I64, P, I64, INT, I64, I64, P] + [DBL] * 4 + [I64] * 4 + [INT, INT, P, P]
    lib.mpf_last_error.restype = ctypes.c_char_p
    return lib


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("n", [5000, (1 << 18) + 4321])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_raw_ctypes_quadratic_3d_matches_oracle(n, dtype):
    """mpf_scatter_quadratic_3d (fixed point) and mpf_gather_quadratic_3d on strided marker
    rows (row stride n + 37) and a strided result, below and above the binning threshold,
    against the oracle fed the same values."""
    print(f"raw ctypes quadratic 3D n={n} {dtype}", flush=True)
    lib = raw_lib()
    nodes, comps = 32, 3
    h = 1.0 / nodes
    rng = np.random.default_rng(n)
    stride = n + 37
    pos_store = torch.zeros((3, stride), dtype=dtype, device="cuda")
    val_store = torch.zeros((comps, stride), dtype=dtype, device="cuda")
    pos_h = rng.random((3, n)).astype(np.float32 if dtype == torch.float32 else np.float64)
    val_h = rng.normal(1.0, 0.1, (comps, n)).astype(pos_h.dtype)
    pos_store[:, :n] = torch.as_tensor(pos_h, device="cuda")
    val_store[:, :n] = torch.as_tensor(val_h, device="cuda")
    field = torch.zeros((nodes, nodes, nodes, comps), dtype=torch.float64, device="cuda")
    code = _lib.F32 if dtype == torch.float32 else _lib.F64
    geo = (0.0, 0.0, 0.0, h, h, h, nodes, nodes, nodes, 1, 1, 1)
    rc = lib.mpf_scatter_quadratic_3d(pos_store.data_ptr(), stride, val_store.data_ptr(), stride, code, n, comps,
                                      field.data_ptr(), *geo, _lib.MODE_FIXED, None, stream())
    assert rc == 0, lib.mpf_last_error()
    want = np.zeros((nodes, nodes, nodes, comps))
    orc.scatter("quadratic", pos_h.astype(np.float64), val_h.astype(np.float64), want, *geo)
    got = field.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    res_store = torch.full((comps, stride), -7.0, dtype=dtype, device="cuda")
    rc = lib.mpf_gather_quadratic_3d(pos_store.data_ptr(), stride, code, n, field.data_ptr(), comps,
                                     res_store.data_ptr(), stride, *geo, None, stream())
    assert rc == 0, lib.mpf_last_error()
    want_g = np.zeros((comps, n))
    orc.gather("quadratic", pos_h.astype(np.float64), got, want_g, *geo)
    res = res_store.cpu().numpy()
    tol = 2e-6 if dtype == torch.float32 else 1e-12
    np.testing.assert_allclose(res[:, :n], want_g, rtol=tol, atol=tol * np.abs(want_g).max())
    assert np.all(res[:, n:] == -7.0)  # the gaps between rows are left alone


def test_raw_ctypes_spline_2d_above_the_threshold_matches_oracle():
    """mpf_scatter_spline_2d (cubic, periodic x, bounded y) above the threshold, where the
    library bins inside the call and runs the 2D tiles, against the oracle."""
    lib = raw_lib()
    n, nodes = (1 << 18) + 100, 64
    h = 1.0 / nodes
    rng = np.random.default_rng(3)
    pos = torch.as_tensor(rng.random((2, n)), device="cuda")
    val = torch.as_tensor(rng.normal(size=(1, n)), device="cuda")
    geo = (0.0, 0.0, h, h, nodes, nodes, 1, 0)
    whole = torch.zeros((nodes, nodes, 1), dtype=torch.float64, device="cuda")
    rc = lib.mpf_scatter_spline_2d(pos.data_ptr(), n, val.data_ptr(), n, _lib.F64, n, 1, whole.data_ptr(), *geo, 3,
                                   _lib.MODE_FIXED, None, stream())
    assert rc == 0, lib.mpf_last_error()
    want = np.zeros((nodes, nodes, 1))
    orc.scatter("spline", pos.cpu().numpy(), val.cpu().numpy(), want, *geo, degree=3)
    assert np.abs(whole.cpu().numpy() - want).max() <= 1e-10 * np.abs(want).max()


def test_binned_entry_points_refuse_a_field_of_another_grid():
    """gather_binned, scatter_forms_binned, gather_forms_binned and the derivative check the
    field against the bins' grid (and device) before any launch."""
    n = 1000
    geo = (0.0, 0.0, 0.0, 1 / 16, 1 / 16, 1 / 32, 16, 16, 32, 1, 1, 1)
    pos = torch.rand((3, n), device="cuda", dtype=torch.float64)
    bins = binning.bin_markers(pos, None, (1, 2), geo)
    wrong = torch.zeros((16, 16, 16, 3), dtype=torch.float64, device="cuda")
    out = torch.empty((3, n), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="a field on this grid"):
        binning.gather_binned(bins, wrong, out)
    with pytest.raises(ValueError, match="a field on this grid"):
        binning.gather_derivative_binned(bins, wrong, out)
    sbins = binning.bin_markers(pos, torch.ones((3, n), dtype=torch.float64, device="cuda"), (2, 2), geo)
    with pytest.raises(ValueError, match="2-form field on this grid"):
        binning.scatter_forms_binned(sbins, wrong, 2)
    gbins = binning.bin_markers(pos, None, (2, 2), geo)
    with pytest.raises(ValueError, match="2-form field on this grid"):
        binning.gather_forms_binned(gbins, wrong, out, 2)
    with pytest.raises(ValueError, match="result must be"):
        binning.gather_binned(bins, torch.zeros((16, 16, 32, 3), dtype=torch.float64, device="cuda"),
                              out[:2])


@pytest.mark.parametrize("mode", ["fixed", "atomic"])
def test_bins_passed_as_positions_give_the_results_of_positions(mode):
    """Field.bin_markers / Weighting.bin: bins passed as positions to Field.scatter and
    Field.gather give the results of the positions themselves (fixed point: bit for bit)."""
    nodes, n = 32, 300000
    grid_args = (3, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), nodes)
    f1 = m.Field(*grid_args, num_components=3, periodic=True)
    f2 = m.Field(*grid_args, num_components=3, periodic=True)
    rng = np.random.default_rng(4)
    pos = torch.as_tensor(rng.random((3, n)), device="cuda", dtype=torch.float32)
    val = torch.as_tensor(rng.normal(size=(3, n)), device="cuda", dtype=torch.float32)
    bins = f1.bin_markers(pos, m.QUADRATIC)
    with K.deposit_mode(mode):
        f1.scatter(bins, val, m.QUADRATIC)
        f2.scatter(pos, val, m.QUADRATIC)
    a, b = f1.local, f2.local
    if mode == "fixed":
        assert torch.equal(a, b)
    else:
        assert (a - b).abs().max().item() <= 1e-12 * b.abs().max().item()
    ga = f1.gather(bins, m.QUADRATIC)
    gb = f1.gather(pos, m.QUADRATIC)
    assert torch.equal(ga, gb)
    with pytest.raises(ValueError, match="another weighting or grid"):
        f1.gather(bins, m.LINEAR)


def test_default_contexts_are_per_stream():
    """Fixed-point deposits (the direct kernels, whose int64 accumulator lives in the default
    context) issued on two streams at once give the results of the same calls made one after
    the other on one stream."""
    n, nodes = 100000, 24
    h = 1.0 / nodes
    geo = (0.0, 0.0, 0.0, h, h, h, nodes, nodes, nodes, 1, 1, 1)
    rng = np.random.default_rng(6)
    pos = [torch.as_tensor(rng.random((3, n)), device="cuda") for _ in range(2)]
    val = [torch.as_tensor(rng.normal(size=(2, n)), device="cuda") for _ in range(2)]
    want = []
    for p, v in zip(pos, val):
        f = torch.zeros((nodes, nodes, nodes, 2), dtype=torch.float64, device="cuda")
        for _ in range(4):
            K.scatter(2, 3, p, v, f, geo, mode=_lib.MODE_FIXED)
        want.append(f)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    got = [torch.zeros_like(want[0]) for _ in range(2)]
    for _ in range(3):
        for f in got:
            f.zero_()
        torch.cuda.synchronize()
        for s, p, v, f in zip(streams, pos, val, got):
            with torch.cuda.stream(s):
                for _ in range(4):
                    K.scatter(2, 3, p, v, f, geo, mode=_lib.MODE_FIXED)
        torch.cuda.synchronize()
        for f, w in zip(got, want):
            assert torch.equal(f, w)
