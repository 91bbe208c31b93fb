"""mpf_scatter_noalloc (the deposit that allocates nothing) against mpf_scatter, through the
C ABI: two calls in a row, the second with a smaller input, each equal to mpf_scatter's
deposit of the same markers (fixed point: bit for bit; float64 atomics: to 1e-12 of the
field, since the addition order differs); a field beyond the scratch is refused."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from mpifield_amd import _lib
    from mpifield_amd import kernels as K


def _arrays(ndim):
    P = ctypes.c_void_p
    return (ctypes.c_double * ndim), (ctypes.c_int64 * ndim), P


def deposit(name, kind, degree, pos, val, field, lower, cell, num, code, mode):
    lib = _lib.load()
    ndim = pos.shape[0]
    D, I, _ = _arrays(ndim)
    lo, cs, nn, cd = D(*lower), D(*cell), I(*num), I(*code)
    dtype = _lib.F32 if pos.dtype == torch.float32 else _lib.F64
    stream = torch.cuda.current_stream().cuda_stream
    common = (kind, degree, ndim, pos.data_ptr(), pos.shape[1], val.data_ptr(), val.shape[1], dtype, pos.shape[1],
              val.shape[0], field.data_ptr(), ctypes.addressof(lo), ctypes.addressof(cs), ctypes.addressof(nn),
              ctypes.addressof(cd), mode, _lib.SCALE_AUTO)
    if name == "mpf_scatter":
        rc = lib.mpf_scatter(*common, None, stream)
    else:
        rc = lib.mpf_scatter_noalloc(*common, stream)
    assert rc == 0, lib.mpf_last_error()


CASES = [
    # kind, degree, ndim, code, ncomp, dtype, nodes
    (1, 2, 3, (1, 1, 1), 3, torch.float32, (64, 64, 64)),
    (2, 3, 2, (2, 0), 1, torch.float64, (50, 40)),
    (0, 1, 1, (1,), 2, torch.float64, (300,)),
]


@pytest.mark.parametrize("mode", ["fixed", "atomic"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_deposit_equals_mpf_scatter_twice(case, mode):
    kind, degree, ndim, code, ncomp, dtype, nodes = CASES[case]
    print(f"noalloc case {case} mode={mode}", flush=True)
    mode_code = _lib.MODE_FIXED if mode == "fixed" else _lib.MODE_ATOMIC
    cell = [1.0 / n for n in nodes]
    lower = [0.0] * ndim
    extent = tuple(n + (degree - 1 if c == 2 else 0) for n, c in zip(nodes, code))
    rng = np.random.default_rng(case)
    # mpf_scatter on the direct kernels (the ones the no-alloc entry point runs) at every
    # size; the second call smaller than the first
    for n in (250000, 7000):
        pos = torch.as_tensor(rng.uniform(-0.05, 1.05, (ndim, n)), device="cuda", dtype=dtype)
        val = torch.as_tensor(rng.normal(1.0, 0.3, (ncomp, n)), device="cuda", dtype=dtype)
        want = torch.zeros((*extent, ncomp), dtype=torch.float64, device="cuda")
        got = torch.zeros_like(want)
        with K.direct_kernels():
            deposit("mpf_scatter", kind, degree, pos, val, want, lower, cell, nodes, code, mode_code)
        deposit("mpf_scatter_noalloc", kind, degree, pos, val, got, lower, cell, nodes, code, mode_code)
        torch.cuda.synchronize()
        if mode == "fixed":
            assert torch.equal(got, want), n
        else:
            assert (got - want).abs().max().item() <= 1e-12 * want.abs().max().item(), n


def test_noalloc_refuses_a_field_beyond_its_scratch():
    lib = _lib.load()
    D, I, _ = _arrays(3)
    lower, cell = D(0.0, 0.0, 0.0), D(1.0, 1.0, 1.0)
    num, code = I(256, 256, 256), I(1, 1, 1)  # 2^24 nodes x 1 component: twice MPF_NOALLOC_MAX_VALUES
    rc = lib.mpf_scatter_noalloc(1, 2, 3, None, 0, None, 0, _lib.F32, 1, 1, None, ctypes.addressof(lower),
                                 ctypes.addressof(cell), ctypes.addressof(num), ctypes.addressof(code),
                                 _lib.MODE_FIXED, _lib.SCALE_AUTO, None)
    assert rc == 1
    assert b"MPF_NOALLOC_MAX_VALUES" in lib.mpf_last_error()
