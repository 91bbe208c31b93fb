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
