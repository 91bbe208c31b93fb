"""Launch variants of the 3D tiled deposit and gather (mpf_tiled_variant): every variant
gives the deposit of variant 0 bit for bit and the same gathered values, on grids whose
last workgroup tile along z is partial (a super-tile of one bin tile) and on bounded axes;
and the deposits on periodic axes whose last tile is short enough that tile 0 meets tile
nt - 2 below the seam (three colours) are bit-reproducible on top of a nonzero field.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from mpifield_amd import _lib, binning

#: (kind, degree, ncomp, nodes, codes); kind 0 linear, 1 quadratic, 2 spline
CASES = [
    (1, 2, 3, (40, 24, 97), (1, 1, 1)),   # 7 bin tiles along z: super-tiles 2+2+2+1; one-cell last tile
    (2, 3, 1, (24, 16, 50), (1, 1, 1)),   # cubic: a two-cell last bin tile (three colours)
    (0, 1, 3, (32, 16, 40), (1, 0, 0)),   # linear, bounded y and z
    (2, 0, 1, (16, 17, 33), (0, 1, 1)),   # nearest node
    (2, 2, 3, (25, 24, 64), (1, 1, 1)),   # spline 2: a one-cell last tile along x
]
KIND_NAME = {0: "linear", 1: "quadratic", 2: "spline"}
N = 200000


def markers(nodes, codes, ncomp, seed):
    rng = np.random.default_rng(seed)
    h = np.array([1.0 / n if c == 1 else 1.0 / (n - 1) for n, c in zip(nodes, codes)])
    pos = rng.uniform(0.0, 1.0, (3, N)).astype(np.float32)
    val = rng.normal(1.0, 0.5, (ncomp, N)).astype(np.float32)
    return (0.0, 0.0, 0.0, *h, *nodes, *codes), pos, val


@pytest.fixture
def lib():
    lib = _lib.load()
    saved = {(op, k, d, c): lib.mpf_tiled_variant(op, k, d, c, -1)
             for op in (0, 1) for k, d, *_ in CASES for c in (1, 3)}
    yield lib
    for (op, k, d, c), v in saved.items():
        lib.mpf_tiled_variant(op, k, d, c, v)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_every_variant_gives_the_same_deposit_and_gather(lib, case):
    kind, degree, ncomp, nodes, codes = CASES[case]
    geo, pos, val = markers(nodes, codes, ncomp, seed=case)
    bins = binning.bin_markers(torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda"),
                               (kind, degree), geo)
    group = 3 if ncomp == 3 and (kind != 2 or degree <= 2) else 1
    f = torch.as_tensor(np.random.default_rng(case + 50).normal(size=(*nodes, ncomp)), device="cuda")
    deposits, gathers = [], []
    for v in range(4):
        print(f"variants case {case} variant {v}", flush=True)
        lib.mpf_tiled_variant(0, kind, degree, group, v)
        lib.mpf_tiled_variant(1, kind, degree, group, v)
        field = torch.zeros((*nodes, ncomp), dtype=torch.float64, device="cuda")
        binning.scatter_binned(bins, field)
        out = torch.empty((ncomp, N), dtype=torch.float32, device="cuda")
        binning.gather_binned(bins, f, out)
        torch.cuda.synchronize()
        deposits.append(field)
        gathers.append(out)
    for v in range(1, 4):
        assert torch.equal(deposits[v], deposits[0]), f"deposit of variant {v}"
        assert torch.equal(gathers[v], gathers[0]), f"gather of variant {v}"
    want = np.zeros((*nodes, ncomp))
    orc.scatter(KIND_NAME[kind], pos.astype(np.float64), val.astype(np.float64), want, *geo, degree=degree)
    got = deposits[0].cpu().numpy()
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    want_g = np.zeros((ncomp, N))
    orc.gather(KIND_NAME[kind], pos.astype(np.float64), f.cpu().numpy(), want_g, *geo, degree=degree)
    res = bins.unsort(gathers[0]).cpu().numpy().astype(np.float64)  # back to the input order
    np.testing.assert_allclose(res, want_g, rtol=2e-6, atol=2e-6 * np.abs(want_g).max())


@pytest.mark.parametrize("case", [0, 1, 4])
def test_short_last_tiles_deposit_reproducibly_onto_a_field(lib, case):
    kind, degree, ncomp, nodes, codes = CASES[case]
    geo, pos, val = markers(nodes, codes, ncomp, seed=case + 9)
    group = 3 if ncomp == 3 and (kind != 2 or degree <= 2) else 1
    base = torch.as_tensor(np.random.default_rng(case).normal(size=(*nodes, ncomp)), device="cuda")
    perm = np.random.default_rng(case + 1).permutation(N)
    for v in (0, 1):
        print(f"reproducible case {case} variant {v}", flush=True)
        lib.mpf_tiled_variant(0, kind, degree, group, v)
        results = []
        for p in (np.arange(N), perm):
            bins = binning.bin_markers(torch.as_tensor(pos[:, p], device="cuda"),
                                       torch.as_tensor(val[:, p], device="cuda"), (kind, degree), geo)
            field = base.clone()
            binning.scatter_binned(bins, field)
            results.append(field)
        assert torch.equal(results[0], results[1])
