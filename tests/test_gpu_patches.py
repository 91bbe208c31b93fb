"""Batched patches: every patch's deposit and gather in one persistent launch equals the
oracle run patch by patch (what a caller of the reference does with scatter_block per
patch); the autotuner's choice changes nothing in the result; HIP-graph replays equal the
eager deposit every time."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    import mpifield_amd as m
    from mpifield_amd import patches as P
    from mpifield_amd.weighting import spline_nodes

NAME = {"linear": "linear", "quadratic": "quadratic"}


def make_batch(seed=0, n_patches=40, ndim=2, dtype=np.float64, sides=None):
    rng = np.random.default_rng(seed)
    batch = P.PatchBatch(ndim=ndim, ncomp=1)
    spec, pos_parts, val_parts = [], [], []
    choices = [m.LINEAR, m.QUADRATIC, m.CUBIC, spline_nodes("spline0", 0)]
    if sides is None:
        sides = (16, 64) if ndim == 2 else (8, 16)
    for i in range(n_patches):
        side = rng.integers(sides[0], sides[1] + 1, ndim)
        cell = rng.uniform(0.01, 0.05, ndim)
        lower = rng.uniform(-1, 1, ndim)
        weighting = choices[i % len(choices)]
        count = int(np.exp(rng.uniform(np.log(100), np.log(20000))))
        periodic = tuple(bool(x) for x in rng.integers(0, 2, ndim))
        offset, sl = batch.add(lower, cell, side, weighting, count, periodic)
        upper = lower + cell * side
        pos = rng.uniform(lower - cell, upper + cell, (count, ndim)).T
        val = rng.normal(1.0, 0.2, (1, count))
        val[:, ::50] = 0.0
        pos_parts.append(pos)
        val_parts.append(val)
        spec.append((lower, cell, side, weighting, periodic, offset, sl))
    pos = np.concatenate(pos_parts, axis=1).astype(dtype)
    val = np.concatenate(val_parts, axis=1).astype(dtype)
    return batch, spec, pos, val


def oracle_patch(weighting, pos, val, lower, cell, side, periodic):
    kind = {"linear": "linear", "quadratic": "quadratic"}.get(weighting.name, "spline")
    g = (*lower, *cell, *side, *(int(p) for p in periodic))
    out = np.zeros((*side, 1))
    orc.scatter(kind, pos, val, out, *g, degree=weighting.degree)
    return out, g, kind


@pytest.mark.parametrize("ndim", [2, 3])
def test_batched_patches_match_the_oracle_patch_by_patch(ndim):
    batch, spec, pos, val = make_batch(seed=ndim, ndim=ndim)
    fields = torch.zeros(batch.field_values, dtype=torch.float64, device="cuda")
    pos_d, val_d = torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda")
    batch.build(chunk=1024)
    batch.scatter(pos_d, val_d, fields)
    result = torch.empty((1, pos.shape[1]), dtype=torch.float64, device="cuda")
    probe = torch.as_tensor(np.random.default_rng(9).normal(size=batch.field_values), device="cuda")
    batch.gather(pos_d, probe, result)
    got = fields.cpu().numpy()
    res = result.cpu().numpy()
    probe_h = probe.cpu().numpy()
    for i, (lower, cell, side, weighting, periodic, offset, sl) in enumerate(spec):
        want, g, kind = oracle_patch(weighting, pos[:, sl], val[:, sl], lower, cell, side, periodic)
        mine = got[offset : offset + want.size].reshape(want.shape)
        assert np.abs(mine - want).max() <= 1e-12 * max(1.0, np.abs(want).max()), i
        f = probe_h[offset : offset + want.size].reshape(want.shape)
        wg = np.zeros((1, sl.stop - sl.start))
        orc.gather(kind, pos[:, sl], f, wg, *g, degree=weighting.degree)
        np.testing.assert_allclose(res[:, sl], wg, rtol=1e-12, atol=1e-12)


def test_autotune_and_graph_replay_keep_the_result():
    batch, spec, pos, val = make_batch(seed=5, n_patches=60, dtype=np.float32)
    pos_d, val_d = torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda")
    batch.build(chunk=4096)
    eager = torch.zeros(batch.field_values, dtype=torch.float64, device="cuda")
    batch.scatter(pos_d, val_d, eager)
    best = P.autotune(batch, pos_d, val_d, eager, candidates=[(1024, 256, 512), (8192, 512, 1024)], repeats=1)
    assert best in [(1024, 256, 512), (8192, 512, 1024)]
    again = torch.zeros_like(eager)
    batch.scatter(pos_d, val_d, again)
    assert torch.equal(again, eager)  # the work split does not change a bit
    fields = torch.zeros_like(eager)
    graph = batch.capture(pos_d, val_d, fields)
    for k in range(3):
        fields.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(fields, eager), k
        _ = (fields * 2).sum().item()  # eager work and a sync between replays


def test_float32_patches_of_16_to_128_nodes_match_the_oracle():
    """The batched-patch workload's shape: 2D patches of 16 to 128 nodes per side, float32
    markers, spline degrees 0 to 3, periodic and bounded axes; each patch against the oracle
    fed the same float32 values (deposit in fixed point, gather in float32)."""
    batch, spec, pos, val = make_batch(seed=21, n_patches=48, ndim=2, dtype=np.float32, sides=(16, 128))
    sizes = [int(np.prod(sp[2])) for sp in spec]
    assert min(sizes) < 48 * 48 and max(sizes) > 100 * 100
    pos_d, val_d = torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda")
    batch.build(chunk=4096)
    fields = torch.zeros(batch.field_values, dtype=torch.float64, device="cuda")
    batch.scatter(pos_d, val_d, fields)
    probe = torch.as_tensor(np.random.default_rng(22).normal(size=batch.field_values), device="cuda")
    result = torch.empty((1, pos.shape[1]), dtype=torch.float32, device="cuda")
    batch.gather(pos_d, probe, result)
    got, res, probe_h = fields.cpu().numpy(), result.cpu().numpy().astype(np.float64), probe.cpu().numpy()
    pos64, val64 = pos.astype(np.float64), val.astype(np.float64)
    for i, (lower, cell, side, weighting, periodic, offset, sl) in enumerate(spec):
        want, g, kind = oracle_patch(weighting, pos64[:, sl], val64[:, sl], lower, cell, side, periodic)
        mine = got[offset : offset + want.size].reshape(want.shape)
        assert np.abs(mine - want).max() <= 1e-10 * max(1.0, np.abs(want).max()), i
        f = probe_h[offset : offset + want.size].reshape(want.shape)
        wg = np.zeros((1, sl.stop - sl.start))
        orc.gather(kind, pos64[:, sl], f, wg, *g, degree=weighting.degree)
        np.testing.assert_allclose(res[:, sl], wg, rtol=2e-6, atol=2e-6 * max(1.0, np.abs(wg).max()))


def test_sorted_patches_give_the_same_integers_and_values():
    """PatchBatch.sort orders each patch's markers by cell: the deposit of the sorted markers
    equals the deposit in the input order bit for bit (each contribution is rounded on its
    own), and the gather at the sorted markers is the input-order gather permuted."""
    batch, spec, pos, val = make_batch(seed=31, n_patches=48, ndim=2, dtype=np.float32, sides=(16, 128))
    pos_d, val_d = torch.as_tensor(pos, device="cuda"), torch.as_tensor(val, device="cuda")
    batch.build(chunk=4096)
    eager = torch.zeros(batch.field_values, dtype=torch.float64, device="cuda")
    batch.scatter(pos_d, val_d, eager)
    spos, sval, order = batch.sort(pos_d, val_d)
    assert torch.equal(spos, pos_d[:, order]) and torch.equal(sval, val_d[:, order])
    for lower, cell, side, weighting, periodic, offset, sl in spec:  # each patch's markers stay in its range
        part = order[sl.start:sl.stop]
        assert int(part.min()) >= sl.start and int(part.max()) < sl.stop
    fields = torch.zeros_like(eager)
    batch.scatter(spos, sval, fields)
    assert torch.equal(fields, eager)
    probe = torch.as_tensor(np.random.default_rng(32).normal(size=batch.field_values), device="cuda")
    want = torch.empty((1, pos.shape[1]), dtype=torch.float32, device="cuda")
    got = torch.empty_like(want)
    batch.gather(pos_d, probe, want)
    batch.gather(spos, probe, got)
    assert torch.equal(got, want[:, order])
