"""Differentiation through gather and scatter (torch.autograd with the adjoint kernels):
gradients against float64 central finite differences, and the adjoint identity
<scatter(w), f> = <w, gather(f)> at 1e-12."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    import mpifield_amd as m
    from mpifield_amd import autograd as AG
    from mpifield_amd import kernels as K
    from mpifield_amd.binning import kind_of

WEIGHTINGS = ["LINEAR", "QUADRATIC", "CUBIC"]


def gaussians(nodes: int, cell: float, seed: int = 5) -> torch.Tensor:
    """A periodic field: the sum of 8 Gaussians with random centres, widths and amplitudes."""
    rng = np.random.default_rng(seed)
    x = np.arange(nodes) * cell
    length = nodes * cell
    grid = np.stack(np.meshgrid(x, x, x, indexing="ij"))
    f = np.zeros((nodes,) * 3)
    for _ in range(8):
        c = rng.uniform(0, length, 3)[:, None, None, None]
        sigma = rng.uniform(0.05, 0.2) * length
        d = (grid - c + length / 2) % length - length / 2
        f += rng.normal() * np.exp(-(d**2).sum(axis=0) / (2 * sigma**2))
    return torch.as_tensor(f[..., None], device="cuda")


def setup(weighting_name, n=3000, nodes=16, periodic=True, seed=0):
    weighting = getattr(m, weighting_name)
    cell = 1.0 / nodes
    grid = m.BlockGrid.from_geometry([0.0] * 3, [cell] * 3, [nodes] * 3, [periodic] * 3)
    rng = np.random.default_rng(seed)
    pos = torch.as_tensor(rng.uniform(0.05, 0.95, (3, n)), device="cuda")
    return weighting, grid, pos, cell


@pytest.mark.parametrize("name", WEIGHTINGS)
def test_gather_gradients_match_finite_differences(name):
    weighting, grid, pos, cell = setup(name)
    field = gaussians(16, cell)
    a = torch.as_tensor(np.random.default_rng(1).normal(size=(1, pos.shape[1])), device="cuda")
    x = pos.clone().requires_grad_(True)
    f = field.clone().requires_grad_(True)
    loss = (AG.gather(x, f, weighting, grid) * a).sum()
    loss.backward()

    def L(xx, ff):
        out = torch.empty((1, xx.shape[1]), dtype=torch.float64, device="cuda")
        K.gather(*kind_of(weighting), xx, ff, out, grid.geometry)
        return float((out * a).sum())

    h = 1e-6 * cell
    for mk in range(0, pos.shape[1], 500):
        for d in range(3):
            xp, xm = pos.clone(), pos.clone()
            xp[d, mk] += h
            xm[d, mk] -= h
            fd = (L(xp, field) - L(xm, field)) / (2 * h)
            assert abs(x.grad[d, mk].item() - fd) <= 1e-5 * max(1.0, abs(fd)), (d, mk, x.grad[d, mk].item(), fd)
    # the field gradient is the deposit of a: the loss is linear in the field
    rng = np.random.default_rng(3)
    for _ in range(16):
        idx = tuple(int(i) for i in rng.integers(0, 16, 3))
        fp, fm = field.clone(), field.clone()
        fp[idx + (0,)] += 1.0
        fm[idx + (0,)] -= 1.0
        fd = (L(pos, fp) - L(pos, fm)) / 2.0
        assert abs(f.grad[idx + (0,)].item() - fd) <= 1e-10 * max(1.0, abs(fd))


@pytest.mark.parametrize("name", WEIGHTINGS)
def test_scatter_gradients_and_adjoint(name):
    weighting, grid, pos, cell = setup(name, seed=2)
    probe = gaussians(16, cell, seed=7)
    v = torch.as_tensor(np.random.default_rng(4).normal(1.0, 0.1, (1, pos.shape[1])), device="cuda")
    with K.deposit_mode("atomic"):
        x = pos.clone().requires_grad_(True)
        vv = v.clone().requires_grad_(True)
        rho = AG.scatter(x, vv, weighting, grid)
        loss = (rho * probe).sum()
        loss.backward()
        # adjoint: <scatter(v), probe> = <v, gather(probe)>
        g = AG.gather(pos, probe, weighting, grid)
        assert abs(loss.item() - float((g * v).sum())) <= 1e-12 * abs(loss.item())
        np.testing.assert_allclose(vv.grad.cpu().numpy(), g.detach().cpu().numpy(), rtol=1e-12, atol=1e-14)

        def L(xx):
            r = AG.scatter(xx, v, weighting, grid)
            return float((r * probe).sum())

        h = 1e-6 * cell
        for mk in range(0, pos.shape[1], 600):
            for d in range(3):
                xp, xm = pos.clone(), pos.clone()
                xp[d, mk] += h
                xm[d, mk] -= h
                fd = (L(xp) - L(xm)) / (2 * h)
                assert abs(x.grad[d, mk].item() - fd) <= 1e-5 * max(1.0, abs(fd)), (d, mk)


def test_cubic_config_shape_runs_in_float32():
    """The differentiable cubic gather-of-scatter of 16M float32 markers on 128^3 (one component)."""
    nodes, n = 128, 1 << 24
    cell = 1.0 / nodes
    grid = m.BlockGrid.from_geometry([0.0] * 3, [cell] * 3, [nodes] * 3, [True] * 3)
    gen = torch.Generator(device="cuda").manual_seed(5)
    pos = torch.rand((3, n), generator=gen, device="cuda", dtype=torch.float32).requires_grad_(True)
    w = torch.ones((1, n), device="cuda", dtype=torch.float32).requires_grad_(True)
    field = gaussians(nodes, cell).requires_grad_(True)
    rho = AG.scatter(pos, w, m.CUBIC, grid)
    g = AG.gather(pos, rho + field, m.CUBIC, grid)
    loss = g.double().sum()
    loss.backward()
    assert torch.isfinite(pos.grad).all() and torch.isfinite(w.grad).all() and torch.isfinite(field.grad).all()
    # d loss / d field = scatter of ones: the total charge
    assert abs(field.grad.sum().item() - n) <= 1e-6 * n


TILED_CASES = [("LINEAR", 1), ("LINEAR", 3), ("QUADRATIC", 1), ("QUADRATIC", 3), ("CUBIC", 1)]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("name,ncomp", TILED_CASES)
def test_binned_gradient_matches_the_direct_derivative(name, ncomp, periodic, dtype):
    """The tiled gradient kernel on the bins equals the direct derivative gather contracted
    with the same coefficients (float64: to rounding; float32: to float32 rounding)."""
    from mpifield_amd import binning

    print(f"binned gradient {name} C={ncomp} periodic={periodic} {dtype}", flush=True)
    weighting = getattr(m, name)
    kind, degree = kind_of(weighting)
    num = (32, 16, 48)
    cell = (1 / 32, 1 / 16, 1 / 48)
    grid = m.BlockGrid.from_geometry([0.0] * 3, list(cell), list(num), [periodic] * 3)
    g = tuple(grid.geometry)
    rng = np.random.default_rng(11)
    n = 50000
    pos = torch.as_tensor(rng.uniform(-0.1, 1.1, (3, n)), device="cuda", dtype=dtype)
    field = torch.as_tensor(rng.normal(size=(*num, ncomp)), device="cuda")
    coef = torch.as_tensor(rng.normal(size=(ncomp, n)), device="cuda", dtype=dtype)
    bins = binning.bin_markers(pos, None, (kind, degree), g)
    got = binning.gather_derivative_binned(bins, field, coef)
    d = AG.gather_derivative(kind, degree, pos.double(), field, g)
    want = (d * coef.double().unsqueeze(0)).sum(dim=1)
    scale = want.abs().max().item()
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    assert (got.double() - want).abs().max().item() <= tol * scale


@pytest.mark.parametrize("mode", ["fixed", "atomic"])
def test_shared_bins_give_the_gradients_of_separate_calls(mode):
    """gather(scatter) above the binning threshold with the positions binned once
    (bin_positions, passed as bins=) gives the values and gradients of the same calls each
    binning by itself."""
    nodes, n = 32, 300000
    cell = 1.0 / nodes
    grid = m.BlockGrid.from_geometry([0.0] * 3, [cell] * 3, [nodes] * 3, [True] * 3)
    rng = np.random.default_rng(21)
    pos0 = torch.as_tensor(rng.uniform(0, 1, (3, n)), device="cuda")
    w0 = torch.as_tensor(rng.normal(1.0, 0.2, (1, n)), device="cuda")
    field0 = gaussians(nodes, cell)
    a = torch.as_tensor(rng.normal(size=(1, n)), device="cuda")

    def run(shared):
        pos, w, field = (t.clone().requires_grad_(True) for t in (pos0, w0, field0))
        bins = AG.bin_positions(pos, m.CUBIC, grid) if shared else None
        assert (bins is not None) == shared
        with K.deposit_mode(mode):
            rho = AG.scatter(pos, w, m.CUBIC, grid, bins=bins)
            g = AG.gather(pos, rho + field, m.CUBIC, grid, bins=bins)
            (g * a).sum().backward()
        return g.detach(), pos.grad, w.grad, field.grad

    shared, separate = run(True), run(False)
    for got, want in zip(shared, separate):
        scale = want.abs().max().item()
        assert (got - want).abs().max().item() <= 1e-9 * scale


def test_positions_written_in_place_are_binned_afresh():
    """Nothing is reused implicitly: after the positions change through .data (which bumps no
    version counter), the next gather follows the new positions."""
    nodes, n = 32, 300000
    cell = 1.0 / nodes
    grid = m.BlockGrid.from_geometry([0.0] * 3, [cell] * 3, [nodes] * 3, [True] * 3)
    rng = np.random.default_rng(22)
    pos = torch.as_tensor(rng.uniform(0, 1, (3, n)), device="cuda")
    field = gaussians(nodes, cell)
    first = AG.gather(pos, field, m.QUADRATIC, grid)
    pos.data.add_(0.37 * cell)
    moved = AG.gather(pos, field, m.QUADRATIC, grid)
    want = torch.empty_like(moved)
    K.gather(1, 0, pos.detach().clone(), field, want, tuple(grid.geometry))
    assert torch.equal(moved, want)
    assert not torch.equal(moved, first)
    with pytest.raises(ValueError, match="another weighting"):
        AG.gather(pos, field, m.LINEAR, grid, bins=AG.bin_positions(pos, m.QUADRATIC, grid))


@pytest.mark.parametrize("degree", [2, 3])
def test_clamped_scatter_is_differentiable_in_its_values(degree):
    """autograd.scatter on clamped axes sizes the field num + degree - 1 per clamped axis: the
    deposit equals the oracle's clamped spline deposit, and the gradient with respect to the
    values is the clamped gather of the upstream gradient (the adjoint)."""
    from oracle import oracle as orc

    shape = m.splines(degree, clamped=True).nodes
    num = (9, 8, 7)
    grid = m.BlockGrid.from_geometry([0.0] * 3, [0.5, 0.4, 0.3], list(num), [False, True, False],
                                     [True, False, True])
    g = tuple(grid.geometry)
    extent = (num[0] + degree - 1, num[1], num[2] + degree - 1)
    rng = np.random.default_rng(degree)
    upper = np.array([0.5, 0.4, 0.3]) * (np.array(num) - 1)
    pos = np.ascontiguousarray(rng.uniform(-0.1, upper + 0.1, (500, 3)).T)
    vals = rng.normal(size=(1, 500))
    w = torch.as_tensor(vals, device="cuda").requires_grad_(True)
    rho = AG.scatter(torch.as_tensor(pos, device="cuda"), w, shape, grid)
    assert tuple(rho.shape) == (*extent, 1)
    want = np.zeros((*extent, 1))
    orc.scatter("spline", pos, vals, want, *g, degree=degree)
    np.testing.assert_allclose(rho.detach().cpu().numpy(), want, rtol=0, atol=1e-10 * np.abs(want).max())
    up = rng.normal(size=(*extent, 1))
    (rho * torch.as_tensor(up, device="cuda")).sum().backward()
    adj = np.zeros((1, 500))
    orc.gather("spline", pos, up, adj, *g, degree=degree)
    np.testing.assert_allclose(w.grad.cpu().numpy(), adj, rtol=1e-12, atol=1e-12 * np.abs(adj).max())
