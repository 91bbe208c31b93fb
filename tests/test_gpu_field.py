"""The Field API on the GPU: the reference's unit tests of Field.gather / scatter /
deposit_density / node_volumes / block_grid / halos, run through mpifield_amd.Field
(tests/unit/test_weighting.py, test_density.py, test_field.py:530-577 of the reference)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    import mpifield_amd as m
    from mpifield_amd import Field


def _field(ndim, periodic=None, components=1, **kwargs):
    return Field(ndim=ndim, lower_bound=(0.0,) * ndim, upper_bound=(1.0,) * ndim,
                 num_gridpoints=(11, 9, 7)[:ndim], num_components=components, periodic=periodic, **kwargs)


@pytest.mark.parametrize("mode", ["fixed", "atomic"])
def test_known_answers(mode):
    with m.deposit_mode(mode):
        field = _field(1)
        dx = float(field.grid_spacing[0])
        field.scatter(np.array([[5 * dx]]), 1.0, m.QUADRATIC)
        field.scatter(np.array([[2.5 * dx]]), 8.0, m.QUADRATIC)
        np.testing.assert_allclose(field.to_numpy()[:, 0], [0, 0, 4, 4, 0.125, 0.75, 0.125, 0, 0, 0, 0], atol=1e-14)
        field = _field(1)
        field.scatter(np.array([[0.0]]), 1.0, m.QUADRATIC)
        np.testing.assert_allclose(field.to_numpy()[:3, 0], [0.875, 0.125, 0.0], atol=1e-14)
        field = _field(1, periodic=(True,))
        field.scatter(np.array([[0.0]]), 1.0, m.QUADRATIC)
        np.testing.assert_allclose(field.to_numpy()[[-1, 0, 1], 0], [0.125, 0.75, 0.125], atol=1e-14)
        field = _field(1)
        field.scatter(np.array([[1.2, -0.3]]), np.array([1.0, 2.0]))
        data = field.to_numpy()[:, 0]
        assert data[-1] == 1.0 and data[0] == 2.0 and data.sum() == 3.0
    f = Field(ndim=1, lower_bound=(0.0,), upper_bound=(1.0,), num_gridpoints=4, periodic=True)
    f.set_field_data(np.array([0.0, 1.0, 2.0, 3.0]))
    np.testing.assert_allclose(f.gather(np.array([[0.875, 1.125]])), [[1.5, 0.5]])


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("name", ["LINEAR", "QUADRATIC", "CUBIC"])
def test_constant_and_linear_fields_are_gathered_exactly_inside(ndim, name):
    """test_weighting.py:92-103."""
    weighting = getattr(m, name)
    field = _field(ndim)
    nodes = np.meshgrid(*(np.linspace(0.0, 1.0, n) for n in field.num_gridpoints), indexing="ij")
    field.set_field_data(2.0 + sum((d + 1) * x for d, x in enumerate(nodes)))
    points = np.random.default_rng(3).uniform(0.2, 0.8, (ndim, 300))
    gathered = field.gather(points, weighting)[0]
    np.testing.assert_allclose(gathered, 2.0 + sum((d + 1) * points[d] for d in range(ndim)), rtol=1e-12)


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_quadratic_deposit_keeps_all_the_charge(ndim, periodic):
    """test_weighting.py:106-118: also for markers at and beyond the edges of a bounded axis."""
    field = _field(ndim, (periodic,) * ndim, components=2)
    rng = np.random.default_rng(4)
    points = rng.uniform(-0.1, 1.1, (ndim, 500))
    points[:, :50] = rng.choice([0.0, 1.0], (ndim, 50))
    values = rng.normal(size=(2, 500))
    field.scatter(points, values, m.QUADRATIC)
    np.testing.assert_allclose(field.local_with_halos.reshape(-1, 2).sum(dim=0).cpu().numpy(), values.sum(axis=1),
                               rtol=1e-10)


@pytest.mark.parametrize("weighting", ["LINEAR", "QUADRATIC", "CUBIC"])
@pytest.mark.parametrize("periodic", [(False, False), (False, True), (True, True)])
def test_density_times_node_volumes_is_the_total(weighting, periodic):
    """test_density.py:22-32."""
    rng = np.random.default_rng(0)
    positions = rng.uniform(0.0, 1.0, (2, 5000))
    values = rng.normal(0.0, 1.0, (3, 5000))
    field = Field(ndim=2, lower_bound=(0, 0), upper_bound=(1, 1), num_gridpoints=(11, 9), num_components=3,
                  periodic=periodic)
    field.deposit_density(positions, values, getattr(m, weighting))
    totals = (field.to_numpy() * field.node_volumes().to_numpy()).sum(axis=(0, 1))
    np.testing.assert_allclose(totals, values.sum(axis=1), rtol=1e-12, atol=1e-12)


def test_density_adds_and_volumes_halve_on_bounded_edges():
    """test_density.py:46-70."""
    def f(periodic):
        return Field(ndim=2, lower_bound=(0, 0), upper_bound=(1, 1), num_gridpoints=(11, 9), periodic=periodic)

    field = f((True, True))
    field.set_field_data(2.0)
    field.deposit_density(np.full((2, 1), 0.5), 0.0)
    np.testing.assert_allclose(field.to_numpy(), 2.0)
    density = f((True, True))
    density.deposit_density(np.full((2, 1), 0.5), 1.0)
    field.deposit_density(np.full((2, 1), 0.5), 1.0)
    np.testing.assert_allclose(field.to_numpy(), 2.0 + density.to_numpy())
    volumes = f((False, True)).node_volumes().to_numpy()[..., 0]
    cell = 0.1 * (1.0 / 9)
    np.testing.assert_allclose(volumes[1:-1], cell)
    np.testing.assert_allclose(volumes[[0, -1]], cell / 2)
    assert volumes.sum() == pytest.approx(1.0)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("name,halo", [("LINEAR", 1), ("QUADRATIC", 2), ("CUBIC", 2)])
def test_halo_block_deposits_and_interpolates_as_the_whole_grid(ndim, name, halo):
    """test_field.py:541-577, with the seam folded and filled on the device."""
    weighting = getattr(m, name)
    shape = (7, 5, 4)[:ndim]
    periodic = (True,) + (False,) * (ndim - 1)
    kwargs = dict(ndim=ndim, lower_bound=(0.0,) * ndim, upper_bound=(1.0,) * ndim, num_gridpoints=shape,
                  periodic=periodic)
    bare = Field(**kwargs)
    padded = Field(**kwargs, num_ghost_points=halo)
    stored = padded.block_grid(cpu=True)
    assert tuple(stored.num_gridpoints) == tuple(n + 2 * halo for n in shape)
    assert stored.periodic == (False,) * ndim
    rng = np.random.default_rng(5)
    positions = rng.uniform(0.0, 1.0, (ndim, 400))
    positions[1:] *= 0.999
    values = rng.uniform(0.5, 1.5, 400)
    bare.scatter(positions, values, weighting)
    padded.scatter(positions, values, weighting)
    padded.exchange_halos()
    np.testing.assert_allclose(padded.local.cpu().numpy(), bare.local.cpu().numpy(), atol=1e-12)
    padded.fill_halos()
    np.testing.assert_allclose(padded.gather(positions, weighting), bare.gather(positions, weighting), atol=1e-12)


def test_block_grid_inside_the_domain_leaves_out_halos_beyond_bounded_edges():
    """test_weighting.py:176-187."""
    field = _field(2, periodic=(True, False), num_ghost_points=2)
    whole = field.block_grid(cpu=True)
    inside = field.block_grid(cpu=True, inside_domain=True)
    assert tuple(whole.num_gridpoints) == (15, 13)
    assert whole.index == (slice(0, 15), slice(0, 13))
    assert tuple(inside.num_gridpoints) == (15, 9)
    assert inside.index == (slice(0, 15), slice(2, 11))
    assert inside.geometry[1] == pytest.approx(0.0)
    assert inside.geometry[0] == pytest.approx(-2 * float(field.grid_spacing[0]))
    assert field.local_with_halos[inside.index].shape == (15, 9, 1)


def test_values_must_match_the_components_and_halo_must_be_wide_enough():
    field = _field(2, components=3)
    with pytest.raises(ValueError, match="Expected 3 components, got 2"):
        field.scatter(np.full((2, 4), 0.5), np.ones((2, 4)))
    field.scatter(np.full((2, 4), 0.5), 1.0)
    np.testing.assert_allclose(field.to_numpy().sum(axis=(0, 1)), [4.0, 4.0, 4.0])
    thin = _field(2, periodic=(True, True), num_ghost_points=1)
    with pytest.raises(ValueError, match="needs num_ghost_points >= 2"):
        thin.scatter(np.full((2, 4), 0.5), 1.0, m.QUADRATIC)


def test_edge_weighting_continuity_through_the_api():
    """EdgeWeighting.deposit_paths_block + divergence_block = change of the node charge."""
    grid = m.BlockGrid.from_geometry((0.0, -1.0), (0.25, 0.5), (9, 8), (True, False))
    rng = np.random.default_rng(1)
    start = rng.uniform((-0.5, -1.5), (2.5, 3.0), (300, 2)).T.copy()
    end = start + rng.normal(scale=0.5, size=start.shape)
    q = rng.normal(size=300)
    for edges_w, nodes_w in ((m.WHITNEY, m.LINEAR), (m.QUADRATIC_EDGES, m.QUADRATIC), (m.CUBIC_EDGES, m.CUBIC)):
        current = np.zeros((9, 8, 2))
        edges_w.deposit_paths_block(start, end, q, current, grid)
        before = np.zeros((9, 8, 1))
        after = np.zeros((9, 8, 1))
        nodes_w.scatter_block(start, q, before, grid)
        nodes_w.scatter_block(end, q, after, grid)
        div = edges_w.divergence_block(current, grid)
        np.testing.assert_allclose(div, (after - before)[..., 0], atol=1e-11)
    with pytest.raises(ValueError, match="an edge field on this grid is"):
        m.WHITNEY.deposit_paths_block(start, end, q, np.zeros((9, 8, 3)), grid)
    # the edge gather through the API, host arrays in and out
    field = np.zeros((2, 300))
    m.WHITNEY.gather_block(start, current, field, grid)
    assert np.isfinite(field).all() and np.abs(field).max() > 0
    with pytest.raises(ValueError, match="an edge field on this grid is"):
        m.WHITNEY.gather_block(start, np.zeros((9, 8, 3)), field, grid)
    no_gather = m.EdgeWeighting("paths only", gather={}, deposit_paths=dict(m.WHITNEY.deposit_paths))
    with pytest.raises(ValueError, match="has no 2D gather kernel"):
        no_gather.gather_block(start, np.zeros((9, 8, 2)), field, grid)
