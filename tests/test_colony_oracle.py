"""The numpy restatement of the reference's biomass deposition and colony
mobility (oracle/colony.py), pinned by the reference's own tests and by
closed forms of the tent kernel's voxel integrals."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.colony import (ColonyError, biomass_volume, colony_mobility, colony_volume_fraction,
                           deposition_weights)


def rods(positions, length=2.0, radius=0.5):
    p = np.asarray(positions, np.float64).reshape(-1, 3)
    n = len(p)
    return dict(position_x=p[:, 0], position_y=p[:, 1], position_z=p[:, 2],
                lengths=np.full(n, length), radii=np.full(n, radius))


def test_deposition_conserves_amount_across_refinement():
    """test_flow_foundations.py:75-88: 100 rods at (0.3, 0.1, -0.2) on 4^3,
    8^3 and 16^3 grids of a 16-unit box: the deposited amount is 100 biomass
    volumes, and the fine grids' voxel integrals aggregate exactly to the
    coarse grid's."""
    expected = 100 * biomass_volume(2.0, 0.5)
    densities = []
    for n in (4, 8, 16):
        h = 16 / n
        origin = (-8 + h / 2,) * 3
        density = colony_volume_fraction((n, n, n), origin, (h, h, h), None,
                                         rods([(0.3, 0.1, -0.2)] * 100), averaging_radius=4)
        assert math.isclose(float(density.sum()) * h ** 3, expected, rel_tol=2e-7)
        densities.append(density)
    fine_to_coarse = densities[2].reshape(4, 4, 4, 4, 4, 4).mean(axis=(1, 3, 5))
    assert float(np.max(np.abs(fine_to_coarse - densities[0]))) < 1e-10
    assert float(densities[2].max()) > 0.9


def test_tent_kernel_voxel_integrals_closed_form():
    """A rod at a voxel centre with radius equal to the spacing: the tent's
    integrals over the three voxels it covers are 1/8, 3/4, 1/8 per axis, and
    the 2-D deposit is their outer product (one layer)."""
    lows, w = deposition_weights((5, 5, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (2.0, 2.0, 0.0),
                                 1.0, None)
    axis = np.array([0.125, 0.75, 0.125])
    assert lows == (1, 1, 0)
    assert np.allclose(w[:, :, 0], np.outer(axis, axis), rtol=0, atol=1e-15)
    # A quarter spacing off centre in x: the scaled edges -1.75, -0.75, 0.25,
    # 1.25 give the tent's cumulative weights 0, 1/32, 23/32, 1.
    lows, w = deposition_weights((5, 5, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (2.25, 2.0, 0.0),
                                 1.0, None)
    assert lows == (1, 1, 0)
    assert np.allclose(w[:, :, 0], np.outer([1 / 32, 22 / 32, 9 / 32], axis), rtol=0, atol=1e-15)


def test_boundary_truncation_renormalises():
    """Near a grid edge the kernel loses support outside the grid and is
    renormalised: the deposit still sums to one biomass volume."""
    d = colony_volume_fraction((6, 6, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 2.0), None,
                               rods([(0.1, 5.2, 0.0)]), averaging_radius=2.0)
    assert math.isclose(float(d.sum()) * 2.0, biomass_volume(2.0, 0.5), rel_tol=1e-14)


def test_obstacle_restricts_to_the_connected_component():
    """A wall of solid voxels through the kernel's support: the deposit stays
    on the side of the strongest weight, none crosses the wall."""
    solid = np.zeros((9, 9, 1), bool)
    solid[5, :, 0] = True
    d = colony_volume_fraction((9, 9, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), solid,
                               rods([(4.2, 4.0, 0.0)]), averaging_radius=3.0)
    assert float(d[6:].sum()) == 0.0 and float(d[5].sum()) == 0.0
    assert math.isclose(float(d.sum()), biomass_volume(2.0, 0.5), rel_tol=1e-14)


def test_cells_outside_the_modelled_volume_are_omitted():
    d = colony_volume_fraction((3, 3, 1), (0.0, 0.0, 0.0), (4.0, 4.0, 4.0), None,
                               rods([(-10.0, 0.0, 0.0), (0.0, 10.0, 0.0)]))
    assert float(d.sum()) == 0.0


def test_colony_mobility_adds_drag_where_cells_pack():
    """test_flow.py:243-274: 40 rods packed on voxel (0,0), one on (1,1), one
    outside, spacing 4, a solid corner; rods of length 3, radius 0.5."""
    dims, spacing = (3, 3, 1), (4.0, 4.0, 4.0)
    obstacles = np.zeros(9, np.uint8)
    obstacles[(2 * 3 + 2)] = 1
    cells = rods([(0.0, 0.0, 0.0)] * 40 + [(4.0, 4.0, 0.0), (-10.0, 0.0, 0.0)], length=3.0)
    m = colony_mobility(dims, (0.0, 0.0, 0.0), spacing, obstacles, cells, base=1.0,
                        drag_coefficient=100.0)
    packed, sparse, empty = m[0], m[4], m[2]
    assert m[8] == 0.0
    assert packed < sparse < empty <= 1.0
    capped = 1.0 / (1.0 + 100.0 * 0.9 ** 2 / (1.0 - 0.9) ** 3)
    assert math.isclose(packed, capped, rel_tol=1e-9)
    layered = colony_mobility(dims, (0.0, 0.0, 0.0), spacing, obstacles, rods([]),
                              base=[0.5] * 9, drag_coefficient=100.0)
    assert layered[4] == 0.5 and layered[8] == 0.0


@pytest.mark.parametrize("kwargs,message", [
    (dict(base=0.0), "base mobility must be finite and positive"),
    (dict(base=[1.0]), "base mobility must hold one value per grid site"),
    (dict(base=[1.0] * 8 + [-1.0]), "base mobility values must be finite and non-negative"),
    (dict(max_volume_fraction=1.0), "maximum volume fraction must lie strictly between zero and one"),
    (dict(drag_coefficient=-1.0), "drag coefficient must be finite and non-negative"),
    (dict(averaging_radius=0.0), "averaging radius must be finite and positive"),
])
def test_colony_mobility_refusals(kwargs, message):
    """test_flow.py:269-274 and the checks of flow.py:282-382."""
    with pytest.raises(ColonyError) as info:
        colony_mobility((3, 3, 1), (0.0, 0.0, 0.0), (4.0, 4.0, 4.0), None, rods([]), **kwargs)
    assert str(info.value) == message


def test_deposition_refusals():
    with pytest.raises(ColonyError, match="deposited positions must be finite"):
        colony_volume_fraction((3, 3, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), None,
                               rods([(math.nan, 0.0, 0.0)]))
    solid = np.ones((3, 3, 1), bool)
    with pytest.raises(ColonyError, match="no connected fluid support"):
        colony_volume_fraction((3, 3, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), solid,
                               rods([(1.0, 1.0, 0.0)]))
    with pytest.raises(ColonyError, match="radius finite and positive"):
        colony_volume_fraction((3, 3, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), None,
                               rods([(1.0, 1.0, 0.0)], radius=0.0))


def test_flow_coupling_mobility_is_the_colony_mobility():
    """colony_flow_coupling's mobility is colony_mobility's for the same
    colony (flow.py:384-424 calls the same _darcy_coefficients)."""
    from oracle.colony import colony_flow_coupling
    rng = np.random.default_rng(2)
    solid = (rng.random((12, 9, 2)) < 0.1).astype(np.uint8).reshape(-1)
    cells = rods(np.column_stack([rng.uniform(0, 11, 30), rng.uniform(0, 8, 30),
                                  rng.uniform(0, 1, 30)]))
    velocities = rng.uniform(-1, 1, (30, 2))
    args = ((12, 9, 2), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), solid, cells)
    mobility, force = colony_flow_coupling(*args, velocities, base=1.5, averaging_radius=2.0)
    assert np.array_equal(mobility, colony_mobility(*args, base=1.5, averaging_radius=2.0))
    assert force.shape == (12 * 9, 2)


def test_a_common_velocity_gives_drag_times_velocity():
    """With every cell moving at one velocity V the column's momentum is its
    biomass times V, so the force is drag * V wherever a column holds
    biomass and zero elsewhere (flow.py:421-423)."""
    from oracle.colony import colony_flow_coupling
    rng = np.random.default_rng(3)
    dims = (10, 8, 1)
    cells = rods(np.column_stack([rng.uniform(2, 7, 25), rng.uniform(2, 5, 25), np.zeros(25)]))
    velocity = (0.75, -1.25)
    mobility, force = colony_flow_coupling(dims, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), None, cells,
                                           np.tile(velocity, (25, 1)), base=1.0,
                                           drag_coefficient=2.0, max_volume_fraction=0.5,
                                           averaging_radius=1.8)
    density = colony_volume_fraction(dims, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), None, cells,
                                     averaging_radius=1.8)
    fraction = np.minimum(density.sum(axis=2), 0.5).reshape(-1)
    drag = 2.0 * fraction ** 2 / (1.0 - fraction) ** 3
    has = density.sum(axis=2).reshape(-1) > 0
    v = np.array([float(np.float32(velocity[0])), float(np.float32(velocity[1]))])
    np.testing.assert_allclose(force[has], drag[has, None] * v, rtol=1e-12)
    assert np.all(force[~has] == 0.0)
    assert has.sum() > 10


def test_flow_coupling_refusals():
    from oracle.colony import colony_flow_coupling
    cells = rods([(1.0, 1.0, 0.0), (math.nan, 1.0, 0.0)])
    args = ((3, 3, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), None, cells)
    with pytest.raises(ColonyError, match="in-plane cell velocity must be finite"):
        colony_flow_coupling(*args, [(math.inf, 0.0), (0.0, 0.0)])
    # A cell's velocity is checked before its position (flow.py:409-412).
    with pytest.raises(ColonyError, match="deposited positions must be finite"):
        colony_flow_coupling(*args, [(0.0, 0.0), (0.0, 0.0)])
    with pytest.raises(ColonyError, match="in-plane cell velocity must be finite"):
        colony_flow_coupling(*args, [(0.0, 0.0), (math.nan, 0.0)])
    with pytest.raises(ColonyError, match="averaging radius"):
        colony_flow_coupling(*args, [(0.0, 0.0)] * 2, averaging_radius=0.0)


def test_relative_drag_in_a_uniform_duct_from_the_colony(oracle):
    """The reference's own test of colony_flow_coupling
    (python/tests/test_flow.py:192-233): one rod of length 3, radius 0.5 at
    (3, 4, 0) in a 7 x 9 duct along y, base 1, drag coefficient 2, cap 0.5,
    radius 1.8; mobility and force from the coupling, the flow from the
    oracle's forced depth-averaged solve (bit-exact with the reference's)."""
    from oracle import FIXED, NO_FLUX, Grid
    from oracle.colony import colony_flow_coupling
    nx, ny = 7, 9
    grid = Grid(nx, ny, 1, 1.0, 1.0, 1.0, None,
                boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    cells = rods([(3.0, 4.0, 0.0)], length=3.0)

    def solve(vx, vy, inlet):
        mobility, force = colony_flow_coupling((nx, ny, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                                               None, cells, [(vx, vy)], base=1.0,
                                               drag_coefficient=2.0, max_volume_fraction=0.5,
                                               averaging_radius=1.8)
        faces, _, _ = oracle.solve_flow(grid, mobility.astype(np.float32), 1, inlet, 1e-6,
                                        50000, body_force=force.astype(np.float32))
        x = faces[:(nx + 1) * ny].reshape(nx + 1, ny)
        y = faces[(nx + 1) * ny:(nx + 1) * ny + nx * (ny + 1)].reshape(nx, ny + 1)
        np.testing.assert_allclose(y.mean(axis=0), inlet, atol=2e-5)
        np.testing.assert_allclose(np.diff(x, axis=0) + np.diff(y, axis=1), 0, atol=5e-5)
        assert np.all(x[[0, -1], :] == 0.0)
        return x, y

    _, slow = solve(0, 0, 2)
    assert slow[3, 4] < 1.5 and slow[1, 4] > 2
    x, matched = solve(0, 2, 2)
    np.testing.assert_allclose(matched, 2, atol=4e-5)
    np.testing.assert_allclose(x, 0, atol=4e-5)
    _, fast = solve(0, 4, 2)
    np.testing.assert_allclose(fast - matched, matched - slow, atol=5e-5)
    x, y = solve(3, 0, 0)
    assert x[3, 4] > 1
    reverse_x, reverse_y = solve(-3, 0, 0)
    np.testing.assert_allclose(reverse_x, -x, atol=1e-5)
    np.testing.assert_allclose(reverse_y, -y, atol=1e-5)
