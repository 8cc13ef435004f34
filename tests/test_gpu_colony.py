"""Biomass deposition and colony mobility on the GPU through the C ABI
(msim_hip_colony_volume_fraction, msim_hip_colony_mobility), against the
numpy restatement of the reference's Python functions (oracle/colony.py).

Tolerance. The device adds each voxel's deposits with FP64 atomics, in an
order that differs from the reference's cell order, and its kernel total and
pow are its own FP64 roundings: fields agree to 1e-12 relative to the field's
largest value. The reference tests' own bounds (conservation 2e-7, the
fine-to-coarse aggregate 1e-10, the capped mobility 1e-9) hold as they are.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import colony as ref
from tests.helpers import random_cells

pytestmark = pytest.mark.gpu

NO_FLUX = 0


def rods(positions, length=2.0, radius=0.5):
    p = np.asarray(positions, np.float32).reshape(-1, 3)
    n = len(p)
    return dict(position_x=p[:, 0].copy(), position_y=p[:, 1].copy(),
                position_z=p[:, 2].copy(), lengths=np.full(n, length, np.float32),
                radii=np.full(n, radius, np.float32))


def close(ours, theirs, rel=1e-12):
    scale = max(float(np.max(np.abs(theirs))), 1e-300)
    return float(np.max(np.abs(np.asarray(ours) - np.asarray(theirs)))) <= rel * scale


def test_deposition_conserves_amount_across_refinement(hip):
    """test_flow_foundations.py:75-88 on the device (100 rods at (0.3, 0.1,
    -0.2), 4^3, 8^3 and 16^3 grids of a 16-unit box), and each field against
    the restatement."""
    expected = 100 * ref.biomass_volume(2.0, np.float32(0.5))
    position = [float(np.float32(v)) for v in (0.3, 0.1, -0.2)]
    densities = []
    for n in (4, 8, 16):
        h = 16 / n
        origin = (-8 + h / 2,) * 3
        grid = msa.Grid(n, n, n, h, h, h, None, (NO_FLUX,) * 6)
        cells = rods([position] * 100)
        density = hip.colony_volume_fraction(grid, cells, averaging_radius=4, origin=origin)
        assert math.isclose(float(density.sum()) * h ** 3, expected, rel_tol=2e-7)
        assert close(density, ref.colony_volume_fraction((n, n, n), origin, (h, h, h), None,
                                                         cells, averaging_radius=4))
        densities.append(density)
    fine_to_coarse = densities[2].reshape(4, 4, 4, 4, 4, 4).mean(axis=(1, 3, 5))
    assert float(np.max(np.abs(fine_to_coarse - densities[0]))) < 1e-10
    assert float(densities[2].max()) > 0.9


def test_colony_mobility_adds_drag_where_cells_pack(hip):
    """test_flow.py:243-274 on the device."""
    grid = msa.Grid(3, 3, 1, 4.0, 4.0, 4.0, np.array([0] * 8 + [1], np.uint8), (NO_FLUX,) * 6)
    cells = rods([(0.0, 0.0, 0.0)] * 40 + [(4.0, 4.0, 0.0), (-10.0, 0.0, 0.0)], length=3.0)
    m = hip.colony_mobility(grid, cells, base=1.0, drag_coefficient=100.0)
    packed, sparse, empty = m[0], m[4], m[2]
    assert m[8] == 0.0
    assert packed < sparse < empty <= 1.0
    capped = 1.0 / (1.0 + 100.0 * 0.9 ** 2 / (1.0 - 0.9) ** 3)
    assert math.isclose(packed, capped, rel_tol=1e-9)
    assert close(m, ref.colony_mobility((3, 3, 1), (0.0, 0.0, 0.0), (4.0, 4.0, 4.0),
                                        np.array([0] * 8 + [1]), cells, base=1.0))
    layered = hip.colony_mobility(grid, rods([]), base=[0.5] * 9, drag_coefficient=100.0)
    assert layered[4] == 0.5 and layered[8] == 0.0


@pytest.mark.parametrize("radius", [1.5, 4.0, 7.0])
def test_random_colonies_on_an_obstacle_grid_match(hip, radius):
    """2000 cells on a 96x64 grid with 15% solid sites (ragged connectivity
    inside many supports) and on three layers of it, deposit and mobility."""
    rng = np.random.default_rng(int(radius * 10))
    nx, ny = 96, 64
    solid = (rng.random(nx * ny) < 0.15).astype(np.uint8)
    cells, _ = random_cells(int(radius), nx, ny, solid, 2000, 1.0, 1.0)
    cells = {k: cells[k] for k in ("position_x", "position_y", "position_z", "lengths", "radii")}
    origin = (0.0, 0.0, 0.0)
    grid = msa.Grid(nx, ny, 1, 1.0, 1.0, 2.0, solid, (NO_FLUX,) * 6)
    ours = hip.colony_volume_fraction(grid, cells, averaging_radius=radius)
    theirs = ref.colony_volume_fraction((nx, ny, 1), origin, (1.0, 1.0, 2.0), solid, cells,
                                        averaging_radius=radius)
    assert close(ours, theirs)
    base = rng.uniform(0.5, 2.0, nx * ny)
    ours = hip.colony_mobility(grid, cells, base=base, averaging_radius=radius)
    theirs = ref.colony_mobility((nx, ny, 1), origin, (1.0, 1.0, 2.0), solid, cells, base=base,
                                 averaging_radius=radius)
    assert close(ours, theirs)
    # Three layers: cells at mid-depth, a solid floor under part of the grid.
    layered = np.repeat(solid.reshape(nx, ny, 1), 3, axis=2)
    layered[: nx // 3, :, 0] = 1
    cells3 = dict(cells, position_z=np.full(2000, 1.0, np.float32))
    grid3 = msa.Grid(nx, ny, 3, 1.0, 1.0, 1.0, layered.reshape(-1), (NO_FLUX,) * 6)
    ours = hip.colony_mobility(grid3, cells3, base=1.0, averaging_radius=radius)
    theirs = ref.colony_mobility((nx, ny, 3), origin, (1.0, 1.0, 1.0), layered.reshape(-1),
                                 cells3, base=1.0, averaging_radius=radius)
    assert close(ours, theirs)


def test_a_large_colony_conserves_its_biomass(hip):
    """2^18 cells on the 1024x512 pillar array: the deposit holds every inside
    cell's biomass volume to FP64 rounding; a sample of 3000 cells against the
    restatement."""
    from microsimulator_amd.workloads import colony, config
    w = config("c2")
    cells, _ = colony(w, 1 << 18)
    grid = msa.Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    density = hip.colony_volume_fraction(grid, cells)
    r = cells["radii"].astype(np.float64)
    total = float(np.sum(np.pi * r * r * (cells["lengths"].astype(np.float64) + 2 * r)))
    voxel = float(np.float32(np.float32(w.dx) * np.float32(w.dy)) * np.float32(w.dz))
    assert math.isclose(float(density.sum()) * voxel, total, rel_tol=1e-11)
    sample = {k: v[:3000] for k, v in cells.items()}
    ours = hip.colony_volume_fraction(grid, sample)
    theirs = ref.colony_volume_fraction((w.nx, w.ny, 1), (0.0, 0.0, 0.0), (w.dx, w.dy, w.dz),
                                        w.obstacles, sample)
    assert close(ours, theirs)


REFUSALS = [
    ("radius", dict(averaging_radius=0.0), "averaging radius must be finite and positive"),
    ("position", dict(position=(math.nan, 1.0, 0.0)), "deposited positions must be finite"),
    ("size", dict(radius=0.0), "length must be finite and nonnegative; radius finite and positive"),
    ("support", dict(walled=True), "biomass deposition has no connected fluid support"),
    ("cap", dict(max_volume_fraction=1.0),
     "maximum volume fraction must lie strictly between zero and one"),
    ("base", dict(base=0.0), "base mobility must be finite and positive"),
    ("base count", dict(base=[1.0]), "base mobility must hold one value per grid site"),
    ("base value", dict(base=[1.0] * 24 + [-1.0]),
     "base mobility values must be finite and non-negative"),
    ("drag", dict(drag_coefficient=-1.0), "drag coefficient must be finite and non-negative"),
    ("support before cap", dict(walled=True, max_volume_fraction=1.0),
     "biomass deposition has no connected fluid support"),
]


@pytest.mark.parametrize("name,change,message", REFUSALS, ids=[r[0] for r in REFUSALS])
def test_refusals_match_the_restatement(hip, name, change, message):
    solid = np.zeros((5, 5, 1), np.uint8)
    if change.get("walled"):
        solid[:] = 1
        solid[4, 4, 0] = 0  # fluid, but outside the first cell's support
    grid = msa.Grid(5, 5, 1, 1.0, 1.0, 1.0, solid.reshape(-1), (NO_FLUX,) * 6)
    cells = rods([change.get("position", (1.0, 1.0, 0.0)), (3.0, 3.0, 0.0)],
                 radius=change.get("radius", 0.5))
    kw = {k: change[k] for k in ("base", "drag_coefficient", "max_volume_fraction",
                                 "averaging_radius") if k in change}
    if "averaging_radius" not in kw:
        kw["averaging_radius"] = 1.0
    with pytest.raises(ref.ColonyError) as theirs:
        ref.colony_mobility((5, 5, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), solid.reshape(-1),
                            cells, **kw)
    with pytest.raises(msa.MsimError) as ours:
        hip.colony_mobility(grid, cells, **kw)
    assert str(theirs.value) == message
    assert str(ours.value) == message


@pytest.mark.parametrize("layers", [1, 3])
def test_flow_coupling_matches_the_restatement(hip, layers):
    """colony_flow_coupling on the device: 1500 cells with random in-plane
    velocities on a 64x48 obstacle grid (and three layers of it), mobility
    and body force against the restatement to 1e-12 of the largest value."""
    rng = np.random.default_rng(10 + layers)
    nx, ny = 64, 48
    solid = (rng.random(nx * ny) < 0.12).astype(np.uint8)
    cells, _ = random_cells(3, nx, ny, solid, 1500, 1.0, 1.0)
    cells = {k: cells[k] for k in ("position_x", "position_y", "position_z", "lengths", "radii")}
    velocities = rng.uniform(-2.0, 2.0, (1500, 2)).astype(np.float32)
    mask = np.repeat(solid.reshape(nx, ny, 1), layers, axis=2).reshape(-1)
    if layers > 1:
        cells = dict(cells, position_z=np.full(1500, 1.0, np.float32))
    grid = msa.Grid(nx, ny, layers, 1.0, 1.0, 1.0, mask, (NO_FLUX,) * 6)
    base = rng.uniform(0.5, 2.0, nx * ny * layers)
    mobility, force = hip.colony_flow_coupling(grid, cells, velocities, base=base,
                                               drag_coefficient=50.0, averaging_radius=2.5)
    m_ref, f_ref = ref.colony_flow_coupling((nx, ny, layers), (0.0, 0.0, 0.0),
                                            (1.0, 1.0, 1.0), mask, cells, velocities, base=base,
                                            drag_coefficient=50.0, averaging_radius=2.5)
    assert close(mobility, m_ref)
    assert close(force, f_ref)
    assert np.count_nonzero(force) > 100


def test_relative_drag_in_a_uniform_duct_from_the_colony(hip):
    """The reference's own test of colony_flow_coupling with the body-forced
    flow (python/tests/test_flow.py:192-233), every step on the device: one
    rod (length 3, radius 0.5) at (3, 4, 0) in a 7 x 9 duct along y, base 1,
    drag coefficient 2, cap 0.5, radius 1.8; the reference's assertions."""
    from tests.helpers import FIXED
    nx, ny = 7, 9
    grid = msa.Grid(nx, ny, 1, 1.0, 1.0, 1.0, None,
                    (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    cells = rods([(3.0, 4.0, 0.0)], length=3.0)

    def solve(vx, vy, inlet):
        mobility, force = hip.colony_flow_coupling(grid, cells, [(vx, vy)], base=1.0,
                                                   drag_coefficient=2.0, max_volume_fraction=0.5,
                                                   averaging_radius=1.8)
        p = msa.DepthAveragedFlowParameters(mean_inlet_speed=inlet, axis=1,
                                            body_force=force.astype(np.float32))
        field = hip.solve_depth_averaged_flow(grid, mobility.astype(np.float32), p).field
        x = field.x_faces.reshape(nx + 1, ny)
        y = field.y_faces.reshape(nx, ny + 1)
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


@pytest.mark.parametrize("velocities,message", [
    ([(math.inf, 0.0), (0.0, 0.0)], "in-plane cell velocity must be finite"),
    ([(0.0, 0.0), (0.0, math.nan)], "in-plane cell velocity must be finite"),
])
def test_flow_coupling_refusals(hip, velocities, message):
    grid = msa.Grid(5, 5, 1, 1.0, 1.0, 1.0, None, (NO_FLUX,) * 6)
    cells = rods([(1.0, 1.0, 0.0), (3.0, 3.0, 0.0)])
    with pytest.raises(ref.ColonyError) as theirs:
        ref.colony_flow_coupling((5, 5, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), None, cells,
                                 velocities, averaging_radius=1.0)
    with pytest.raises(msa.MsimError) as ours:
        hip.colony_flow_coupling(grid, cells, velocities, averaging_radius=1.0)
    assert str(theirs.value) == message
    assert str(ours.value) == message
    # The context still deposits after a refusal.
    hip.colony_flow_coupling(grid, cells, [(0.0, 0.0)] * 2, averaging_radius=1.0)
