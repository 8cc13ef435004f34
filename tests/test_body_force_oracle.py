"""The oracle's depth-averaged solve with a body force
(DepthAveragedFlowParameters::body_force, cpp/include/microsimulator/flow.hpp:24-25;
solve_shallow_axis_flow, cpp/core/flow_system.hpp:1420-1447; body_flux, 395-420;
scale_velocity's driven term, 1368-1415), pinned three ways:
- bit for bit against the reference's own solve_depth_averaged_flow_cpu
  (oracle/_ref, live) on random grids of one and several layers;
- the reference's refusals and messages;
- the physics its own test states (python/tests/test_flow.py:192-233): with
  drag d, mobility b/(1 + b d) and force d V, a medium moving at the plug
  speed leaves the plug flow untouched; the field is affine in V; reversing V
  with no imposed throughput reverses the field; every field is
  divergence-free.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import Grid, OracleError
from tests.helpers import FIXED, LAYERED_CASES, NO_FLUX, layered_case, random_case


def random_force(seed, columns, scale=0.5):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, (columns, 2)).astype(np.float32)


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_forced_solve_matches_reference_live(oracle, reference, axis, seed):
    nx, ny = (40, 23) if axis == 0 else (23, 40)
    solid, mobility = random_case(seed, nx, ny)
    s = solid.reshape(nx, ny)
    if axis == 0:
        s[0, :] = s[-1, :] = 0
    else:
        s[:, 0] = s[:, -1] = 0
    solid = s.reshape(-1)
    kinds = (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX) if axis == 0 else \
        (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX)
    grid = Grid(nx, ny, 1, 0.7, 1.3, 2.0, solid, boundary=kinds)
    force = random_force(seed, nx * ny)
    for speed in (3.0, 0.0, -1.5):
        f_o, _, rep = oracle.solve_flow(grid, mobility, axis, speed, 2e-7, 20000,
                                        body_force=force)
        f_r, rr = reference.solve_flow(grid, mobility, axis, speed, 2e-7, 20000,
                                       body_force=force)
        assert rep.iterations == rr["iterations"], speed
        assert rep.max_speed == rr["max_speed"]
        assert np.array_equal(f_o, f_r), speed


@pytest.mark.parametrize("axis", [0, 1])
def test_forced_layered_solve_matches_reference_live(oracle, reference, axis):
    """Several layers: one force per column; the driven flux is lifted like
    the pressure-driven one (vertical faces by continuity)."""
    for nx, ny, nz, seed, mob in LAYERED_CASES:
        obstacles, mobility, boundary = layered_case(nx, ny, nz, seed, mob, axis)
        grid = Grid(nx, ny, nz, 1.0, 0.5, 0.7, obstacles, boundary=boundary)
        force = random_force(seed + 10, nx * ny, 2.0)
        a, _, rep = oracle.solve_flow(grid, mobility, axis, 2.0, 1e-6, 50000, body_force=force)
        b, info = reference.solve_flow(grid, mobility, axis, 2.0, 1e-6, 50000, body_force=force)
        assert np.array_equal(a, b), (nx, ny, nz, axis)
        assert rep.iterations == info["iterations"] and rep.max_speed == info["max_speed"]


def test_forced_solve_refusals_match_reference(oracle, reference):
    nx, ny = 6, 5
    grid = Grid(nx, ny, 1, 1.0, 1.0, 1.0, None)
    cases = [
        (1.0, np.zeros((nx * ny - 1, 2), np.float32), "one in-plane body force per column"),
        (1.0, np.array([[np.nan, 0.0]] * (nx * ny), np.float32), "body force must be finite"),
        (0.0, None, "nonzero without a body force"),
    ]
    for speed, force, message in cases:
        with pytest.raises(OracleError, match=message):
            oracle.solve_flow(grid, None, 0, speed, 1e-6, 1000, body_force=force)
        with pytest.raises(OracleError, match=message):
            reference.solve_flow(grid, None, 0, speed, 1e-6, 1000, body_force=force)
    # A body force makes a zero throughput valid.
    oracle.solve_flow(grid, None, 0, 0.0, 1e-6, 1000,
                      body_force=np.ones((nx * ny, 2), np.float32))


def _duct_solve(oracle, drag, vx, vy, inlet):
    """The reference test's 7 x 9 duct along y with relative drag: mobility
    1 / (1 + d) and force d V per column."""
    nx, ny = 7, 9
    grid = Grid(nx, ny, 1, 1.0, 1.0, 1.0, None,
                boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    mobility = (1.0 / (1.0 + drag)).astype(np.float32)
    force = np.stack([drag * vx, drag * vy], axis=1).astype(np.float32)
    faces, _, _ = oracle.solve_flow(grid, mobility, 1, inlet, 1e-10, 10000, body_force=force)
    x = faces[:(nx + 1) * ny].reshape(nx + 1, ny)
    y = faces[(nx + 1) * ny:(nx + 1) * ny + nx * (ny + 1)].reshape(nx, ny + 1)
    assert np.all(x[[0, -1], :] == 0.0)
    np.testing.assert_allclose(y[:, 0].mean(), inlet, atol=2e-5)
    np.testing.assert_allclose(np.diff(x, axis=0) + np.diff(y, axis=1), 0, atol=5e-5)
    return x, y


def test_relative_drag_in_a_uniform_duct(oracle):
    rng = np.random.default_rng(4)
    drag = np.zeros((7, 9))
    drag[2:5, 3:6] = rng.uniform(0.5, 3.0, (3, 3))  # a patch of biomass, discontinuous
    drag = drag.reshape(-1)
    # A stationary patch diverts flow to its sides.
    _, slow = _duct_solve(oracle, drag, 0.0, 0.0, 2.0)
    assert slow[3, 4] < 1.5 and slow[0, 4] > 2.0
    # At the plug speed relative drag vanishes everywhere.
    x, matched = _duct_solve(oracle, drag, 0.0, 2.0, 2.0)
    np.testing.assert_allclose(matched, 2.0, atol=4e-5)
    np.testing.assert_allclose(x, 0.0, atol=4e-5)
    # The field is affine in the medium's velocity.
    _, fast = _duct_solve(oracle, drag, 0.0, 4.0, 2.0)
    np.testing.assert_allclose(fast - matched, matched - slow, atol=5e-5)
    # Transverse motion with no imposed throughput drives a transverse flow,
    # and reversing it reverses the field.
    x, y = _duct_solve(oracle, drag, 3.0, 0.0, 0.0)
    assert x[3, 4] > 0.5
    rx, ry = _duct_solve(oracle, drag, -3.0, 0.0, 0.0)
    np.testing.assert_allclose(rx, -x, atol=1e-5)
    np.testing.assert_allclose(ry, -y, atol=1e-5)


def test_zero_force_is_the_unforced_solve(oracle):
    nx, ny = 30, 17
    solid, mobility = random_case(7, nx, ny)
    s = solid.reshape(nx, ny)
    s[0, :] = s[-1, :] = 0
    grid = Grid(nx, ny, 1, 1.0, 1.0, 1.0, s.reshape(-1))
    plain, _, rp = oracle.solve_flow(grid, mobility, 0, 2.0, 1e-9, 20000)
    zero, _, rz = oracle.solve_flow(grid, mobility, 0, 2.0, 1e-9, 20000,
                                    body_force=np.zeros((nx * ny, 2), np.float32))
    # The forced system's right-hand side is zero: its solve takes no
    # iterations and adds exactly nothing.
    assert rz.iterations == rp.iterations
    assert np.array_equal(plain, zero)
