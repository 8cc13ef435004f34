"""Body-forced depth-averaged flow on the GPU through the C ABI
(DepthAveragedFlowParameters::body_force, flow.hpp:24-25; the forced second
solve and the driven term of scale_velocity, flow_system.hpp:1368-1447),
against the oracle, which is bit-exact with the reference here
(tests/test_body_force_oracle.py).

Tolerances are those of the unforced solve: the FP64 Jacobi-PCG form holds its
iterations to +-2 per solve of the oracle's (flow_conformance_test.cpp:259-263)
and the faces to 1e-6 x max speed (:124); the mixed and multilevel forms hold
the FP64 true residual to the requested 1e-8 and the faces to 1e-5 x max
speed of the oracle's FP64 field.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import Grid as OGrid
from tests.helpers import FIXED, LAYERED_CASES, NO_FLUX, layered_case, random_case

pytestmark = pytest.mark.gpu


def case(seed, axis, nx=96, ny=61):
    """A random sheet whose fluid is one connected region. A pocket of fluid
    cut off from the boundary makes the forced system only semidefinite
    there (its right-hand side is the body flux's divergence, not zero), and
    Jacobi-PCG then breaks down below a few 1e-8, the reference's as well."""
    from scipy import ndimage
    if axis == 1:
        nx, ny = ny, nx
    solid, mobility = random_case(seed, nx, ny)
    s = solid.reshape(nx, ny)
    if axis == 0:
        s[0, :] = s[-1, :] = 0
    else:
        s[:, 0] = s[:, -1] = 0
    labels, count = ndimage.label(s == 0)
    largest = np.argmax(np.bincount(labels.reshape(-1))[1:]) + 1
    s[labels != largest] = 1
    mobility = mobility.reshape(nx, ny)
    mobility[s != 0] = 0.0
    mobility = mobility.reshape(-1)
    kinds = (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX) if axis == 0 else \
        (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX)
    obstacles = s.reshape(-1)
    force = np.random.default_rng(seed + 100).uniform(-0.5, 0.5, (nx * ny, 2)).astype(
        np.float32)
    return (msa.Grid(nx, ny, 1, 0.7, 1.3, 2.0, obstacles, kinds),
            OGrid(nx, ny, 1, 0.7, 1.3, 2.0, obstacles, kinds), mobility, force)


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("speed", [3.0, 0.0, -1.5])
def test_forced_fp64_solve_matches_the_oracle(hip, oracle, axis, speed):
    grid, ogrid, mobility, force = case(11 + axis, axis)
    print(f"forced FP64: axis {axis}, speed {speed}", flush=True)
    p = msa.DepthAveragedFlowParameters(mean_inlet_speed=speed, axis=axis,
                                        relative_tolerance=1e-8, max_iterations=200000,
                                        solver=msa.SOLVER_JACOBI_FP64, body_force=force)
    res = hip.solve_depth_averaged_flow(grid, mobility, p)
    ref, _, rep = oracle.solve_flow(ogrid, mobility, axis, speed, 1e-8, 200000,
                                    body_force=force)
    faces = res.field.concatenated()
    print("iterations", res.report.iterations, rep.iterations, "max speed",
          res.report.max_speed, rep.max_speed, flush=True)
    assert abs(res.report.iterations - rep.iterations) <= 4  # +-2 per solve
    assert np.abs(faces - ref).max() <= 1e-6 * rep.max_speed
    assert res.report.max_speed == pytest.approx(rep.max_speed, rel=1e-6)


@pytest.mark.parametrize("solver", [msa.SOLVER_MIXED, msa.SOLVER_MULTIGRID])
def test_forced_mixed_and_multilevel_forms(hip, oracle, solver):
    grid, ogrid, mobility, force = case(21, 0, 256, 128)
    p = msa.DepthAveragedFlowParameters(mean_inlet_speed=2.0, axis=0, relative_tolerance=1e-8,
                                        max_iterations=200000, solver=solver, body_force=force)
    res = hip.solve_depth_averaged_flow(grid, mobility, p)
    ref, _, rep = oracle.solve_flow(ogrid, mobility, 0, 2.0, 1e-12, 400000, body_force=force)
    faces = res.field.concatenated()
    print("solver", solver, "iterations", res.report.iterations, "true residual",
          res.report.true_relative_residual, flush=True)
    assert res.report.true_relative_residual <= 1e-8
    assert np.abs(faces - ref).max() <= 1e-5 * rep.max_speed


@pytest.mark.parametrize("axis", [0, 1])
def test_forced_layered_solve_matches_the_oracle(hip, oracle, axis):
    for nx, ny, nz, seed, mob in LAYERED_CASES:
        obstacles, mobility, boundary = layered_case(nx, ny, nz, seed, mob, axis)
        grid = msa.Grid(nx, ny, nz, 1.0, 0.5, 0.7, obstacles, boundary)
        ogrid = OGrid(nx, ny, nz, 1.0, 0.5, 0.7, obstacles, boundary)
        force = np.random.default_rng(seed + 10).uniform(-2.0, 2.0, (nx * ny, 2)).astype(
            np.float32)
        p = msa.DepthAveragedFlowParameters(mean_inlet_speed=2.0, axis=axis,
                                            relative_tolerance=1e-9, max_iterations=200000,
                                            body_force=force)
        res = hip.solve_depth_averaged_flow(grid, mobility, p)
        ref, _, rep = oracle.solve_flow(ogrid, mobility, axis, 2.0, 1e-9, 200000,
                                        body_force=force)
        faces = res.field.concatenated()
        assert np.abs(faces - ref).max() <= 1e-6 * rep.max_speed, (nx, ny, nz, axis)
        assert abs(res.report.iterations - rep.iterations) <= 4


def test_forced_solve_refusals(hip, oracle):
    nx, ny = 12, 9
    grid = msa.Grid(nx, ny, 1, 1.0, 1.0, 1.0, None)
    ogrid = OGrid(nx, ny, 1, 1.0, 1.0, 1.0, None)
    from oracle import OracleError
    cases = [(1.0, np.zeros((nx * ny - 1, 2), np.float32)),
             (1.0, np.array([[np.nan, 0.0]] * (nx * ny), np.float32)),
             (0.0, None)]
    for speed, force in cases:
        with pytest.raises(OracleError) as expected:
            oracle.solve_flow(ogrid, None, 0, speed, 1e-6, 1000, body_force=force)
        p = msa.DepthAveragedFlowParameters(mean_inlet_speed=speed, relative_tolerance=1e-6,
                                            body_force=force)
        with pytest.raises(msa.InvalidArgument) as got:
            hip.solve_depth_averaged_flow(grid, None, p)
        assert str(got.value) in str(expected.value)


def test_relative_drag_in_a_uniform_duct(hip):
    """The reference test's duct (python/tests/test_flow.py:192-233): drag d,
    mobility 1/(1 + d), force d V. The plug speed leaves the plug flow; the
    field is affine in V; reversing V with no throughput reverses it."""
    nx, ny = 7, 9
    grid = msa.Grid(nx, ny, 1, 1.0, 1.0, 1.0, None,
                    (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    rng = np.random.default_rng(4)
    drag = np.zeros((nx, ny))
    drag[2:5, 3:6] = rng.uniform(0.5, 3.0, (3, 3))
    drag = drag.reshape(-1)
    mobility = (1.0 / (1.0 + drag)).astype(np.float32)

    def solve(vx, vy, inlet):
        force = np.stack([drag * vx, drag * vy], axis=1).astype(np.float32)
        p = msa.DepthAveragedFlowParameters(mean_inlet_speed=inlet, axis=1,
                                            relative_tolerance=1e-10, body_force=force)
        field = hip.solve_depth_averaged_flow(grid, mobility, p).field
        x = field.x_faces.reshape(nx + 1, ny)
        y = field.y_faces.reshape(nx, ny + 1)
        assert np.all(x[[0, -1], :] == 0.0)
        np.testing.assert_allclose(y[:, 0].mean(), inlet, atol=2e-5)
        np.testing.assert_allclose(np.diff(x, axis=0) + np.diff(y, axis=1), 0, atol=5e-5)
        return x, y

    _, slow = solve(0.0, 0.0, 2.0)
    assert slow[3, 4] < 1.5 and slow[0, 4] > 2.0
    x, matched = solve(0.0, 2.0, 2.0)
    np.testing.assert_allclose(matched, 2.0, atol=4e-5)
    np.testing.assert_allclose(x, 0.0, atol=4e-5)
    _, fast = solve(0.0, 4.0, 2.0)
    np.testing.assert_allclose(fast - matched, matched - slow, atol=5e-5)
    x, y = solve(3.0, 0.0, 0.0)
    assert x[3, 4] > 0.5
    rx, ry = solve(-3.0, 0.0, 0.0)
    np.testing.assert_allclose(rx, -x, atol=1e-5)
    np.testing.assert_allclose(ry, -y, atol=1e-5)


def test_forced_solve_keeps_the_pressure_driven_pressure(hip):
    """flow_download after a forced solve returns the pressure-driven solve's
    pressure, and a later unforced solve of the same system is the plain one."""
    grid, _, mobility, force = case(31, 0, 64, 40)
    hip.flow_prepare(grid, mobility, 0)
    plain = msa.DepthAveragedFlowParameters(mean_inlet_speed=1.0, relative_tolerance=1e-9,
                                            solver=msa.SOLVER_MULTIGRID)
    hip.flow_solve(plain)
    p0, f0 = hip.flow_download()
    forced = msa.DepthAveragedFlowParameters(mean_inlet_speed=1.0, relative_tolerance=1e-9,
                                             solver=msa.SOLVER_MULTIGRID, body_force=force)
    hip.flow_solve(forced)
    p1, f1 = hip.flow_download()
    assert np.array_equal(p0, p1)
    assert not np.array_equal(f0.concatenated(), f1.concatenated())
    hip.flow_solve(plain)
    p2, f2 = hip.flow_download()
    assert np.array_equal(p0, p2)
    assert np.array_equal(f0.concatenated(), f2.concatenated())
