"""Port-driven depth-averaged flow on the GPU through the C ABI
(msim_hip_solve_depth_averaged_port_flow), against the reference's outputs
(committed fixture) and the CPU oracle, which is bit-exact with the
reference on these cases.

Tolerances. Each FP64 Jacobi-PCG solve differs from the reference's only in
the order of its dot-product sums, so the summed iteration count is held to
within 2 per solve (the conformance suite's bound for one solve,
tests/conformance/flow_conformance_test.cpp:259-263); faces, port fluxes and
mean speeds to the suite's close() (8e-4 absolute + 8e-4 relative, :35-40)
and faces also to 1e-5 of the largest speed; port pressures to 8e-4 of the
first port's (:215-216). The combination itself (FP64 weights over FP32
fields, the 1x1..4x4 elimination) repeats the reference's arithmetic, so on
bit-equal solves it would be bit-equal.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import Grid as OGrid
from oracle import OracleError
from tests.helpers import (DUCT, PORT_CASES, PRESSURE, SIDES_FIXED, SPEED, X_LOWER, Y_LOWER,
                           Y_UPPER, depth_port_case, golden, largest_divergence, port,
                           random_port_case, winding_channel_case)
from tests.test_port_oracle import refusals

pytestmark = pytest.mark.gpu


def close(actual, expected):
    return np.abs(np.asarray(actual) - np.asarray(expected)) <= 8.0e-4 + 8.0e-4 * np.abs(
        np.asarray(expected))


def mports(ports):
    return [msa.FlowPort(f, tuple(b), tuple(e), k, v) for f, b, e, k, v in ports]


def solve(hip, g, mobility, ports, **kw):
    grid = msa.Grid(g["nx"], g["ny"], g["nz"], g["dx"], g["dy"], g["dz"], g["obstacles"],
                    g["boundary"])
    params = msa.DepthAveragedPortFlowParameters(**kw)
    res = hip.solve_depth_averaged_port_flow(grid, mports(ports), mobility, params)
    states = np.array([[s.pressure, s.flux, s.mean_speed] for s in res.ports], np.float32)
    return res.field.concatenated(), states, res.report


def check_against(faces, states, report, ref_faces, ref_states, ref_solves, ref_iterations,
                  max_speed):
    assert report.solves == ref_solves
    assert abs(report.iterations - ref_iterations) <= 2 * ref_solves
    assert report.relative_residual <= 1.1e-6
    assert np.all(close(faces, ref_faces))
    assert np.abs(faces - ref_faces).max() <= 1e-5 * max_speed
    assert np.all(close(states[:, 1], ref_states[:, 1]))
    assert np.all(close(states[:, 2], ref_states[:, 2]))
    assert np.all(np.abs(states[:, 0] - ref_states[:, 0]) <= 8e-4 * abs(ref_states[0, 0]) + 1e-12)


@pytest.mark.parametrize("name", ["depth", "winding"])
def test_conformance_port_cases_match_reference_fixture(hip, name):
    d = golden("port_flow.npz")
    g, mobility, ports = depth_port_case() if name == "depth" else winding_channel_case()
    print("case", name, flush=True)
    faces, states, report = solve(hip, g, mobility, ports)
    ref_faces = d[f"{name}_faces"]
    max_speed = float(np.abs(ref_faces).max())
    print("  solves", report.solves, "iterations", report.iterations,
          int(d[f"{name}_iterations"]), "error / max speed",
          np.abs(faces - ref_faces).max() / max_speed, flush=True)
    check_against(faces, states, report, ref_faces, d[f"{name}_states"],
                  int(d[f"{name}_solves"]), int(d[f"{name}_iterations"]), max_speed)
    assert report.max_speed == pytest.approx(max_speed, rel=1e-5)
    # What the conformance suite asserts of the candidate itself.
    for k in range(len(ports)):
        if ports[k][3] == SPEED:
            assert close(states[k, 2], ports[k][4])
    if name == "winding":
        assert close(states[0, 1], -states[1, 1])


@pytest.mark.parametrize("seed,nx,ny,nz,with_mobility", PORT_CASES)
def test_random_port_cases_match_the_oracle(hip, oracle, seed, nx, ny, nz, with_mobility):
    g, mobility, ports = random_port_case(seed, nx, ny, nz, with_mobility)
    print("case", (seed, nx, ny, nz, with_mobility), flush=True)
    faces, states, report = solve(hip, g, mobility, ports)
    ref_faces, ref_states, ref = oracle.port_flow(OGrid(**g), mobility, ports)
    max_speed = float(ref["max_speed"])
    print("  solves", report.solves, "iterations", report.iterations, ref["iterations"],
          "error / max speed", np.abs(faces - ref_faces).max() / max_speed, flush=True)
    check_against(faces, states, report, ref_faces, ref_states, ref["solves"],
                  ref["iterations"], max_speed)
    assert report.max_speed == pytest.approx(max_speed, rel=1e-5)


def test_known_answers_on_the_device(hip):
    """tests/cpp/flow_test.cpp:168-281 on the device: stated inlet speeds and
    fluxes, the closed face between two inlets, a perpendicular inlet and its
    reversal, and the pressure-driven duct's 0.8."""
    g = dict(nx=6, ny=8, nz=2, dx=0.5, dy=1.0, dz=0.25, obstacles=None, boundary=SIDES_FIXED)
    ports = [port(Y_LOWER, 0, 2, 2, SPEED, 3.0), port(Y_LOWER, 4, 6, 2, SPEED, 1.0),
             port(Y_UPPER, 0, 6, 2, PRESSURE, 0.0)]
    faces, st, rep = solve(hip, g, None, ports)
    assert rep.solves == 2 and rep.iterations > 0 and rep.relative_residual <= 1e-6
    assert abs(st[0, 2] - 3.0) <= 1e-5 and abs(st[1, 2] - 1.0) <= 1e-5
    assert abs(st[0, 1] - 1.5) <= 1e-5 and abs(st[1, 1] - 0.5) <= 1e-5
    assert abs(st[2, 1] + 2.0) <= 1e-5
    assert st[0, 0] > st[1, 0] > st[2, 0] and st[2, 0] == 0.0
    nxf, nyf = 7 * 8 * 2, 6 * 9 * 2
    xf, yf, zf = faces[:nxf], faces[nxf:nxf + nyf], faces[nxf + nyf:]
    assert largest_divergence(6, 8, 2, 0.5, 1.0, 0.25, np.zeros(96), xf, yf, zf) <= 2e-5
    assert yf[(2 * 9 + 0) * 2 + 0] == 0.0 and yf[(3 * 9 + 0) * 2 + 1] == 0.0

    g = dict(nx=5, ny=7, nz=1, dx=1.0, dy=1.0, dz=1.0, obstacles=None, boundary=SIDES_FIXED)
    for speed in (2.0, -2.0):
        ports = [port(X_LOWER, 1, 3, 1, SPEED, speed), port(Y_UPPER, 0, 5, 1, PRESSURE, 0.0)]
        faces, st, rep = solve(hip, g, None, ports)
        assert rep.solves == 1
        assert abs(st[0, 1] - 2.0 * speed) <= 1e-5 and abs(st[1, 1] + 2.0 * speed) <= 1e-5
        assert np.sign(st[0, 0]) == np.sign(speed)
        assert faces[0] == 0.0 and np.sign(faces[1]) == np.sign(speed)

    g = dict(nx=3, ny=5, nz=1, dx=1.0, dy=0.5, dz=1.0, obstacles=None, boundary=DUCT)
    ports = [port(Y_LOWER, 0, 3, 1, PRESSURE, 3.0), port(Y_UPPER, 0, 3, 1, PRESSURE, 1.0)]
    faces, st, rep = solve(hip, g, None, ports)
    assert rep.solves == 1
    assert np.all(np.abs(faces[4 * 5:4 * 5 + 3 * 6] - 0.8) <= 1e-5)
    assert st[0, 0] == 3.0 and st[1, 0] == 1.0


@pytest.mark.parametrize("case", range(len(refusals())))
def test_refusals_match_the_oracle(hip, oracle, case):
    grid, ports, _ = refusals()[case]
    with pytest.raises(OracleError) as ref:
        oracle.port_flow(grid, None, ports)
    g = msa.Grid(grid.nx, grid.ny, grid.nz, grid.dx, grid.dy, grid.dz, grid.obstacles,
                 grid.boundary)
    with pytest.raises(msa.MsimError) as got:
        hip.solve_depth_averaged_port_flow(g, mports(ports))
    assert str(got.value) == ref.value.message
    assert (ref.value.code == 1) == isinstance(got.value, msa.InvalidArgument)


@pytest.mark.parametrize("solver", [msa.SOLVER_MIXED, msa.SOLVER_MULTIGRID])
def test_solver_forms_on_a_port_flow(hip, oracle, solver):
    """The mixed-precision and multilevel forms of each pressure solve, on a
    256x128 sheet with four ports (four solves), against the oracle's FP64
    field, all at 2e-7 (the tolerance floor is float epsilon)."""
    g, mobility, ports = random_port_case(11, 256, 128, 1, True)
    ref_faces, ref_states, ref = oracle.port_flow(OGrid(**g), mobility, ports,
                                                  tolerance=2e-7, max_iterations=200000)
    faces, states, report = solve(hip, g, mobility, ports, solver=solver,
                                  relative_tolerance=2e-7, max_iterations=200000)
    max_speed = float(ref["max_speed"])
    err = np.abs(faces - ref_faces).max() / max_speed
    print("solver", solver, "solves", report.solves, "iterations", report.iterations,
          "oracle", ref["iterations"], "error / max speed", err, "seconds", report.seconds,
          flush=True)
    assert report.solves == ref["solves"] == 4
    assert report.true_relative_residual <= 4e-7
    assert err <= 1e-4
    assert np.all(close(states, ref_states))


def test_full_size_port_flow_properties(hip):
    """2048 x 2048 with four ports, multilevel solves: every speed port carries
    its stated mean speed, what enters leaves, and the field is divergence
    free to the solve's accuracy."""
    g, mobility, ports = random_port_case(31, 2048, 2048, 1, True)
    ports = [(X_LOWER, (0, 0), (1024, 1), PRESSURE, 0.0),
             (Y_LOWER, (100, 0), (900, 1), SPEED, 2.0),
             (Y_UPPER, (1200, 0), (2000, 1), SPEED, -1.5),
             (1, (300, 0), (1900, 1), PRESSURE, 0.5)]
    faces, states, report = solve(hip, g, mobility, ports, solver=msa.SOLVER_MULTIGRID,
                                  relative_tolerance=2e-7)
    print("solves", report.solves, "iterations", report.iterations, "seconds", report.seconds,
          "states", states.tolist(), flush=True)
    assert report.solves == 3 and report.true_relative_residual <= 4e-7
    assert abs(states[1, 2] - 2.0) <= 1e-5 * 2.0 and abs(states[2, 2] + 1.5) <= 1e-5 * 1.5
    inflow = float(np.abs(states[:, 1]).sum())
    assert abs(float(states[:, 1].astype(np.float64).sum())) <= 1e-4 * inflow
    nx = ny = 2048
    nxf, nyf = (nx + 1) * ny, nx * (ny + 1)
    div = largest_divergence(nx, ny, 1, g["dx"], g["dy"], g["dz"], g["obstacles"],
                             faces[:nxf], faces[nxf:nxf + nyf], faces[nxf + nyf:])
    print("largest divergence / max speed", div / report.max_speed, flush=True)
    assert div <= 1e-3 * report.max_speed / min(g["dx"], g["dy"])
