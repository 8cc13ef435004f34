"""The oracle's port-driven flow (mo_port_flow_solve) pinned against the
reference's solve_depth_averaged_port_flow_cpu: bit for bit through the live
oracle/_ref build and the committed fixture, and against the known answers
and refusals of the reference's own tests (tests/cpp/flow_test.cpp:168-282,
tests/conformance/flow_conformance_test.cpp:167-271).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import FIXED, NO_FLUX, Grid, OracleError
from tests.helpers import (DUCT, PORT_CASES, PRESSURE, SIDES_FIXED, SPEED, X_LOWER, X_UPPER,
                           Y_LOWER, Y_UPPER, Z_UPPER, depth_port_case, golden,
                           largest_divergence, port, random_port_case, winding_channel_case)


def close(actual, expected):
    """flow_conformance_test.cpp:35-40."""
    return abs(actual - expected) <= 8.0e-4 + 8.0e-4 * abs(expected)


def split(grid: Grid, faces):
    nfx, nfy, _ = grid.face_counts()
    return faces[:nfx], faces[nfx:nfx + nfy], faces[nfx + nfy:]


def y_face(grid: Grid, x, y, z):
    return (x * (grid.ny + 1) + y) * grid.nz + z


# ------------------------------------------------------------ live _ref --
@pytest.mark.parametrize("seed,nx,ny,nz,with_mobility", PORT_CASES)
def test_random_port_flow_matches_reference_live(oracle, reference, seed, nx, ny, nz,
                                                 with_mobility):
    g, mobility, ports = random_port_case(seed, nx, ny, nz, with_mobility)
    grid = Grid(**g)
    faces, states, report = oracle.port_flow(grid, mobility, ports)
    f_r, s_r, r_r = reference.port_flow(grid, mobility, ports)
    assert np.array_equal(faces, f_r)
    assert np.array_equal(states, s_r)
    for key in ("solves", "iterations", "relative_residual", "max_speed"):
        assert report[key] == r_r[key], key


@pytest.mark.parametrize("case", ["depth_ports", "winding_channel"])
def test_conformance_port_cases_match_reference_live(oracle, reference, case):
    g, mobility, ports = depth_port_case() if case == "depth_ports" else winding_channel_case()
    grid = Grid(**g)
    faces, states, report = oracle.port_flow(grid, mobility, ports)
    f_r, s_r, r_r = reference.port_flow(grid, mobility, ports)
    assert np.array_equal(faces, f_r) and np.array_equal(states, s_r)
    assert report["iterations"] == r_r["iterations"] and report["solves"] == r_r["solves"]


def test_refusals_match_reference_live(oracle, reference):
    """Every refusal of the random cases carries the reference's kind and
    message."""
    seen = 0
    for seed in range(120):
        rng = np.random.default_rng(seed + 1000)
        g, mobility, ports = random_port_case(seed, int(rng.integers(3, 20)),
                                              int(rng.integers(3, 20)), [1, 2][seed % 2])
        grid = Grid(**g)
        errors = []
        for lib in (oracle, reference):
            try:
                lib.port_flow(grid, mobility, ports)
                errors.append(None)
            except OracleError as e:
                errors.append((e.code, e.message))
        assert errors[0] == errors[1], seed
        seen += errors[0] is not None
    assert seen > 0


# --------------------------------------------------------- golden fixture --
def test_port_cases_match_reference_fixture(oracle):
    d = golden("port_flow.npz")
    for name, builder in (("depth", depth_port_case), ("winding", winding_channel_case)):
        g, mobility, ports = builder()
        faces, states, report = oracle.port_flow(Grid(**g), mobility, ports)
        assert np.array_equal(faces, d[f"{name}_faces"]), name
        assert np.array_equal(states, d[f"{name}_states"]), name
        assert report["iterations"] == int(d[f"{name}_iterations"]), name
        assert report["solves"] == int(d[f"{name}_solves"]), name


def test_conformance_port_case_properties(oracle):
    """What the conformance suite asserts of the reference itself."""
    g, mobility, ports = depth_port_case()
    _, states, report = oracle.port_flow(Grid(**g), mobility, ports)
    assert report["solves"] == 3 and report["relative_residual"] <= 1.1e-6
    for k in range(3):
        assert close(states[k][2], ports[k][4])
    g, mobility, ports = winding_channel_case()
    _, states, report = oracle.port_flow(Grid(**g), mobility, ports)
    assert report["iterations"] > 200
    assert close(states[0][1], -states[1][1]) and close(states[0][2], ports[0][4])


# ------------------------------------------------------------ known answers --
def open_box(nx, ny, nz, dx=1.0, dy=1.0, dz=1.0, obstacles=None):
    return Grid(nx, ny, nz, dx, dy, dz, obstacles, SIDES_FIXED)


def test_two_inlets_feed_one_outlet(oracle):
    """flow_test.cpp:168-203: each inlet opens four faces of 0.5 x 0.25."""
    grid = open_box(6, 8, 2, 0.5, 1.0, 0.25)
    ports = [port(Y_LOWER, 0, 2, 2, SPEED, 3.0), port(Y_LOWER, 4, 6, 2, SPEED, 1.0),
             port(Y_UPPER, 0, 6, 2, PRESSURE, 0.0)]
    faces, st, report = oracle.port_flow(grid, None, ports)
    assert report["solves"] == 2 and report["iterations"] > 0
    assert report["relative_residual"] <= 1.0e-6
    assert abs(st[0][2] - 3.0) <= 1e-5 and abs(st[1][2] - 1.0) <= 1e-5
    assert abs(st[0][1] - 1.5) <= 1e-5 and abs(st[1][1] - 0.5) <= 1e-5
    assert abs(st[2][1] + 2.0) <= 1e-5 and abs(st[2][2] + 2.0 / 1.5) <= 1e-5
    assert st[0][0] > st[1][0] > st[2][0] and st[2][0] == 0.0
    xf, yf, zf = split(grid, faces)
    assert largest_divergence(6, 8, 2, 0.5, 1.0, 0.25, np.zeros(96, np.uint8), xf, yf, zf) <= 2e-5
    assert yf[y_face(grid, 2, 0, 0)] == 0.0 and yf[y_face(grid, 3, 0, 1)] == 0.0
    assert yf[y_face(grid, 0, 0, 0)] > 0.0


def test_perpendicular_inlet_turns_the_flow(oracle):
    """flow_test.cpp:205-228, and the same port drawing fluid out."""
    grid = open_box(5, 7, 1)
    ports = [port(X_LOWER, 1, 3, 1, SPEED, 2.0), port(Y_UPPER, 0, 5, 1, PRESSURE, 0.0)]
    faces, st, report = oracle.port_flow(grid, None, ports)
    assert report["solves"] == 1
    assert abs(st[0][1] - 4.0) <= 1e-5 and abs(st[1][1] + 4.0) <= 1e-5
    xf, yf, zf = split(grid, faces)
    assert largest_divergence(5, 7, 1, 1.0, 1.0, 1.0, np.zeros(35, np.uint8), xf, yf, zf) <= 2e-5
    assert xf[1] > 0.0 and xf[2] > 0.0 and xf[0] == 0.0
    assert yf[y_face(grid, 2, 7, 0)] > 0.0
    ports[0] = port(X_LOWER, 1, 3, 1, SPEED, -2.0)
    faces, st, _ = oracle.port_flow(grid, None, ports)
    assert abs(st[0][1] + 4.0) <= 1e-5 and st[0][0] < 0.0
    assert split(grid, faces)[1][y_face(grid, 2, 7, 0)] < 0.0


def test_whole_face_ports_state_the_single_axis_flow(oracle):
    """flow_test.cpp:230-264: ports over the whole lower and upper y faces
    give the single-axis solve's field, around an obstacle."""
    nx, ny, nz = 5, 8, 2
    obstacles = np.zeros((nx, ny, nz), np.uint8)
    mobility = np.zeros((nx, ny, nz), np.float32)
    for x in range(nx):
        obstacles[x, 3:5, :] = 1 if x == 2 else 0
        mobility[x] = np.float32(0.5) + np.float32(0.25) * np.float32(x)
    grid = Grid(nx, ny, nz, 0.8, 1.2, 0.6, obstacles.reshape(-1), DUCT)
    stated, _, srep = oracle.solve_flow(grid, mobility.reshape(-1), 1, 5.0, 1e-6, 50000)
    ports = [port(Y_LOWER, 0, 5, 2, SPEED, 5.0), port(Y_UPPER, 0, 5, 2, PRESSURE, 0.0)]
    through, _, rep = oracle.port_flow(grid, mobility.reshape(-1), ports)
    assert np.abs(through - stated).max() <= 2e-5 * srep.max_speed
    assert abs(rep["max_speed"] - srep.max_speed) <= 2e-5 * srep.max_speed


def test_pressure_ports_alone_drive_the_duct(oracle):
    """flow_test.cpp:266-281: unit mobility, a duct of length 2.5 between
    pressures 3 and 1 carries speed 0.8 on every y face."""
    grid = Grid(3, 5, 1, 1.0, 0.5, 1.0, None, DUCT)
    ports = [port(Y_LOWER, 0, 3, 1, PRESSURE, 3.0), port(Y_UPPER, 0, 3, 1, PRESSURE, 1.0)]
    faces, st, report = oracle.port_flow(grid, None, ports)
    assert report["solves"] == 1
    assert np.all(np.abs(split(grid, faces)[1] - 0.8) <= 1e-5)
    assert st[0][0] == 3.0 and st[1][0] == 1.0
    assert abs(st[0][2] - 0.8) <= 1e-5 and abs(st[1][2] + 0.8) <= 1e-5


def refusals():
    """flow_test.cpp:283-331: (grid, ports, status) that the reference
    refuses; status 1 is invalid_argument, 2 runtime_error."""
    grid = open_box(4, 6, 2)
    inlet = port(Y_LOWER, 0, 2, 2, SPEED, 1.0)
    outlet = port(Y_UPPER, 0, 4, 2, PRESSURE, 0.0)
    capped = Grid(4, 6, 2, 1.0, 1.0, 1.0, None, SIDES_FIXED[:5] + (FIXED,))
    closed = Grid(4, 6, 2, 1.0, 1.0, 1.0, None, (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    blocked = np.zeros((4, 6, 2), np.uint8)
    blocked[0:2, 0, :] = 1
    walled = np.zeros((4, 6, 2), np.uint8)
    walled[:, 3, :] = 1
    return [
        (grid, [], 1),
        (grid, [inlet], 1),
        (grid, [outlet], 1),
        (grid, [inlet, port(Y_LOWER, 1, 3, 2, SPEED, 1.0), outlet], 1),
        (grid, [port(Y_LOWER, 0, 5, 2, SPEED, 1.0), outlet], 1),
        (grid, [port(Y_LOWER, 2, 2, 2, SPEED, 1.0), outlet], 1),
        (grid, [port(Y_LOWER, 0, 2, 1, SPEED, 1.0), outlet], 1),
        (capped, [(Z_UPPER, (0, 0), (2, 2), SPEED, 1.0), outlet], 1),
        (closed, [port(X_LOWER, 0, 2, 2, SPEED, 1.0), outlet], 1),
        (open_box(4, 6, 2, obstacles=blocked.reshape(-1)), [inlet, outlet], 1),
        (open_box(4, 6, 2, obstacles=walled.reshape(-1)), [inlet, outlet], 2),
        (grid, [port(Y_LOWER, 0, 2, 2, 7, 1.0), outlet], 1),
        (grid, [port(Y_LOWER, 0, 2, 2, SPEED, float("nan")), outlet], 1),
        (grid, [outlet, port(X_UPPER, 0, 6, 2, PRESSURE, 0.0)], 1),
        (grid, [(6, (0, 0), (1, 1), SPEED, 1.0), outlet], 1),
    ]


@pytest.mark.parametrize("case", range(len(refusals())))
def test_refusals(oracle, case):
    grid, ports, status = refusals()[case]
    with pytest.raises(OracleError) as info:
        oracle.port_flow(grid, None, ports)
    assert info.value.code == status, info.value.message


@pytest.mark.parametrize("case", range(len(refusals())))
def test_refusals_match_reference_messages(oracle, reference, case):
    grid, ports, _ = refusals()[case]
    messages = []
    for lib in (oracle, reference):
        with pytest.raises(OracleError) as info:
            lib.port_flow(grid, None, ports)
        messages.append((info.value.code, info.value.message))
    assert messages[0] == messages[1]
