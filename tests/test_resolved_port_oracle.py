"""Pins the oracle's port-driven resolved flow (mo_resolved_port_flow_solve, a
restatement of solve_resolved_port_flow_cpu, cpp/cpu/cpu_flow.cpp:474-479,
over detail::solve_resolved_port_flow, cpp/core/flow_system.hpp:1482-1548,
shallow_port_pressure, 1318-1338, and superpose_port_flows, 1146-1252):

- bit for bit against the reference built from its sources (oracle/_ref):
  faces, port states and every report field, on ducts driven by one or more
  speed ports against pressure ports, with and without the depth-averaged
  pressure shift (a z-face port, a port short of its columns' depth and a
  column broken by a solid voxel each leave shallow flow undefined), held
  pressures only, obstacles and drag;
- what the physics fixes: a speed port delivers its stated mean speed, the
  port fluxes balance, the field is linear in the stated speeds, and a duct
  driven through whole-face ports is the single-axis resolved flow;
- the refusals of ResolvedPortFlowParameters::validate (cpu_flow.cpp:179-185)
  and validate_flow_ports (194-286), with the reference's messages.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import FIXED, NO_FLUX, PERIODIC, Grid, OracleError

SPEED, PRESSURE = 0, 1
X_LOWER, X_UPPER, Y_LOWER, Y_UPPER, Z_LOWER, Z_UPPER = range(6)


def duct(nx, ny, nz, spacing=(1.0, 1.0, 1.0), obstacles=None, kinds=None):
    kinds = kinds or (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX)
    return Grid(nx, ny, nz, *spacing, obstacles, tuple(kinds))


def pillars(nx, ny, nz, fraction, seed):
    """Solid columns through the full depth (shallow flow describes them),
    none on the x end planes."""
    rng = np.random.default_rng(seed)
    solid = np.repeat((rng.random((nx, ny, 1)) < fraction).astype(np.uint8), nz, axis=2)
    solid[0] = solid[-1] = 0
    return solid.reshape(-1)


def port_cases():
    rng = np.random.default_rng(11)
    voxels = (rng.random((9, 6, 4)) < 0.1).astype(np.uint8)
    voxels[0] = voxels[-1] = 0
    broken = np.zeros((6, 5, 4), np.uint8)
    broken[3, 2, 1] = 1
    drag = (rng.random(9 * 6 * 4) * 3).astype(np.float32)
    return {
        "whole faces": (duct(8, 5, 4), None,
                        [(X_LOWER, (0, 0), (5, 4), SPEED, 1.0),
                         (X_UPPER, (0, 0), (5, 4), PRESSURE, 0.0)]),
        "two inlets over pillars": (
            duct(10, 6, 3, (1.0, 0.8, 1.2), pillars(10, 6, 3, 0.15, 2)),
            (rng.random(10 * 6 * 3) * 2).astype(np.float32),
            [(X_LOWER, (0, 0), (3, 3), SPEED, 1.0), (X_LOWER, (3, 0), (6, 3), SPEED, 2.0),
             (X_UPPER, (0, 0), (6, 3), PRESSURE, 0.5)]),
        "inlet and outlet on y faces": (
            duct(6, 9, 3, kinds=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX)), None,
            [(Y_LOWER, (1, 0), (4, 3), SPEED, -0.7), (Y_UPPER, (2, 0), (6, 3), PRESSURE, 1.0)]),
        "voxel obstacles": (duct(9, 6, 4, (1.0, 0.8, 1.2), voxels.reshape(-1)), drag,
                            [(X_LOWER, (0, 0), (3, 4), SPEED, 1.0),
                             (X_LOWER, (3, 0), (6, 4), SPEED, 2.0),
                             (X_UPPER, (0, 0), (6, 4), PRESSURE, 0.5)]),
        "z face inlet": (duct(6, 6, 5, kinds=(FIXED, NO_FLUX, NO_FLUX, NO_FLUX, FIXED, NO_FLUX)),
                         None, [(Z_LOWER, (1, 1), (4, 4), SPEED, 1.0),
                                (X_LOWER, (0, 0), (6, 5), PRESSURE, 0.0)]),
        "partial depth": (duct(7, 5, 4), None, [(X_LOWER, (0, 0), (5, 2), SPEED, 1.0),
                                                (X_UPPER, (0, 0), (5, 4), PRESSURE, 0.0)]),
        "pressures only": (duct(7, 5, 3), None, [(X_LOWER, (0, 0), (5, 3), PRESSURE, 2.0),
                                                 (X_UPPER, (0, 0), (5, 3), PRESSURE, 0.5)]),
        "broken column": (duct(6, 5, 4, obstacles=broken.reshape(-1)), None,
                          [(X_LOWER, (0, 0), (5, 4), SPEED, 1.0),
                           (X_UPPER, (0, 0), (5, 4), PRESSURE, 0.0)]),
    }


# Whether the depth-averaged pressure shifts every solve of a case.
SHIFTED = {"whole faces": True, "two inlets over pillars": True,
           "inlet and outlet on y faces": True, "voxel obstacles": False,
           "z face inlet": False, "partial depth": False, "pressures only": True,
           "broken column": False}


@pytest.mark.parametrize("name", list(port_cases()))
def test_port_flows_match_the_reference(oracle, reference, name):
    g, drag, ports = port_cases()[name]
    fa, sa, ra = oracle.resolved_port_flow(g, drag, ports)
    fb, sb, rb = reference.resolved_port_flow(g, drag, ports)
    print(name, rb, "shifted", ra["shifted_solves"], flush=True)
    assert np.array_equal(fa, fb)
    assert np.array_equal(sa, sb)
    assert {k: ra[k] for k in rb} == rb
    assert (ra["shifted_solves"] == ra["solves"]) == SHIFTED[name]
    assert ra["shifted_solves"] in (0, ra["solves"])


@pytest.mark.parametrize("name", [n for n in port_cases() if n != "pressures only"])
def test_speed_ports_deliver_their_speed_and_fluxes_balance(oracle, name):
    g, drag, ports = port_cases()[name]
    faces, states, rep = oracle.resolved_port_flow(g, drag, ports)
    for port, (pressure, flux, mean_speed) in zip(ports, states):
        if port[3] == SPEED:
            assert mean_speed == pytest.approx(port[4], rel=1e-5)
    # Inflow is positive: the ports' fluxes sum to zero (the field is
    # divergence-free and closed elsewhere).
    total = float(np.abs(states[:, 1]).sum())
    assert abs(float(states[:, 1].sum())) <= 2e-5 * total
    assert rep["divergence_rms"] < 2e-5 * max(1.0, rep["max_speed"])


def test_field_is_linear_in_the_stated_speeds(oracle):
    g, drag, ports = port_cases()["two inlets over pillars"]
    ports = [p[:4] + ((0.0,) if p[3] == PRESSURE else (p[4],)) for p in ports]
    doubled = [p[:4] + (2.0 * p[4],) for p in ports]
    f1, _, r1 = oracle.resolved_port_flow(g, drag, ports)
    f2, _, _ = oracle.resolved_port_flow(g, drag, doubled)
    assert np.abs(f2 - 2.0 * f1).max() <= 1e-5 * r1["max_speed"]


def test_whole_face_ports_are_the_single_axis_flow(oracle):
    """A speed inlet over the whole lower x face and pressure 0 over the upper
    one give the single-axis resolved flow at the same mean speed (both hold
    each end at one pressure). Both solve to the tightest tolerance the
    parameters allow (1.2e-7), so the fields agree to 5e-6 of the largest
    speed (at 1e-6 they differ by up to 2.4e-5)."""
    g = duct(10, 6, 4, (1.0, 0.9, 1.1), pillars(10, 6, 4, 0.1, 5))
    drag = (np.random.default_rng(5).random(10 * 6 * 4) * 4).astype(np.float32)
    ports = [(X_LOWER, (0, 0), (6, 4), SPEED, 1.5), (X_UPPER, (0, 0), (6, 4), PRESSURE, 0.0)]
    port_faces, states, rep = oracle.resolved_port_flow(g, drag, ports, 1.2e-7, 5000, 1.2e-7)
    axis_faces, axis_rep = oracle.resolved_flow(g, drag, 0, 1.5, 1.2e-7, 500, 1.2e-7)
    assert np.abs(port_faces - axis_faces).max() <= 5e-6 * axis_rep["max_speed"]
    assert rep["min_gap_voxels"] <= axis_rep["min_gap_voxels"] or axis_rep["min_gap_voxels"] == 0


def refusal_cases():
    g = duct(6, 5, 3)
    ok = [(X_LOWER, (0, 0), (5, 3), SPEED, 1.0), (X_UPPER, (0, 0), (5, 3), PRESSURE, 0.0)]
    wall = np.zeros((6, 5, 3), np.uint8)
    wall[3] = 1
    blocked = np.zeros((6, 5, 3), np.uint8)
    blocked[0] = 1
    base = dict(grid=g, drag=None, ports=ok, tol=1e-6, outer=5000, inner_tol=1e-6, inner=50000)
    return [
        ("loose tolerance", dict(base, tol=1e-9)),
        ("inner tolerance", dict(base, inner_tol=float("inf"))),
        ("no iterations", dict(base, outer=0)),
        ("no ports", dict(base, ports=[])),
        ("periodic", dict(base, grid=duct(6, 5, 3, kinds=(FIXED, FIXED, PERIODIC, PERIODIC,
                                                           NO_FLUX, NO_FLUX)))),
        ("closed face", dict(base, ports=ok + [(Y_LOWER, (0, 0), (6, 3), PRESSURE, 0.0)])),
        ("overlap", dict(base, ports=ok + [(X_LOWER, (2, 0), (5, 3), PRESSURE, 0.0)])),
        ("empty patch", dict(base, ports=[(X_LOWER, (2, 0), (2, 3), SPEED, 1.0)] + ok[1:])),
        ("outside the face", dict(base, ports=[(X_LOWER, (0, 0), (6, 3), SPEED, 1.0)] + ok[1:])),
        ("blocked port", dict(base, grid=duct(6, 5, 3, obstacles=blocked.reshape(-1)))),
        ("no pressure port", dict(base, ports=[ok[0]])),
        ("nothing flows", dict(base, ports=[(X_LOWER, (0, 0), (5, 3), PRESSURE, 1.0),
                                            (X_UPPER, (0, 0), (5, 3), PRESSURE, 1.0)])),
        ("unreachable", dict(base, grid=duct(6, 5, 3, obstacles=wall.reshape(-1)))),
        ("nan value", dict(base, ports=[(X_LOWER, (0, 0), (5, 3), SPEED, float("nan"))] + ok[1:])),
        ("negative drag", dict(base, drag=np.full(90, -1.0, np.float32))),
        ("few outer iterations", dict(base, outer=2)),
    ]


@pytest.mark.parametrize("name,case", refusal_cases(), ids=[n for n, _ in refusal_cases()])
def test_refusals_match_the_reference(oracle, reference, name, case):
    args = (case["grid"], case["drag"], case["ports"], case["tol"], case["outer"],
            case["inner_tol"], case["inner"])
    with pytest.raises(OracleError) as ours:
        oracle.resolved_port_flow(*args)
    with pytest.raises(OracleError) as theirs:
        reference.resolved_port_flow(*args)
    print(name, ":", theirs.value, flush=True)
    assert ours.value.message == theirs.value.message
    if theirs.value.code != 2:
        assert ours.value.code == theirs.value.code
