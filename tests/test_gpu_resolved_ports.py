"""Port-driven resolved flow on the GPU through the C ABI
(msim_hip_solve_resolved_port_flow) against the oracle, which is bit-exact
with the reference's solve_resolved_port_flow_cpu on the same cases
(tests/test_resolved_port_oracle.py).

Every block solve runs the reference's algorithm in FP64 on the device;
the depth-averaged pressure that shifts a solve is the device's FP64
Jacobi-PCG solve of the same ports. Only the order of the dot products'
sums differs, and with it where each solve crosses its relative tolerance
of 1e-6. Tolerances: the same number of block solves; outer iterations
within 5 % (and 2) of the oracle's; faces within 5e-5 of the largest speed;
speed ports' mean speeds within 1e-5 relative, the driven pressures within
1e-4; the same gap; the divergence RMS within twice the oracle's (or 1e-6).
The refusals carry the oracle's messages.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import OracleError
from tests.test_resolved_port_oracle import SPEED, port_cases, refusal_cases

pytestmark = pytest.mark.gpu


def device_grid(og):
    return msa.Grid(og.nx, og.ny, og.nz, og.dx, og.dy, og.dz, og.obstacles, og.boundary)


def device_ports(ports):
    return [msa.FlowPort(face, tuple(begin), tuple(end), kind, value)
            for face, begin, end, kind, value in ports]


def params(tol=1e-6, outer=5000, inner_tol=1e-6, inner=50000):
    return msa.ResolvedPortFlowParameters(relative_tolerance=tol, max_outer_iterations=outer,
                                          inner_relative_tolerance=inner_tol,
                                          max_inner_iterations=inner)


@pytest.mark.parametrize("name", list(port_cases()))
def test_port_flows_on_device(hip, oracle, name):
    og, drag, ports = port_cases()[name]
    result = hip.solve_resolved_port_flow(device_grid(og), device_ports(ports), drag, params())
    ref, states, rep = oracle.resolved_port_flow(og, drag, ports)
    got = result.field.concatenated()
    err = np.abs(got - ref).max() / rep["max_speed"]
    r = result.report
    print(name, "solves", r.solves, rep["solves"], "outer", r.outer_iterations,
          rep["outer_iterations"], "inner", r.inner_iterations, rep["inner_iterations"],
          "error / max speed", err, "device", r.seconds, "s", flush=True)
    assert r.solves == rep["solves"]
    assert abs(r.outer_iterations - rep["outer_iterations"]) <= max(
        2, 0.05 * rep["outer_iterations"])
    assert err <= 5e-5
    assert r.min_gap_voxels == rep["min_gap_voxels"]
    assert r.relative_residual <= 1e-6
    assert r.divergence_rms <= max(2.0 * rep["divergence_rms"], 1e-6)
    for port, state, want in zip(ports, result.ports, states):
        if port[3] == SPEED:
            assert state.mean_speed == pytest.approx(float(want[2]), rel=1e-5)
            assert state.pressure == pytest.approx(float(want[0]), rel=1e-4, abs=1e-6)
        else:
            assert state.pressure == float(want[0])
        assert state.flux == pytest.approx(float(want[1]), rel=1e-4, abs=1e-5 * rep["max_speed"])


def test_port_flow_is_the_axis_flow_on_device(hip):
    """Whole-face ports give the single-axis resolved flow at the same mean
    speed, both solved on the device to the tightest tolerance the parameters
    allow (1.2e-7): within 5e-6 of the largest speed."""
    from tests.test_resolved_port_oracle import duct, pillars
    og = duct(12, 8, 5, (1.0, 0.9, 1.1), pillars(12, 8, 5, 0.1, 7))
    drag = (np.random.default_rng(7).random(12 * 8 * 5) * 4).astype(np.float32)
    grid = device_grid(og)
    ports = [msa.FlowPort(msa.X_LOWER, (0, 0), (8, 5), msa.SPEED, 1.5),
             msa.FlowPort(msa.X_UPPER, (0, 0), (8, 5), msa.PRESSURE, 0.0)]
    by_ports = hip.solve_resolved_port_flow(grid, ports, drag, params(1.2e-7, 5000, 1.2e-7))
    field, rep = hip.solve_resolved_flow(grid, drag, msa.ResolvedFlowParameters(
        mean_inlet_speed=1.5, axis=0, relative_tolerance=1.2e-7, inner_relative_tolerance=1.2e-7))
    diff = np.abs(by_ports.field.concatenated() - field.concatenated()).max()
    print("port flow against axis flow", diff / rep.max_speed, flush=True)
    assert diff <= 5e-6 * rep.max_speed
    assert by_ports.ports[0].mean_speed == pytest.approx(1.5, rel=1e-5)


@pytest.mark.parametrize("name,case", refusal_cases(), ids=[n for n, _ in refusal_cases()])
def test_refusals_on_device(hip, oracle, name, case):
    og = case["grid"]
    with pytest.raises(OracleError) as expected:
        oracle.resolved_port_flow(og, case["drag"], case["ports"], case["tol"], case["outer"],
                                  case["inner_tol"], case["inner"])
    with pytest.raises(msa.MsimError) as got:
        hip.solve_resolved_port_flow(device_grid(og), device_ports(case["ports"]), case["drag"],
                                     params(case["tol"], case["outer"], case["inner_tol"],
                                            case["inner"]))
    print(name, ":", got.value, flush=True)
    message = expected.value.message
    if name == "few outer iterations":
        # The residual where the solve stops differs in its last digits.
        assert str(got.value).split(":")[0] == message.split(":")[0]
    else:
        assert str(got.value) == message
    # The context still solves after a refusal.
    good, drag, ports = port_cases()["whole faces"]
    hip.solve_resolved_port_flow(device_grid(good), device_ports(ports), drag, params())
