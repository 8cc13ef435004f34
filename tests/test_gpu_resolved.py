"""Resolved Stokes-Brinkman flow on the GPU through the C ABI
(msim_hip_solve_resolved_flow) against the oracle, which is bit-exact with
the reference's solve_resolved_flow_cpu (tests/test_resolved_oracle.py).

The device runs the reference's algorithm in FP64 (flexible GMRES with
restarts of 40 and twice-modified Gram-Schmidt, the momentum block by
Jacobi-PCG); only the order of the dot products' sums differs. Tolerances:
outer iterations within 1 and inner iterations within 2 % of the oracle's;
faces within 2e-5 of the largest speed (the solves stop at a relative
residual of 1e-6, so a stop one iteration apart moves the field by about
that much); the report's divergence RMS and gap as the oracle's. The
refusals carry the oracle's messages.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import OracleError
from tests.test_resolved_oracle import conformance_case, duct, random_case, refusal_cases

pytestmark = pytest.mark.gpu


def device_grid(og):
    return msa.Grid(og.nx, og.ny, og.nz, og.dx, og.dy, og.dz, og.obstacles, og.boundary)


def params(axis, speed, tol=1e-6, outer=500, inner_tol=1e-6, inner=50000):
    return msa.ResolvedFlowParameters(mean_inlet_speed=speed, axis=axis, relative_tolerance=tol,
                                      max_outer_iterations=outer, inner_relative_tolerance=inner_tol,
                                      max_inner_iterations=inner)


def compare(hip, oracle, og, drag, axis, speed):
    field, rep = hip.solve_resolved_flow(device_grid(og), drag, params(axis, speed))
    ref, orep = oracle.resolved_flow(og, drag, axis, speed, 1e-6, 500, 1e-6, 50000)
    got = field.concatenated()
    err = np.abs(got - ref).max() / orep["max_speed"]
    print((og.nx, og.ny, og.nz), "axis", axis, "outer", rep.outer_iterations,
          orep["outer_iterations"], "inner", rep.inner_iterations, orep["inner_iterations"],
          "error / max speed", err, "device", rep.seconds, "s", flush=True)
    assert abs(rep.outer_iterations - orep["outer_iterations"]) <= 1
    assert abs(rep.inner_iterations - orep["inner_iterations"]) <= 0.02 * orep["inner_iterations"] + 2
    assert err <= 2e-5
    assert rep.min_gap_voxels == orep["min_gap_voxels"]
    assert rep.relative_residual <= 1e-6
    assert rep.divergence_rms == pytest.approx(orep["divergence_rms"], rel=1e-2, abs=1e-6)
    return rep


def test_conformance_duct_on_device(hip, oracle):
    og, drag = conformance_case()
    compare(hip, oracle, og, drag, 1, 2.0)


@pytest.mark.parametrize("seed", range(12))
def test_random_grids_on_device(hip, oracle, seed):
    og, drag, axis = random_case(seed)
    compare(hip, oracle, og, drag, axis, 1.5)


def test_plane_poiseuille_on_device(hip):
    """flow_test.cpp:146-168 on the device."""
    nx = 8
    og = duct(nx, 6, 1, (1.0 / nx, 0.25, 1.0), axis=1)
    field, rep = hip.solve_resolved_flow(device_grid(og), None, params(1, 1.0))
    y = field.y_faces.reshape(nx, 7)
    position = (np.arange(nx) + 0.5) / nx
    exact = 6.0 * position * (1.0 - position)
    assert np.abs(y[:, 3] - exact).max() / 1.5 < 0.02
    assert rep.outer_iterations > 0 and rep.inner_iterations > 0
    assert rep.divergence_rms < 2e-5
    assert rep.min_gap_voxels == nx


def test_larger_grid_on_device(hip, oracle):
    """A 48 x 24 x 12 channel with seeded obstacles and drag along x."""
    rng = np.random.default_rng(17)
    nx, ny, nz = 48, 24, 12
    solid = (rng.random((nx, ny, nz)) < 0.08).astype(np.uint8)
    solid[0] = solid[-1] = 0
    og = duct(nx, ny, nz, (1.0, 1.0, 1.0), 0, solid.reshape(-1))
    drag = (rng.random(nx * ny * nz) * 2.0).astype(np.float32)
    compare(hip, oracle, og, drag, 0, 1.0)


@pytest.mark.parametrize("name,case", refusal_cases(), ids=[n for n, _ in refusal_cases()])
def test_refusals_on_device(hip, oracle, name, case):
    og = case["grid"]
    with pytest.raises(OracleError) as expected:
        oracle.resolved_flow(og, case["drag"], case["axis"], case["speed"], case["tol"],
                             case["outer"], case["inner_tol"], case["inner"])
    with pytest.raises(msa.MsimError) as got:
        hip.solve_resolved_flow(device_grid(og), case["drag"],
                                params(case["axis"], case["speed"], case["tol"], case["outer"],
                                       case["inner_tol"], case["inner"]))
    print(name, ":", got.value, flush=True)
    assert str(got.value) == expected.value.message
    # The context still solves after a refusal.
    good, drag = conformance_case()
    hip.solve_resolved_flow(device_grid(good), drag, params(1, 2.0))
