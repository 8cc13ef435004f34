"""Depth-averaged flow on the GPU through the C ABI, against the CPU oracle
and the reference's golden fixtures.

Tolerances: assembly and the FP64 operator are bit-exact (same float/double
operations, no contraction). The FP64 PCG differs from the reference only in
the order of its dot-product sums, so the iteration count is held to +-2 of
the reference's (tests/conformance/flow_conformance_test.cpp:259-263) and the
face velocities to 1e-6 x max speed (its ill-conditioned gate, :124). The
mixed-precision solve is held to the FP64 true residual 1e-8, and its
velocities to 1e-5 x max speed of the FP64 oracle's.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import Grid as OGrid
from tests.helpers import (FIXED, LAYERED_CASES, NO_FLUX, PERIODIC, c1_inputs, divergence,
                           illcond_inputs, layered_case, random_case, split_faces)

pytestmark = pytest.mark.gpu


def grids(nx, ny, dz, obstacles, boundary=(FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX),
          dx=1.0, dy=1.0):
    return (msa.Grid(nx, ny, 1, dx, dy, dz, obstacles, boundary),
            OGrid(nx, ny, 1, dx, dy, dz, obstacles, boundary))


def params(**kw):
    base = dict(mean_inlet_speed=100.0, axis=0, relative_tolerance=1e-10, max_iterations=200000)
    base.update(kw)
    return msa.DepthAveragedFlowParameters(**base)


def test_device_present(hip):
    info = hip.info()
    assert info.kind == "hip" and info.native
    assert msa.device_count() >= 1
    print("device:", info.device, flush=True)


@pytest.mark.parametrize("case", ["c1", "illcond", "random"])
def test_assembly_is_bit_exact(hip, oracle, case):
    if case == "c1":
        c = c1_inputs()
        g, og = grids(c["nx"], c["ny"], c["dz"], c["obstacles"])
        mob, axis = c["mobility"], 0
    elif case == "illcond":
        c = illcond_inputs()
        g, og = grids(64, 64, 1.0, c["obstacles"], (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX,
                                                    NO_FLUX))
        mob, axis = c["mobility"], 1
    else:
        solid, mob = random_case(21, 37, 300)
        g, og = grids(37, 300, 2.5, solid, dx=0.5, dy=1.5)
        axis = 0
    hip.flow_prepare(g, mob, axis)
    ours = hip.flow_arrays()
    ref = oracle.assemble(og, mob, axis)
    for key in ("conductance", "diagonal", "rhs", "bc"):
        assert np.array_equal(ours[key], ref[key]), key


@pytest.mark.parametrize("shape", [(256, 128), (37, 300), (5, 1000)])
def test_operator_is_bit_exact(hip, oracle, shape):
    nx, ny = shape
    solid, mob = random_case(7, nx, ny)
    g, og = grids(nx, ny, 1.0, solid, dx=0.8, dy=1.1)
    hip.flow_prepare(g, mob, 0)
    x = np.random.default_rng(1).standard_normal(nx * ny)
    assert np.array_equal(hip.flow_apply(x), oracle.apply(og, mob, x))


def test_fp64_pcg_matches_reference_on_channel_trap(hip):
    c = c1_inputs()
    d = c["data"]
    g, _ = grids(c["nx"], c["ny"], c["dz"], c["obstacles"])
    res = hip.solve_depth_averaged_flow(g, c["mobility"], params(relative_tolerance=1.5e-7))
    ref = d["faces"]
    faces = res.field.concatenated()
    print("iterations", res.report.iterations, "reference", int(d["iterations"]), flush=True)
    assert abs(res.report.iterations - int(d["iterations"])) <= 2
    assert np.abs(faces - ref).max() <= 1e-6 * float(d["max_speed"])
    assert res.report.relative_residual <= 1.5e-7
    assert abs(res.report.max_speed - float(d["max_speed"])) <= 1e-5 * float(d["max_speed"])


def test_fp64_pcg_to_1e10(hip, oracle):
    c = c1_inputs()
    g, og = grids(c["nx"], c["ny"], c["dz"], c["obstacles"])
    hip.flow_prepare(g, c["mobility"], 0)
    rep = hip.flow_solve(params())
    pressure, field = hip.flow_download()
    ref_faces, ref_p, ref_rep = oracle.solve_flow(og, c["mobility"], 0, 100.0, 1e-10, 200000)
    assert np.array_equal(ref_p, c["data"]["pressure"])  # oracle = reference here
    print("iterations", rep.iterations, ref_rep.iterations, "true", rep.true_relative_residual,
          flush=True)
    assert abs(rep.iterations - ref_rep.iterations) <= 2
    assert rep.true_relative_residual <= 2e-10
    assert np.abs(pressure - ref_p).max() <= 1e-7
    assert np.abs(field.concatenated() - ref_faces).max() <= 1e-6 * ref_rep.max_speed
    # Inactive sites stay exactly zero.
    assert np.all(pressure[c["obstacles"] != 0] == 0.0)


def test_fp64_pcg_ill_conditioned_conformance_case(hip):
    c = illcond_inputs()
    d = c["data"]
    g, _ = grids(64, 64, 1.0, c["obstacles"], (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    res = hip.solve_depth_averaged_flow(g, c["mobility"], params(
        mean_inlet_speed=1.0, axis=1, relative_tolerance=1e-6))
    assert abs(res.report.iterations - int(d["iterations"])) <= 2
    assert np.abs(res.field.concatenated() - d["faces"]).max() <= 1e-6 * float(d["max_speed"])


@pytest.mark.parametrize("case", ["c1", "illcond"])
def test_mixed_precision_reaches_1e8(hip, oracle, case):
    if case == "c1":
        c = c1_inputs()
        g, og = grids(c["nx"], c["ny"], c["dz"], c["obstacles"])
        mob, axis, speed = c["mobility"], 0, 100.0
    else:
        c = illcond_inputs()
        kinds = (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX)
        g, og = grids(64, 64, 1.0, c["obstacles"], kinds)
        mob, axis, speed = c["mobility"], 1, 1.0
    hip.flow_prepare(g, mob, axis)
    rep = hip.flow_solve(params(solver=msa.SOLVER_MIXED, relative_tolerance=1e-8, axis=axis,
                                mean_inlet_speed=speed))
    pressure, field = hip.flow_download()
    ref_faces, ref_p, ref_rep = oracle.solve_flow(og, mob, axis, speed, 1e-12, 200000)
    err_faces = np.abs(field.concatenated() - ref_faces).max() / ref_rep.max_speed
    err_p = np.abs(pressure - ref_p).max() / np.abs(ref_p).max()
    print(case, "iterations", rep.iterations, "refinements", rep.refinements, "true",
          rep.true_relative_residual, "velocity error / max speed", err_faces,
          "pressure error", err_p, flush=True)
    assert rep.true_relative_residual <= 1e-8
    assert err_faces <= 1e-5


def test_closed_forms_on_gpu(hip):
    duct = msa.Grid(4, 8, 1, boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    res = hip.solve_depth_averaged_flow(duct, None, params(mean_inlet_speed=5.0, axis=1,
                                                           relative_tolerance=1e-6))
    assert np.all(np.abs(res.field.y_faces - 5.0) <= 1e-5)
    assert np.all(np.abs(res.field.x_faces) <= 2e-5) and np.all(res.field.z_faces == 0)
    split = msa.Grid(2, 6, 1, boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    mob = np.where(np.arange(12) // 6 == 0, 1.0, 3.0).astype(np.float32)
    res = hip.solve_depth_averaged_flow(split, mob, params(mean_inlet_speed=4.0, axis=1,
                                                           relative_tolerance=1e-6))
    yf = res.field.y_faces.reshape(2, 7)
    assert abs(yf[1, 3] / yf[0, 3] - 3.0) <= 2e-5


def test_full_size_pillar_array_properties(hip):
    """1024x512 with ~300 pillars: both solver forms reach the FP64 true
    residual 1e-8, agree with each other, and give a divergence-free field
    whose open-inlet mean is the requested speed."""
    from microsimulator_amd.workloads import config
    w = config("c2")
    g = msa.Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    hip.flow_prepare(g, w.mobility, 0)
    fp64 = hip.flow_solve(params(relative_tolerance=1e-8))
    p64, f64 = hip.flow_download()
    mixed = hip.flow_solve(params(solver=msa.SOLVER_MIXED, relative_tolerance=1e-8))
    p32, f32 = hip.flow_download()
    print("fp64", fp64.iterations, fp64.seconds, "mixed", mixed.iterations, mixed.refinements,
          mixed.seconds, flush=True)
    assert fp64.true_relative_residual <= 1.1e-8 and mixed.true_relative_residual <= 1e-8
    a, b = f64.concatenated(), f32.concatenated()
    assert np.abs(a - b).max() <= 1e-4 * fp64.max_speed
    div = divergence(a, w.nx, w.ny)
    assert np.abs(div[w.obstacles == 0]).max() <= 1e-4 * fp64.max_speed
    inlet = f64.x_faces[: w.ny][(w.obstacles[0] == 0)]
    assert abs(inlet.mean() - 100.0) <= 1e-3


def test_error_paths(hip):
    blocked = np.ones((8, 6), np.uint8)
    blocked[:, 2] = 0
    blocked[5, 2] = 1
    g = msa.Grid(8, 6, 1, obstacles=blocked.reshape(-1))
    with pytest.raises(msa.MsimError, match="outlet is unreachable"):
        hip.solve_depth_averaged_flow(g, None, params())
    duct = msa.Grid(4, 8, 1, boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    with pytest.raises(msa.NotConverged):
        hip.solve_depth_averaged_flow(duct, None, params(axis=1, max_iterations=1))
    periodic = msa.Grid(4, 8, 1, boundary=(PERIODIC, PERIODIC, FIXED, FIXED, NO_FLUX, NO_FLUX))
    with pytest.raises(msa.InvalidArgument):
        hip.solve_depth_averaged_flow(periodic, None, params(axis=1))
    with pytest.raises(msa.InvalidArgument):
        hip.solve_depth_averaged_flow(duct, -np.ones(32, np.float32), params(axis=1))
    overhang = np.zeros((4, 8, 3), np.uint8)
    overhang[2, 3, 1] = 1
    with pytest.raises(msa.InvalidArgument, match="contiguous fluid columns"):
        hip.solve_depth_averaged_flow(msa.Grid(4, 8, 3, obstacles=overhang.reshape(-1)), None,
                                      params())
    with pytest.raises(msa.Unsupported):
        hip.advance_species()
    # The context still works after the failures.
    res = hip.solve_depth_averaged_flow(duct, None, params(axis=1, mean_inlet_speed=5.0,
                                                           relative_tolerance=1e-6))
    assert np.all(np.abs(res.field.y_faces - 5.0) <= 1e-5)


def _multilevel_reference(oracle, og, mob, nx, ny, dx, dy):
    from tests import mg_reference as mr
    a = oracle.assemble(og, mob, 0)
    m = a["conductance"].reshape(nx, ny).astype(np.float32)

    def harm(p, q):
        s = p + q
        return np.where(s > 0, np.float32(2) * p * q / np.where(s > 0, s, 1), 0).astype(np.float32)
    cxl = np.zeros((nx, ny), np.float32)
    cxl[1:] = harm(m[1:], m[:-1])
    cyl = np.zeros((nx, ny), np.float32)
    cyl[:, 1:] = harm(m[:, 1:], m[:, :-1])
    return mr.Multilevel(mr.level0(cxl, cyl, a["bc"], nx, ny, dx, dy))


@pytest.mark.parametrize("shape", [(256, 128), (301, 77), (1024, 512), (20, 13)])
def test_multilevel_cycle_matches_cpu_restatement(hip, oracle, shape):
    """One device cycle (FP32) against the float64 restatement."""
    nx, ny = shape
    solid, mob = random_case(13, nx, ny, solid_fraction=0.1)
    g, og = grids(nx, ny, 1.0, solid, dx=1.0, dy=1.5)
    hip.flow_prepare(g, mob, 0)
    ml = _multilevel_reference(oracle, og, mob, nx, ny, 1.0, 1.5)
    r = np.random.default_rng(2).standard_normal(nx * ny).astype(np.float32)
    r[solid != 0] = 0.0
    z = hip.flow_precondition(r).astype(np.float64)
    ref = ml.cycle(r.astype(np.float64).reshape(nx, ny)).ravel()
    err = np.abs(z - ref).max() / np.abs(ref).max()
    print(shape, "levels", len(ml.levels), "relative error", err, flush=True)
    assert err <= 2e-4


@pytest.mark.parametrize("case", ["c1", "illcond"])
def test_multigrid_solver_reaches_1e8(hip, oracle, case):
    if case == "c1":
        c = c1_inputs()
        g, og = grids(c["nx"], c["ny"], c["dz"], c["obstacles"])
        mob, axis, speed = c["mobility"], 0, 100.0
    else:
        c = illcond_inputs()
        kinds = (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX)
        g, og = grids(64, 64, 1.0, c["obstacles"], kinds)
        mob, axis, speed = c["mobility"], 1, 1.0
    hip.flow_prepare(g, mob, axis)
    rep = hip.flow_solve(params(solver=msa.SOLVER_MULTIGRID, relative_tolerance=1e-8, axis=axis,
                                mean_inlet_speed=speed))
    _, field = hip.flow_download()
    ref_faces, _, ref_rep = oracle.solve_flow(og, mob, axis, speed, 1e-12, 200000)
    err = np.abs(field.concatenated() - ref_faces).max() / ref_rep.max_speed
    print(case, "multigrid iterations", rep.iterations, "refinements", rep.refinements,
          "true", rep.true_relative_residual, "velocity error / max speed", err, flush=True)
    assert rep.true_relative_residual <= 1e-8
    assert err <= 1e-5
    assert rep.iterations <= 150


def test_multigrid_full_size(hip):
    from microsimulator_amd.workloads import config
    w = config("c2")
    g = msa.Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    hip.flow_prepare(g, w.mobility, 0)
    fp64 = hip.flow_solve(params(relative_tolerance=1e-8))
    _, f64 = hip.flow_download()
    mg = hip.flow_solve(params(solver=msa.SOLVER_MULTIGRID, relative_tolerance=1e-8))
    _, fmg = hip.flow_download()
    print("fp64", fp64.iterations, fp64.seconds, "multigrid", mg.iterations, mg.refinements,
          mg.seconds, flush=True)
    assert mg.true_relative_residual <= 1e-8
    assert np.abs(fmg.concatenated() - f64.concatenated()).max() <= 1e-4 * fp64.max_speed
    assert mg.iterations <= 150


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("solver", [msa.SOLVER_JACOBI_FP64, msa.SOLVER_MULTIGRID])
def test_layered_grids_match_the_oracle(hip, oracle, axis, solver):
    """Grids of several layers: the device solves the grid of columns and
    lifts the column flux onto the voxel faces (vertical faces by
    continuity); the host scales it. Against the oracle (bit-exact with the
    reference on these grids): FP64 Jacobi-PCG within 2 iterations and 1e-6 of
    the largest speed, the multilevel form within 1e-5."""
    for nx, ny, nz, seed, mob in LAYERED_CASES:
        obstacles, mobility, boundary = layered_case(nx, ny, nz, seed, mob, axis)
        g = msa.Grid(nx, ny, nz, 1.0, 0.5, 0.7, obstacles, boundary)
        og = OGrid(nx, ny, nz, 1.0, 0.5, 0.7, obstacles, boundary)
        print("case", (nx, ny, nz, seed, mob), "axis", axis, "solver", solver, flush=True)
        res = hip.solve_depth_averaged_flow(g, mobility, params(
            axis=axis, mean_inlet_speed=2.0, relative_tolerance=1e-10, solver=solver))
        ref, _, ref_rep = oracle.solve_flow(og, mobility, axis, 2.0, 1e-10, 200000)
        got = res.field.concatenated()
        assert got.size == ref.size
        err = np.abs(got - ref).max() / ref_rep.max_speed
        print("  velocity error / max speed", err, "iterations", res.report.iterations,
              ref_rep.iterations, flush=True)
        if solver == msa.SOLVER_JACOBI_FP64:
            assert abs(res.report.iterations - ref_rep.iterations) <= 2
            assert err <= 1e-6
        else:
            assert err <= 1e-5
        assert res.report.max_speed == pytest.approx(ref_rep.max_speed, rel=1e-5)

