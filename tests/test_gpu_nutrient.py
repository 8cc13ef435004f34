"""The device-resident nutrient loop (msim_hip_nutrient_run): per step the
occupancy's mobility by the reference's Brinkman closure, the pressure solve
and velocity, one FP64 transport step with Monod uptake and its ledger row,
and the occupancy grown by the uptake (the steps of the reference's
scripts/run_nutrient_benchmarks.py:96-120 on the grid).

Against a CPU restatement of the same loop on the oracle: the oracle's FP64
Jacobi-PCG solve to 1e-12 (the device solves by the multilevel form to
1e-10, so the fields agree to about 1e-8 of the largest speed), the oracle's
FP64 transport step, and the growth in numpy in the device's order. Levels and
occupancy are held to 1e-6 of their largest value, the per-step mass balance to
1e-12; at 4096 x 2048 with the learned solver every step's FP64 true residual
must meet 1e-8 without the fallback.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from microsimulator_amd import learned
from microsimulator_amd.workloads import channel_trap, config
from oracle import Grid as OGrid

pytestmark = pytest.mark.gpu

C_DRAG, CAP = 100.0, 0.9


def closure(base, phi, solid):
    f = np.minimum(phi.astype(np.float64), CAP)
    drag = C_DRAG * f * f / ((1.0 - f) * (1.0 - f) * (1.0 - f))
    b = base.astype(np.float64)
    return np.where(solid, 0.0, b / (1.0 + b * drag)).astype(np.float32)


def grow(phi, levels, vmax, K, yield_dt, solid):
    f = phi.astype(np.float64)
    uptake = np.zeros_like(f)
    n = f.size
    for s in range(len(vmax)):
        c = levels[s * n:(s + 1) * n]
        uptake = uptake + np.float64(vmax[s]) * f * c / (np.float64(K[s]) + c)
    out = (f + yield_dt * uptake).astype(np.float32)
    return np.where(solid, phi, out)


def test_nutrient_loop_matches_the_restatement(hip, oracle):
    w = channel_trap(128, 64, seed=9)
    nx, ny = w.nx, w.ny
    solid = w.obstacles.reshape(-1) != 0
    base = np.where(solid, 0.0, w.dz * w.dz / 12.0).astype(np.float32)
    phi = w.phi.reshape(-1).copy()
    grid = msa.Grid(nx, ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    ogrid = OGrid(nx, ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    S = 3
    diffusion = np.array([100.0, 500.0, 1000.0], np.float32)
    vmax = np.array([1.0, 0.5, 2.0], np.float32)
    K = np.array([0.1, 0.2, 0.3], np.float32)
    bv = np.zeros(6 * S, np.float32)
    bv[0::6] = 1.0
    speed = 50.0
    flow = msa.DepthAveragedFlowParameters(mean_inlet_speed=speed, axis=0,
                                           relative_tolerance=1e-10, max_iterations=100000,
                                           solver=msa.SOLVER_MULTIGRID)
    hip.flow_prepare(grid, closure(base, phi, solid), 0)
    rep = hip.flow_solve(flow)
    spec = msa.SignalGridSpec(grid, diffusion, bv, phi=phi, monod_vmax=vmax, monod_k=K)
    # Half the inlet level everywhere in the fluid, so uptake starts at once.
    levels = np.tile(np.where(solid, 0.0, 0.5), S)
    hip.adr_upload(spec, levels, precision=64, use_flow_field=True)
    bound = 2 * 1000.0 * 2 + 2 * rep.max_speed * 2 + float(vmax.max()) / 0.1 * 1.0
    dt = 0.4 / bound
    steps = 5
    yield_ = 500.0  # strong enough that the occupancy moves in five steps
    ledger, records = hip.nutrient_run(flow, base, dt, steps, C_DRAG, CAP, yield_)
    got_levels = hip.adr_download()
    got_phi = hip.adr_occupancy()
    # The same loop on the oracle.
    lv = levels.copy()
    ph = phi.copy()
    for _ in range(steps):
        faces, _, _ = oracle.solve_flow(ogrid, closure(base, ph, solid), 0, speed, 1e-12, 400000)
        before = lv.copy()
        lv, led = oracle.adr_step_f64(ogrid, lv, dt, 1, diffusion=diffusion, boundary_values=bv,
                                      faces=faces, phi=ph, vmax=vmax, monod_k=K)
        ph = grow(ph, before, vmax, K, np.float64(yield_) * np.float64(dt), solid)
    print("levels", np.abs(got_levels - lv).max(), lv.max(), "phi",
          np.abs(got_phi - ph).max(), ph.max(), flush=True)
    assert np.abs(got_levels - lv).max() <= 1e-6 * lv.max()
    assert np.abs(got_phi - ph).max() <= 1e-6 * ph.max()
    assert got_phi.max() > phi.max()  # the biomass grew
    for r in records:
        assert r["iterations"] > 0 and r["true_relative_residual"] <= 1e-10
        assert r["mass_balance"] <= 1e-12
    balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
    assert np.abs(balance).max() <= 1e-12 * ledger[:, :, 1].max()
    # The next run's ledger carries on from this one's mass.
    ledger2, _ = hip.nutrient_run(flow, base, dt, 1, C_DRAG, CAP, yield_)
    assert np.allclose(ledger2[0, :, 0], ledger[-1, :, 1], rtol=0, atol=0)


def test_nutrient_loop_refusals(hip):
    w = channel_trap(64, 32, seed=2)
    grid = msa.Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    flow = msa.DepthAveragedFlowParameters(mean_inlet_speed=10.0, relative_tolerance=1e-8,
                                           solver=msa.SOLVER_MULTIGRID)
    hip.flow_prepare(grid, w.mobility, 0)
    hip.flow_solve(flow)
    # FP32 transport is refused: the loop keeps the FP64 ledger.
    spec = msa.SignalGridSpec(grid, [100.0], phi=w.phi, monod_vmax=[1.0], monod_k=[0.1])
    hip.adr_upload(spec, np.zeros(grid.sites, np.float32), precision=32, use_flow_field=True)
    with pytest.raises(msa.InvalidArgument, match="FP64 forward-Euler"):
        hip.nutrient_run(flow, w.mobility, 1e-5, 1)
    hip.adr_upload(spec, np.zeros(grid.sites), precision=64, use_flow_field=True)
    with pytest.raises(msa.InvalidArgument, match="explicit stability bound"):
        hip.nutrient_run(flow, w.mobility, 10.0, 1)
    with pytest.raises(msa.InvalidArgument, match="max volume fraction"):
        hip.nutrient_run(flow, w.mobility, 1e-6, 1, max_volume_fraction=1.0)


def test_nutrient_loop_at_4096x2048_with_the_learned_solver(hip):
    """100 chained steps of the full-size loop: every step's solve meets the
    FP64 true residual 1e-8 without the fallback, every step's mass balances
    to 1e-12 in FP64, and each step starts from the mass the last one left."""
    w = config("c5")
    grid = msa.Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    solid = w.obstacles.reshape(-1) != 0
    base = np.where(solid, 0.0, w.dz * w.dz / 12.0).astype(np.float32)
    hip.learned_set_weights(learned.load_weights())
    flow = msa.DepthAveragedFlowParameters(mean_inlet_speed=100.0, axis=0,
                                           relative_tolerance=1e-8, max_iterations=2000,
                                           solver=msa.SOLVER_LEARNED, check_interval=8)
    hip.flow_prepare(grid, closure(base, w.phi.reshape(-1), solid), 0)
    rep = hip.flow_solve(flow)
    S = 3
    bv = np.zeros(6 * S, np.float32)
    bv[0::6] = 1.0
    spec = msa.SignalGridSpec(grid, [100.0, 500.0, 1000.0], bv, phi=w.phi,
                              monod_vmax=[1.0] * S, monod_k=[0.1] * S)
    hip.adr_upload(spec, np.zeros(S * grid.sites), precision=64, use_flow_field=True)
    dt = 0.5 / (2 * 1000.0 * 2 + 2 * rep.max_speed * 2 + 1.0 / 0.1 * 0.7)
    ledger, records = hip.nutrient_run(flow, base, dt, 100)
    assert len(records) == 100
    print(records[0], records[-1], flush=True)
    for r in records:
        assert not r["fell_back"]
        assert r["true_relative_residual"] <= 1e-8
        assert r["mass_balance"] <= 1e-12
    np.testing.assert_array_equal(ledger[1:, :, 0], ledger[:-1, :, 1])
    assert np.all(ledger[-1, :, 1] > 0.0)  # the inlet feeds every solute
