"""Benchmark: pressure solves/s at relative residual 1e-8 and transport
cell-updates/s on one MI355X, on the 4096 x 2048 pillar-array channel
(~4k seeded pillars r ~ U[8,20], phi from 500 seeded Gaussian blobs).

A step is one pressure solve from the assembled system resident in HBM to an
FP64 true residual ||b - Ax||/||b|| <= 1e-8, including the final true-residual
check and the velocity reconstruction (lift and scaling). The headline solver
form (--solver, default learned) is FP32 flexible CG preconditioned by the
multilevel cycle with the learned finest-level smoother (bfloat16 MFMA
convolutions, weights microsimulator_amd/learned_smoother.npz), inside FP64
iterative refinement. Timed beside it on the same GPU: the same cycle without
the network (multigrid), the plain FP64 Jacobi-PCG (the reference's
algorithm) and FP32 Jacobi-PCG with FP64 refinement.
Transport: 3 solutes (D = 100, 500, 1000 um^2/s) with Monod uptake, FP64
state, advected by the solved field, FP64 mass ledger every step.

Prints ONE JSON line. Usage:
    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "pressure solves/s at rel-res 1e-8 (learned vs plain PCG); ADR Gcell-upd/s"


def log(*args):
    print(*args, file=sys.stderr, flush=True)


# Compulsory HBM bytes per grid site of the multilevel solver's finest-level
# kernels (FP32 vectors, FP32 coefficients; the V(2,2) cycle):
#   k_mg_down   reads r, dinv, wx, wy, bc; writes the pre-smoothed iterate and
#               the coarse residual (1/4 site)
#   k_mg_up     reads r, the pre-smoothed iterate, q, dinv, wx, wy, bc and the
#               coarse correction (1/4); writes z
#   k_fcg_apply reads z, p_old, x, cxl, cyl, bc; writes p_new, q, x
#   k_fcg_update reads r, q; writes r
#   k_learn_up  the same traffic as k_mg_up
MG_BYTES = {"k_mg_down": 5 * 4 + 4 + 1, "k_mg_up": 7 * 4 + 1 + 4, "k_fcg_apply": 6 * 4 + 3 * 4,
            "k_fcg_update": 3 * 4, "k_learn_up": 7 * 4 + 1 + 4}

# Algorithmic matrix flops per site of the learned smoother (3x3 convolutions
# 4 -> 16, 16 -> 16, 16 -> 1, two flops per multiply-add), halo excluded.
LEARNED_FLOPS = 2 * 9 * (4 * 16 + 16 * 16 + 16 * 1)
BF16_PEAK_TFLOPS = 2500.0


# HBM bytes per launch measured with rocprofv3 FETCH_SIZE and WRITE_SIZE
# (scaled by a 256 MiB copy's readings) on this workload; the commands and
# figures are in profiles/krylov_and_resolved_kernels_traffic.md.
TRAFFIC_MB = {"k_learn_up": 751.4, "k_fcg_apply": 315.7, "k_mg_up": 530.1, "k_mg_down": 210.9,
              "k_fcg_update": 100.9, "k_cg_apply<float>": 353.7}


def apply_bytes(precision: int) -> int:
    v = 8 if precision == 64 else 4
    reads = 3 * v + 4 * 4  # r, p_old, x (vectors); diag, cxl, cyl, bc (float)
    writes = 3 * v  # p_new, q, x
    return reads + writes


def update_bytes(precision: int) -> int:
    v = 8 if precision == 64 else 4
    return 2 * v + 4 + v  # read r, q, diag; write r


def adr_bytes_per_site(signals: int, state_bytes: int) -> int:
    """Levels in and out per signal, two face velocities, mask, phi."""
    return signals * 2 * state_bytes + 2 * 4 + 1 + 4


def cpu_baseline(w, fp64_iterations: int, threads: int, budget_s: float):
    """The reference's own FP64 Jacobi-PCG (oracle/_ref, built from its
    sources) on this host, timed over a bounded number of iterations of the
    same system; solves/s = 1 / (seconds per iteration x the iterations the
    same algorithm needs on this system, from the GPU's FP64 solve)."""
    from oracle import Grid as OGrid
    from oracle import Reference
    grid = OGrid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    if Reference.available():
        ref = Reference()
        ref.set_threads(threads)
        system = ref.flow_system(grid, w.mobility, 0)
        system.time_iterations(5)  # thread pool and pages warm
        t1 = system.time_iterations(20)
        k = int(max(10, min(8000, budget_s / max(t1 / 20, 1e-6))))
        t = system.time_iterations(k)
        per_iteration = t / k
        return dict(value=1.0 / (per_iteration * fp64_iterations), unit="solves/s",
                    cores=ref.threads(), kind="reference",
                    sample=(f"{k} FP64 Jacobi-PCG iterations of the reference's CPU solver "
                            f"(detail::solve_depth_averaged_pressure) on the same 4096x2048 "
                            f"system in {t:.2f} s, scaled by the {fp64_iterations} iterations "
                            f"the same algorithm takes to 1e-8"),
                    seconds_per_iteration=per_iteration,
                    extrapolation_check=extrapolation_check(ref))
    from oracle import Oracle, OracleError
    oracle = Oracle()
    k = 20
    t0 = time.perf_counter()
    try:
        oracle.solve_flow(grid, w.mobility, 0, 100.0, 1e-30, k)
    except OracleError:
        pass
    t = time.perf_counter() - t0
    return dict(value=1.0 / (t / k * fp64_iterations), unit="solves/s", cores=1, kind="port",
                sample=f"{k} iterations of the C oracle (assembly included) in {t:.2f} s",
                seconds_per_iteration=t / k)


def extrapolation_check(ref):
    """The scaling above checked where a whole solve fits the budget: the
    reference's complete solve_depth_averaged_flow_cpu of the 1024 x 512
    pillar array (assembly, every iteration, velocity, scaling) at 1.2e-7,
    the tightest tolerance it accepts (float epsilon, cpu_flow.cpp:125-130),
    timed, against a tenth of its iterations timed and scaled the same way."""
    from oracle import Grid as OGrid
    from microsimulator_amd.workloads import config
    w2 = config("c2")
    grid = OGrid(w2.nx, w2.ny, 1, w2.dx, w2.dy, w2.dz, w2.obstacles)
    t0 = time.perf_counter()
    _, rep = ref.solve_flow(grid, w2.mobility, 0, 100.0, 1.2e-7, 400000)
    whole = time.perf_counter() - t0
    iterations = rep["iterations"]
    system = ref.flow_system(grid, w2.mobility, 0)
    system.time_iterations(5)
    k = max(10, iterations // 10)
    sample = system.time_iterations(k)
    scaled = sample / k * iterations
    return dict(workload="1024x512 pillar array", tolerance=1.2e-7, iterations=iterations,
                whole_solve_s=round(whole, 3), sample_iterations=k,
                scaled_s=round(scaled, 3), whole_over_scaled=round(whole / scaled, 3))


def cpu_adr_baseline(w, faces, threads: int, steps: int):
    """The reference's explicit transport (advance_signal_grid_cpu) for 3
    solutes with the Monod loss as its affine reaction, `steps` steps."""
    from oracle import FIXED, NO_FLUX, Reference
    from oracle import Grid as OGrid
    if not Reference.available():
        return None
    ref = Reference()
    ref.set_threads(threads)
    grid = OGrid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles,
                 boundary=(FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX))
    S = 3
    n = w.sites
    bv = np.zeros(6 * S, np.float32)
    bv[0::6] = 1.0
    loss = np.tile((np.float32(1.0) * w.phi.reshape(-1)) / np.float32(0.1 + 0.5), S)
    levels = np.full(S * n, 0.5, np.float32)
    levels[np.tile(w.obstacles.reshape(-1), S) != 0] = 0.0
    _, secs = ref.advance_signal_grid(grid, levels, 1e-6, steps,
                                      diffusion=np.array([100.0, 500.0, 1000.0], np.float32),
                                      boundary_values=bv, faces=faces, loss=loss)
    return dict(value=w.fluid * S * steps / secs / 1e9, unit="Gcell-upd/s",
                cores=ref.threads(), kind="reference",
                sample=f"{steps} steps of advance_signal_grid_cpu, 3 solutes, in {secs:.2f} s")


def implicit_bench(msa, hip, grid, w, bv, diffusion, dt, steps: int):
    """Crank-Nicolson steps of 3 solutes with Monod uptake over the solved
    field, from zero levels, one adr_run per step so that every step's
    Jacobi iterations are counted; dt is four times the explicit bound."""
    S = len(diffusion)
    spec = msa.SignalGridSpec(grid, diffusion, bv, phi=w.phi, monod_vmax=[1.0] * S,
                              monod_k=[0.1] * S, integration=msa.CRANK_NICOLSON)
    hip.adr_upload(spec, np.zeros(S * w.sites, np.float32), precision=32, use_flow_field=True)
    hip.adr_run(dt, 1, ledger=False)  # warm
    hip.adr_upload(spec, np.zeros(S * w.sites, np.float32), precision=32, use_flow_field=True)
    seconds, iterations, first = 0.0, 0, None
    for _ in range(steps):
        hip.adr_run(dt, 1, ledger=False)
        rep = hip.adr_report()
        seconds += hip.adr_seconds()
        iterations += rep.iterations
        first = rep.iterations if first is None else first
    # Per iteration the fused sweep reads each level, its right-hand side and
    # lagged loss and writes the next iterate (16 B per level), and reads the
    # site's two face velocities and mask (9 B).
    bytes_per_site = 16 * S + 9
    sweeps = iterations + steps
    out = {"integration": "crank_nicolson", "solutes": S, "steps": steps, "dt": dt,
           "jacobi_iterations": iterations, "first_step_iterations": first,
           "gcell_upd_per_s": round(w.fluid * S * steps / seconds / 1e9, 3),
           "seconds_per_step": round(seconds / steps, 5),
           "achieved_gbs_sweeps": round(bytes_per_site * w.sites * sweeps / seconds / 1e9, 1),
           "bytes_per_site_per_sweep": bytes_per_site}
    log(f"implicit transport: {out['gcell_upd_per_s']} Gcell-upd/s, {iterations} Jacobi "
        f"iterations over {steps} steps ({seconds / steps * 1e3:.2f} ms per step)")
    return out


def cpu_implicit_baseline(w, faces, threads: int, dt: float):
    """One Crank-Nicolson step of the reference (advance_signal_grid_cpu) from
    zero levels, 3 solutes, the Monod uptake of zero levels as its affine
    loss: the same first step the device takes."""
    from oracle import CRANK_NICOLSON, FIXED, NO_FLUX, Reference
    from oracle import Grid as OGrid
    if not Reference.available():
        return None
    ref = Reference()
    ref.set_threads(threads)
    grid = OGrid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles,
                 boundary=(FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX))
    S = 3
    bv = np.zeros(6 * S, np.float32)
    bv[0::6] = 1.0
    loss = np.tile((np.float32(1.0) * w.phi.reshape(-1)) / np.float32(0.1), S).astype(np.float32)
    loss[np.tile(w.obstacles.reshape(-1), S) != 0] = 0.0
    levels = np.zeros(S * w.sites, np.float32)
    _, secs, rep = ref.advance_signal_grid(grid, levels, dt, 1,
                                           diffusion=np.array([100.0, 500.0, 1000.0], np.float32),
                                           boundary_values=bv, faces=faces, loss=loss,
                                           integration=CRANK_NICOLSON, with_report=True)
    return dict(value=w.fluid * S / secs / 1e9, unit="Gcell-upd/s", cores=ref.threads(),
                kind="reference", iterations=rep["iterations"],
                sample=f"1 Crank-Nicolson step of advance_signal_grid_cpu, 3 solutes, "
                       f"{rep['iterations']} Jacobi iterations, in {secs:.2f} s")


def coupled_setup(w, count):
    """The coupled workload on the channel: nutrient at 1 in the fluid (AHL and
    by-product at 0), a colony of `count` cells, LuxI and GFP at 1."""
    from microsimulator_amd.workloads import colony, quorum_plan
    cells, previous = colony(w, count)
    S = 3
    levels = np.zeros(S * w.sites, np.float32)
    levels[:w.sites] = np.where(w.obstacles.reshape(-1) != 0, 0.0, 1.0)
    species = np.ones(2 * count, np.float32)
    return cells, previous, quorum_plan(), levels, species


def coupled_bench(msa, hip, grid, w, field, bv, diffusion, dt, steps: int, count: int):
    """Device-resident coupled steps (advance_coupled) over the solved field:
    the colony samples the three solutes, runs its rate plan, updates its two
    species and returns its exchange through the stencils; the grid steps by
    forward Euler at dt, then a few by Crank-Nicolson at 4 dt."""
    cells, previous, plan, levels, species = coupled_setup(w, count)
    out = {"cells": count, "solutes": 3, "species": 2, "instructions": len(plan[1])}
    for name, integration, step_dt, n in (("forward_euler", msa.FORWARD_EULER, dt, steps),
                                          ("crank_nicolson", msa.CRANK_NICOLSON, 4.0 * dt,
                                           max(steps // 5, 1))):
        spec = msa.SignalGridSpec(grid, diffusion, bv, field, integration=integration)
        hip.coupled_upload(spec, levels, cells, species, plan)
        hip.coupled_run(step_dt, 1, previous)  # warm
        hip.coupled_upload(spec, levels, cells, species, plan)
        rep = hip.coupled_run(step_dt, n, previous)
        secs = hip.coupled_seconds()
        out[name] = {"steps": n, "dt": step_dt, "seconds_per_step": round(secs / n, 6),
                     "steps_per_s": round(n / secs, 2),
                     "mcell_per_s": round(count * n / secs / 1e6, 1),
                     "grid_gcell_upd_per_s": round(w.fluid * 3 * n / secs / 1e9, 3)}
        if integration != msa.FORWARD_EULER:
            out[name]["last_step_iterations"] = rep.iterations
        log(f"coupled {name}: {count} cells, {secs / n * 1e3:.3f} ms per step "
            f"({count * n / secs / 1e6:.0f} Mcell/s)")
    # The colony's mobility (biomass deposition with the tent kernel of radius
    # 4, connected-fluid restriction, Kozeny-Carman drag): the flow input the
    # next pressure solve would take.
    hip.colony_mobility(grid, cells, base=1.0)  # warm
    t0 = time.perf_counter()
    hip.colony_mobility(grid, cells, base=1.0)
    call = time.perf_counter() - t0
    secs = hip.colony_seconds()
    out["colony_mobility"] = {"averaging_radius": 4.0, "device_seconds": round(secs, 6),
                              "mcell_per_s": round(count / secs / 1e6, 1),
                              "call_seconds_with_transfers": round(call, 4)}
    log(f"colony mobility: {count} cells, {secs * 1e3:.3f} ms on the device "
        f"({call * 1e3:.1f} ms per call with transfers)")
    return out


def cpu_colony_baseline(w, count: int, sample: int = 20000):
    """The numpy restatement of the reference's colony_volume_fraction (its
    Python loop over cells) on a sample of the colony, single-threaded."""
    from oracle import colony as ref
    from microsimulator_amd.workloads import colony
    cells, _ = colony(w, count)
    part = {k: v[:sample] for k, v in cells.items()}
    t0 = time.perf_counter()
    ref.colony_volume_fraction((w.nx, w.ny, 1), (0.0, 0.0, 0.0), (w.dx, w.dy, w.dz), w.obstacles,
                               part)
    secs = time.perf_counter() - t0
    return dict(value=round(sample / secs / 1e6, 5), unit="Mcell/s", cores=1, kind="port",
                sample=f"colony_volume_fraction of {sample} of the {count} cells (numpy "
                       f"restatement of the reference's Python loop) in {secs:.2f} s")


def cpu_coupled_baseline(w, faces, threads: int, dt: float, count: int):
    """One forward-Euler coupled step of the reference (advance_coupled_cpu)
    on the same colony and grid, timed around the call only."""
    from oracle import FIXED, NO_FLUX, Reference
    from oracle import Grid as OGrid
    if not Reference.available():
        return None
    ref = Reference()
    ref.set_threads(threads)
    cells, previous, plan, levels, species = coupled_setup(w, count)
    grid = OGrid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles,
                 boundary=(FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX))
    bv = np.zeros(18, np.float32)
    bv[0::6] = 1.0
    ref.advance_coupled(grid, levels, cells, species, plan, previous, dt,
                        diffusion=np.array([100.0, 500.0, 1000.0], np.float32),
                        boundary_values=bv, faces=faces)
    secs = ref.last_seconds
    return dict(value=round(1.0 / secs, 4), unit="coupled steps/s", cores=ref.threads(),
                kind="reference", mcell_per_s=round(count / secs / 1e6, 3),
                sample=f"1 forward-Euler step of advance_coupled_cpu, {count} cells, 3 solutes, "
                       f"in {secs:.2f} s")


def batch_sweep_bench(msa, hip, steps: int):
    """The batched sweep: 256 seeded instances drawn from 64x64, 128x96,
    256x128 and 512x256, each solved (FP64 true residual 1e-8), then `steps`
    graph-captured flow-check + velocity + FP32 transport steps."""
    from microsimulator_amd.workloads import batch_sweep
    inst = batch_sweep(256)
    fluid = sum(int((i.obstacles == 0).sum()) for i in inst)
    sites = sum(i.sites for i in inst)
    hip.batch_upload(inst, relative_tolerance=1e-8, max_iterations=60000)
    t0 = time.perf_counter()
    it, res, fl = hip.batch_solve_flow()
    flow_s = time.perf_counter() - t0
    worst = 0.0
    for k, i in enumerate(inst):
        f = hip.batch_download(k)
        speed = max(float(np.abs(f["x_faces"]).max()), float(np.abs(f["y_faces"]).max()))
        worst = max(worst, 2.0 * i.diffusion * (1 / i.dx ** 2 + 1 / i.dy ** 2) + 2.0 * speed / i.dx
                    + float(i.phi.max()) / 0.1)
    dt = 0.9 / worst
    shape, secs = hip.batch_autotune(dt, steps=3)
    hip.batch_run(dt, steps, graph=True, ledger=False)  # capture and warm
    led, run_s = hip.batch_run(dt, steps, graph=True)
    _, eager_s = hip.batch_run(dt, steps, graph=False, ledger=False)
    flags = hip.batch_flags()
    log(f"batch: {len(inst)} instances, {sites} sites, flow {flow_s:.3f} s (max {int(it.max())} "
        f"FP32 iterations), shape {hip.batch_shapes()[shape]}, {steps} graph steps "
        f"{run_s * 1e3:.2f} ms ({fluid * steps / run_s / 1e9:.2f} Gcell-upd/s), eager "
        f"{eager_s * 1e3:.2f} ms")
    return {
        "instances": len(inst), "sites": sites, "fluid_sites": fluid,
        "sizes": sorted({f"{i.nx}x{i.ny}" for i in inst}),
        "flow_solve_seconds": round(flow_s, 4), "flow_max_fp32_iterations": int(it.max()),
        "flow_max_true_relative_residual": float(res.max()),
        "autotuned_shape": {"threads": hip.batch_shapes()[shape][0],
                            "row_width": hip.batch_shapes()[shape][1],
                            "sites_per_item": hip.batch_shapes()[shape][2]},
        "autotune_step_us": [round(x * 1e6, 1) for x in secs],
        "steps": steps, "dt": dt,
        "graph_ms_per_step": round(run_s * 1e3 / steps, 4),
        "eager_ms_per_step": round(eager_s * 1e3 / steps, 4),
        "gcell_upd_per_s": round(fluid * steps / run_s / 1e9, 3),
        "worst_step_pressure_residual": float(led[:, :, 3].max()),
        "flags": int(flags.max()),
    }


def nutrient_loop_bench(msa, hip, grid, w, flow, steps: int):
    """The device-resident nutrient loop at the workload: per step the
    occupancy's Brinkman mobility, the headline pressure solve and velocity,
    one 3-solute FP64 transport step with Monod uptake and its mass ledger,
    and the occupancy grown by the uptake."""
    S = 3
    solid = w.obstacles.reshape(-1) != 0
    base = np.where(solid, 0.0, w.dz * w.dz / 12.0).astype(np.float32)
    hip.flow_prepare(grid, w.mobility, 0)
    rep = hip.flow_solve(flow)
    bv = np.zeros(6 * S, np.float32)
    bv[0::6] = 1.0
    spec = msa.SignalGridSpec(grid, [100.0, 500.0, 1000.0], bv, phi=w.phi, monod_vmax=[1.0] * S,
                              monod_k=[0.1] * S)
    hip.adr_upload(spec, np.zeros(S * w.sites), precision=64, use_flow_field=True)
    # 0.5 of the explicit bound with the first field; a tenth of the uptake
    # becomes biomass.
    dt = 0.5 / (2 * 1000.0 * 2 + 4.0 * rep.max_speed + 1.0 * float(w.phi.max()) / 0.1 * 1.5)
    t0 = time.perf_counter()
    ledger, records = hip.nutrient_run(flow, base, dt, steps, 100.0, 0.9, 0.1)
    elapsed = time.perf_counter() - t0
    solve_ms = [1e3 * r["solve_seconds"] for r in records]
    log(f"nutrient loop: {steps} steps in {elapsed:.2f} s ({steps / elapsed:.2f} steps/s), solve "
        f"{np.mean(solve_ms):.1f} ms/step, worst balance "
        f"{max(r['mass_balance'] for r in records):.2e}")
    return {
        "steps": steps, "steps_per_s": round(steps / elapsed, 3), "dt": dt,
        "solver": flow.solver, "mean_solve_ms": round(float(np.mean(solve_ms)), 3),
        "mean_step_ms": round(1e3 * float(np.mean([r["seconds"] for r in records])), 3),
        "mean_iterations": round(float(np.mean([r["iterations"] for r in records])), 2),
        "worst_true_relative_residual": max(r["true_relative_residual"] for r in records),
        "fell_back": any(r["fell_back"] for r in records),
        "worst_step_mass_balance": max(r["mass_balance"] for r in records),
        "max_phi_after": records[-1]["max_phi"], "solutes": S,
    }


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def port_flow_bench(msa, w, hip, cpu_seconds_per_iteration):
    """A port-driven flow on the same channel: two speed inlets splitting the
    lower x face (mean speeds 1 and 2) and the whole upper x face held at
    pressure 0, so two pressure solves per flow, combined on the device.
    Timed with the multilevel form and once with the reference's FP64
    Jacobi-PCG, both at the reference's default tolerance 1e-6. The CPU
    figure is the reference's seconds per FP64 iteration on this system (from
    the CPU baseline) times the FP64 iterations the two solves take."""
    grid = msa.Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    half = w.ny // 2
    ports = [msa.FlowPort(msa.X_LOWER, (0, 0), (half, 1), msa.SPEED, 1.0),
             msa.FlowPort(msa.X_LOWER, (half, 0), (w.ny, 1), msa.SPEED, 2.0),
             msa.FlowPort(msa.X_UPPER, (0, 0), (w.ny, 1), msa.PRESSURE, 0.0)]
    mg = msa.DepthAveragedPortFlowParameters(solver=msa.SOLVER_MULTIGRID)
    hip.solve_depth_averaged_port_flow(grid, ports, w.mobility, mg)  # warm
    count = 5
    t0 = time.perf_counter()
    runs = [hip.solve_depth_averaged_port_flow(grid, ports, w.mobility, mg) for _ in range(count)]
    mg_rate = count / (time.perf_counter() - t0)
    fp64 = msa.DepthAveragedPortFlowParameters(solver=msa.SOLVER_JACOBI_FP64,
                                               max_iterations=400000)
    t0 = time.perf_counter()
    ref_form = hip.solve_depth_averaged_port_flow(grid, ports, w.mobility, fp64)
    fp64_rate = 1.0 / (time.perf_counter() - t0)
    last = runs[-1]
    speeds = [round(p.mean_speed, 6) for p in last.ports]
    log(f"port flow: multigrid {mg_rate:.2f} flows/s ({last.report.iterations} it over "
        f"{last.report.solves} solves), fp64 {fp64_rate:.3f} flows/s "
        f"({ref_form.report.iterations} it), port mean speeds {speeds}")
    out = {
        "workload": f"same channel, ports: x_lower halves at speeds 1 and 2, x_upper at p=0",
        "solves_per_flow": last.report.solves,
        "flows_per_s_multigrid": round(mg_rate, 3),
        "multigrid_iterations": last.report.iterations,
        "multigrid_true_relative_residual": last.report.true_relative_residual,
        "flows_per_s_fp64_jacobi_pcg": round(fp64_rate, 4),
        "fp64_iterations": ref_form.report.iterations,
        "port_mean_speeds": speeds,
        "port_pressures": [round(p.pressure, 4) for p in last.ports],
        "max_speed": last.report.max_speed,
    }
    if cpu_seconds_per_iteration:
        cpu_s = cpu_seconds_per_iteration * ref_form.report.iterations
        out["cpu_reference_flows_per_s"] = round(1.0 / cpu_s, 5)
        out["cpu_reference_note"] = ("seconds per FP64 iteration of the reference's CPU solver "
                                     "on this system x the FP64 iterations of both solves")
    return out


def resolved_channel(msa, nx, ny, nz, seed=17):
    """A channel along x with 8 % seeded solid voxels (none on the inlet and
    outlet planes) and seeded Brinkman drag in [0, 2)."""
    rng = np.random.default_rng(seed)
    solid = (rng.random((nx, ny, nz)) < 0.08).astype(np.uint8)
    solid[0] = solid[-1] = 0
    kinds = (msa.FIXED, msa.FIXED, msa.NO_FLUX, msa.NO_FLUX, msa.NO_FLUX, msa.NO_FLUX)
    grid = msa.Grid(nx, ny, nz, 1.0, 1.0, 1.0, solid.reshape(-1), kinds)
    drag = (rng.random(nx * ny * nz) * 2.0).astype(np.float32)
    return grid, drag


def resolved_flow_bench(msa, hip, threads: int, with_cpu: bool):
    """The resolved Stokes-Brinkman solve (FP64 flexible GMRES over Jacobi-PCG
    momentum solves, the reference's algorithm) on a 64x32x16 channel, timed
    on the device and, on the same system, the reference's own CPU solver
    (oracle/_ref) over a whole solve; and on a 128x64x32 channel on the
    device only."""
    out = {}
    for name, dims, cpu in (("channel_64x32x16", (64, 32, 16), with_cpu),
                            ("channel_128x64x32", (128, 64, 32), False)):
        grid, drag = resolved_channel(msa, *dims)
        # The outer iterations grow with the grid: 404 at 64x32x16.
        params = msa.ResolvedFlowParameters(mean_inlet_speed=1.0, axis=0,
                                            max_outer_iterations=4000)
        field, rep = hip.solve_resolved_flow(grid, drag, params)  # buffers and code warm
        count = 3 if grid.nx <= 64 else 1
        t0 = time.perf_counter()
        for _ in range(count):
            field, rep = hip.solve_resolved_flow(grid, drag, params)
        rate = count / (time.perf_counter() - t0)
        entry = {"workload": f"{dims[0]}x{dims[1]}x{dims[2]} channel along x, 8% solid voxels, "
                             f"drag in [0, 2), mean inlet speed 1, tolerances 1e-6, "
                             f"at most 4000 outer iterations",
                 "faces": int(field.concatenated().size), "solves_per_s": round(rate, 4),
                 "outer_iterations": rep.outer_iterations,
                 "inner_iterations": rep.inner_iterations,
                 "device_seconds": round(rep.seconds, 4),
                 "relative_residual": rep.relative_residual,
                 "divergence_rms": rep.divergence_rms}
        log(f"resolved flow {name}: {rate:.3f} solves/s, {rep.outer_iterations} outer / "
            f"{rep.inner_iterations} inner iterations, device {rep.seconds:.3f} s")
        if cpu:
            from oracle import Grid as OGrid
            from oracle import Reference
            if Reference.available():
                ref = Reference()
                ref.set_threads(threads)
                og = OGrid(grid.nx, grid.ny, grid.nz, grid.dx, grid.dy, grid.dz, grid.obstacles,
                           grid.boundary)
                t0 = time.perf_counter()
                faces, rrep = ref.resolved_flow(og, drag, 0, 1.0, 1e-6, 4000, 1e-6, 50000)
                secs = time.perf_counter() - t0
                diff = float(np.abs(field.concatenated() - faces).max() / rrep["max_speed"])
                entry["cpu_reference"] = {
                    "solves_per_s": round(1.0 / secs, 5), "cores": ref.threads(),
                    "kind": "reference", "outer_iterations": rrep["outer_iterations"],
                    "inner_iterations": rrep["inner_iterations"],
                    "sample": "one whole solve of the same system by the reference's "
                              "solve_resolved_flow_cpu"}
                entry["max_field_difference_over_max_speed"] = diff
                log(f"resolved flow reference CPU: {secs:.2f} s on {ref.threads()} threads, "
                    f"{rrep['outer_iterations']} outer / {rrep['inner_iterations']} inner, "
                    f"field difference {diff:.2e}")
        out[name] = entry
        last = (grid, field)
    return out, last


def layered_transport_bench(msa, hip, grid, field, threads: int, with_cpu: bool, steps: int = 50):
    """Explicit FP32 transport of 3 solutes with Monod uptake on a grid of
    several layers, over the resolved flow's field on the same grid (the
    three-axis kernel k_adr3_step), and the reference's advance_signal_grid_cpu
    on the same grid and field for a few steps (Monod as the affine loss at
    level 0.5, the same work per site)."""
    S = 3
    n = grid.nx * grid.ny * grid.nz
    rng = np.random.default_rng(23)
    solid = np.asarray(grid.obstacles, np.uint8).reshape(-1)
    phi = (0.6 * rng.random(n)).astype(np.float32)
    phi[solid != 0] = 0.0
    bv = np.zeros(6 * S, np.float32)
    bv[0::6] = 1.0
    diffusion = np.array([1.0, 5.0, 10.0], np.float32)
    spec = msa.SignalGridSpec(grid, diffusion, bv, field, phi=phi, monod_vmax=[1.0] * S,
                              monod_k=[0.1] * S)
    outflow = 2.0 * sum(float(np.abs(f).max()) for f in (field.x_faces, field.y_faces,
                                                          field.z_faces))
    bound = 2.0 * 10.0 * 3 + outflow + float(phi.max()) / 0.1
    dt = 0.9 / bound
    levels = np.zeros(S * n, np.float32)
    hip.adr_upload(spec, levels, precision=32)
    hip.adr_run(dt, 2, ledger=False)  # warm
    hip.adr_upload(spec, levels, precision=32)
    hip.adr_run(dt, steps, ledger=False)
    secs = hip.adr_seconds()
    fluid = int((solid == 0).sum())
    gcups = fluid * S * steps / secs / 1e9
    bytes_per_site = 8 * S + 3 * 4 + 4 + 1  # levels in and out, three faces, phi, mask
    out = {"workload": f"{grid.nx}x{grid.ny}x{grid.nz} channel (8% solid voxels), the resolved "
                       f"flow's field, 3 solutes, D in {{1, 5, 10}}, Monod uptake, FP32, "
                       f"forward Euler at 0.9 of the explicit bound",
           "steps": steps, "dt": dt, "gcell_upd_per_s": round(gcups, 3),
           "achieved_gbs": round(bytes_per_site * n / (secs / steps) / 1e9, 1),
           "bytes_per_site": bytes_per_site}
    log(f"layered transport {grid.nx}x{grid.ny}x{grid.nz}: {gcups:.2f} Gcell-upd/s "
        f"({out['achieved_gbs']:.0f} GB/s)")
    if with_cpu:
        from oracle import Grid as OGrid
        from oracle import Reference
        if Reference.available():
            ref = Reference()
            ref.set_threads(threads)
            og = OGrid(grid.nx, grid.ny, grid.nz, grid.dx, grid.dy, grid.dz, grid.obstacles,
                       grid.boundary)
            loss = np.tile(phi / np.float32(0.1 + 0.5), S)  # zero at obstacles, as phi is
            faces = field.concatenated()
            cpu_steps = 3
            start = np.tile(np.where(solid != 0, 0.0, 0.5).astype(np.float32), S)
            _, cpu_secs = ref.advance_signal_grid(og, start, dt, cpu_steps, diffusion=diffusion,
                                                  boundary_values=bv, faces=faces, loss=loss)
            out["cpu_reference"] = {
                "gcell_upd_per_s": round(fluid * S * cpu_steps / cpu_secs / 1e9, 4),
                "cores": ref.threads(), "kind": "reference",
                "sample": f"{cpu_steps} steps of advance_signal_grid_cpu on the same grid and "
                          f"field, Monod as the affine loss at level 0.5, in {cpu_secs:.2f} s"}
            log(f"layered transport reference CPU: "
                f"{out['cpu_reference']['gcell_upd_per_s']} Gcell-upd/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--solver", default="learned", choices=("learned", "multigrid", "mixed"))
    ap.add_argument("--adr-steps", type=int, default=100)
    ap.add_argument("--fp64-solves", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--kernel-iterations", type=int, default=0,
                    help="profiling run: only N iterations of each solver form, then exit")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="profiling calibration: copy this many MiB device to device, 3 times")
    ap.add_argument("--no-ports", action="store_true", help="skip the port-driven flow")
    ap.add_argument("--no-resolved", action="store_true",
                    help="skip the resolved Stokes-Brinkman flow")
    ap.add_argument("--implicit-steps", type=int, default=5,
                    help="Crank-Nicolson transport steps to time (0: skip)")
    ap.add_argument("--coupled-steps", type=int, default=20,
                    help="device-resident coupled steps to time (0: skip)")
    ap.add_argument("--coupled-cells", type=int, default=1 << 20)
    ap.add_argument("--batch-steps", type=int, default=50,
                    help="graph-captured steps of the 256-instance batched sweep (0: skip)")
    ap.add_argument("--loop-steps", type=int, default=100,
                    help="steps of the device-resident nutrient loop (solve, velocity, "
                         "3-solute FP64 transport, occupancy growth) at the workload (0: skip)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.copy_probe} MiB x 3")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported by this build "
              "(the path is single-GPU; multi-GPU sharding is design only).", flush=True)
        sys.exit(2)

    import microsimulator_amd as msa
    from microsimulator_amd.workloads import config

    t0 = time.perf_counter()
    w = config(args.workload)
    log(f"workload {args.workload}: {w.nx}x{w.ny}, fluid {w.fluid}, pillars "
        f"{w.meta.get('pillars')}, generated in {time.perf_counter() - t0:.1f} s")
    grid = msa.Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    hip = msa.HipBackend(0)
    log("device", hip.info().device)
    hip.flow_prepare(grid, w.mobility, 0)
    from microsimulator_amd import learned
    hip.learned_set_weights(learned.load_weights())
    tol = 1e-8
    speed = 100.0
    def solver_params(kind, profile=False):
        solver = {"learned": msa.SOLVER_LEARNED, "multigrid": msa.SOLVER_MULTIGRID,
                  "mixed": msa.SOLVER_MIXED, "fp64": msa.SOLVER_JACOBI_FP64}[kind]
        return msa.DepthAveragedFlowParameters(mean_inlet_speed=speed, axis=0,
                                               relative_tolerance=tol, max_iterations=400000,
                                               solver=solver, profile=profile)
    if args.kernel_iterations:
        for kind in ("learned", "multigrid", "mixed", "fp64"):
            p = solver_params(kind)
            p.max_iterations = args.kernel_iterations
            try:
                hip.flow_solve(p)
            except msa.NotConverged:
                pass
        # A few transport steps of each form over a solved field.
        rep = hip.flow_solve(solver_params("multigrid"))
        S = 3
        bv = np.zeros(6 * S, np.float32)
        bv[0::6] = 1.0
        for scheme, prec in ((msa.UPWIND, 64), (msa.UPWIND, 32), (msa.VAN_LEER, 32)):
            spec = msa.SignalGridSpec(grid, [100.0, 500.0, 1000.0], bv, phi=w.phi,
                                      monod_vmax=[1.0] * S, monod_k=[0.1] * S,
                                      advection_scheme=scheme)
            hip.adr_upload(spec, np.zeros(S * w.sites), precision=prec, use_flow_field=True)
            hip.adr_run(1e-6, 5, ledger=prec == 64)
        implicit = msa.SignalGridSpec(grid, [100.0, 500.0, 1000.0], bv, phi=w.phi,
                                      monod_vmax=[1.0] * S, monod_k=[0.1] * S,
                                      integration=msa.CRANK_NICOLSON)
        hip.adr_upload(implicit, np.zeros(S * w.sites, np.float32), precision=32,
                       use_flow_field=True)
        hip.adr_run(8e-4, 2, ledger=False)
        if args.coupled_steps > 0:
            _, field_k = hip.flow_download()
            cells, previous, plan, levels, species = coupled_setup(w, args.coupled_cells)
            coupled_spec = msa.SignalGridSpec(grid, [100.0, 500.0, 1000.0], bv, field_k)
            hip.coupled_upload(coupled_spec, levels, cells, species, plan)
            hip.coupled_run(1e-6, 3, previous)
        # The resolved flow's kernels: 20 outer iterations on the 128x64x32 channel.
        grid3, drag3 = resolved_channel(msa, 128, 64, 32)
        try:
            hip.solve_resolved_flow(grid3, drag3, msa.ResolvedFlowParameters(
                mean_inlet_speed=1.0, axis=0, max_outer_iterations=20))
        except msa.NotConverged:
            pass
        log(f"ran {args.kernel_iterations} iterations of each solver form, 5 transport steps "
            f"of each explicit form, 2 implicit steps, 3 coupled steps and 20 outer iterations "
            f"of the resolved flow")
        return

    headline = solver_params(args.solver)
    for k in range(args.warmup):
        rep = hip.flow_solve(headline)
        log(f"warmup {k}: {rep.iterations} it, {rep.refinements} refinements, "
            f"{rep.seconds:.4f} s, true {rep.true_relative_residual:.2e}")
    reports = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        reports.append(hip.flow_solve(headline))
    elapsed = time.perf_counter() - t0
    for k, rep in enumerate(reports):
        log(f"step {k}: {rep.iterations} it, {rep.refinements} ref, {rep.seconds:.4f} s, "
            f"true {rep.true_relative_residual:.2e}")
    solves_per_s = args.steps / elapsed
    worst_true = max(r.true_relative_residual for r in reports)

    # The other forms on the same GPU: the cycle without the network, the
    # reference's algorithm (FP64 Jacobi-PCG) and FP32 Jacobi-PCG with FP64
    # refinement.
    others = {}
    for kind in ("multigrid", "fp64", "mixed"):
        if kind == args.solver:
            continue
        count = max(args.steps // 2, 1) if kind == "multigrid" else args.fp64_solves
        if kind == "multigrid":
            hip.flow_solve(solver_params(kind))  # warm
        t0 = time.perf_counter()
        reps = [hip.flow_solve(solver_params(kind)) for _ in range(count)]
        others[kind] = (count / (time.perf_counter() - t0), reps[-1])
        log(f"{kind}: {reps[-1].iterations} it, {reps[-1].seconds:.3f} s, "
            f"true {reps[-1].true_relative_residual:.2e}")
    fp64_solves_per_s, fp64_rep = others["fp64"]

    # One profiled solve of the headline form: every kernel of the finest
    # level timed with HIP events on the context's stream.
    prof = hip.flow_solve(solver_params(args.solver, profile=True))
    sites = w.sites
    mfma = None
    if args.solver in ("multigrid", "learned"):
        c = max(prof.cycles, 1)
        up = "k_learn_up" if args.solver == "learned" else "k_mg_up"
        per_launch = {"k_mg_down": prof.cycle_down_seconds / c,
                      up: prof.cycle_up_seconds / c,
                      "k_fcg_apply": prof.apply_seconds / max(prof.apply_launches, 1),
                      "k_fcg_update": prof.update_seconds / max(prof.update_launches, 1)}
        rest = prof.cycle_rest_seconds / c
        for name, sec in per_launch.items():
            log(f"{name}: avg {sec * 1e6:.1f} us, {MG_BYTES[name] * sites / sec / 1e9:.0f} GB/s")
        log(f"coarser levels + tail: avg {rest * 1e6:.1f} us per cycle; {prof.cycles} cycles")
        if args.solver == "learned":
            sec = per_launch["k_learn_up"]
            tflops = LEARNED_FLOPS * sites / sec / 1e12
            mfma = {"kernel": "k_learn_up", "bound": "mfma", "achieved": round(tflops, 2),
                    "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(tflops / BF16_PEAK_TFLOPS, 4),
                    "flops_per_site": LEARNED_FLOPS, "avg_launch_us": round(sec * 1e6, 2)}
            log(f"k_learn_up: {tflops:.1f} TFLOP/s of bf16 MFMA ({tflops / BF16_PEAK_TFLOPS:.3f}"
                f" of dense peak)")
        dominant = max(per_launch, key=per_launch.get)
        apply_avg = per_launch[dominant]
        bytes_per_site = MG_BYTES[dominant]
        kernel_name = dominant
        launches = prof.cycles
    else:
        apply_avg = prof.apply_seconds / max(prof.apply_launches, 1)
        bytes_per_site = apply_bytes(32)
        kernel_name = "k_cg_apply<float>"
        launches = prof.apply_launches
    achieved = bytes_per_site * sites / apply_avg / 1e9
    log(f"dominant {kernel_name}: avg {apply_avg * 1e6:.1f} us, {achieved:.0f} GB/s")
    pressure, field = hip.flow_download()

    # Transport over the solved field, FP64 state and ledger.
    S = 3
    bv = np.zeros(6 * S, np.float32)
    bv[0::6] = 1.0
    diffusion = np.array([100.0, 500.0, 1000.0], np.float32)
    spec = msa.SignalGridSpec(grid, diffusion, bv, phi=w.phi, monod_vmax=[1.0] * S,
                              monod_k=[0.1] * S)
    hip.adr_upload(spec, np.zeros(S * sites), precision=64, use_flow_field=True)
    # 0.9 of the explicit bound: 2 D sum(1/h^2) + greatest outflow + max loss.
    outflow_bound = 2.0 * float(np.abs(field.x_faces).max() + np.abs(field.y_faces).max())
    bound = 2 * 1000.0 * 2 + outflow_bound + 1.0 * float(w.phi.max()) / 0.1
    dt = 0.9 / bound
    ledger = hip.adr_run(dt, args.adr_steps)
    adr_s = hip.adr_seconds()
    balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
    scale = np.maximum(ledger[:, :, 1], 1e-300)
    worst_balance = float(np.max(np.abs(balance) / scale))
    adr_gcups = w.fluid * S * args.adr_steps / adr_s / 1e9
    adr_gbs = adr_bytes_per_site(S, 8) * sites / (adr_s / args.adr_steps) / 1e9
    hip.adr_upload(spec, np.zeros(S * sites, np.float32), precision=32, use_flow_field=True)
    hip.adr_run(dt, args.adr_steps, ledger=False)
    adr32_s = hip.adr_seconds()
    adr32_gcups = w.fluid * S * args.adr_steps / adr32_s / 1e9
    # The limited (van Leer) scheme: half the advective bound, the settle after
    # each step.
    limited = msa.SignalGridSpec(grid, diffusion, bv, phi=w.phi, monod_vmax=[1.0] * S,
                                 monod_k=[0.1] * S, advection_scheme=msa.VAN_LEER)
    dt_vl = 0.9 / (bound + outflow_bound)
    hip.adr_upload(limited, np.zeros(S * sites, np.float32), precision=32, use_flow_field=True)
    hip.adr_run(dt_vl, args.adr_steps, ledger=False)
    adr_vl_s = hip.adr_seconds()
    adr_vl_gcups = w.fluid * S * args.adr_steps / adr_vl_s / 1e9
    implicit = None
    if args.implicit_steps > 0:
        implicit = implicit_bench(msa, hip, grid, w, bv, diffusion, 4.0 * dt, args.implicit_steps)
    coupled = None
    if args.coupled_steps > 0:
        coupled = coupled_bench(msa, hip, grid, w, field, bv, diffusion, dt, args.coupled_steps,
                                args.coupled_cells)
    log(f"transport fp64: {adr_gcups:.2f} Gcell-upd/s ({adr_gbs:.0f} GB/s), worst per-step "
        f"imbalance {worst_balance:.2e}; fp32: {adr32_gcups:.2f} Gcell-upd/s; fp32 van Leer: "
        f"{adr_vl_gcups:.2f} Gcell-upd/s")

    line = {
        "metric": METRIC,
        "value": round(solves_per_s, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"pillar-array channel {w.nx}x{w.ny}, {w.meta.get('pillars')} pillars "
                         f"r~U[8,20], phi from {w.meta.get('blobs')} Gaussian blobs, "
                         f"Dirichlet 1/0 along x, rel-res 1e-8 (FP64 true residual)"),
            "solver": {
                "learned": ("FP32 flexible CG, multilevel preconditioner (2x2 aggregation, "
                            "V(2,2)) whose finest-level post-smoothing is one damped-Jacobi step "
                            "plus the learned bf16-MFMA correction, inside FP64 iterative "
                            "refinement; Jacobi-PCG fallback at 200 iterations"),
                "multigrid": ("FP32 flexible CG with the multilevel preconditioner (2x2 "
                              "aggregation, V(2,2)) inside FP64 iterative refinement"),
                "mixed": "FP32 Jacobi-PCG inside FP64 iterative refinement"}[args.solver],
            "sites": sites,
            "fluid_sites": w.fluid,
        },
        "solve": {
            "iterations": reports[-1].iterations,
            "refinements": reports[-1].refinements,
            "worst_true_relative_residual": worst_true,
            "device_seconds": reports[-1].seconds,
            "fell_back": any(r.fell_back for r in reports),
        },
        "fp64_jacobi_pcg": {
            "solves_per_s": round(fp64_solves_per_s, 5),
            "iterations": fp64_rep.iterations,
            "true_relative_residual": fp64_rep.true_relative_residual,
            "speedup_of_headline": round(solves_per_s / fp64_solves_per_s, 3),
        },
        "other_forms": {k: {"solves_per_s": round(v[0], 5), "iterations": v[1].iterations,
                            "true_relative_residual": v[1].true_relative_residual}
                        for k, v in others.items()},
        "adr": {
            "gcell_upd_per_s_fp64": round(adr_gcups, 3),
            "gcell_upd_per_s_fp32": round(adr32_gcups, 3),
            "gcell_upd_per_s_fp32_van_leer": round(adr_vl_gcups, 3),
            "dt_van_leer": dt_vl,
            "solutes": S,
            "steps": args.adr_steps,
            "dt": dt,
            "worst_step_mass_imbalance": worst_balance,
            "achieved_gbs_fp64": round(adr_gbs, 1),
        },
        "roofline": {
            "kernel": kernel_name,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (round(TRAFFIC_MB[kernel_name] * 1e6) if kernel_name in TRAFFIC_MB
                        else None),
            "traffic_unit": "bytes per launch",
            "bytes_per_site": bytes_per_site,
            "sites_per_launch": sites,
            "avg_launch_us": round(apply_avg * 1e6, 2),
            "launches": launches,
        },
    }
    if implicit is not None:
        line["adr_implicit"] = implicit
    if coupled is not None:
        line["coupled"] = coupled
    if mfma is not None:
        line["roofline_mfma"] = mfma
    if args.loop_steps > 0:
        line["nutrient_loop"] = nutrient_loop_bench(msa, hip, grid, w, headline, args.loop_steps)
    # The achievable HBM bandwidth of a plain device copy, and the dominant
    # kernel against it.
    ceiling = hip.copy_bandwidth(1 << 30, 10)
    line["roofline"]["copy_ceiling"] = round(ceiling, 1)
    line["roofline"]["frac_of_copy_ceiling"] = round(achieved / ceiling, 4)
    log(f"device copy: {ceiling:.0f} GB/s (read + written); dominant kernel "
        f"{achieved / ceiling:.3f} of it")
    if args.batch_steps > 0:
        line["batch"] = batch_sweep_bench(msa, hip, args.batch_steps)
    if args.solver == "learned":
        with np.load(learned.WEIGHTS_FILE, allow_pickle=False) as f:
            meta = json.loads(str(f["meta"]))
        line["learned"] = {
            "iterations_vs_multigrid": [reports[-1].iterations,
                                        others["multigrid"][1].iterations],
            # Recorded when the committed weights were trained; not timed here.
            "training_record": {
                "pairs": meta.get("total_pairs"), "geometries": meta.get("geometries"),
                "adam_steps": meta.get("steps"), "train_device_seconds": meta.get("train_seconds"),
                "harvest_seconds": meta.get("harvest_seconds"),
                "source": "microsimulator_amd/learned_smoother.npz metadata"},
        }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads ...")
        line["cpu_baseline"] = cpu_baseline(w, fp64_rep.iterations, threads, args.cpu_budget)
        line["cpu_baseline"]["cpu_model"] = cpu_model()
        line["cpu_baseline"]["host_cpus"] = os.cpu_count()
        faces = field.concatenated()
        adr_cpu = cpu_adr_baseline(w, faces, threads, steps=3)
        if adr_cpu is not None:
            line["cpu_baseline_adr"] = adr_cpu
        if implicit is not None:
            implicit_cpu = cpu_implicit_baseline(w, faces, threads, implicit["dt"])
            if implicit_cpu is not None:
                line["cpu_baseline_adr_implicit"] = implicit_cpu
        if coupled is not None:
            coupled_cpu = cpu_coupled_baseline(w, faces, threads, dt, args.coupled_cells)
            if coupled_cpu is not None:
                line["cpu_baseline_coupled"] = coupled_cpu
            line["cpu_baseline_colony"] = cpu_colony_baseline(w, args.coupled_cells)
    if not args.no_ports:
        per_iteration = line.get("cpu_baseline", {}).get("seconds_per_iteration")
        line["port_flow"] = port_flow_bench(msa, w, hip, per_iteration)
    if not args.no_resolved:
        threads = min(16, os.cpu_count() or 1)
        line["resolved_flow"], (grid3, field3) = resolved_flow_bench(msa, hip, threads,
                                                                     not args.no_cpu)
        line["adr_layered"] = layered_transport_bench(msa, hip, grid3, field3, threads,
                                                      not args.no_cpu)
    hip.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
