"""Weights and training driver of the learned smoother of the LEARNED solver.

The smoother (include/msim_hip.h, microsimulator_amd/csrc/learned.hip) is a
three-layer 3x3 convolutional network that adds s N(r') to the damped-Jacobi
post-smoother of the finest multilevel level. It has no counterpart in the
reference, whose only preconditioner is Jacobi (cpp/cpu/cpu_flow.cpp:65-69).

Training data come from the solver itself: for each seeded 512 x 256
geometry, a flexible-CG solve of its pressure system gives one residual per
iteration, an accurate solve gives the error of each, and the device turns
every (r, e) into the pair (r', e') the post-smoother sees. Training runs on
the device (bfloat16 MFMA forward and backward kernels, Adam).

    python -m microsimulator_amd.learned --out learned_smoother.npz

writes the weights, the loss curve and an iteration comparison on held-out
geometries. The weights the LEARNED solver loads by default are
``learned_smoother.npz`` beside this file.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import (LEARNED_WEIGHTS, SOLVER_LEARNED, SOLVER_MULTIGRID, DepthAveragedFlowParameters,
               Grid, HipBackend)
from .workloads import channel_trap, pillar_array

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_FILE = os.path.join(HERE, "learned_smoother.npz")
C0, C1 = 4, 16
SIZES = (9 * C0 * C1, 9 * C1 * C1, 9 * C1)


def initial_weights(seed: int = 0) -> np.ndarray:
    """W1, W2 ~ N(0, 1 / fan-in), W3 = 0: the untrained smoother is damped Jacobi."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, 1.0 / np.sqrt(9 * C0), SIZES[0])
    w2 = rng.normal(0.0, 1.0 / np.sqrt(9 * C1), SIZES[1])
    return np.concatenate([w1, w2, np.zeros(SIZES[2])]).astype(np.float32)


def load_weights(path: str = WEIGHTS_FILE) -> np.ndarray:
    """The trained weights (float32, LEARNED_WEIGHTS values)."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: train with `python -m microsimulator_amd.learned`")
    with np.load(path, allow_pickle=False) as f:
        w = np.asarray(f["weights"], np.float32).reshape(-1)
    if w.size != LEARNED_WEIGHTS:
        raise ValueError(f"{path} holds {w.size} weights, expected {LEARNED_WEIGHTS}")
    return w


def training_workload(index: int, nx: int = 512, ny: int = 256):
    """Geometry `index` of the training family: even indices are channel +
    trap devices, odd ones pillar arrays, each seeded by its index."""
    seed = 10_000 + index
    if index % 2 == 0:
        return channel_trap(nx, ny, seed=seed)
    return pillar_array(nx, ny, seed=seed, pillars=75, radius=(8, 20), blobs=10, depth=10.0)


def grid_of(w) -> Grid:
    return Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)


def harvest(hip: HipBackend, workloads, pairs: int, use_learned: bool = False, log=print):
    """(operators [G, 4, nx, ny], pair geometry [P], r' [P, nx, ny], e' [P, nx, ny])."""
    ops, geo, rs, es = [], [], [], []
    for g, w in enumerate(workloads):
        hip.flow_prepare(grid_of(w), w.mobility, axis=0)
        o = hip.learned_operator()
        r, e = hip.learned_harvest(pairs, use_learned=use_learned)
        ops.append(np.stack([o["wx"], o["wy"], o["bc"], o["dinv"]]))
        geo.extend([g] * len(r))
        rs.append(r)
        es.append(e)
        log(f"harvest {g + 1}/{len(workloads)} {w.name}: {len(r)} pairs", flush=True)
    return (np.stack(ops), np.asarray(geo, np.uint32), np.concatenate(rs), np.concatenate(es))


def iterations(hip: HipBackend, w, solver: int, tolerance: float = 1e-8):
    hip.flow_prepare(grid_of(w), w.mobility, axis=0)
    p = DepthAveragedFlowParameters(mean_inlet_speed=100.0, relative_tolerance=tolerance,
                                    solver=solver, check_interval=8)
    hip.flow_solve(p)  # warm
    t0 = time.perf_counter()
    rep = hip.flow_solve(p)
    return rep, time.perf_counter() - t0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--geometries", type=int, default=48)
    ap.add_argument("--pairs", type=int, default=24, help="pairs per geometry")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--chunk", type=int, default=250, help="steps between progress lines")
    ap.add_argument("--rounds", type=int, default=2,
                    help="harvest rounds; later rounds harvest with the learned smoother")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=WEIGHTS_FILE)
    ap.add_argument("--evaluate-large", action="store_true",
                    help="also compare on the 4096 x 2048 pillar array")
    args = ap.parse_args(argv)

    log = print
    with HipBackend(0) as hip:
        workloads = [training_workload(i) for i in range(args.geometries)]
        w = initial_weights(args.seed)
        hip.learned_set_weights(w)
        history = []
        sets = None
        harvest_seconds = train_seconds = 0.0
        for rnd in range(args.rounds):
            t0 = time.perf_counter()
            ops, geo, r, e = harvest(hip, workloads, args.pairs, use_learned=rnd > 0, log=log)
            harvest_seconds += time.perf_counter() - t0
            if sets is None:
                sets = (ops, geo, r, e)
            else:
                sets = (ops, np.concatenate([sets[1], geo]), np.concatenate([sets[2], r]),
                        np.concatenate([sets[3], e]))
            ops, geo, r, e = sets
            hip.learned_upload(ops, geo, r, e)
            log(f"round {rnd + 1}: {len(r)} pairs on {len(ops)} geometries", flush=True)
            steps = args.steps if rnd == 0 else args.steps // 2
            done = 0
            while done < steps:
                k = min(args.chunk, steps - done)
                losses, seconds = hip.learned_train(k, batch=args.batch, learning_rate=args.lr,
                                                    seed=args.seed * 1000 + rnd * 100 + done)
                train_seconds += seconds
                history.extend(losses.tolist())
                done += k
                log(f"round {rnd + 1} step {done}/{steps}: loss {losses.mean():.4f} "
                    f"({seconds * 1e3 / k:.2f} ms/step)", flush=True)
            w = hip.learned_weights()
        held_out = [channel_trap(512, 256, seed=4), pillar_array(512, 256, seed=99, pillars=75,
                                                                 radius=(8, 20), blobs=10,
                                                                 depth=10.0)]
        if args.evaluate_large:
            from .workloads import config
            held_out.append(config("c5"))
        evaluation = []
        for wl in held_out:
            mg, t_mg = iterations(hip, wl, SOLVER_MULTIGRID)
            ln, t_ln = iterations(hip, wl, SOLVER_LEARNED)
            evaluation.append(dict(workload=wl.name, multigrid=mg.iterations, learned=ln.iterations,
                                   multigrid_ms=t_mg * 1e3, learned_ms=t_ln * 1e3,
                                   fell_back=ln.fell_back))
            log(f"held-out {wl.name} {wl.nx}x{wl.ny}: multigrid {mg.iterations} it {t_mg * 1e3:.2f} ms, "
                f"learned {ln.iterations} it {t_ln * 1e3:.2f} ms", flush=True)
        meta = dict(geometries=args.geometries, pairs_per_geometry=args.pairs,
                    total_pairs=int(len(sets[2])), steps=args.steps, batch=args.batch,
                    lr=args.lr, rounds=args.rounds, harvest_seconds=harvest_seconds,
                    train_seconds=train_seconds, evaluation=evaluation)
        np.savez(args.out, weights=w, losses=np.asarray(history, np.float32),
                 meta=np.array(json.dumps(meta)))
        log(json.dumps(meta), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
