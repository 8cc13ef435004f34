"""The batched sweep on the GPU through the C ABI.

- Pressure: FP32 Jacobi-PCG corrections (FP64 dot products) inside FP64
  refinement of each instance's FP64 pressure, to FP64 true relative residual
  1e-10, against the FP64 oracle's solve of the same instance (mo_flow_solve,
  unscaled, times the inlet pressure, to 1e-13): the device's residual equals
  the one recomputed with the oracle's operator (mo_flow_apply) to 1e-6
  relative, and face velocities agree within 1e-5 of the largest speed.
- Transport: one batched step is the single-grid FP32 step bit for bit (the
  oracle's mo_adr_step_f32 on the faces the batched step used).
- Graph: replays of a captured run, with eager work between them, equal the
  eager run bit for bit.
- Ledger: the FP64 masses it records are the masses of the levels, and the
  pressure residual each step checks is the solve's.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from microsimulator_amd.workloads import batch_instance
from oracle import Grid as OGrid

pytestmark = pytest.mark.gpu

SEED = 11


def small_batch():
    inst = [batch_instance(k, SEED) for k in range(7)]
    assert len({(i.nx, i.ny) for i in inst}) >= 3  # uneven sizes
    return inst


def stable_dt(hip, instances):
    worst = 0.0
    for k, inst in enumerate(instances):
        f = hip.batch_download(k)
        speed = max(float(np.abs(f["x_faces"]).max()), float(np.abs(f["y_faces"]).max()))
        phi = 0.0 if inst.phi is None else float(np.max(inst.phi))
        bound = (2.0 * inst.diffusion * (1.0 / inst.dx ** 2 + 1.0 / inst.dy ** 2)
                 + 2.0 * speed / min(inst.dx, inst.dy) + 1.0 * phi / 0.1)
        worst = max(worst, bound)
    return 0.5 / worst


def ogrid(inst):
    return OGrid(inst.nx, inst.ny, 1, inst.dx, inst.dy, inst.dz, inst.obstacles)


def test_batch_flow_matches_the_oracle(hip, oracle):
    inst = small_batch()
    hip.batch_upload(inst, relative_tolerance=1e-10, max_iterations=40000)
    it, res, fl = hip.batch_solve_flow()
    print("iterations", it.tolist(), "residuals", res.tolist(), flush=True)
    assert np.all(fl == 0)
    assert np.all(res <= 1e-10)
    for k, i in enumerate(inst):
        g = ogrid(i)
        f = hip.batch_download(k)
        # The right-hand side for inlet pressure p: p times the assembled one, in FP64.
        b = oracle.assemble(g, i.mobility, 0)["rhs"].astype(np.float64) * np.float64(np.float32(i.inlet_pressure))
        true = np.linalg.norm(b - oracle.apply(g, i.mobility, f["pressure"])) / np.linalg.norm(b)
        faces, _, _ = oracle.solve_flow(g, i.mobility, 0, 1.0, 1e-13, 400000, scale=False)
        ref = faces.astype(np.float64) * i.inlet_pressure
        got = np.concatenate([f["x_faces"], f["y_faces"]]).astype(np.float64)
        n = got.size
        err = np.abs(got - ref[:n]).max() / np.abs(ref[:n]).max()
        print(k, (i.nx, i.ny), "FP64 true residual", true, "velocity error / max speed", err,
              flush=True)
        assert true <= 1e-10
        assert true == pytest.approx(float(res[k]), rel=1e-6, abs=1e-15)
        assert err <= 1e-5
    # A second solve starts from the converged pressures.
    it2, _, _ = hip.batch_solve_flow()
    assert np.all(it2 <= 2)


def test_batch_step_is_the_single_grid_step(hip, oracle):
    inst = small_batch()
    rng = np.random.default_rng(4)
    for i in inst:
        i.levels = rng.uniform(0.0, 1.0, i.sites).astype(np.float32) * (i.obstacles.reshape(-1) == 0)
    hip.batch_upload(inst, relative_tolerance=1e-8, max_iterations=40000)
    hip.batch_solve_flow()
    dt = stable_dt(hip, inst)
    hip.batch_run(dt, 1, graph=False)
    for k, i in enumerate(inst):
        f = hip.batch_download(k)
        bv = np.zeros(6, np.float32)
        bv[0] = i.reservoir
        zf = np.zeros(i.sites * 2, np.float32)
        ref = oracle.adr_step_f32(ogrid(i), i.levels, dt, diffusion=[i.diffusion],
                                  boundary_values=bv,
                                  faces=np.concatenate([f["x_faces"], f["y_faces"], zf]),
                                  phi=i.phi, vmax=1.0, monod_k=0.1)
        mismatch = int(np.count_nonzero(f["levels"] != ref))
        print(k, (i.nx, i.ny), "mismatching sites", mismatch, flush=True)
        assert mismatch == 0


def test_graph_replays_equal_the_eager_run(hip):
    inst = small_batch()
    hip.batch_upload(inst, relative_tolerance=1e-8, max_iterations=40000)
    hip.batch_solve_flow()
    dt = stable_dt(hip, inst)
    eager_led, _ = hip.batch_run(dt, 6, graph=False)
    eager_led2, _ = hip.batch_run(dt, 6, graph=False)
    eager = [hip.batch_download(k)["levels"] for k in range(len(inst))]

    hip.batch_upload(inst, relative_tolerance=1e-8, max_iterations=40000)
    hip.batch_solve_flow()
    led1, _ = hip.batch_run(dt, 6, graph=True)
    mid = [hip.batch_download(k)["levels"] for k in range(len(inst))]  # eager work between
    led2, _ = hip.batch_run(dt, 6, graph=True)                         # the same graph again
    final = [hip.batch_download(k)["levels"] for k in range(len(inst))]
    np.testing.assert_array_equal(led1, eager_led)
    np.testing.assert_array_equal(led2, eager_led2)
    for a, b in zip(final, eager):
        np.testing.assert_array_equal(a, b)
    assert any(np.any(a != b) for a, b in zip(mid, final))


def test_ledger_masses_are_the_levels(hip):
    inst = small_batch()
    hip.batch_upload(inst, relative_tolerance=1e-8, max_iterations=40000)
    hip.batch_solve_flow()
    dt = stable_dt(hip, inst)
    led, _ = hip.batch_run(dt, 5, graph=True)
    assert np.all(hip.batch_flags() == 0)
    assert np.all(led[:, :, 3] <= 1e-8)  # the pressure each step checked
    # Mass after one step is mass before the next, bit for bit.
    np.testing.assert_array_equal(led[1:, :, 0], led[:-1, :, 1])
    for k, i in enumerate(inst):
        lv = hip.batch_download(k)["levels"].astype(np.float64)
        mass = lv[i.obstacles.reshape(-1) == 0].sum() * i.dx * i.dy * i.dz
        assert led[-1, k, 1] == pytest.approx(mass, rel=1e-12)
        assert led[0, k, 0] == 0.0  # started empty
        assert led[-1, k, 1] > 0.0  # the reservoir feeds it


def test_autotune_picks_one_of_the_shapes(hip):
    inst = small_batch()
    hip.batch_upload(inst, relative_tolerance=1e-8, max_iterations=40000)
    hip.batch_solve_flow()
    dt = stable_dt(hip, inst)
    eager, _ = hip.batch_run(dt, 3, graph=False)
    ref = [hip.batch_download(k)["levels"] for k in range(len(inst))]
    hip.batch_upload(inst, relative_tolerance=1e-8, max_iterations=40000)
    hip.batch_solve_flow()
    shape, secs = hip.batch_autotune(dt, steps=2)
    print("shape", hip.batch_shapes()[shape], "per-step seconds", secs.tolist(), flush=True)
    assert 0 <= shape < len(hip.batch_shapes())
    assert np.all(secs > 0.0)
    assert secs[shape] == secs.min()
    hip.batch_run(dt, 3, graph=True)
    for k, a in enumerate(ref):  # any shape computes every site the same way
        np.testing.assert_array_equal(hip.batch_download(k)["levels"], a)


def test_batch_rejects_bad_instances(hip):
    inst = small_batch()[:2]
    inst[1].mobility = inst[1].mobility.copy()
    inst[1].mobility[3] = -1.0
    with pytest.raises(msa.InvalidArgument):
        hip.batch_upload(inst)
    with pytest.raises(msa.InvalidArgument):
        hip.batch_upload(inst[:1], relative_tolerance=0.0)
