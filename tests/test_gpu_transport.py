"""The explicit transport step on the GPU through the C ABI.

FP32 steps are bit-exact with the oracle, which is bit-exact with the
reference (same float operations, no contraction). FP64 steps agree with the
oracle's FP64 face-flux step to 1e-13 relative and conserve mass to 1e-12
relative per step.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import VAN_LEER
from oracle import Grid as OGrid
from tests.helpers import (FIXED, NO_FLUX, PATCH_KINDS, PERIODIC, TRANSPORT_KINDS,
                           TRANSPORT_SHAPES, c1_inputs, golden, random_patches,
                           random_transport_case, split_faces)

pytestmark = pytest.mark.gpu


def c1_spec(revision=0):
    c = c1_inputs()
    d = golden("c1_adr.npz")
    faces = c["data"]["faces"]
    xf, yf, zf = split_faces(faces, c["nx"], c["ny"])
    grid = msa.Grid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"])
    spec = msa.SignalGridSpec(grid, d["diffusion"], d["boundary_values"],
                              msa.VelocityField(xf, yf, zf), phi=c["phi"],
                              monod_vmax=[float(d["vmax"])], monod_k=[float(d["K"])],
                              revision=revision)
    return spec, c, d


def test_monod_transport_matches_reference_fixture(hip):
    spec, c, d = c1_spec(revision=101)
    levels = np.zeros(spec.grid.sites, np.float32)
    for _ in range(int(d["steps"])):
        levels, rep = hip.advance_signal_grid(spec, levels, float(d["dt"]))
    assert rep.converged
    assert np.array_equal(levels, d["levels"])


def test_device_resident_run_equals_single_steps(hip):
    spec, c, d = c1_spec(revision=102)
    levels = np.zeros(spec.grid.sites, np.float32)
    hip.adr_upload(spec, levels, precision=32)
    hip.adr_run(float(d["dt"]), int(d["steps"]), ledger=False)
    assert np.array_equal(hip.adr_download(), d["levels"])


@pytest.mark.parametrize("kinds", [
    (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX),
    (FIXED, NO_FLUX, PERIODIC, PERIODIC, NO_FLUX, NO_FLUX),
    (PERIODIC, PERIODIC, FIXED, FIXED, NO_FLUX, NO_FLUX),
])
def test_affine_transport_is_bit_exact(hip, oracle, kinds):
    rng = np.random.default_rng(5)
    nx, ny, S = 40, 70, 3
    solid = (rng.random(nx * ny) < 0.15).astype(np.uint8)
    levels = rng.random(S * nx * ny).astype(np.float32)
    mask = np.tile(solid, S) != 0
    levels[mask] = 0
    loss = (rng.random(S * nx * ny) * 0.5).astype(np.float32)
    source = (rng.random(S * nx * ny) * 0.2).astype(np.float32)
    loss[mask] = 0
    source[mask] = 0
    bv = rng.random(6 * S).astype(np.float32)
    diffusion = np.array([0.3, 0.05, 0.1], np.float32)
    grid = msa.Grid(nx, ny, 1, 1.0, 0.5, 1.0, solid, kinds)
    spec = msa.SignalGridSpec(grid, diffusion, bv, source_rates=source, loss_rates=loss)
    ours = levels
    for _ in range(3):
        ours, _ = hip.advance_signal_grid(spec, ours, 0.05)
    og = OGrid(nx, ny, 1, 1.0, 0.5, 1.0, solid, kinds)
    ref = oracle.adr_step_f32(og, levels, 0.05, steps=3, diffusion=diffusion,
                              boundary_values=bv, source=source, loss=loss)
    assert np.array_equal(ours, ref)


# Shapes cover one and several column strips, a last strip of one column,
# strips of 64, 128 and 256 columns, and axes of extent one.
SHAPES = TRANSPORT_SHAPES
KINDS = TRANSPORT_KINDS


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kinds", KINDS)
def test_marching_step_on_every_shape(hip, oracle, shape, kinds):
    """FP32 bit for bit and FP64 to 1e-13 against the oracle, with a face
    field, affine rates and Monod uptake, on strips of every width."""
    nx, ny = shape
    S = 2
    d = random_transport_case(nx, ny, S, kinds, seed=nx * 1000 + ny)
    grid = msa.Grid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], kinds)
    spec = msa.SignalGridSpec(grid, d["diffusion"], d["bv"],
                              msa.VelocityField(d["xf"], d["yf"], d["zf"]),
                              source_rates=d["source"], loss_rates=d["loss"], phi=d["phi"],
                              monod_vmax=[0.7, 0.3], monod_k=[0.2, 0.5])
    og = OGrid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], kinds)
    kw = dict(diffusion=d["diffusion"], boundary_values=d["bv"], source=d["source"],
              loss=d["loss"], faces=np.concatenate([d["xf"], d["yf"], d["zf"]]), phi=d["phi"],
              vmax=[0.7, 0.3], monod_k=[0.2, 0.5])
    dt = 0.02
    print("case", shape, kinds, flush=True)
    lv32 = d["levels"].astype(np.float32)
    hip.adr_upload(spec, lv32, precision=32)
    hip.adr_run(dt, 3, ledger=False)
    ref32 = oracle.adr_step_f32(og, lv32, dt, steps=3, **kw)
    assert np.array_equal(hip.adr_download(), ref32)
    hip.adr_upload(spec, d["levels"], precision=64)
    ledger = hip.adr_run(dt, 3)
    ref64, ref_ledger = oracle.adr_step_f64(og, d["levels"], dt, steps=3, **kw)
    ours = hip.adr_download()
    assert np.abs(ours - ref64).max() <= 1e-13 * np.abs(ref64).max()
    assert np.allclose(ledger, ref_ledger, rtol=1e-12, atol=1e-300)
    balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
    assert np.abs(balance).max() <= 1e-12 * ledger[:, :, 1].max()


VAN_LEER_SHAPES = TRANSPORT_SHAPES + [(2, 30), (30, 2), (3, 17)]


@pytest.mark.parametrize("shape", VAN_LEER_SHAPES)
@pytest.mark.parametrize("kinds", KINDS)
def test_van_leer_step_on_every_shape(hip, oracle, shape, kinds):
    """The limited (van Leer) step: FP32 bit for bit with the oracle (which is
    bit-exact with the reference's CPU step), FP64 to 1e-13 with a balanced
    ledger, on every boundary kind, axes of two and three sites and strips of
    every width."""
    nx, ny = shape
    d = random_transport_case(nx, ny, 2, kinds, seed=nx * 1000 + ny)
    grid = msa.Grid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], kinds)
    spec = msa.SignalGridSpec(grid, d["diffusion"], d["bv"],
                              msa.VelocityField(d["xf"], d["yf"], d["zf"]),
                              source_rates=d["source"], loss_rates=d["loss"], phi=d["phi"],
                              monod_vmax=[0.7, 0.3], monod_k=[0.2, 0.5],
                              advection_scheme=msa.VAN_LEER)
    og = OGrid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], kinds)
    kw = dict(diffusion=d["diffusion"], boundary_values=d["bv"], source=d["source"],
              loss=d["loss"], faces=np.concatenate([d["xf"], d["yf"], d["zf"]]), phi=d["phi"],
              vmax=[0.7, 0.3], monod_k=[0.2, 0.5], scheme=VAN_LEER)
    dt = 0.02
    print("case", shape, kinds, flush=True)
    lv32 = d["levels"].astype(np.float32)
    ref32 = oracle.adr_step_f32(og, lv32, dt, steps=3, **kw)
    hip.adr_upload(spec, lv32, precision=32)
    hip.adr_run(dt, 3, ledger=False)
    assert np.array_equal(hip.adr_download(), ref32)
    one, _ = hip.advance_signal_grid(spec, lv32, dt)
    assert np.array_equal(one, oracle.adr_step_f32(og, lv32, dt, **kw))
    hip.adr_upload(spec, d["levels"], precision=64)
    ledger = hip.adr_run(dt, 3)
    ref64, ref_ledger = oracle.adr_step_f64(og, d["levels"], dt, steps=3, **kw)
    ours = hip.adr_download()
    assert np.abs(ours - ref64).max() <= 1e-13 * np.abs(ref64).max()
    assert np.allclose(ledger, ref_ledger, rtol=1e-12, atol=1e-300)
    balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
    assert np.abs(balance).max() <= 1e-12 * ledger[:, :, 1].max()


def test_van_leer_front_and_bound(hip, oracle):
    """A step front carried 40 steps at half the limited bound matches the
    oracle bit for bit and stays within [0, 1]; a step between the limited
    and the upwind bounds is refused only under van Leer."""
    nx, ny = 300, 70
    lv = np.zeros(nx * ny, np.float32)
    lv.reshape(nx, ny)[:40] = 1.0
    xf = np.ones((nx + 1) * ny, np.float32)
    xf[:ny] = 0.0
    xf[nx * ny:] = 0.0
    yf = np.zeros(nx * (ny + 1), np.float32)
    zf = np.zeros(2 * nx * ny, np.float32)
    grid = msa.Grid(nx, ny, 1, 1.0, 1.0, 1.0, None, (NO_FLUX,) * 6)
    spec = msa.SignalGridSpec(grid, [0.0], velocity=msa.VelocityField(xf, yf, zf),
                              advection_scheme=msa.VAN_LEER)
    og = OGrid(nx, ny, 1, 1.0, 1.0, 1.0, None, (NO_FLUX,) * 6)
    ref = oracle.adr_step_f32(og, lv, 0.49, steps=40, scheme=VAN_LEER, diffusion=0.0,
                              faces=np.concatenate([xf, yf, zf]))
    hip.adr_upload(spec, lv, precision=32)
    hip.adr_run(0.49, 40, ledger=False)
    ours = hip.adr_download()
    assert np.array_equal(ours, ref)
    assert ours.min() >= 0.0 and ours.max() <= 1.0
    with pytest.raises(msa.InvalidArgument, match="explicit stability bound"):
        hip.advance_signal_grid(spec, lv, 0.51)
    upwind = msa.SignalGridSpec(grid, [0.0], velocity=msa.VelocityField(xf, yf, zf))
    hip.advance_signal_grid(upwind, lv, 0.51)


@pytest.mark.parametrize("kinds", PATCH_KINDS)
def test_boundary_patches_on_device(hip, oracle, kinds):
    """Reservoir patches of fixed faces: FP32 bit for bit with the oracle
    (bit-exact with the reference's values_at), upwind and van Leer; FP64 to
    1e-13 with the patches' supply in a balanced ledger."""
    nx, ny, S = 130, 70, 2
    d = random_transport_case(nx, ny, S, kinds, seed=5)
    patches = random_patches(kinds, nx, ny, S, seed=1)
    og = OGrid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], kinds)
    grid = msa.Grid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], kinds)
    kw = dict(diffusion=d["diffusion"], boundary_values=d["bv"], source=d["source"],
              loss=d["loss"], faces=np.concatenate([d["xf"], d["yf"], d["zf"]]))
    for scheme in (msa.UPWIND, msa.VAN_LEER):
        print("case", kinds, "scheme", scheme, flush=True)
        spec = msa.SignalGridSpec(grid, d["diffusion"], d["bv"],
                                  msa.VelocityField(d["xf"], d["yf"], d["zf"]),
                                  source_rates=d["source"], loss_rates=d["loss"],
                                  advection_scheme=scheme, patches=patches)
        lv32 = d["levels"].astype(np.float32)
        hip.adr_upload(spec, lv32, precision=32)
        hip.adr_run(0.02, 3, ledger=False)
        ref32 = oracle.adr_step_f32(og, lv32, 0.02, steps=3, scheme=scheme, patches=patches, **kw)
        assert np.array_equal(hip.adr_download(), ref32)
        hip.adr_upload(spec, d["levels"], precision=64)
        ledger = hip.adr_run(0.02, 3)
        ref64, ref_ledger = oracle.adr_step_f64(og, d["levels"], 0.02, steps=3, scheme=scheme,
                                                patches=patches, **kw)
        assert np.abs(hip.adr_download() - ref64).max() <= 1e-13 * np.abs(ref64).max()
        assert np.allclose(ledger, ref_ledger, rtol=1e-12, atol=1e-300)


def test_boundary_patch_refusals(hip):
    """The reference's patch checks (GridBoundary::validate and the face
    extents, signals.cpp:449-479, 734-744), with its messages."""
    grid = msa.Grid(10, 6, 1, boundary=(FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX))
    cases = [((2, (0, 0), (3, 1), [0.5]), "non-fixed signal grid boundaries cannot have values"),
             ((0, (2, 0), (2, 1), [0.5]), "must hold at least one site"),
             ((0, (0, 0), (3, 1), [-1.0]), "must be finite and non-negative"),
             ((1, (4, 0), (7, 1), [0.5]), "must lie on their face")]
    for patch, message in cases:
        spec = msa.SignalGridSpec(grid, [0.1], [1.0, 0, 0, 0, 0, 0], patches=(patch,))
        with pytest.raises(msa.InvalidArgument, match=message):
            hip.advance_signal_grid(spec, np.zeros(60, np.float32), 0.1)


def test_explicit_bound_refuses_and_leaves_levels(hip):
    line = msa.Grid(3, 1, 1, boundary=(NO_FLUX,) * 6)
    spec = msa.SignalGridSpec(line, [1.0])
    levels, _ = hip.advance_signal_grid(spec, [0.0, 1.0, 0.0], 0.25)
    assert np.allclose(levels, [0.25, 0.5, 0.25], atol=1e-6)
    with pytest.raises(msa.InvalidArgument, match="explicit stability bound"):
        hip.advance_signal_grid(spec, levels, 0.51)
    one = msa.Grid(1, 1, 1, boundary=(NO_FLUX,) * 6)
    spec = msa.SignalGridSpec(one, [0.0], source_rates=[2.0], loss_rates=[0.5])
    c, _ = hip.advance_signal_grid(spec, [1.0], 0.5)
    assert abs(c[0] - 1.75) <= 1e-6
    with pytest.raises(msa.InvalidArgument):
        hip.advance_signal_grid(spec, [-1.0], 0.5)


def test_fp64_transport_conserves_and_matches_oracle(hip, oracle):
    spec, c, d = c1_spec(revision=103)
    hip.adr_upload(spec, np.zeros(spec.grid.sites), precision=64)
    ledger = hip.adr_run(float(d["dt"]), 20)
    ours = hip.adr_download()
    og = OGrid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"])
    ref, ref_ledger = oracle.adr_step_f64(og, np.zeros(spec.grid.sites), float(d["dt"]),
                                          steps=20, diffusion=d["diffusion"],
                                          boundary_values=d["boundary_values"],
                                          faces=c["data"]["faces"], phi=c["phi"],
                                          vmax=d["vmax"], monod_k=d["K"])
    assert np.abs(ours - ref).max() <= 1e-13 * np.abs(ref).max()
    balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
    print("worst per-step imbalance / mass", np.abs(balance).max() / ledger[:, :, 1].max(),
          flush=True)
    assert np.abs(balance).max() <= 1e-12 * ledger[:, :, 1].max()
    assert np.allclose(ledger, ref_ledger, rtol=1e-12, atol=1e-300)


def test_coupled_flow_then_transport_on_device(hip):
    """Flow solve, then three solutes advected by the field left on the device,
    FP64 state, mass ledger every step."""
    from microsimulator_amd.workloads import config
    w = config("c2")
    g = msa.Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    hip.flow_prepare(g, w.mobility, 0)
    rep = hip.flow_solve(msa.DepthAveragedFlowParameters(mean_inlet_speed=100.0, axis=0,
                                                         relative_tolerance=1e-8,
                                                         solver=msa.SOLVER_MIXED))
    S = 3
    bv = np.zeros(6 * S, np.float32)
    bv[0::6] = 1.0
    diffusion = np.array([100.0, 500.0, 1000.0], np.float32)
    spec = msa.SignalGridSpec(g, diffusion, bv, phi=w.phi, monod_vmax=[1.0] * S,
                              monod_k=[0.1] * S)
    hip.adr_upload(spec, np.zeros(S * g.sites), precision=64, use_flow_field=True)
    bound = 2 * 1000.0 * 2 + 2 * rep.max_speed * 2 + 7.0 / 0.1 * 1.0
    dt = 0.5 / bound
    ledger = hip.adr_run(dt, 50)
    levels = hip.adr_download().reshape(S, -1)
    balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
    assert np.abs(balance).max() <= 1e-12 * ledger[:, :, 1].max()
    assert np.all(levels >= 0) and np.all(levels <= 1.0 + 1e-12)
    assert np.all(levels[:, w.obstacles.reshape(-1) != 0] == 0)


def test_transport_over_the_device_field_refused_after_another_solve(hip):
    """A transport reading the flow field in place is refused once a later
    solve (same grid, ten times the speed) has rewritten that field: its
    explicit bound was taken from the old one."""
    from microsimulator_amd.workloads import channel_trap
    w = channel_trap(128, 64, seed=3)
    g = msa.Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    hip.flow_prepare(g, w.mobility, 0)
    params = msa.DepthAveragedFlowParameters(mean_inlet_speed=10.0, axis=0,
                                             relative_tolerance=1e-8,
                                             solver=msa.SOLVER_MULTIGRID)
    rep = hip.flow_solve(params)
    bv = np.zeros(6, np.float32)
    bv[0] = 1.0
    spec = msa.SignalGridSpec(g, [500.0], bv)
    hip.adr_upload(spec, np.zeros(g.sites, np.float32), precision=32, use_flow_field=True)
    dt = 0.5 / (2 * 500.0 * 2 + 2 * rep.max_speed * 2)
    hip.adr_run(dt, 2, ledger=False)
    params.mean_inlet_speed = 100.0
    hip.flow_solve(params)
    with pytest.raises(msa.InvalidArgument, match="no longer the transport's"):
        hip.adr_run(dt, 1, ledger=False)
    # A fresh upload over the new field runs.
    hip.adr_upload(spec, np.zeros(g.sites, np.float32), precision=32, use_flow_field=True)
    hip.adr_run(dt / 10, 1, ledger=False)
