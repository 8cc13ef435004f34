"""The coupled cell/grid step on the GPU through the C ABI
(msim_hip_advance_coupled and the device-resident coupled_* calls), against
the CPU oracle, which is bit-exact with the reference's advance_coupled_cpu.

Tolerances. The stencils, samples, dilution, every arithmetic and comparison
instruction, the per-site source sums (in the caller's cell order, merged from
each site's lower-corner runs of the spatially ordered cells) and the
transport are the reference's float operations in its order, so grid levels,
species and the implicit report are compared bit for bit. The exponential, logarithm and power are not: the device evaluates
them in FP64 and rounds to float, the reference calls the C library's float
functions (within one ulp of each other), so plans that use them hold species
to a relative 1e-5 (the grid stays exact wherever the exchange does not use
them, as in the plans here).
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import Grid as OGrid
from oracle import OracleError
from tests.helpers import (NO_FLUX, OP_CONSTANT, OP_MULTIPLY, OP_SIGNAL, OP_SPECIES,
                           TRANSPORT_KINDS, random_cells, random_transport_case, uptake_plan)
from tests.test_coupled_oracle import refusal_base, refusal_cases

pytestmark = pytest.mark.gpu

INTEGRATIONS = [msa.FORWARD_EULER, msa.CRANK_NICOLSON, msa.BACKWARD_EULER]
SHAPES = [(20, 15, 2), (33, 9, 1), (12, 40, 3), (1, 25, 1), (64, 48, 2)]


def spec_of(grid, c, scheme, integration, faces=True):
    velocity = msa.VelocityField(c["xf"], c["yf"], c["zf"]) if faces else None
    return msa.SignalGridSpec(grid, c["diffusion"], c["bv"], velocity, source_rates=c["source"],
                              loss_rates=c["loss"], advection_scheme=scheme,
                              integration=integration)


def oracle_kw(c, scheme, integration, faces=True):
    kw = dict(diffusion=c["diffusion"], boundary_values=c["bv"], source=c["source"],
              loss=c["loss"], scheme=scheme, integration=integration)
    if faces:
        kw["faces"] = np.concatenate([c["xf"], c["yf"], c["zf"]])
    return kw


def same_report(rep, ref):
    assert rep.converged == ref["converged"]
    assert rep.iterations == ref["iterations"]
    assert abs(rep.residual_rms - ref["residual_rms"]) <= 2.0 ** -23 * abs(ref["residual_rms"])


def random_case(seed, nx, ny, S, kinds, integration, transcendental):
    c = random_transport_case(nx, ny, S, kinds, seed + 10 * integration)
    cells, prev = random_cells(seed, nx, ny, c["solid"], 40 + 7 * seed, 1.0, 0.8)
    K = 1 + seed % 3
    plan = uptake_plan(S, K, transcendental=transcendental)
    species = np.random.default_rng(seed).random(len(prev) * K).astype(np.float32)
    return c, cells, prev, plan, species


@pytest.mark.parametrize("integration", INTEGRATIONS)
@pytest.mark.parametrize("scheme", [msa.UPWIND, msa.VAN_LEER])
@pytest.mark.parametrize("kinds", TRANSPORT_KINDS)
def test_coupled_steps_match_the_oracle(hip, oracle, kinds, scheme, integration):
    for seed, (nx, ny, S) in enumerate(SHAPES):
        transcendental = seed % 2 == 0
        c, cells, prev, plan, species = random_case(seed, nx, ny, S, kinds, integration,
                                                    transcendental)
        dt = 0.02 if integration == msa.FORWARD_EULER else 0.1
        print("case", kinds, scheme, integration, (nx, ny, S), len(prev), flush=True)
        grid = msa.Grid(nx, ny, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
        og = OGrid(nx, ny, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
        ref_levels, ref_species, ref = oracle.advance_coupled(
            og, c["levels"], cells, species, plan, prev, dt, **oracle_kw(c, scheme, integration))
        levels, sp, rep = hip.advance_coupled(spec_of(grid, c, scheme, integration),
                                              c["levels"], cells, species, plan, prev, dt)
        assert np.array_equal(levels, ref_levels)
        if transcendental:
            assert np.allclose(sp, ref_species, rtol=1e-5, atol=1e-7)
        else:
            assert np.array_equal(sp, ref_species)
        same_report(rep, ref)


def test_bilinear_sample_and_scatter_known_answer(hip, oracle):
    """coupled_rates_test.cpp:45-72 on one layer: a cell at the centre of a
    2x2 box (spacing 2 x 1, level 4) samples 4, so species 10 + 0.5 * 4 = 12;
    it returns -2 / voxel volume 2 = -1 per unit time through weights 1/4:
    every level 4 - 0.5 / 4 = 3.875, grid amount 32 - 1 = 31."""
    grid = msa.Grid(2, 2, 1, 2.0, 1.0, 1.0, None, (NO_FLUX,) * 6)
    spec = msa.SignalGridSpec(grid, [0.0], None)
    cell = dict(position_x=[1.0], position_y=[0.5], position_z=[0.0], lengths=[2.0],
                radii=[0.5], growth_rates=[0.0], cell_types=[0])
    plan = (1, [(OP_SIGNAL, 0, 0, 0, 0.0), (OP_CONSTANT, 0, 0, 0, -2.0)], [0], [1])
    levels, sp, _ = hip.advance_coupled(spec, np.full(4, 4.0, np.float32), cell, [10.0], plan,
                                        [2.0], 0.5)
    assert sp[0] == 12.0
    assert np.all(levels == 3.875) and float(levels.sum()) * 2.0 == 31.0
    ref_levels, ref_sp, _ = oracle.advance_coupled(
        OGrid(2, 2, 1, 2.0, 1.0, 1.0, None, (NO_FLUX,) * 6), np.full(4, 4.0), cell, [10.0], plan,
        [2.0], 0.5, diffusion=(0.0,))
    assert np.array_equal(levels, ref_levels) and np.array_equal(sp, ref_sp)


def test_crank_nicolson_includes_coupled_sources(hip):
    """coupled_rates_test.cpp: one site of level 1 under Crank-Nicolson, a
    cell releasing 2 per unit time for a unit step: level 3."""
    grid = msa.Grid(1, 1, 1, 1.0, 1.0, 1.0, None, (NO_FLUX,) * 6)
    spec = msa.SignalGridSpec(grid, [0.0], None, integration=msa.CRANK_NICOLSON)
    cell = dict(position_x=[0.0], position_y=[0.0], position_z=[0.0], lengths=[2.0],
                radii=[0.5], growth_rates=[0.0], cell_types=[0])
    plan = (0, [(OP_CONSTANT, 0, 0, 0, 2.0)], [], [0])
    levels, _, rep = hip.advance_coupled(spec, np.array([1.0], np.float32), cell, [], plan, [2.0],
                                         1.0)
    assert rep.converged and abs(levels[0] - 3.0) <= 1e-6


def crowded_case():
    """make_crowded_reference (coupled_rates_conformance_test.cpp:120-178) on
    one layer: 400 cells on a few sites of a 6x4 grid with an obstacle, on
    lattice lines and on the upper faces (one-site stencil axes)."""
    nx, ny = 6, 4
    solid = np.zeros(nx * ny, np.uint8)
    solid[3 * ny + 2] = 1
    f32 = np.float32
    index = np.arange(2 * nx * ny)
    levels = (1.0 + f32(0.01) * (index % 13).astype(f32)).astype(f32)
    levels[np.tile(solid, 2) != 0] = 0.0
    ox, oy = f32(-1.0), f32(0.5)
    ex, ey = f32(1.5) * f32(5), f32(0.75) * f32(3)
    xs, ys = [], []
    for i in range(400):
        if i % 8 == 0:
            xs.append(ox + f32(1.5) * f32((i // 8) % 6))
            ys.append(oy + ey * f32(i % 5) / f32(4))
        elif i % 8 == 1:
            xs.append(ox + ex)
            ys.append(oy + ey)
        else:
            xs.append(ox + f32(1.5) * (f32(1) + f32(i % 37) / f32(37)))
            ys.append(oy + f32(0.75) * f32(i % 29) / f32(29))
    k = np.arange(400)
    cells = dict(position_x=np.array(xs, f32), position_y=np.array(ys, f32),
                 position_z=np.zeros(400, f32),
                 lengths=(f32(1.0) + f32(0.01) * (k % 31).astype(f32)).astype(f32),
                 radii=np.full(400, 0.3, f32),
                 growth_rates=(f32(0.01) * (k % 5).astype(f32)).astype(f32),
                 cell_types=np.zeros(400, np.int32))
    species = (f32(0.5) + f32(0.01) * (k % 29).astype(f32)).astype(f32)
    ins = [(OP_SPECIES, 0, 0, 0, 0.0), (OP_SIGNAL, 0, 0, 0, 0.0), (OP_CONSTANT, 0, 0, 0, 0.02),
           (OP_CONSTANT, 0, 0, 0, -0.01), (OP_MULTIPLY, 1, 2, 0, 0.0),
           (OP_MULTIPLY, 0, 2, 0, 0.0), (OP_MULTIPLY, 0, 3, 0, 0.0)]
    plan = (1, ins, [4], [5, 6])
    prev = (cells["lengths"] * f32(0.98)).astype(f32)
    return nx, ny, solid, levels, cells, species, plan, prev, (-1.0, 0.5, 0.0)


@pytest.mark.parametrize("integration", INTEGRATIONS)
def test_crowded_colony_matches_the_oracle(hip, oracle, integration):
    nx, ny, solid, levels, cells, species, plan, prev, origin = crowded_case()
    for scheme in (msa.UPWIND, msa.VAN_LEER):
        grid = msa.Grid(nx, ny, 1, 1.5, 0.75, 1.0, solid, (NO_FLUX,) * 6)
        spec = msa.SignalGridSpec(grid, [0.05, 0.02], None, advection_scheme=scheme,
                                  integration=integration)
        ours = hip.advance_coupled(spec, levels, cells, species, plan, prev, 0.05, origin)
        ref = oracle.advance_coupled(OGrid(nx, ny, 1, 1.5, 0.75, 1.0, solid, (NO_FLUX,) * 6),
                                     levels, cells, species, plan, prev, 0.05, origin=origin,
                                     diffusion=[0.05, 0.02], scheme=scheme,
                                     integration=integration)
        assert np.array_equal(ours[0], ref[0]) and np.array_equal(ours[1], ref[1])
        same_report(ours[2], ref[2])


def test_large_colony_matches_the_oracle(hip, oracle):
    """A 256x192 grid with 20000 cells, many sharing sites: the sort, the runs
    and the per-site sums at size."""
    nx, ny, S = 256, 192, 2
    kinds = TRANSPORT_KINDS[1]
    c = random_transport_case(nx, ny, S, kinds, 77)
    cells, prev = random_cells(77, nx, ny, c["solid"], 20000)
    plan = uptake_plan(S, 2, transcendental=False)
    species = np.random.default_rng(5).random(2 * 20000).astype(np.float32)
    for integration in INTEGRATIONS:
        dt = 0.02 if integration == msa.FORWARD_EULER else 0.1
        grid = msa.Grid(nx, ny, 1, 1.0, 1.0, 1.0, c["solid"], kinds)
        og = OGrid(nx, ny, 1, 1.0, 1.0, 1.0, c["solid"], kinds)
        print("integration", integration, flush=True)
        ref = oracle.advance_coupled(og, c["levels"], cells, species, plan, prev, dt,
                                     **oracle_kw(c, msa.UPWIND, integration))
        ours = hip.advance_coupled(spec_of(grid, c, msa.UPWIND, integration), c["levels"], cells,
                                   species, plan, prev, dt)
        assert np.array_equal(ours[0], ref[0]) and np.array_equal(ours[1], ref[1])
        same_report(ours[2], ref[2])


def test_no_cells_is_a_transport_step(hip):
    """A colony of no cells adds no sources: the step is advance_signal_grid's."""
    kinds = TRANSPORT_KINDS[0]
    c = random_transport_case(30, 20, 2, kinds, 8)
    grid = msa.Grid(30, 20, 1, 1.0, 1.0, 1.0, c["solid"], kinds)
    spec = spec_of(grid, c, msa.VAN_LEER, msa.FORWARD_EULER)
    empty = dict(position_x=[], position_y=[], position_z=[], lengths=[], radii=[],
                 growth_rates=[], cell_types=[])
    levels, sp, _ = hip.advance_coupled(spec, c["levels"], empty, [], uptake_plan(2, 0), [], 0.02)
    plain, _ = hip.advance_signal_grid(spec, c["levels"], 0.02)
    assert sp.size == 0 and np.array_equal(levels, plain)


@pytest.mark.parametrize("integration", INTEGRATIONS)
def test_zero_step_keeps_the_grid_and_dilutes_species(hip, oracle, integration):
    kinds = TRANSPORT_KINDS[0]
    c, cells, prev, plan, species = random_case(3, 16, 12, 2, kinds, integration, False)
    grid = msa.Grid(16, 12, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
    ours = hip.advance_coupled(spec_of(grid, c, msa.UPWIND, integration), c["levels"], cells,
                               species, plan, prev, 0.0)
    ref = oracle.advance_coupled(OGrid(16, 12, 1, 1.0, 0.8, 1.0, c["solid"], kinds), c["levels"],
                                 cells, species, plan, prev, 0.0,
                                 **oracle_kw(c, msa.UPWIND, integration))
    assert np.array_equal(ours[0], c["levels"].astype(np.float32))
    assert np.array_equal(ours[0], ref[0]) and np.array_equal(ours[1], ref[1])


@pytest.mark.parametrize("name,change", refusal_cases(), ids=[n for n, _ in refusal_cases()])
def test_refusals_match_the_oracle(hip, oracle, name, change):
    """Each refusal of the reference, with its message; nothing changes."""
    case = refusal_base()
    change(case)
    c = case["c"]
    og = OGrid(14, 11, 1, 1.0, 1.0, 1.0, c["solid"], case["kinds"])
    with pytest.raises(OracleError) as ref:
        oracle.advance_coupled(og, c["levels"], case["cells"], case["species"], case["plan"],
                               case["prev"], case["dt"], origin=case["origin"],
                               **oracle_kw(c, msa.UPWIND, msa.FORWARD_EULER))
    grid = msa.Grid(14, 11, 1, 1.0, 1.0, 1.0, c["solid"], case["kinds"])
    with pytest.raises(msa.MsimError) as got:
        hip.advance_coupled(spec_of(grid, c, msa.UPWIND, msa.FORWARD_EULER), c["levels"],
                            case["cells"], case["species"], case["plan"], case["prev"],
                            case["dt"], case["origin"])
    assert str(got.value) == ref.value.message
    assert got.value.status == ref.value.code


def test_cell_inside_an_obstacle_is_refused(hip):
    solid = np.zeros(16, np.uint8)
    solid[[5, 6, 9, 10]] = 1
    grid = msa.Grid(4, 4, 1, 1.0, 1.0, 1.0, solid, (NO_FLUX,) * 6)
    spec = msa.SignalGridSpec(grid, [0.1], None)
    levels = np.where(solid != 0, 0.0, 1.0).astype(np.float32)
    cell = dict(position_x=[1.5], position_y=[1.5], position_z=[0.0], lengths=[2.0],
                radii=[0.5], growth_rates=[0.0], cell_types=[0])
    with pytest.raises(msa.InvalidArgument) as got:
        hip.advance_coupled(spec, levels, cell, [1.0], (1, [(OP_SIGNAL, 0, 0, 0, 0.0)], [0], [0]),
                            [2.0], 0.1)
    assert str(got.value) == "signal sample position is inside a grid obstacle"


@pytest.mark.parametrize("integration", INTEGRATIONS)
def test_device_resident_steps_match_the_oracle(hip, oracle, integration):
    """coupled_run's steps against the oracle step by step: the first with the
    given previous lengths, the later ones with the cells' own."""
    kinds = TRANSPORT_KINDS[2]
    nx, ny, S = 48, 40, 2
    c, cells, prev, plan, species = random_case(4, nx, ny, S, kinds, integration, False)
    grid = msa.Grid(nx, ny, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
    og = OGrid(nx, ny, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
    dt = 0.02 if integration == msa.FORWARD_EULER else 0.1
    for scheme in (msa.UPWIND, msa.VAN_LEER):
        spec = spec_of(grid, c, scheme, integration)
        hip.coupled_upload(spec, c["levels"], cells, species, plan)
        rep = hip.coupled_run(dt, 4, prev)
        levels, sp = hip.coupled_download()
        ref_levels, ref_sp = c["levels"].astype(np.float32), species
        for step in range(4):
            ref_levels, ref_sp, ref = oracle.advance_coupled(
                og, ref_levels, cells, ref_sp, plan, prev if step == 0 else cells["lengths"], dt,
                **oracle_kw(c, scheme, integration))
        assert np.array_equal(levels, ref_levels) and np.array_equal(sp, ref_sp)
        same_report(rep, ref)


def test_refused_step_in_a_run_keeps_its_starting_state(hip, oracle):
    """A release so strong that the levels near the cells go negative after a
    few forward-Euler steps under uptake: the run stops at that step with the
    reference's message, holding the levels and species it started from."""
    nx, ny = 24, 20
    kinds = TRANSPORT_KINDS[0]
    c = random_transport_case(nx, ny, 1, kinds, 12)
    cells, prev = random_cells(12, nx, ny, c["solid"], 60)
    ins = [(OP_SPECIES, 0, 0, 0, 0.0), (OP_CONSTANT, 0, 0, 0, -3.0), (OP_CONSTANT, 0, 0, 0, 0.1)]
    plan = (1, ins, [2], [1])
    species = np.ones(60, np.float32)
    og = OGrid(nx, ny, 1, 1.0, 1.0, 1.0, c["solid"], kinds)
    kw = oracle_kw(c, msa.UPWIND, msa.FORWARD_EULER, faces=False)
    states = [(c["levels"].astype(np.float32), species)]
    message = None
    for step in range(40):
        try:
            lv, sp, _ = oracle.advance_coupled(og, states[-1][0], cells, states[-1][1], plan,
                                               cells["lengths"], 0.05, **kw)
        except OracleError as e:
            message = e.message
            break
        states.append((lv, sp))
    assert message is not None and 1 <= len(states) - 1 < 39
    failed = len(states) - 1
    grid = msa.Grid(nx, ny, 1, 1.0, 1.0, 1.0, c["solid"], kinds)
    spec = spec_of(grid, c, msa.UPWIND, msa.FORWARD_EULER, faces=False)
    hip.coupled_upload(spec, c["levels"], cells, species, plan)
    with pytest.raises(msa.MsimError) as got:
        hip.coupled_run(0.05, failed + 3)
    assert str(got.value) == message
    levels, sp = hip.coupled_download()
    assert np.array_equal(levels, states[failed][0]) and np.array_equal(sp, states[failed][1])


def test_a_new_grid_drops_the_coupled_state(hip):
    """Uploading another signal grid leaves no coupled state to run on: the
    run is refused instead of reading state sized for the old grid."""
    kinds = TRANSPORT_KINDS[0]
    c, cells, prev, plan, species = random_case(1, 20, 15, 2, kinds, msa.FORWARD_EULER, False)
    grid = msa.Grid(20, 15, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
    hip.coupled_upload(spec_of(grid, c, msa.UPWIND, msa.FORWARD_EULER), c["levels"], cells,
                       species, plan)
    hip.coupled_run(0.02, 1, prev)
    big = random_transport_case(64, 48, 2, kinds, 3)
    other = msa.Grid(64, 48, 1, 1.0, 1.0, 1.0, big["solid"], kinds)
    hip.adr_upload(spec_of(other, big, msa.UPWIND, msa.FORWARD_EULER),
                   big["levels"].astype(np.float32), precision=32)
    with pytest.raises(msa.InvalidArgument) as info:
        hip.coupled_run(0.02, 1)
    assert str(info.value) == "no coupled state on the device"


def test_device_matches_the_committed_fixture(hip):
    """The reference's own outputs (tests/golden/coupled.npz) for a one-layer
    version of its conformance colony, every integration and scheme: levels
    and species bit for bit, the implicit report as the oracle's."""
    import os
    from tests.helpers import coupled_colony_case
    path = os.path.join(os.path.dirname(__file__), "golden", "coupled.npz")
    g, levels, cells, species, plan, previous, transport = coupled_colony_case()
    with np.load(path, allow_pickle=False) as f:
        fx = {k: f[k] for k in f.files}
    grid = msa.Grid(g["nx"], g["ny"], 1, g["dx"], g["dy"], g["dz"], g["obstacles"], g["boundary"])
    nfx, nfy = (g["nx"] + 1) * g["ny"], g["nx"] * (g["ny"] + 1)
    faces = transport["faces"]
    velocity = msa.VelocityField(faces[:nfx], faces[nfx:nfx + nfy], faces[nfx + nfy:])
    for integration in INTEGRATIONS:
        for scheme in (msa.UPWIND, msa.VAN_LEER):
            key = f"i{integration}s{scheme}"
            print("case", key, flush=True)
            spec = msa.SignalGridSpec(grid, transport["diffusion"], transport["boundary_values"],
                                      velocity, source_rates=transport["source"],
                                      loss_rates=transport["loss"], advection_scheme=scheme,
                                      integration=integration)
            lv, sp, rep = hip.advance_coupled(spec, levels, cells, species, plan, previous, 0.05,
                                              (-3.0, -2.0, 0.0))
            assert np.array_equal(lv, fx[f"{key}_levels"]), key
            assert np.array_equal(sp, fx[f"{key}_species"]), key
            assert rep.converged == bool(fx[f"{key}_report"][0])
            assert rep.iterations == int(fx[f"{key}_report"][1])
            ref_rms = float(fx[f"{key}_report"][2])
            assert abs(rep.residual_rms - ref_rms) <= 2.0 ** -23 * abs(ref_rms)
