"""The oracle's coupled cell/grid step (mo_advance_coupled_f32) pinned against
the reference's advance_coupled_cpu (cpp/cpu/cpu_coupled.cpp:125-231): bit for
bit (grid levels, cell species, implicit report) through the live oracle/_ref
build, on the reference's own conformance colonies
(tests/conformance/coupled_rates_conformance_test.cpp:18-178), and against the
known answers of tests/cpp/coupled_rates_test.cpp.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import (BACKWARD_EULER, CRANK_NICOLSON, FORWARD_EULER, NO_FLUX, PERIODIC, UPWIND,
                    VAN_LEER, Grid, OracleError)
from tests.helpers import (OP_ADD, OP_CONSTANT, OP_LOGARITHM, OP_MULTIPLY, OP_SIGNAL, OP_SPECIES,
                           TRANSPORT_KINDS, random_cells, random_transport_case, uptake_plan)

FIXED = 2
INTEGRATIONS = [FORWARD_EULER, CRANK_NICOLSON, BACKWARD_EULER]


def both(oracle, reference, *args, **kw):
    """The oracle's and the reference's results, or their (code, message)."""
    out = []
    for side in (oracle, reference):
        try:
            out.append(side.advance_coupled(*args, **kw))
        except OracleError as e:
            out.append((e.code, e.message))
    return out


def same(a, b):
    return (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2])


# ----------------------------------------------------------- known answers --
def single_signal_box(shape, spacing, level):
    nx, ny, nz = shape
    return Grid(nx, ny, nz, *spacing, None, (NO_FLUX,) * 6), np.full(nx * ny * nz, level,
                                                                     np.float32)


def one_cell(x, y, z, length=2.0, radius=0.5, growth=0.0):
    f = lambda v: np.array([v], np.float32)  # noqa: E731
    return dict(position_x=f(x), position_y=f(y), position_z=f(z), lengths=f(length),
                radii=f(radius), growth_rates=f(growth), cell_types=np.zeros(1, np.int32))


def test_sample_and_scatter_share_trilinear_weights(oracle):
    """coupled_rates_test.cpp:45-72: a cell at the centre of a 2x2x2 box of
    level 4 samples 4 (species 10 + 0.5 * 4 = 12) and returns -2 * 0.5 = -1
    spread over the eight sites by the same weights (1/8 each of a voxel
    volume of 2): every level 4 - 1/16 = 3.9375, grid amount 63."""
    grid, levels = single_signal_box((2, 2, 2), (2.0, 1.0, 1.0), 4.0)
    cell = one_cell(1.0, 0.5, 0.5)
    plan = (1, [(OP_SIGNAL, 0, 0, 0, 0.0), (OP_CONSTANT, 0, 0, 0, -2.0)], [0], [1])
    lv, sp, _ = oracle.advance_coupled(grid, levels, cell, [10.0], plan, cell["lengths"], 0.5,
                                       diffusion=(0.0,))
    assert abs(sp[0] - 12.0) <= 1e-6
    assert np.all(np.abs(lv - 3.9375) <= 1e-6)
    assert abs(float(lv.sum()) * 2.0 - 63.0) <= 1e-5


def test_crank_nicolson_includes_coupled_sources(oracle):
    """coupled_rates_test.cpp: one site of level 1 under Crank-Nicolson, a
    cell releasing 2 per unit time for a unit step: level 3."""
    grid, levels = single_signal_box((1, 1, 1), (1.0, 1.0, 1.0), 1.0)
    cell = one_cell(0.0, 0.0, 0.0)
    plan = (0, [(OP_CONSTANT, 0, 0, 0, 2.0)], [], [0])
    lv, _, rep = oracle.advance_coupled(grid, levels, cell, [], plan, cell["lengths"], 1.0,
                                        diffusion=(0.0,), integration=CRANK_NICOLSON)
    assert rep["converged"] and abs(lv[0] - 3.0) <= 1e-6


def test_dilution_follows_the_effective_volume(oracle):
    """A cell whose length grew from L0 to L1 holds species c * V(L0)/V(L1)
    with V = pi r^2 (L + 2r) before its rates act (here none: zero rate)."""
    grid, levels = single_signal_box((3, 3, 1), (1.0, 1.0, 1.0), 1.0)
    cell = one_cell(1.0, 1.0, 0.0, length=3.0, radius=0.5)
    plan = (1, [(OP_CONSTANT, 0, 0, 0, 0.0)], [0], [0])
    _, sp, _ = oracle.advance_coupled(grid, levels, cell, [6.0], plan, [1.0], 0.1,
                                      diffusion=(0.0,))
    expected = 6.0 * (0.25 * (1.0 + 1.0)) / (0.25 * (3.0 + 1.0))
    assert abs(sp[0] - expected) <= 1e-6


def test_obstacle_connectivity_renormalises_the_stencil(oracle):
    """A cell whose bilinear stencil has one solid corner samples only the
    three fluid corners, renormalised: sites of level 1, 2, 3 (and a solid
    one) weighted equally at the voxel centre read (1 + 2 + 3) / 3."""
    solid = np.array([0, 0, 0, 1], np.uint8)  # site (1, 1) is solid
    grid = Grid(2, 2, 1, 1.0, 1.0, 1.0, solid, (NO_FLUX,) * 6)
    levels = np.array([1.0, 2.0, 3.0, 0.0], np.float32)
    cell = one_cell(0.5, 0.5, 0.0)
    plan = (1, [(OP_SIGNAL, 0, 0, 0, 0.0), (OP_CONSTANT, 0, 0, 0, 0.0)], [0], [1])
    _, sp, _ = oracle.advance_coupled(grid, levels, cell, [0.0], plan, cell["lengths"], 1.0,
                                      diffusion=(0.0,))
    assert abs(sp[0] - 2.0) <= 1e-6


# ------------------------------------------------------------ live _ref --
@pytest.mark.parametrize("integration", INTEGRATIONS)
@pytest.mark.parametrize("scheme", [UPWIND, VAN_LEER])
@pytest.mark.parametrize("kinds", TRANSPORT_KINDS)
def test_coupled_steps_match_reference_live(oracle, reference, kinds, scheme, integration):
    """Random obstacle grids, face fields and colonies under every boundary
    kind, both schemes and the three integrations, with the plan that uses
    every operation: grid levels, species and the report bit for bit."""
    for seed, (nx, ny, S) in enumerate([(20, 15, 2), (33, 9, 1), (12, 40, 3), (1, 25, 1)]):
        c = random_transport_case(nx, ny, S, kinds, seed + 10 * integration)
        grid = Grid(nx, ny, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
        cells, prev = random_cells(seed, nx, ny, c["solid"], 40 + 7 * seed, 1.0, 0.8)
        K = 1 + seed % 3
        plan = uptake_plan(S, K, transcendental=seed % 2 == 0)
        species = np.random.default_rng(seed).random(len(prev) * K).astype(np.float32)
        dt = 0.02 if integration == FORWARD_EULER else 0.1
        kw = dict(diffusion=c["diffusion"], boundary_values=c["bv"],
                  faces=np.concatenate([c["xf"], c["yf"], c["zf"]]), source=c["source"],
                  loss=c["loss"], scheme=scheme, integration=integration)
        a, b = both(oracle, reference, grid, c["levels"], cells, species, plan, prev, dt, **kw)
        assert len(a) == 3 and len(b) == 3, (nx, ny, a if len(a) == 2 else b)
        assert same(a, b), (nx, ny)


def conformance_colony(masked, integration, scheme):
    """make_reference (coupled_rates_conformance_test.cpp:18-118) without its
    reservoir patches and with the uniform advection written as face
    velocities: a 9x7x5 grid of two signals, 513 cells of three species."""
    nx, ny, nz = 9, 7, 5
    kinds = (PERIODIC, PERIODIC, NO_FLUX, NO_FLUX, FIXED, FIXED)
    solid = None
    site = lambda x, y, z: (x * ny + y) * nz + z  # noqa: E731
    n = nx * ny * nz
    if masked:
        solid = np.zeros(n, np.uint8)
        solid[site(4, 3, 2)] = 1
        solid[site(2, 1, 1)] = 1
    grid = Grid(nx, ny, nz, 0.75, 1.25, 2.0, solid, kinds)
    bv = np.zeros(12, np.float32)
    bv[[4, 5]] = [0.9, 1.2]
    bv[[10, 11]] = [1.1, 1.0]
    index = np.arange(2 * n)
    source = (np.float32(0.004) * (index % 7).astype(np.float32)).astype(np.float32)
    loss = (np.float32(0.003) * (index % 11).astype(np.float32)).astype(np.float32)
    levels = (1.0 + np.float32(0.002) * (index % 113).astype(np.float32)).astype(np.float32)
    if masked:
        dead = np.tile(solid, 2) != 0
        source[dead] = loss[dead] = levels[dead] = 0.0
    # Uniform advection of signal 0 as faces; the grid's velocity is shared
    # by the signals here.
    nfx, nfy, nfz = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    faces = np.concatenate([np.full(nfx, 0.05, np.float32), np.full(nfy, -0.02, np.float32),
                            np.full(nfz, 0.01, np.float32)])
    F = faces[:nfx].reshape(nx + 1, ny, nz)
    Y = faces[nfx:nfx + nfy].reshape(nx, ny + 1, nz)
    Y[:, 0, :] = Y[:, -1, :] = 0.0
    if masked:
        for (x, y, z) in ((4, 3, 2), (2, 1, 1)):
            F[x, y, z] = F[x + 1, y, z] = 0.0
            Y[x, y, z] = Y[x, y + 1, z] = 0.0
            Z = faces[nfx + nfy:].reshape(nx, ny, nz + 1)
            Z[x, y, z] = Z[x, y, z + 1] = 0.0
    k = np.arange(513)
    fx = (k % 17).astype(np.float32) / np.float32(16)
    fy = ((k * 7) % 19).astype(np.float32) / np.float32(18)
    fz = ((k * 11) % 23).astype(np.float32) / np.float32(22)
    origin = (-3.0, -2.0, -1.0)
    f32 = np.float32
    cells = dict(
        position_x=(f32(-3.0) + fx * f32(0.75) * f32(nx - 1)).astype(np.float32),
        position_y=(f32(-2.0) + fy * f32(1.25) * f32(ny - 1)).astype(np.float32),
        position_z=(f32(-1.0) + fz * f32(2.0) * f32(nz - 1)).astype(np.float32),
        lengths=(f32(1.5) + f32(0.01) * (k % 31).astype(np.float32)).astype(np.float32),
        radii=(f32(0.3) + f32(0.005) * (k % 7).astype(np.float32)).astype(np.float32),
        growth_rates=(f32(0.01) * (k % 5).astype(np.float32)).astype(np.float32),
        cell_types=(k % 4).astype(np.int32))
    species = np.stack([0.5 + f32(0.01) * (k % 29), 0.75 + f32(0.02) * (k % 17),
                        1.0 + f32(0.015) * (k % 13)], 1).astype(np.float32).reshape(-1)
    ins = [(OP_SPECIES, 0, 0, 0, 0.0), (OP_SPECIES, 1, 0, 0, 0.0), (OP_SPECIES, 2, 0, 0, 0.0),
           (OP_SIGNAL, 0, 0, 0, 0.0), (OP_SIGNAL, 1, 0, 0, 0.0), (OP_CONSTANT, 0, 0, 0, 0.02),
           (OP_CONSTANT, 0, 0, 0, -0.01), (OP_MULTIPLY, 3, 5, 0, 0.0), (OP_ADD, 0, 7, 0, 0.0),
           (OP_MULTIPLY, 1, 6, 0, 0.0), (OP_ADD, 2, 9, 0, 0.0), (OP_MULTIPLY, 0, 5, 0, 0.0),
           (OP_MULTIPLY, 4, 6, 0, 0.0)]
    plan = (3, ins, [8, 9, 10], [11, 12])
    # A previous length a little shorter, so the dilution is not one.
    prev = (cells["lengths"] * f32(0.97)).astype(np.float32)
    kw = dict(origin=origin, diffusion=[0.08, 0.03], boundary_values=bv, faces=faces,
              source=source, loss=loss, scheme=scheme, integration=integration)
    return grid, levels, cells, species, plan, prev, kw


@pytest.mark.parametrize("integration", INTEGRATIONS)
@pytest.mark.parametrize("masked", [False, True])
def test_conformance_colony_matches_reference_live(oracle, reference, masked, integration):
    for scheme in (UPWIND, VAN_LEER):
        grid, levels, cells, species, plan, prev, kw = conformance_colony(masked, integration,
                                                                          scheme)
        a, b = both(oracle, reference, grid, levels, cells, species, plan, prev, 0.05, **kw)
        assert len(a) == 3, a
        assert same(a, b)


def test_crowded_colony_matches_reference_live(oracle, reference):
    """make_crowded_reference (coupled_rates_conformance_test.cpp:120-178):
    400 cells on a few sites of a 6x1x4 grid, on lattice planes and on the
    upper faces, where each axis's stencil holds a single site."""
    nx, ny, nz = 6, 1, 4
    solid = np.zeros(nx * ny * nz, np.uint8)
    solid[3 * nz + 2] = 1
    grid = Grid(nx, ny, nz, 1.5, 1.0, 0.75, solid, (NO_FLUX,) * 6)
    n = nx * ny * nz
    index = np.arange(2 * n)
    levels = (1.0 + np.float32(0.01) * (index % 13).astype(np.float32)).astype(np.float32)
    levels[np.tile(solid, 2) != 0] = 0.0
    f32 = np.float32
    ox, oz = f32(-1.0), f32(0.5)
    ex, ez = f32(1.5) * f32(5), f32(0.75) * f32(3)
    xs, ys, zs = [], [], []
    for i in range(400):
        if i % 8 == 0:
            xs.append(ox + f32(1.5) * f32((i // 8) % 6))
            ys.append(0.25)
            zs.append(oz + ez * f32(i % 5) / f32(4))
        elif i % 8 == 1:
            xs.append(ox + ex)
            ys.append(0.0)
            zs.append(oz + ez)
        else:
            xs.append(ox + f32(1.5) * (f32(1) + f32(i % 37) / f32(37)))
            ys.append(0.1)
            zs.append(oz + f32(0.75) * f32(i % 29) / f32(29))
    k = np.arange(400)
    cells = dict(position_x=np.array(xs, np.float32), position_y=np.array(ys, np.float32),
                 position_z=np.array(zs, np.float32),
                 lengths=(f32(1.0) + f32(0.01) * (k % 31).astype(np.float32)).astype(np.float32),
                 radii=np.full(400, 0.3, np.float32),
                 growth_rates=(f32(0.01) * (k % 5).astype(np.float32)).astype(np.float32),
                 cell_types=np.zeros(400, np.int32))
    species = (f32(0.5) + f32(0.01) * (k % 29).astype(np.float32)).astype(np.float32)
    ins = [(OP_SPECIES, 0, 0, 0, 0.0), (OP_SIGNAL, 0, 0, 0, 0.0), (OP_CONSTANT, 0, 0, 0, 0.02),
           (OP_CONSTANT, 0, 0, 0, -0.01), (OP_MULTIPLY, 1, 2, 0, 0.0),
           (OP_MULTIPLY, 0, 2, 0, 0.0), (OP_MULTIPLY, 0, 3, 0, 0.0)]
    plan = (1, ins, [4], [5, 6])
    prev = (cells["lengths"] * f32(0.98)).astype(np.float32)
    for integration in INTEGRATIONS:
        a, b = both(oracle, reference, grid, levels, cells, species, plan, prev, 0.05,
                    origin=(-1.0, 0.0, 0.5), diffusion=[0.05, 0.02], integration=integration)
        assert len(a) == 3, a
        assert same(a, b)


def test_zero_step_keeps_the_grid_and_dilutes_species(oracle, reference):
    """dt = 0: no transport and no exchange, the species are still diluted and
    the volume change rate reads zero."""
    kinds = TRANSPORT_KINDS[0]
    c = random_transport_case(16, 12, 2, kinds, 3)
    grid = Grid(16, 12, 1, 1.0, 1.0, 1.0, c["solid"], kinds)
    cells, prev = random_cells(3, 16, 12, c["solid"], 30)
    plan = uptake_plan(2, 2)
    species = np.random.default_rng(1).random(60).astype(np.float32)
    for integration in INTEGRATIONS:
        a, b = both(oracle, reference, grid, c["levels"], cells, species, plan, prev, 0.0,
                    diffusion=c["diffusion"], boundary_values=c["bv"], integration=integration)
        assert len(a) == 3 and same(a, b)
        assert np.array_equal(a[0], c["levels"].astype(np.float32))


# --------------------------------------------------------------- refusals --
def refusal_cases():
    """(name, change) pairs: each change edits a valid case and makes the step
    refused. Shared with the device tests."""
    def at(field, index, value):
        def change(case):
            case["cells"][field] = case["cells"][field].copy()
            case["cells"][field][index] = value
        return change

    def plan_edit(fn):
        def change(case):
            K, ins, so, go = case["plan"]
            case["plan"] = fn(K, list(ins), list(so), list(go))
        return change

    def setter(key, value):
        def change(case):
            case[key] = value
        return change

    def species_nan(case):
        case["species"] = case["species"].copy()
        case["species"][3] = np.nan

    def prev_negative(case):
        case["prev"] = case["prev"].copy()
        case["prev"][1] = -1.0

    def log_of_zero(K, ins, so, go):
        ins.append((OP_CONSTANT, 0, 0, 0, 0.0))
        ins.append((OP_LOGARITHM, len(ins) - 1, 0, 0, 0.0))
        return K, ins, so, [len(ins) - 1] + go[1:]

    return [
        ("outside x", at("position_x", 2, 40.0)),
        ("outside y", at("position_y", 0, -0.5)),
        ("nonfinite position", at("position_y", 0, np.inf)),
        ("negative length", at("lengths", 1, -1.0)),
        ("zero radius", at("radii", 4, 0.0)),
        ("nonfinite species", species_nan),
        ("negative previous length", prev_negative),
        ("negative dt", setter("dt", -0.1)),
        ("nonfinite dt", setter("dt", np.nan)),
        ("nonfinite origin", setter("origin", (np.nan, 0.0, 0.0))),
        ("unknown species", plan_edit(lambda K, i, s, g: (K, i + [(OP_SPECIES, K, 0, 0, 0.0)],
                                                          s, g))),
        ("unknown signal", plan_edit(lambda K, i, s, g: (K, i + [(OP_SIGNAL, 9, 0, 0, 0.0)],
                                                         s, g))),
        ("forward input", plan_edit(lambda K, i, s, g: (K, i + [(OP_ADD, 0, len(i), 0, 0.0)],
                                                        s, g))),
        ("unknown operation", plan_edit(lambda K, i, s, g: (K, i + [(77, 0, 0, 0, 0.0)], s, g))),
        ("nonfinite constant", plan_edit(lambda K, i, s, g: (K, i + [(OP_CONSTANT, 0, 0, 0,
                                                                      np.inf)], s, g))),
        ("species output", plan_edit(lambda K, i, s, g: (K, i, [len(i)] + s[1:], g))),
        ("signal output", plan_edit(lambda K, i, s, g: (K, i, s, [len(i)] + g[1:]))),
        ("signal count", plan_edit(lambda K, i, s, g: (K, i, s, g + [0]))),
        ("non-finite instruction", plan_edit(log_of_zero)),
        ("explicit bound", setter("dt", 50.0)),
    ]


def refusal_base():
    kinds = (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX)
    c = random_transport_case(14, 11, 2, kinds, 21)
    solid = c["solid"].copy()
    cells, prev = random_cells(21, 14, 11, solid, 12)
    return dict(c=c, kinds=kinds, cells=cells, prev=prev, plan=uptake_plan(2, 2),
                species=np.random.default_rng(2).random(24).astype(np.float32), dt=0.02,
                origin=(0.0, 0.0, 0.0))


@pytest.mark.parametrize("name,change", refusal_cases(), ids=[n for n, _ in refusal_cases()])
def test_refusals_match_reference(oracle, reference, name, change):
    case = refusal_base()
    change(case)
    c = case["c"]
    grid = Grid(14, 11, 1, 1.0, 1.0, 1.0, c["solid"], case["kinds"])
    a, b = both(oracle, reference, grid, c["levels"], case["cells"], case["species"],
                case["plan"], case["prev"], case["dt"], origin=case["origin"],
                diffusion=c["diffusion"], boundary_values=c["bv"],
                faces=np.concatenate([c["xf"], c["yf"], c["zf"]]))
    assert len(a) == 2, name
    assert a == b


def test_cell_inside_an_obstacle_is_refused(oracle, reference):
    solid = np.zeros(16, np.uint8)
    solid[[5, 6, 9, 10]] = 1  # the 2x2 block in the middle of a 4x4 grid
    grid = Grid(4, 4, 1, 1.0, 1.0, 1.0, solid, (NO_FLUX,) * 6)
    levels = np.where(solid != 0, 0.0, 1.0).astype(np.float32)
    cell = one_cell(1.5, 1.5, 0.0)
    plan = (1, [(OP_SIGNAL, 0, 0, 0, 0.0)], [0], [0])
    a, b = both(oracle, reference, grid, levels, cell, [1.0], plan, cell["lengths"], 0.1,
                diffusion=(0.1,))
    assert a == b == (1, "signal sample position is inside a grid obstacle")


def test_oracle_matches_the_committed_fixture(oracle):
    """tests/golden/coupled.npz holds the reference's advance_coupled_cpu on a
    one-layer version of its conformance colony: the oracle reproduces it bit
    for bit under every integration and scheme."""
    import os
    from tests.helpers import coupled_colony_case
    path = os.path.join(os.path.dirname(__file__), "golden", "coupled.npz")
    g, levels, cells, species, plan, previous, transport = coupled_colony_case()
    with np.load(path, allow_pickle=False) as f:
        fx = {k: f[k] for k in f.files}
    assert np.array_equal(fx["levels"], levels) and np.array_equal(fx["species"], species)
    assert all(np.array_equal(fx[f"cell_{k}"], v) for k, v in cells.items())
    for integration in INTEGRATIONS:
        for scheme in (UPWIND, VAN_LEER):
            key = f"i{integration}s{scheme}"
            lv, sp, rep = oracle.advance_coupled(Grid(**g), levels, cells, species, plan, previous,
                                                 0.05, origin=(-3.0, -2.0, 0.0), scheme=scheme,
                                                 integration=integration, **transport)
            assert np.array_equal(lv, fx[f"{key}_levels"]), key
            assert np.array_equal(sp, fx[f"{key}_species"]), key
            assert [rep["converged"], rep["iterations"]] == list(fx[f"{key}_report"][:2]), key
            assert np.float32(rep["residual_rms"]) == np.float32(fx[f"{key}_report"][2]), key
