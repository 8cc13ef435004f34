"""Coupled cell/grid steps on grids of several layers on the GPU, through the
C ABI, against the oracle, which is bit-exact with the reference's
advance_coupled_cpu on these inputs (tests/test_coupled_oracle.py):

- the reference conformance suite's colony (coupled_rates_conformance_test.cpp:
  18-118): 513 cells of three species on a 9 x 7 x 5 grid of two signals,
  with and without obstacles, every integration kind and both schemes;
- random layered grids (obstacles, a velocity field on all three axes, affine
  reactions) with random colonies whose trilinear stencils reach fluid.

The trilinear stencils (eight corners, the face-connected component of the
strongest fluid one), the samples, the plan, the per-site source sums in the
caller's cell order and the three-axis transport are the reference's float
operations in its order, so levels, species and the implicit report are
compared bit for bit (the plans here use no exp, log or pow).
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import Grid as OGrid
from oracle import UPWIND, VAN_LEER
from tests.helpers import FIXED, NO_FLUX, PERIODIC, uptake_plan
from tests.test_coupled_oracle import conformance_colony

pytestmark = pytest.mark.gpu

INTEGRATIONS = [msa.FORWARD_EULER, msa.CRANK_NICOLSON, msa.BACKWARD_EULER]


def device_spec(grid, kw, scheme, integration):
    faces = kw.get("faces")
    velocity = None
    if faces is not None:
        nfx, nfy, _ = grid.face_counts()
        velocity = msa.VelocityField(faces[:nfx], faces[nfx:nfx + nfy], faces[nfx + nfy:])
    return msa.SignalGridSpec(grid, kw["diffusion"], kw["boundary_values"], velocity,
                              source_rates=kw.get("source"), loss_rates=kw.get("loss"),
                              advection_scheme=scheme, integration=integration)


def same_report(rep, ref):
    assert rep.converged == ref["converged"]
    assert rep.iterations == ref["iterations"]
    assert abs(rep.residual_rms - ref["residual_rms"]) <= 2.0 ** -23 * abs(ref["residual_rms"])


def device_grid(og):
    return msa.Grid(og.nx, og.ny, og.nz, og.dx, og.dy, og.dz, og.obstacles, og.boundary)


@pytest.mark.parametrize("integration", INTEGRATIONS)
@pytest.mark.parametrize("masked", [False, True])
def test_conformance_colony_on_device(hip, oracle, masked, integration):
    for scheme in (UPWIND, VAN_LEER):
        og, levels, cells, species, plan, prev, kw = conformance_colony(masked, integration,
                                                                        scheme)
        print(f"conformance colony: masked {masked}, integration {integration}, scheme {scheme}",
              flush=True)
        origin = kw.pop("origin")
        ref_levels, ref_species, ref = oracle.advance_coupled(og, levels, cells, species, plan,
                                                              prev, 0.05, origin=origin, **kw)
        spec = device_spec(device_grid(og), kw, scheme, integration)
        got_levels, got_species, rep = hip.advance_coupled(spec, levels, cells, species, plan,
                                                           prev, 0.05, origin)
        assert np.array_equal(got_levels, ref_levels)
        assert np.array_equal(got_species, ref_species)
        same_report(rep, ref)


def layered_cells(seed, nx, ny, nz, solid, count, spacing, origin):
    """Cells inside the grid whose stencils hold a fluid corner of positive
    weight: a fluid lower corner and fractions below one."""
    rng = np.random.default_rng(seed)
    s = np.asarray(solid).reshape(nx, ny, nz)
    n = (nx, ny, nz)
    corners = np.argwhere((s == 0)
                          & (np.arange(nx)[:, None, None] < max(nx - 1, 1))
                          & (np.arange(ny)[None, :, None] < max(ny - 1, 1))
                          & (np.arange(nz)[None, None, :] < max(nz - 1, 1)))
    pick = corners[rng.integers(len(corners), size=count)]
    frac = rng.uniform(0.0, 0.999, (count, 3))
    pos = []
    for a in range(3):
        p = (pick[:, a] + (frac[:, a] if n[a] > 1 else 0.0)) * spacing[a] + origin[a]
        pos.append(p.astype(np.float32))
    lengths = rng.uniform(1.0, 4.0, count).astype(np.float32)
    cells = dict(position_x=pos[0], position_y=pos[1], position_z=pos[2], lengths=lengths,
                 radii=rng.uniform(0.3, 0.8, count).astype(np.float32),
                 growth_rates=rng.uniform(0.5, 1.5, count).astype(np.float32),
                 cell_types=rng.integers(0, 4, count).astype(np.int32))
    return cells, (lengths * rng.uniform(0.85, 1.0, count)).astype(np.float32)


def random_layered(nx, ny, nz, S, kinds, seed):
    """Obstacles, levels, affine rates and a field on all three axes, zero on
    closed faces and equal across a periodic boundary."""
    rng = np.random.default_rng(seed)
    solid = rng.random((nx, ny, nz)) < 0.12
    levels = (0.5 + 0.5 * rng.random((S, nx, ny, nz))).astype(np.float32)
    levels[:, solid] = 0.0
    faces = []
    for axis, shape in enumerate([(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]):
        F = rng.uniform(-0.4, 0.4, shape).astype(np.float32)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = slice(0, -1)
        hi[axis] = slice(1, None)
        F[tuple(lo)][solid] = 0.0
        F[tuple(hi)][solid] = 0.0
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[axis] = 0
        last[axis] = -1
        if kinds[2 * axis] == PERIODIC:
            both = (F[tuple(first)] != 0) & (F[tuple(last)] != 0)
            F[tuple(first)] = np.where(both, F[tuple(first)], 0.0)
            F[tuple(last)] = F[tuple(first)]
        else:
            if kinds[2 * axis] == NO_FLUX:
                F[tuple(first)] = 0.0
            if kinds[2 * axis + 1] == NO_FLUX:
                F[tuple(last)] = 0.0
        faces.append(F.reshape(-1))
    source = (0.02 * rng.random(S * nx * ny * nz)).astype(np.float32)
    loss = (0.05 * rng.random(S * nx * ny * nz)).astype(np.float32)
    dead = np.tile(solid.reshape(-1), S)
    source[dead] = loss[dead] = 0.0
    kw = dict(diffusion=rng.uniform(0.05, 0.3, S).astype(np.float32),
              boundary_values=rng.random(6 * S).astype(np.float32),
              faces=np.concatenate(faces), source=source, loss=loss)
    return solid.astype(np.uint8).reshape(-1), levels.reshape(-1), kw


LAYERED_KINDS = [
    (FIXED, FIXED, NO_FLUX, NO_FLUX, FIXED, NO_FLUX),
    (PERIODIC, PERIODIC, FIXED, FIXED, PERIODIC, PERIODIC),
    (FIXED, NO_FLUX, PERIODIC, PERIODIC, NO_FLUX, FIXED),
]


@pytest.mark.parametrize("integration", INTEGRATIONS)
@pytest.mark.parametrize("scheme", [UPWIND, VAN_LEER])
@pytest.mark.parametrize("kinds", LAYERED_KINDS)
def test_random_layered_colonies_on_device(hip, oracle, kinds, scheme, integration):
    for seed, (nx, ny, nz, S) in enumerate([(12, 9, 4, 2), (5, 17, 3, 1), (20, 6, 7, 3)]):
        solid, levels, kw = random_layered(nx, ny, nz, S, kinds, seed=7 * seed + integration)
        spacing = (1.0, 0.8, 1.3)
        origin = (0.5, -1.0, 2.0)
        cells, prev = layered_cells(seed, nx, ny, nz, solid, 60 + 25 * seed, spacing, origin)
        K = 1 + seed % 3
        plan = uptake_plan(S, K, transcendental=False)
        species = np.random.default_rng(seed).random(len(prev) * K).astype(np.float32)
        dt = 0.02 if integration == msa.FORWARD_EULER else 0.1
        og = OGrid(nx, ny, nz, *spacing, solid, boundary=kinds)
        print("case", kinds, scheme, integration, (nx, ny, nz, S), len(prev), flush=True)
        ref_levels, ref_species, ref = oracle.advance_coupled(
            og, levels, cells, species, plan, prev, dt, origin=origin, scheme=scheme,
            integration=integration, **kw)
        spec = device_spec(device_grid(og), kw, scheme, integration)
        got_levels, got_species, rep = hip.advance_coupled(spec, levels, cells, species, plan,
                                                           prev, dt, origin)
        assert np.array_equal(got_levels, ref_levels)
        assert np.array_equal(got_species, ref_species)
        same_report(rep, ref)


def test_layered_device_resident_steps(hip, oracle):
    """Three device-resident steps equal three one-call steps of the oracle,
    each from the last one's levels and species."""
    kinds = LAYERED_KINDS[0]
    solid, levels, kw = random_layered(10, 8, 4, 2, kinds, seed=5)
    spacing, origin = (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    cells, prev = layered_cells(3, 10, 8, 4, solid, 120, spacing, origin)
    plan = uptake_plan(2, 2, transcendental=False)
    species = np.random.default_rng(3).random(len(prev) * 2).astype(np.float32)
    og = OGrid(10, 8, 4, *spacing, solid, boundary=kinds)
    spec = device_spec(device_grid(og), kw, UPWIND, msa.FORWARD_EULER)
    dt = 0.02
    ref_levels, ref_species, p = levels, species, prev
    for _ in range(3):
        ref_levels, ref_species, _ = oracle.advance_coupled(og, ref_levels, cells, ref_species,
                                                            plan, p, dt, origin=origin, **kw)
        p = cells["lengths"]
    hip.coupled_upload(spec, levels, cells, species, plan, origin)
    hip.coupled_run(dt, 3, previous_lengths=prev)
    got_levels, got_species = hip.coupled_download()
    assert np.array_equal(got_levels, ref_levels)
    assert np.array_equal(got_species, ref_species)
