"""Transport on grids of several layers on the GPU, through the C ABI, against
the oracle (bit-exact with the reference there,
tests/test_layered_transport_oracle.py): the reference conformance suite's
9 x 7 x 5 grids (tests/conformance/signal_grid_conformance_test.cpp:14-293)
under every integration kind and scheme, and random layered grids with
obstacles, a velocity field on all three axes, Monod uptake and reservoir
patches. FP32 levels are held to the reference's bits (the suite itself
allows 5e-6); implicit steps also to its iteration counts."""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import Grid as OGrid
from oracle import UPWIND, VAN_LEER
from tests.helpers import (CONFORMANCE_KINDS, CONFORMANCE_STEPS, FIXED, NO_FLUX, PERIODIC,
                           conformance_spec)

pytestmark = pytest.mark.gpu


def _grids(shape, spacing, kinds, obstacles):
    nx, ny, nz = shape
    dx, dy, dz = spacing
    return (msa.Grid(nx, ny, nz, dx, dy, dz, obstacles, kinds),
            OGrid(nx, ny, nz, dx, dy, dz, obstacles, boundary=kinds))


def _device_spec(grid, kw, scheme, integration):
    faces = kw.get("faces")
    velocity = None
    if faces is not None:
        nfx, nfy, _ = grid.face_counts()
        velocity = msa.VelocityField(faces[:nfx], faces[nfx:nfx + nfy], faces[nfx + nfy:])
    return msa.SignalGridSpec(grid, kw["diffusion"], kw["boundary_values"], velocity,
                              source_rates=kw.get("source"), loss_rates=kw.get("loss"),
                              phi=kw.get("phi"), monod_vmax=kw.get("vmax"),
                              monod_k=kw.get("monod_k"), advection=kw.get("advection"),
                              advection_scheme=scheme, patches=kw.get("patches", ()),
                              integration=integration)


@pytest.mark.parametrize("kind", CONFORMANCE_KINDS)
@pytest.mark.parametrize("integration,dt", CONFORMANCE_STEPS)
@pytest.mark.parametrize("scheme", [UPWIND, VAN_LEER])
def test_conformance_grid_on_device(hip, oracle, kind, integration, dt, scheme):
    c = conformance_spec(kind)
    grid, ogrid = _grids(c["shape"], c["spacing"], c["kinds"], c["obstacles"])
    spec = _device_spec(grid, c["kw"], scheme, integration)
    print(f"conformance {kind}: integration {integration}, scheme {scheme}", flush=True)
    levels = c["levels"]
    for _ in range(2):
        levels, rep = hip.advance_signal_grid(spec, levels, dt)
    if integration == 0:
        ref = oracle.adr_step_f32(ogrid, c["levels"], dt, steps=2, scheme=scheme, **c["kw"])
    else:
        ref, orep = oracle.adr_step_implicit(ogrid, c["levels"], dt, steps=2, scheme=scheme,
                                             integration=integration, **c["kw"])
        assert rep.converged and rep.iterations == orep["iterations"]
    assert np.array_equal(levels, ref)


def random_layered_case(nx, ny, nz, S, kinds, seed):
    """Obstacles, levels, Monod occupancy and a random field on all three
    axes that is zero on closed faces and equal on the two sides of a
    periodic boundary."""
    rng = np.random.default_rng(seed)
    solid = (rng.random((nx, ny, nz)) < 0.1)
    levels = (0.5 + 0.5 * rng.random((S, nx, ny, nz))).astype(np.float32)
    levels[:, solid] = 0.0
    faces = []
    for axis, shape in enumerate([(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]):
        F = rng.uniform(-0.5, 0.5, shape).astype(np.float32)
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
    phi = rng.random(nx * ny * nz).astype(np.float32)
    kw = dict(diffusion=rng.uniform(0.05, 0.3, S).astype(np.float32),
              boundary_values=rng.random(6 * S).astype(np.float32),
              faces=np.concatenate(faces), phi=phi, vmax=np.full(S, 0.4, np.float32),
              monod_k=np.full(S, 0.5, np.float32))
    return solid.astype(np.uint8).reshape(-1), levels.reshape(-1), kw


LAYERED_KINDS = [
    (FIXED, FIXED, NO_FLUX, NO_FLUX, FIXED, NO_FLUX),
    (PERIODIC, PERIODIC, FIXED, FIXED, PERIODIC, PERIODIC),
    (FIXED, NO_FLUX, PERIODIC, PERIODIC, NO_FLUX, FIXED),
]


@pytest.mark.parametrize("kinds", LAYERED_KINDS)
@pytest.mark.parametrize("scheme", [UPWIND, VAN_LEER])
@pytest.mark.parametrize("shape", [(12, 9, 4), (5, 17, 3), (30, 2, 6)])
def test_random_layered_grids_on_device(hip, oracle, kinds, scheme, shape):
    nx, ny, nz = shape
    S = 3
    solid, levels, kw = random_layered_case(nx, ny, nz, S, kinds, seed=nx * 7 + nz)
    grid, ogrid = _grids(shape, (1.0, 0.8, 1.3), kinds, solid)
    spec = _device_spec(grid, kw, scheme, 0)
    dt = 0.05
    got = levels
    for _ in range(3):
        got, _ = hip.advance_signal_grid(spec, got, dt)
    ref = oracle.adr_step_f32(ogrid, levels, dt, steps=3, scheme=scheme, **kw)
    assert np.array_equal(got, ref)
    # The device-resident run takes the same steps.
    hip.adr_upload(spec, levels, precision=32)
    hip.adr_run(dt, 3, ledger=False)
    assert np.array_equal(hip.adr_download(), ref)


@pytest.mark.parametrize("integration", [1, 2])
def test_random_layered_implicit_on_device(hip, oracle, integration):
    kinds = LAYERED_KINDS[0]
    solid, levels, kw = random_layered_case(14, 10, 5, 2, kinds, seed=3)
    grid, ogrid = _grids((14, 10, 5), (1.0, 1.0, 1.0), kinds, solid)
    spec = _device_spec(grid, kw, UPWIND, integration)
    got, rep = hip.advance_signal_grid(spec, levels, 0.8)
    ref, orep = oracle.adr_step_implicit(ogrid, levels, 0.8, steps=1, integration=integration,
                                         **kw)
    assert rep.iterations == orep["iterations"]
    assert np.array_equal(got, ref)


def test_layered_refusals(hip):
    c = conformance_spec("plain")
    grid, _ = _grids(c["shape"], c["spacing"], c["kinds"], c["obstacles"])
    spec = _device_spec(grid, c["kw"], UPWIND, 0)
    # The explicit bound over three axes, with the reference's message.
    with pytest.raises(msa.InvalidArgument, match="explicit stability bound"):
        hip.advance_signal_grid(spec, c["levels"], 5.0)
    with pytest.raises(msa.Unsupported):
        hip.adr_upload(spec, c["levels"].astype(np.float64), precision=64)
    # A velocity field with constant advection is refused (signals.cpp:652-657).
    nfx, nfy, nfz = grid.face_counts()
    both = msa.SignalGridSpec(grid, c["kw"]["diffusion"], c["kw"]["boundary_values"],
                              msa.VelocityField(np.zeros(nfx), np.zeros(nfy), np.zeros(nfz)),
                              advection=c["kw"]["advection"])
    with pytest.raises(msa.InvalidArgument, match="zero constant advection"):
        hip.advance_signal_grid(both, c["levels"], 0.02)
