"""Implicit (Crank-Nicolson and backward Euler) transport steps on the GPU
through the C ABI, against the CPU oracle, which is bit-exact with the
reference's advance_signal_grid_cpu on these cases.

Tolerances. Every operation on a level is the reference's float operation in
its order, so levels are compared bit for bit and iteration counts exactly.
The residual RMS is a sum of squares in FP64: the device adds it in a
different order from the reference's pieces of 2048, so it is held to one
float rounding (2^-23 relative); a difference there could move the stopping
iteration only if the RMS fell within that rounding of the threshold.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from oracle import Grid as OGrid
from oracle import OracleError
from tests.helpers import (FIXED, NO_FLUX, PATCH_KINDS, TRANSPORT_KINDS, TRANSPORT_SHAPES,
                           c1_inputs, random_patches, random_transport_case, split_faces)

pytestmark = pytest.mark.gpu

IMPLICIT = [msa.CRANK_NICOLSON, msa.BACKWARD_EULER]


def spec_of(grid, c, scheme, integration, patches=(), solver=(10_000, 1e-7, 1e-5),
            faces=True, phi=None, vmax=None, monod_k=None, affine=True):
    nx, ny = grid.nx, grid.ny
    velocity = msa.VelocityField(c["xf"], c["yf"], c["zf"]) if faces else None
    return msa.SignalGridSpec(
        grid, c["diffusion"], c["bv"], velocity,
        source_rates=c["source"] if affine else None, loss_rates=c["loss"] if affine else None,
        phi=phi, monod_vmax=vmax, monod_k=monod_k, advection_scheme=scheme, patches=patches,
        integration=integration, solver_max_iterations=solver[0],
        solver_absolute_tolerance=solver[1], solver_relative_tolerance=solver[2])


def oracle_kw(c, scheme, integration, patches=(), solver=(10_000, 1e-7, 1e-5), faces=True,
              phi=None, vmax=None, monod_k=None, affine=True):
    kw = dict(diffusion=c["diffusion"], boundary_values=c["bv"], scheme=scheme,
              integration=integration, patches=patches, solver=solver)
    if faces:
        kw["faces"] = np.concatenate([c["xf"], c["yf"], c["zf"]])
    if affine:
        kw.update(source=c["source"], loss=c["loss"])
    if phi is not None:
        kw.update(phi=phi, vmax=vmax, monod_k=monod_k)
    return kw


def same_report(rep, ref):
    assert rep.converged == ref["converged"]
    assert rep.iterations == ref["iterations"]
    assert abs(rep.residual_rms - ref["residual_rms"]) <= 2.0 ** -23 * abs(ref["residual_rms"])


@pytest.mark.parametrize("integration", IMPLICIT)
@pytest.mark.parametrize("scheme", [msa.UPWIND, msa.VAN_LEER])
@pytest.mark.parametrize("kinds", TRANSPORT_KINDS)
def test_implicit_steps_match_the_oracle(hip, oracle, kinds, scheme, integration):
    for nx, ny in TRANSPORT_SHAPES:
        S = 1 + (nx + ny) % 3
        c = random_transport_case(nx, ny, S, kinds, nx * 7 + ny)
        grid = msa.Grid(nx, ny, 1, 0.7, 1.1, 1.0, c["solid"], kinds)
        print("case", kinds, scheme, integration, (nx, ny, S), flush=True)
        spec = spec_of(grid, c, scheme, integration)
        levels = c["levels"].astype(np.float32)
        ref_levels, ref = oracle.adr_step_implicit(OGrid(nx, ny, 1, 0.7, 1.1, 1.0, c["solid"],
                                                         kinds), levels, 0.4,
                                                   **oracle_kw(c, scheme, integration))
        ours, rep = hip.advance_signal_grid(spec, levels, 0.4)
        assert np.array_equal(ours, ref_levels)
        same_report(rep, ref)


@pytest.mark.parametrize("integration", IMPLICIT)
def test_implicit_patches_and_monod_match_the_oracle(hip, oracle, integration):
    for kinds in PATCH_KINDS:
        nx, ny, S = 41, 33, 2
        c = random_transport_case(nx, ny, S, kinds, 17)
        patches = random_patches(kinds, nx, ny, S, 18)
        phi = np.random.default_rng(4).random(nx * ny).astype(np.float32)
        phi[c["solid"] != 0] = 0.0
        grid = msa.Grid(nx, ny, 1, 1.0, 0.9, 1.0, c["solid"], kinds)
        og = OGrid(nx, ny, 1, 1.0, 0.9, 1.0, c["solid"], kinds)
        print("case", kinds, integration, flush=True)
        for scheme in (msa.UPWIND, msa.VAN_LEER):
            kw = dict(patches=patches, phi=phi, vmax=[0.6, 1.2], monod_k=[0.1, 0.3])
            levels = c["levels"].astype(np.float32)
            ref_levels, ref = oracle.adr_step_implicit(og, levels, 0.25,
                                                       **oracle_kw(c, scheme, integration, **kw))
            ours, rep = hip.advance_signal_grid(spec_of(grid, c, scheme, integration, **kw),
                                                levels, 0.25)
            assert np.array_equal(ours, ref_levels)
            same_report(rep, ref)


def test_known_answers_on_the_device(hip):
    """tests/cpp/signal_grid_test.cpp:62-75 and 104-117."""
    line = msa.Grid(1, 1, 1, 1.0, 1.0, 1.0, None, (NO_FLUX,) * 6)
    spec = msa.SignalGridSpec(line, [0.0], None, source_rates=[2.0], loss_rates=[0.5],
                              integration=msa.CRANK_NICOLSON)
    lv, rep = hip.advance_signal_grid(spec, np.array([1.0], np.float32), 1.0)
    assert rep.converged and abs(lv[0] - 2.2) <= 1e-5
    line = msa.Grid(3, 1, 1, 1.0, 1.0, 1.0, None, (NO_FLUX,) * 6)
    spec = msa.SignalGridSpec(line, [1.0], None, integration=msa.CRANK_NICOLSON,
                              solver_relative_tolerance=1e-6)
    lv, rep = hip.advance_signal_grid(spec, np.array([0.0, 1.0, 0.0], np.float32), 1.0)
    assert rep.converged and rep.iterations > 0 and rep.residual_rms <= 2e-6
    assert np.allclose(lv, [0.4, 0.2, 0.4], atol=1e-5)


def test_unconverged_step_is_refused_and_levels_kept(hip):
    """signal_grid_test.cpp:119-133: the reference's message, and a refused
    device-resident step leaves the levels it started from."""
    line = msa.Grid(9, 1, 1, 1.0, 1.0, 1.0, None, (NO_FLUX,) * 6)
    spec = msa.SignalGridSpec(line, [1.0], None, integration=msa.CRANK_NICOLSON,
                              solver_max_iterations=1, solver_absolute_tolerance=1e-12,
                              solver_relative_tolerance=0.0)
    levels = np.zeros(9, np.float32)
    levels[4] = 1.0
    with pytest.raises(msa.NotConverged) as info:
        hip.advance_signal_grid(spec, levels, 2.0)
    assert str(info.value) == "Implicit signal solve did not converge after 1 iterations"
    hip.adr_upload(spec, levels, precision=32)
    with pytest.raises(msa.NotConverged):
        hip.adr_run(2.0, 1, ledger=False)
    assert np.array_equal(hip.adr_download(), levels)


def test_refused_negative_levels_match_the_oracle(hip, oracle):
    kinds = (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX)
    c = random_transport_case(30, 21, 2, kinds, 5)
    grid = msa.Grid(30, 21, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
    og = OGrid(30, 21, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
    with pytest.raises(OracleError) as ref:
        oracle.adr_step_implicit(og, c["levels"], 0.8,
                                 **oracle_kw(c, msa.VAN_LEER, msa.CRANK_NICOLSON))
    levels = c["levels"].astype(np.float32)
    with pytest.raises(msa.InvalidArgument) as got:
        hip.advance_signal_grid(spec_of(grid, c, msa.VAN_LEER, msa.CRANK_NICOLSON), levels, 0.8)
    assert str(got.value) == ref.value.message


@pytest.mark.parametrize("integration", IMPLICIT)
def test_device_resident_implicit_steps_match_the_oracle(hip, oracle, integration):
    """Several implicit steps in one adr_run over the channel-and-trap flow
    field with Monod uptake, against the oracle step by step."""
    c1 = c1_inputs()
    nx, ny = c1["nx"], c1["ny"]
    xf, yf, zf = split_faces(c1["data"]["faces"], nx, ny)
    grid = msa.Grid(nx, ny, 1, 1.0, 1.0, c1["dz"], c1["obstacles"])
    og = OGrid(nx, ny, 1, 1.0, 1.0, c1["dz"], c1["obstacles"])
    bv = np.zeros(6, np.float32)
    bv[0] = 1.0
    spec = msa.SignalGridSpec(grid, [500.0], bv, msa.VelocityField(xf, yf, zf), phi=c1["phi"],
                              monod_vmax=[1.0], monod_k=[0.1], integration=integration)
    levels = np.zeros(nx * ny, np.float32)
    dt = 2.0e-3  # four times the explicit bound
    hip.adr_upload(spec, levels, precision=32)
    hip.adr_run(dt, 5, ledger=False)
    ours, rep = hip.adr_download(), hip.adr_report()
    ref_levels, ref = oracle.adr_step_implicit(
        og, levels, dt, steps=5, diffusion=[500.0], boundary_values=bv,
        faces=c1["data"]["faces"], phi=c1["phi"], vmax=1.0, monod_k=0.1, integration=integration)
    print("iterations of the last step", rep.iterations, ref["iterations"], flush=True)
    assert np.array_equal(ours, ref_levels)
    same_report(rep, ref)


@pytest.mark.parametrize("integration", IMPLICIT)
@pytest.mark.parametrize("scheme", [msa.UPWIND, msa.VAN_LEER])
def test_device_matches_the_committed_fixture(hip, scheme, integration):
    """Three implicit steps with Monod uptake on the channel + trap 256x128
    over the reference's flow field, against the reference's vectors in
    tests/golden/c1_implicit.npz: levels and iteration count exactly, the
    residual RMS to one float rounding."""
    from tests.helpers import golden
    c = c1_inputs()
    adr, d = golden("c1_adr.npz"), golden("c1_implicit.npz")
    xf, yf, zf = split_faces(c["data"]["faces"], c["nx"], c["ny"])
    grid = msa.Grid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"])
    spec = msa.SignalGridSpec(grid, adr["diffusion"], adr["boundary_values"],
                              msa.VelocityField(xf, yf, zf), phi=c["phi"],
                              monod_vmax=[float(adr["vmax"])], monod_k=[float(adr["K"])],
                              advection_scheme=scheme, integration=integration)
    print("case", scheme, integration, flush=True)
    levels = adr["levels"].astype(np.float32)
    for _ in range(int(d["steps"])):
        levels, rep = hip.advance_signal_grid(spec, levels, float(d["dt"]))
    key = f"i{integration}s{scheme}"
    assert np.array_equal(levels, d[f"{key}_levels"])
    converged, iterations, rms = d[f"{key}_report"]
    same_report(rep, dict(converged=bool(converged), iterations=int(iterations),
                          residual_rms=float(rms)))


def test_forward_euler_van_leer_matches_the_committed_fixture(hip):
    """Three forward-Euler van Leer steps with Monod on the same case, at the
    forward-Euler fixture's dt: the reference's levels bit for bit."""
    from tests.helpers import golden
    c = c1_inputs()
    adr, d = golden("c1_adr.npz"), golden("c1_implicit.npz")
    xf, yf, zf = split_faces(c["data"]["faces"], c["nx"], c["ny"])
    grid = msa.Grid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"])
    spec = msa.SignalGridSpec(grid, adr["diffusion"], adr["boundary_values"],
                              msa.VelocityField(xf, yf, zf), phi=c["phi"],
                              monod_vmax=[float(adr["vmax"])], monod_k=[float(adr["K"])],
                              advection_scheme=msa.VAN_LEER)
    levels = adr["levels"].astype(np.float32)
    for _ in range(3):
        levels, rep = hip.advance_signal_grid(spec, levels, float(adr["dt"]))
    assert rep.converged
    assert np.array_equal(levels, d["i0s1_levels"])
