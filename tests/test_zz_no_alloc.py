"""The allocation-free transport step (msim_hip_advance_signal_grid_static)
against msim_hip_advance_signal_grid, through the C ABI.

The static entry point stages everything in one file-scope device array and
allocates nothing; its results must equal the allocating entry point's bit for
bit (FP32 levels), on two calls in a row whose second grid is smaller, so the
second call runs over what the first left in the array.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from tests.helpers import FIXED, NO_FLUX, PERIODIC, random_transport_case

pytestmark = pytest.mark.gpu

KINDS = (FIXED, FIXED, PERIODIC, PERIODIC, NO_FLUX, NO_FLUX)


def _spec(nx, ny, S, seed, scheme, reaction):
    c = random_transport_case(nx, ny, S, KINDS, seed)
    grid = msa.Grid(nx, ny, 1, 1.0, 1.0, 1.0, c["solid"], KINDS)
    kw = {}
    if reaction == "affine":
        kw = dict(loss_rates=c["loss"], source_rates=c["source"])
    elif reaction == "monod":
        kw = dict(phi=c["phi"], monod_vmax=np.full(S, 0.3, np.float32),
                  monod_k=np.full(S, 0.5, np.float32))
    spec = msa.SignalGridSpec(grid, c["diffusion"], c["bv"],
                              msa.VelocityField(c["xf"], c["yf"], c["zf"]),
                              advection_scheme=scheme, **kw)
    return spec, c["levels"].astype(np.float32)


@pytest.mark.parametrize("scheme,reaction", [(msa.UPWIND, "monod"), (msa.VAN_LEER, "affine"),
                                             (msa.UPWIND, None)])
def test_static_step_equals_the_allocating_step(hip, scheme, reaction):
    # Two calls in a row: the second grid is smaller and has fewer signals.
    for nx, ny, S, seed in ((300, 260, 3, 11), (70, 45, 2, 12)):
        spec, levels = _spec(nx, ny, S, seed, scheme, reaction)
        print(f"static step: {nx}x{ny}, {S} signals, scheme {scheme}, {reaction}", flush=True)
        dt = 0.02
        ref, ref_rep = hip.advance_signal_grid(spec, levels, dt)
        got, rep = hip.advance_signal_grid_static(spec, levels, dt)
        assert np.array_equal(got, ref)
        assert rep == ref_rep
        # A second step from the result, as a caller's loop takes it.
        ref2, _ = hip.advance_signal_grid(spec, ref, dt)
        got2, _ = hip.advance_signal_grid_static(spec, got, dt)
        assert np.array_equal(got2, ref2)


def test_static_step_refuses_what_the_allocating_step_refuses(hip):
    spec, levels = _spec(64, 48, 2, 21, msa.UPWIND, "affine")
    # The explicit bound, with the reference's message.
    with pytest.raises(msa.InvalidArgument, match="explicit stability bound"):
        hip.advance_signal_grid_static(spec, levels, 50.0)
    with pytest.raises(msa.InvalidArgument, match="explicit stability bound"):
        hip.advance_signal_grid(spec, levels, 50.0)
    # A negative level.
    bad = levels.copy()
    bad[np.argmax(bad)] = -1.0
    with pytest.raises(msa.InvalidArgument, match="finite and non-negative"):
        hip.advance_signal_grid_static(spec, bad, 0.02)
    # After refusals the next call still equals the allocating step.
    ref, _ = hip.advance_signal_grid(spec, levels, 0.02)
    got, _ = hip.advance_signal_grid_static(spec, levels, 0.02)
    assert np.array_equal(got, ref)


def test_static_step_limits(hip):
    lib = msa.library()
    limit = int(lib.msim_hip_static_scratch_bytes())
    assert limit == 64 << 20
    # A grid whose two level buffers alone exceed the scratch is refused
    # before any device work.
    nx, ny, S = 2048, 2048, 3
    grid = msa.Grid(nx, ny, 1, 1.0, 1.0, 1.0, None, KINDS)
    spec = msa.SignalGridSpec(grid, np.full(S, 0.1, np.float32))
    with pytest.raises(msa.InvalidArgument, match="bytes of device scratch"):
        hip.advance_signal_grid_static(spec, np.zeros(S * nx * ny, np.float32), 0.1)
    # Implicit steps and reservoir patches are the allocating entry point's.
    small, levels = _spec(32, 24, 1, 5, msa.UPWIND, None)
    small.integration = msa.CRANK_NICOLSON
    with pytest.raises(msa.Unsupported):
        hip.advance_signal_grid_static(small, levels, 0.1)
