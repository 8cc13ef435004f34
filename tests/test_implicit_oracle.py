"""The oracle's implicit transport step (mo_adr_step_implicit_f32) pinned
against the reference's advance_signal_grid_cpu under Crank-Nicolson and
backward Euler: bit for bit (levels, iterations, residual RMS) through the
live oracle/_ref build, and against the known answers of the reference's own
tests (tests/cpp/signal_grid_test.cpp:62-75, 104-133, 147-160).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import (BACKWARD_EULER, CRANK_NICOLSON, NO_FLUX, UPWIND, VAN_LEER, Grid,
                    OracleError)
from tests.helpers import TRANSPORT_KINDS, random_patches, random_transport_case

IMPLICIT = [CRANK_NICOLSON, BACKWARD_EULER]


def line(n):
    """line_spec (signal_grid_test.cpp:17-30): n sites along x, unit spacing,
    no-flux faces."""
    return Grid(n, 1, 1, 1.0, 1.0, 1.0, None, (NO_FLUX,) * 6)


# ----------------------------------------------------------- known answers --
def test_crank_nicolson_reaction_known_answer(oracle):
    """c' - (s - l c')/2 = c + (s - l c)/2 with c = 1, s = 2, l = 0.5, dt = 1:
    c' = 2.75 / 1.25 = 2.2 (signal_grid_test.cpp:62-75)."""
    lv, rep = oracle.adr_step_implicit(line(1), [1.0], 1.0, diffusion=(0.0,), source=[2.0],
                                       loss=[0.5], integration=CRANK_NICOLSON)
    assert rep["converged"] and abs(lv[0] - 2.2) <= 1e-5


def test_crank_nicolson_diffusion_known_answer(oracle):
    """A unit pulse on three no-flux sites, D = 1, dt = 1: 0.4, 0.2, 0.4
    (signal_grid_test.cpp:104-117)."""
    lv, rep = oracle.adr_step_implicit(line(3), [0.0, 1.0, 0.0], 1.0, diffusion=(1.0,),
                                       integration=CRANK_NICOLSON, solver=(10000, 1e-7, 1e-6))
    assert rep["converged"] and rep["iterations"] > 0 and rep["residual_rms"] <= 2e-6
    assert np.allclose(lv, [0.4, 0.2, 0.4], atol=1e-5)
    assert abs(float(lv.sum()) - 1.0) <= 1e-5


def test_backward_euler_diffusion_closed_form(oracle):
    """Backward Euler on two no-flux sites: (1 + 2k) c0' - k c1' ... with
    k = D dt = 1 gives c0' = (1 + k) / (1 + 2k) = 2/3 from c = (1, 0)."""
    lv, rep = oracle.adr_step_implicit(line(2), [1.0, 0.0], 1.0, diffusion=(1.0,),
                                       integration=BACKWARD_EULER, solver=(10000, 1e-9, 1e-9))
    assert rep["converged"]
    assert abs(lv[0] - 2.0 / 3.0) <= 1e-5 and abs(lv[1] - 1.0 / 3.0) <= 1e-5


def test_unconverged_solve_is_refused(oracle):
    """One iteration is not enough at a 1e-12 absolute tolerance
    (signal_grid_test.cpp:119-133)."""
    levels = np.zeros(9, np.float32)
    levels[4] = 1.0
    with pytest.raises(OracleError) as info:
        oracle.adr_step_implicit(line(9), levels, 2.0, diffusion=(1.0,),
                                 integration=CRANK_NICOLSON, solver=(1, 1e-12, 0.0))
    assert info.value.code == 3
    assert info.value.message == "Implicit signal solve did not converge after 1 iterations"


def test_obstacle_splits_the_line(oracle):
    """An obstacle site separates two segments, each keeping its mass
    (signal_grid_test.cpp:147-160)."""
    grid = Grid(4, 1, 1, 1.0, 1.0, 1.0, np.array([0, 0, 1, 0], np.uint8), (NO_FLUX,) * 6)
    lv, rep = oracle.adr_step_implicit(grid, [2.0, 0.0, 0.0, 5.0], 1.0, diffusion=(1.0,),
                                       integration=CRANK_NICOLSON)
    assert rep["converged"]
    assert abs(lv[0] + lv[1] - 2.0) <= 1e-4 and lv[2] == 0.0 and abs(lv[3] - 5.0) <= 1e-4


@pytest.mark.parametrize("case", ["max_iterations", "tolerances", "kind"])
def test_solver_parameters_are_validated(oracle, case):
    kw = dict(diffusion=(1.0,), integration=CRANK_NICOLSON)
    if case == "max_iterations":
        kw["solver"], message = (0, 1e-7, 1e-5), "signal solver iteration limit must be positive"
    elif case == "tolerances":
        kw["solver"], message = (10, 0.0, 0.0), "signal solver tolerances must be finite and non-negative"
    else:
        kw["integration"], message = 7, "unknown signal integration kind"
    with pytest.raises(OracleError) as info:
        oracle.adr_step_implicit(line(3), [0.0, 1.0, 0.0], 1.0, **kw)
    assert info.value.message == message


# ------------------------------------------------------------ live _ref --
def implicit_kw(c, faces, scheme, integration):
    return dict(diffusion=c["diffusion"], boundary_values=c["bv"], faces=faces,
                source=c["source"], loss=c["loss"], scheme=scheme, integration=integration)


@pytest.mark.parametrize("integration", IMPLICIT)
@pytest.mark.parametrize("scheme", [UPWIND, VAN_LEER])
@pytest.mark.parametrize("kinds", TRANSPORT_KINDS)
def test_implicit_steps_match_reference_live(oracle, reference, kinds, scheme, integration):
    """Every boundary kind, both schemes and both implicit kinds, on ragged
    shapes and axes of extent one, over two steps of a large dt: levels,
    iterations and residual RMS bit for bit."""
    for nx, ny in [(17, 23), (40, 9), (1, 30), (25, 1), (3, 3)]:
        S = 1 + (nx % 3)
        c = random_transport_case(nx, ny, S, kinds, nx * 31 + ny)
        grid = Grid(nx, ny, 1, 0.7, 1.1, 1.0, c["solid"], kinds)
        faces = np.concatenate([c["xf"], c["yf"], c["zf"]])
        kw = implicit_kw(c, faces, scheme, integration)
        a, rep = oracle.adr_step_implicit(grid, c["levels"], 0.5, steps=2, **kw)
        b, _, ref = reference.advance_signal_grid(grid, c["levels"], 0.5, steps=2,
                                                  with_report=True, **kw)
        assert np.array_equal(a, b), (nx, ny)
        assert rep == ref, (nx, ny)


@pytest.mark.parametrize("integration", IMPLICIT)
def test_implicit_patches_match_reference_live(oracle, reference, integration):
    for kinds in [(2, 2, 0, 0, 0, 0), (2, 2, 2, 2, 0, 0)]:
        c = random_transport_case(30, 21, 2, kinds, 5)
        patches = random_patches(kinds, 30, 21, 2, 6)
        grid = Grid(30, 21, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
        faces = np.concatenate([c["xf"], c["yf"], c["zf"]])
        kw = implicit_kw(c, faces, VAN_LEER, integration)
        a, rep = oracle.adr_step_implicit(grid, c["levels"], 0.3, patches=patches, **kw)
        b, _, ref = reference.advance_signal_grid(grid, c["levels"], 0.3, patches=patches,
                                                  with_report=True, **kw)
        assert np.array_equal(a, b) and rep == ref


def test_negative_implicit_levels_are_refused_as_the_reference_does(oracle, reference):
    """Crank-Nicolson with van Leer at a long step leaves levels below zero
    beyond rounding; both refuse the step with the same message."""
    kinds = (2, 2, 0, 0, 0, 0)
    c = random_transport_case(30, 21, 2, kinds, 5)
    grid = Grid(30, 21, 1, 1.0, 0.8, 1.0, c["solid"], kinds)
    faces = np.concatenate([c["xf"], c["yf"], c["zf"]])
    kw = implicit_kw(c, faces, VAN_LEER, CRANK_NICOLSON)
    messages = []
    for step in (lambda: oracle.adr_step_implicit(grid, c["levels"], 0.8, **kw),
                 lambda: reference.advance_signal_grid(grid, c["levels"], 0.8, **kw)):
        with pytest.raises(OracleError) as info:
            step()
        messages.append((info.value.code, info.value.message))
    assert messages[0] == messages[1] == (1, "signal grid levels must be finite and non-negative")


def test_implicit_monod_is_the_lagged_affine_loss(oracle, reference):
    """Monod uptake in an implicit step is the loss vmax*phi/(K+c) of the
    step's starting levels added to the affine loss: the reference step with
    that loss array gives the same levels bit for bit."""
    nx, ny, S = 26, 19, 2
    kinds = TRANSPORT_KINDS[0]
    c = random_transport_case(nx, ny, S, kinds, 9)
    grid = Grid(nx, ny, 1, 1.0, 1.0, 1.0, c["solid"], kinds)
    faces = np.concatenate([c["xf"], c["yf"], c["zf"]])
    vmax, K = np.float32(0.7), np.float32(0.2)
    phi = np.random.default_rng(3).random(nx * ny).astype(np.float32)
    phi[c["solid"] != 0] = 0.0
    levels = c["levels"].astype(np.float32)
    uptake = (vmax * np.tile(phi, S)) / (K + levels)
    lagged = (c["loss"] + uptake).astype(np.float32)
    for integration in IMPLICIT:
        a, rep = oracle.adr_step_implicit(
            grid, levels, 0.5, diffusion=c["diffusion"], boundary_values=c["bv"], faces=faces,
            source=c["source"], loss=c["loss"], phi=phi, vmax=vmax, monod_k=K,
            integration=integration)
        b, _, ref = reference.advance_signal_grid(
            grid, levels, 0.5, diffusion=c["diffusion"], boundary_values=c["bv"], faces=faces,
            source=c["source"], loss=lagged, integration=integration, with_report=True)
        assert np.array_equal(a, b) and rep == ref


def test_three_dimensional_backward_euler_matches_reference_live(oracle, reference):
    """The oracle's implicit step is general in z: a 4x8x4 box between fixed
    y faces (signal_grid_test.cpp:585-610's grid), a few steps."""
    grid = Grid(4, 8, 4, 2.0, 2.0, 0.75, None,
                (NO_FLUX, NO_FLUX, 2, 2, NO_FLUX, NO_FLUX))
    bv = np.zeros(6, np.float32)
    bv[2] = 10.0
    levels = np.full(4 * 8 * 4, 10.0, np.float32)
    a, rep = oracle.adr_step_implicit(grid, levels, 0.02, steps=3, diffusion=(40.0,),
                                      boundary_values=bv, integration=BACKWARD_EULER)
    b, _, ref = reference.advance_signal_grid(grid, levels, 0.02, steps=3, diffusion=(40.0,),
                                              boundary_values=bv, integration=BACKWARD_EULER,
                                              with_report=True)
    assert np.array_equal(a, b) and rep == ref and rep["iterations"] < 200


@pytest.mark.parametrize("integration", IMPLICIT)
@pytest.mark.parametrize("scheme", [UPWIND, VAN_LEER])
def test_implicit_steps_match_the_committed_fixture(oracle, scheme, integration):
    """The channel + trap 256x128 over the reference's flow field, three
    implicit steps with Monod uptake from the forward-Euler fixture's levels:
    levels and the last report equal the reference's vectors in
    tests/golden/c1_implicit.npz (made by tests/golden/make_golden.py)."""
    from tests.helpers import c1_inputs, golden
    c = c1_inputs()
    adr, d = golden("c1_adr.npz"), golden("c1_implicit.npz")
    grid = Grid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"])
    levels, rep = oracle.adr_step_implicit(
        grid, adr["levels"], float(d["dt"]), steps=int(d["steps"]), diffusion=adr["diffusion"],
        boundary_values=adr["boundary_values"], faces=c["data"]["faces"], phi=c["phi"],
        vmax=adr["vmax"], monod_k=adr["K"], scheme=scheme, integration=integration)
    key = f"i{integration}s{scheme}"
    assert np.array_equal(levels, d[f"{key}_levels"])
    assert [rep["converged"], rep["iterations"], rep["residual_rms"]] == \
        d[f"{key}_report"].tolist()


def test_forward_euler_van_leer_matches_the_committed_fixture(oracle):
    """Three forward-Euler van Leer steps with Monod on the same case, at the
    forward-Euler fixture's dt: the reference's levels bit for bit."""
    from tests.helpers import c1_inputs, golden
    c = c1_inputs()
    adr, d = golden("c1_adr.npz"), golden("c1_implicit.npz")
    grid = Grid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"])
    levels = oracle.adr_step_f32(
        grid, adr["levels"], float(adr["dt"]), steps=3, diffusion=adr["diffusion"],
        boundary_values=adr["boundary_values"], faces=c["data"]["faces"], phi=c["phi"],
        vmax=adr["vmax"], monod_k=adr["K"], scheme=VAN_LEER)
    assert np.array_equal(levels, d["i0s1_levels"])
