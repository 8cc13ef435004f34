"""Pins the CPU oracle (oracle/msim_oracle.c): bit for bit against the
reference's own CPU path (committed golden fixtures from oracle/_ref, and live
oracle/_ref where it was built), and against closed forms, invariants and
brute force where the mathematics fixes the answer."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import FIXED, NO_FLUX, PERIODIC, VAN_LEER, Grid, OracleError
from tests.helpers import (LAYERED_CASES, PATCH_KINDS, TRANSPORT_KINDS, TRANSPORT_SHAPES,
                           c1_inputs, divergence, golden, illcond_inputs, layered_case, monod_loss,
                           random_case, random_patches, random_transport_case, split_faces)


def c1_grid():
    c = c1_inputs()
    return Grid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"]), c


# ---------------------------------------------------------------- golden --
def test_c1_flow_matches_reference_fixture(oracle):
    grid, c = c1_grid()
    d = c["data"]
    faces, pressure, rep = oracle.solve_flow(grid, c["mobility"], 0, 100.0, 1.5e-7, 50000)
    assert rep.iterations == int(d["iterations"])
    assert np.array_equal(faces, d["faces"])
    assert np.float32(rep.max_speed) == np.float32(d["max_speed"])


def test_c1_pressure_1e10_matches_reference_fixture(oracle):
    grid, c = c1_grid()
    _, pressure, rep = oracle.solve_flow(grid, c["mobility"], 0, 100.0, 1e-10, 100000)
    assert np.array_equal(pressure, c["data"]["pressure"])
    assert rep.relative_residual <= 1e-10


def test_c1_assembly_matches_reference_fixture(oracle):
    grid, c = c1_grid()
    a = oracle.assemble(grid, c["mobility"], 0)
    for key, ref_key in (("diagonal", "diagonal"), ("rhs", "rhs"), ("bc", "bc")):
        assert np.array_equal(a[key], c["data"][ref_key]), key


def test_c1_monod_transport_matches_reference_fixture(oracle):
    grid, c = c1_grid()
    d = golden("c1_adr.npz")
    faces = c["data"]["faces"]
    levels = oracle.adr_step_f32(grid, np.zeros(grid.sites, np.float32), float(d["dt"]),
                                 steps=int(d["steps"]), diffusion=d["diffusion"],
                                 boundary_values=d["boundary_values"], faces=faces, phi=c["phi"],
                                 vmax=d["vmax"], monod_k=d["K"])
    assert np.array_equal(levels, d["levels"])


def test_ill_conditioned_case_matches_reference_fixture(oracle):
    c = illcond_inputs()
    grid = Grid(64, 64, 1, 1.0, 1.0, 1.0, c["obstacles"],
                boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    faces, _, rep = oracle.solve_flow(grid, c["mobility"], 1, 1.0, 1e-6, 50000)
    assert rep.iterations == int(c["data"]["iterations"])
    assert np.array_equal(faces, c["data"]["faces"])


def test_small_cases_match_reference_fixture(oracle):
    d = golden("small_cases.npz")
    grid = Grid(5, 8, 1, 1.0, 1.0, 1.0, d["obstacles"],
                boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    faces, _, rep = oracle.solve_flow(grid, d["mobility"], 1, 1.0, 1e-6, 50000)
    assert rep.iterations == int(d["iterations"])
    assert np.array_equal(faces, d["faces"])
    line = Grid(3, 1, 1, boundary=(NO_FLUX,) * 6)
    levels = oracle.adr_step_f32(line, [0, 1, 0], 0.25, diffusion=1.0)
    assert np.array_equal(levels, d["line_levels"])


# ------------------------------------------------------------ live _ref --
@pytest.mark.parametrize("seed,nx,ny,axis", [(1, 17, 9, 0), (2, 9, 23, 1), (3, 40, 33, 0)])
def test_random_flow_matches_reference_live(oracle, reference, seed, nx, ny, axis):
    solid, mobility = random_case(seed, nx, ny)
    if axis == 1:
        solid = solid.reshape(nx, ny)
        solid[nx // 2, :] = 0
        solid = solid.reshape(-1)
        mobility = mobility.reshape(nx, ny)
        mobility[nx // 2, :] = 0.5
        mobility = mobility.reshape(-1)
    kinds = (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX) if axis == 0 else \
        (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX)
    grid = Grid(nx, ny, 1, 0.7, 1.3, 2.0, solid, boundary=kinds)
    f_o, _, rep = oracle.solve_flow(grid, mobility, axis, 3.0, 2e-7, 20000)
    f_r, rr = reference.solve_flow(grid, mobility, axis, 3.0, 2e-7, 20000)
    assert rep.iterations == rr["iterations"]
    assert np.array_equal(f_o, f_r)


@pytest.mark.parametrize("kinds", [
    (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX),
    (FIXED, NO_FLUX, PERIODIC, PERIODIC, NO_FLUX, NO_FLUX),
    (PERIODIC, PERIODIC, FIXED, FIXED, NO_FLUX, NO_FLUX),
])
def test_random_transport_matches_reference_live(oracle, reference, kinds):
    rng = np.random.default_rng(11)
    nx, ny, S = 12, 10, 2
    solid = (rng.random(nx * ny) < 0.15).astype(np.uint8)
    grid = Grid(nx, ny, 1, 1.0, 0.5, 1.0, solid, boundary=kinds)
    levels = rng.random(S * nx * ny).astype(np.float32)
    levels[np.tile(solid, S) != 0] = 0
    loss = (rng.random(S * nx * ny) * 0.5).astype(np.float32)
    source = (rng.random(S * nx * ny) * 0.2).astype(np.float32)
    loss[np.tile(solid, S) != 0] = 0
    source[np.tile(solid, S) != 0] = 0
    bv = rng.random(6 * S).astype(np.float32)
    diffusion = np.array([0.3, 0.05], np.float32)
    a = oracle.adr_step_f32(grid, levels, 0.05, steps=3, diffusion=diffusion,
                            boundary_values=bv, source=source, loss=loss)
    b, _ = reference.advance_signal_grid(grid, levels, 0.05, 3, diffusion=diffusion,
                                         boundary_values=bv, source=source, loss=loss)
    assert np.array_equal(a, b)


def test_apply_matches_reference_live(oracle, reference):
    solid, mobility = random_case(5, 31, 19)
    grid = Grid(31, 19, 1, 1.0, 1.0, 1.0, solid)
    x = np.random.default_rng(0).standard_normal(31 * 19)
    system = reference.flow_system(grid, mobility, 0)
    assert np.array_equal(oracle.apply(grid, mobility, x), system.apply(x))
    ref_arrays = system.arrays()
    ours = oracle.assemble(grid, mobility, 0)
    for key in ("conductance", "diagonal", "rhs", "bc", "face_conductance", "face_pressure"):
        assert np.array_equal(ours[key], ref_arrays[key]), key


# ---------------------------------------------------- closed forms, brute --
def duct(nx, ny, solid=None):
    """Flow along y between fixed y faces (flow_test.cpp:14-25, one layer)."""
    return Grid(nx, ny, 1, 1.0, 1.0, 1.0, solid,
                boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))


def test_uniform_duct_is_plug_flow(oracle):
    """flow_test.cpp:115-129: every y face carries the mean speed 5."""
    grid = duct(4, 8)
    faces, _, rep = oracle.solve_flow(grid, None, 1, 5.0, 1e-6, 1000)
    xf, yf, zf = split_faces(faces, 4, 8)
    assert np.all(np.abs(yf - 5.0) <= 1e-5)
    assert np.all(np.abs(xf) <= 2e-5) and np.all(zf == 0)


def test_mobility_split_gives_flux_ratio(oracle):
    """flow_test.cpp:131-144: two strips of mobility 1 and 3."""
    grid = duct(2, 6)
    mobility = np.where(np.arange(12) // 6 == 0, 1.0, 3.0).astype(np.float32)
    faces, _, _ = oracle.solve_flow(grid, mobility, 1, 4.0, 1e-6, 1000)
    _, yf, _ = split_faces(faces, 2, 6)
    yf = yf.reshape(2, 7)
    assert abs(yf[1, 3] / yf[0, 3] - 3.0) <= 2e-5


def test_cubic_gap_conductance(oracle):
    """Hele-Shaw: a strip 4x deeper carries 64x the flux (mobility H^3/12)."""
    grid = duct(2, 6)
    depth = np.where(np.arange(12) // 6 == 0, 1.0, 4.0)
    mobility = (depth ** 3 / 12.0).astype(np.float32)
    faces, _, _ = oracle.solve_flow(grid, mobility, 1, 1.0, 1e-8, 1000)
    _, yf, _ = split_faces(faces, 2, 6)
    yf = yf.reshape(2, 7)
    assert abs(yf[1, 3] / yf[0, 3] - 64.0) <= 64.0 * 3e-5


def test_operator_against_independent_assembly(oracle):
    """The dense operator from the oracle's apply equals a textbook
    finite-volume assembly written here from the mobility: symmetric, and
    positive definite on the active sites."""
    nx, ny = 6, 5
    solid, mobility = random_case(9, nx, ny, solid_fraction=0.15)
    grid = Grid(nx, ny, 1, 1.0, 2.0, 1.0, solid)
    n = nx * ny
    A = np.column_stack([oracle.apply(grid, mobility, np.eye(n)[k]) for k in range(n)])
    m = mobility.astype(np.float64).reshape(nx, ny)
    h = (1.0, 2.0)
    B = np.zeros((n, n))

    def harm(a, b):
        return 2 * a * b / (a + b) if a + b > 0 else 0.0
    active = np.zeros(n, bool)
    for x in range(nx):
        for y in range(ny):
            i = x * ny + y
            if m[x, y] == 0:
                continue
            for (dx_, dy_), hh in (((-1, 0), h[0]), ((1, 0), h[0]), ((0, -1), h[1]),
                                   ((0, 1), h[1])):
                xx, yy = x + dx_, y + dy_
                if 0 <= xx < nx and 0 <= yy < ny:
                    c = harm(m[x, y], m[xx, yy]) / hh ** 2
                    B[i, i] += c
                    B[i, xx * ny + yy] -= c
                elif dx_ != 0:  # open ends of the x axis, pressure half a spacing out
                    B[i, i] += 2 * m[x, y] / hh ** 2
            active[i] = B[i, i] != 0
    act = np.ix_(active, active)
    assert np.allclose(A[act], B[act], rtol=2e-6, atol=1e-12)
    assert np.allclose(A, A.T, rtol=0, atol=1e-12)
    assert np.all(np.linalg.eigvalsh(A[act]) > 0)


def test_pcg_matches_dense_cholesky(oracle):
    """Brute force on a tiny grid: the PCG pressure solves A p = b."""
    nx, ny = 7, 6
    solid, mobility = random_case(4, nx, ny, solid_fraction=0.1)
    grid = Grid(nx, ny, 1, 1.0, 1.0, 1.0, solid)
    n = nx * ny
    A = np.column_stack([oracle.apply(grid, mobility, np.eye(n)[k]) for k in range(n)])
    b = oracle.assemble(grid, mobility, 0)["rhs"].astype(np.float64)
    active = np.diag(A) != 0
    Aa = A[np.ix_(active, active)]
    L = np.linalg.cholesky(Aa)
    exact = np.linalg.solve(L.T, np.linalg.solve(L, b[active]))
    _, pressure, rep = oracle.solve_flow(grid, mobility, 0, 1.0, 1e-13, 1000)
    assert np.allclose(pressure[active], exact, rtol=1e-10, atol=1e-12)
    assert rep.true_relative_residual <= 1e-12


def test_velocity_is_divergence_free(oracle):
    grid, c = c1_grid()
    faces, _, rep = oracle.solve_flow(grid, c["mobility"], 0, 100.0, 1e-10, 100000)
    div = divergence(faces, c["nx"], c["ny"])
    fluid = c["obstacles"].reshape(c["nx"], c["ny"]) == 0
    assert np.abs(div[fluid]).max() <= 1e-6 * rep.max_speed


def test_diffusion_stencil_and_reaction_closed_forms(oracle):
    """signal_grid_test.cpp:44-60 and 90-103."""
    line = Grid(3, 1, 1, boundary=(NO_FLUX,) * 6)
    levels = oracle.adr_step_f32(line, [0, 1, 0], 0.25, diffusion=1.0)
    assert np.allclose(levels, [0.25, 0.5, 0.25], atol=1e-6)
    assert abs(levels.sum() - 1.0) <= 1e-6
    with pytest.raises(OracleError) as err:
        oracle.adr_step_f32(line, levels, 0.51, diffusion=1.0)
    assert err.value.code == 1
    one = Grid(1, 1, 1, boundary=(NO_FLUX,) * 6)
    c = oracle.adr_step_f32(one, [1.0], 0.5, diffusion=0.0, source=[2.0], loss=[0.5])
    assert abs(c[0] - 1.75) <= 1e-6
    with pytest.raises(OracleError):
        oracle.adr_step_f32(one, [1.0], 0.51, diffusion=0.0, source=[2.0], loss=[2.0])


def test_monod_equals_affine_loss_under_forward_euler(oracle):
    """The Monod rate -vmax*phi*c/(K+c) is the affine loss vmax*phi/(K+c_n)
    applied to c_n: identical bits."""
    grid, c = c1_grid()
    faces = c["data"]["faces"]
    rng = np.random.default_rng(3)
    levels = (rng.random(grid.sites) * 0.5).astype(np.float32)
    levels[c["obstacles"] != 0] = 0
    bv = np.zeros(6, np.float32)
    bv[0] = 1.0
    a = oracle.adr_step_f32(grid, levels, 4e-4, diffusion=500.0, boundary_values=bv,
                            faces=faces, phi=c["phi"], vmax=1.0, monod_k=0.1)
    loss = monod_loss(1.0, c["phi"], 0.1, levels)
    b = oracle.adr_step_f32(grid, levels, 4e-4, diffusion=500.0, boundary_values=bv,
                            faces=faces, loss=loss)
    assert np.array_equal(a, b)


def test_fp64_transport_conserves_and_agrees(oracle):
    grid, c = c1_grid()
    faces = c["data"]["faces"]
    bv = np.zeros(6, np.float32)
    bv[0] = 1.0
    kw = dict(diffusion=500.0, boundary_values=bv, faces=faces, phi=c["phi"], vmax=1.0,
              monod_k=0.1)
    l32 = oracle.adr_step_f32(grid, np.zeros(grid.sites, np.float32), 4e-4, steps=20, **kw)
    l64, ledger = oracle.adr_step_f64(grid, np.zeros(grid.sites), 4e-4, steps=20, **kw)
    assert np.abs(l64 - l32).max() <= 1e-5
    balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
    assert np.abs(balance).max() <= 1e-12 * ledger[:, :, 1].max()


def test_fp64_periodic_faces_join_partner_sites(oracle):
    """On a grid periodic in x and y with a face field crossing both periodic
    boundaries, the FP64 face-flux step agrees with the FP32 step (bit-exact
    with the reference, which wraps periodic neighbours) and keeps the mass
    exactly: nothing enters or leaves."""
    rng = np.random.default_rng(8)
    nx, ny = 14, 11
    solid = (rng.random(nx * ny) < 0.1).astype(np.uint8)
    kinds = (PERIODIC, PERIODIC, PERIODIC, PERIODIC, NO_FLUX, NO_FLUX)
    grid = Grid(nx, ny, 1, 1.0, 0.5, 1.0, solid, boundary=kinds)
    xf = rng.uniform(0.2, 1.0, (nx + 1) * ny).astype(np.float32)  # flow across x = 0
    xf[nx * ny:] = xf[:ny]
    yf = rng.uniform(-1.0, 1.0, (nx, ny + 1)).astype(np.float32)
    yf[:, ny] = yf[:, 0]
    faces = np.concatenate([xf, yf.reshape(-1), np.zeros(2 * nx * ny, np.float32)])
    levels = 0.5 + 0.5 * rng.random(nx * ny)
    levels[solid != 0] = 0.0
    kw = dict(diffusion=0.1, faces=faces)
    l32 = oracle.adr_step_f32(grid, levels.astype(np.float32), 0.02, steps=5, **kw)
    l64, ledger = oracle.adr_step_f64(grid, levels, 0.02, steps=5, **kw)
    assert np.abs(l64 - l32).max() <= 1e-6
    assert np.all(ledger[:, :, 2] == 0.0) and np.all(ledger[:, :, 3] == 0.0)
    assert np.abs(ledger[:, :, 1] - ledger[:, :, 0]).max() <= 1e-13 * ledger[0, 0, 0]
    # The periodic faces carry mass: closing them changes the result.
    closed = Grid(nx, ny, 1, 1.0, 0.5, 1.0, solid, boundary=(NO_FLUX,) * 6)
    lc, _ = oracle.adr_step_f64(closed, levels, 0.02, steps=5, **kw)
    assert np.abs(lc - l64).max() > 1e-3


@pytest.mark.parametrize("kinds", TRANSPORT_KINDS)
def test_fp64_step_agrees_with_fp32_step(oracle, kinds):
    """The FP64 face-flux step against the FP32 step (bit-exact with the
    reference) on every boundary kind, ragged shapes and axes of extent one,
    with a face field, affine rates and Monod uptake; and its ledger balances."""
    for nx, ny in TRANSPORT_SHAPES:
        d = random_transport_case(nx, ny, 2, kinds, seed=nx * 1000 + ny)
        grid = Grid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], boundary=kinds)
        kw = dict(diffusion=d["diffusion"], boundary_values=d["bv"], source=d["source"],
                  loss=d["loss"], faces=np.concatenate([d["xf"], d["yf"], d["zf"]]),
                  phi=d["phi"], vmax=[0.7, 0.3], monod_k=[0.2, 0.5])
        a = oracle.adr_step_f32(grid, d["levels"].astype(np.float32), 0.02, steps=3, **kw)
        b, ledger = oracle.adr_step_f64(grid, d["levels"], 0.02, steps=3, **kw)
        assert np.abs(a - b).max() <= 1e-6, (nx, ny)
        balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
        assert np.abs(balance).max() <= 1e-12 * ledger[:, :, 1].max()


VAN_LEER_SHAPES = TRANSPORT_SHAPES + [(2, 30), (30, 2), (3, 17)]


def van_leer_kw(d, monod=False):
    kw = dict(diffusion=d["diffusion"], boundary_values=d["bv"], source=d["source"],
              loss=d["loss"], faces=np.concatenate([d["xf"], d["yf"], d["zf"]]))
    if monod:
        kw.update(phi=d["phi"], vmax=[0.7, 0.3], monod_k=[0.2, 0.5])
    return kw


@pytest.mark.parametrize("kinds", TRANSPORT_KINDS)
def test_van_leer_transport_matches_reference_live(oracle, reference, kinds):
    """The limited (van Leer) step, bit for bit with advance_signal_grid_cpu,
    on every boundary kind, axes of two and three sites (where a periodic
    axis stops or starts taking corrections) and ragged shapes."""
    for nx, ny in VAN_LEER_SHAPES:
        d = random_transport_case(nx, ny, 2, kinds, seed=nx * 1000 + ny)
        grid = Grid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], boundary=kinds)
        lv = d["levels"].astype(np.float32)
        a = oracle.adr_step_f32(grid, lv, 0.02, steps=3, scheme=VAN_LEER, **van_leer_kw(d))
        b, _ = reference.advance_signal_grid(grid, lv, 0.02, 3, scheme=VAN_LEER, **van_leer_kw(d))
        assert np.array_equal(a, b), (nx, ny)
        upwind = oracle.adr_step_f32(grid, lv, 0.02, steps=3, **van_leer_kw(d))
        assert np.abs(a - upwind).max() > 1e-4  # the correction is active


def test_van_leer_front_matches_reference_live(oracle, reference):
    """A step front carried at half the limited scheme's bound for 40 steps:
    levels stay within [0, 1] and match the reference bit for bit."""
    nx, ny = 60, 8
    lv = np.zeros(nx * ny, np.float32)
    lv.reshape(nx, ny)[:15] = 1.0
    xf = np.ones((nx + 1) * ny, np.float32)
    xf[:ny] = 0.0
    xf[nx * ny:] = 0.0
    faces = np.concatenate([xf, np.zeros(nx * (ny + 1), np.float32),
                            np.zeros(2 * nx * ny, np.float32)])
    grid = Grid(nx, ny, 1, 1.0, 1.0, 1.0, None, boundary=(NO_FLUX,) * 6)
    a = oracle.adr_step_f32(grid, lv, 0.49, steps=40, scheme=VAN_LEER, diffusion=0.0, faces=faces)
    b, _ = reference.advance_signal_grid(grid, lv, 0.49, 40, scheme=VAN_LEER, diffusion=0.0,
                                         faces=faces)
    assert np.array_equal(a, b)
    assert a.min() >= 0.0 and a.max() <= 1.0
    # Mass is conserved exactly up to float rounding on a closed domain.
    assert abs(float(a.astype(np.float64).sum()) - 15 * ny) <= 1e-4


def test_van_leer_halves_the_advective_bound(oracle, reference):
    """dt between the limited and the upwind bounds: only the limited step is
    refused, by both."""
    nx, ny = 20, 6
    xf = np.full((nx + 1) * ny, 2.0, np.float32)
    xf[:ny] = 0.0
    xf[nx * ny:] = 0.0
    faces = np.concatenate([xf, np.zeros(nx * (ny + 1), np.float32),
                            np.zeros(2 * nx * ny, np.float32)])
    grid = Grid(nx, ny, 1, 1.0, 1.0, 1.0, None, boundary=(NO_FLUX,) * 6)
    lv = np.ones(nx * ny, np.float32)
    dt = 0.4  # upwind bound 0.5, limited 0.25
    oracle.adr_step_f32(grid, lv, dt, diffusion=0.0, faces=faces)
    reference.advance_signal_grid(grid, lv, dt, 1, diffusion=0.0, faces=faces)
    with pytest.raises(OracleError, match="explicit stability bound"):
        oracle.adr_step_f32(grid, lv, dt, scheme=VAN_LEER, diffusion=0.0, faces=faces)
    with pytest.raises(OracleError, match="explicit stability bound"):
        reference.advance_signal_grid(grid, lv, dt, 1, scheme=VAN_LEER, diffusion=0.0,
                                      faces=faces)


@pytest.mark.parametrize("kinds", TRANSPORT_KINDS)
def test_fp64_van_leer_agrees_and_conserves(oracle, kinds):
    """The FP64 face-flux form of the limited step: each face's correction is
    one expression of the sites around it, so the ledger balances; it agrees
    with the FP32 step (bit-exact with the reference) to FP32 rounding."""
    for nx, ny in VAN_LEER_SHAPES:
        d = random_transport_case(nx, ny, 2, kinds, seed=nx * 1000 + ny)
        grid = Grid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], boundary=kinds)
        kw = van_leer_kw(d, monod=True)
        a = oracle.adr_step_f32(grid, d["levels"].astype(np.float32), 0.02, steps=3,
                                scheme=VAN_LEER, **kw)
        b, ledger = oracle.adr_step_f64(grid, d["levels"], 0.02, steps=3, scheme=VAN_LEER, **kw)
        assert np.abs(a - b).max() <= 1e-6, (nx, ny)
        balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
        assert np.abs(balance).max() <= 1e-12 * ledger[:, :, 1].max()


@pytest.mark.parametrize("axis", [0, 1])
def test_layered_shallow_flow_matches_reference_live(oracle, reference, axis):
    """Grids of several layers: reduction to columns (heights, the (H/Hmax)^2
    default, conductance), the solve on the grid of columns, the lift onto
    the voxel faces with vertical faces by continuity, and the scaling over
    the voxel inlet faces, bit for bit with the reference."""
    for nx, ny, nz, seed, mob in LAYERED_CASES:
        obstacles, mobility, boundary = layered_case(nx, ny, nz, seed, mob, axis)
        grid = Grid(nx, ny, nz, 1.0, 0.5, 0.7, obstacles, boundary=boundary)
        a, _, rep = oracle.solve_flow(grid, mobility, axis, 2.0, 1e-6, 50000)
        b, info = reference.solve_flow(grid, mobility, axis, 2.0, 1e-6, 50000)
        assert np.array_equal(a, b), (nx, ny, nz, axis)
        assert rep.iterations == info["iterations"] and rep.max_speed == info["max_speed"]
        nfx, nfy, _ = grid.face_counts()
        assert np.abs(a[nfx + nfy:]).max() > 0.0  # the vertical faces carry flow


def test_layered_shallow_flow_refusals_match_reference(oracle, reference):
    """An overhang, a raised floor and a mobility that varies down a column
    are refused with the reference's messages."""
    nx, ny, nz = 8, 6, 3
    base = np.zeros((nx, ny, nz), np.uint8)
    overhang = base.copy()
    overhang[3, 2, 1] = 1  # fluid above a solid voxel above fluid
    floor = base.copy()
    floor[4, 4, 0] = 1  # a column that starts above the common floor
    varying = np.ones((nx, ny, nz), np.float32)
    varying[2, 2, 1] = 2.0
    cases = [(overhang, None, "contiguous fluid columns"), (floor, None, "common planar floor"),
             (base, varying.reshape(-1), "constant through each depth column")]
    for obstacles, mobility, message in cases:
        grid = Grid(nx, ny, nz, 1.0, 1.0, 1.0, obstacles.reshape(-1))
        with pytest.raises(OracleError, match=message):
            oracle.solve_flow(grid, mobility, 0, 1.0, 1e-6, 1000)
        with pytest.raises(OracleError, match=message):
            reference.solve_flow(grid, mobility, 0, 1.0, 1e-6, 1000)


@pytest.mark.parametrize("kinds", PATCH_KINDS)
def test_boundary_patches_match_reference_live(oracle, reference, kinds):
    """Reservoir patches on fixed faces (GridBoundary::values_at: the last
    patch holding a site wins), upwind and van Leer, bit for bit with the
    reference; the patches change the result."""
    nx, ny, S = 30, 20, 2
    d = random_transport_case(nx, ny, S, kinds, seed=5)
    grid = Grid(nx, ny, 1, 1.0, 0.5, 1.0, d["solid"], boundary=kinds)
    patches = random_patches(kinds, nx, ny, S, seed=1)
    kw = dict(diffusion=d["diffusion"], boundary_values=d["bv"], source=d["source"],
              loss=d["loss"], faces=np.concatenate([d["xf"], d["yf"], d["zf"]]))
    lv = d["levels"].astype(np.float32)
    for scheme in (0, VAN_LEER):
        a = oracle.adr_step_f32(grid, lv, 0.02, steps=3, scheme=scheme, patches=patches, **kw)
        b, _ = reference.advance_signal_grid(grid, lv, 0.02, 3, scheme=scheme, patches=patches,
                                             **kw)
        assert np.array_equal(a, b)
        plain = oracle.adr_step_f32(grid, lv, 0.02, steps=3, scheme=scheme, **kw)
        assert np.abs(a - plain).max() > 1e-3
        c, ledger = oracle.adr_step_f64(grid, d["levels"], 0.02, steps=3, scheme=scheme,
                                        patches=patches, **kw)
        assert np.abs(a - c).max() <= 1e-6
        balance = ledger[:, :, 1] - ledger[:, :, 0] - ledger[:, :, 2] + ledger[:, :, 3]
        assert np.abs(balance).max() <= 1e-12 * ledger[:, :, 1].max()


def test_flow_error_paths(oracle):
    blocked = np.ones((8, 6), np.uint8)
    blocked[:, 2] = 0
    blocked[5, 2] = 1  # the channel is cut: no through-flow
    grid = Grid(8, 6, 1, 1.0, 1.0, 1.0, blocked.reshape(-1))
    with pytest.raises(OracleError) as err:
        oracle.solve_flow(grid, None, 0, 1.0)
    assert err.value.code == 2
    grid = duct(4, 8)
    with pytest.raises(OracleError) as err:
        oracle.solve_flow(grid, None, 1, 1.0, 1e-10, 1)
    assert err.value.code == 3
    periodic = Grid(4, 8, 1, boundary=(PERIODIC, PERIODIC, FIXED, FIXED, NO_FLUX, NO_FLUX))
    with pytest.raises(OracleError) as err:
        oracle.solve_flow(periodic, None, 1, 1.0)
    assert err.value.code == 1
    with pytest.raises(OracleError) as err:
        oracle.solve_flow(duct(4, 8), -np.ones(32, np.float32), 1, 1.0)
    assert err.value.code == 1
