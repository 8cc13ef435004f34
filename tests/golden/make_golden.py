"""Writes the golden fixtures of tests/golden/ from the reference's own CPU
path (oracle/_ref/libmsim_ref.so, built from the reference sources by
`make -C oracle ref`). Run in a container that holds the reference tree:

    python tests/golden/make_golden.py

Fixtures (inputs and the reference's outputs, nothing else):
  c1_flow.npz      channel + trap 256x128, seed 42: solve_depth_averaged_flow_cpu
                   at tolerance 1.5e-7 (the reference floor is float epsilon),
                   mean inlet speed 100; and detail::solve_depth_averaged_pressure
                   at 1e-10 (no floor there).
  c1_adr.npz       20 forward-Euler steps of one nutrient over that field, D=500,
                   inlet 1, Monod (vmax 1, K 0.1) as the affine loss it equals
                   under forward Euler, recomputed from the levels each step.
  illcond_flow.npz the reference conformance suite's 64x64 pillar sheet with
                   mobility 10^U(-3,0) from its LCG (seed 12345)
                   (tests/conformance/flow_conformance_test.cpp:94-116).
  small_cases.npz  a 5x8 obstacle case with mobility 0.4+0.1x
                   (flow_conformance_test.cpp:62-88 pattern, one layer) and a
                   3-site diffusion line (signal_grid_test.cpp:90-103).
  port_flow.npz    solve_depth_averaged_port_flow_cpu on the conformance
                   suite's two port cases (flow_conformance_test.cpp:167-271,
                   built by tests/helpers.py): faces, port states (pressure,
                   flux, mean speed), solves and iterations.
  coupled.npz      advance_coupled_cpu on a one-layer version of the
                   conformance suite's coupled colony (513 cells, 3 species,
                   2 signals; coupled_rates_conformance_test.cpp:18-118, built
                   by tests/helpers.py), one step of dt 0.05 under each
                   integration kind and scheme: its inputs, the grid levels,
                   species and the implicit report.
  c1_implicit.npz  three Crank-Nicolson and backward-Euler steps (upwind and van
                   Leer) of ten times c1_adr's dt over the c1 field from
                   c1_adr's levels, Monod as the lagged loss; levels and the
                   last report. Also three forward-Euler van Leer steps at
                   c1_adr's dt (i0s1_levels).

Usage: python tests/golden/make_golden.py [fixture ...] (default: all).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import FIXED, NO_FLUX, Grid, Reference, build  # noqa: E402
from microsimulator_amd.workloads import config  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def c1():
    w = config("c1")
    grid = Grid(w.nx, w.ny, 1, w.dx, w.dy, w.dz, w.obstacles)
    ref = Reference()
    faces, rep = ref.solve_flow(grid, w.mobility, 0, 100.0, 1.5e-7, 50000)
    system = ref.flow_system(grid, w.mobility, 0)
    pressure, _ = system.pressure(1e-10, 100000)
    arrays = system.arrays()
    np.savez_compressed(
        os.path.join(OUT, "c1_flow.npz"), obstacles=np.packbits(w.obstacles.reshape(-1)),
        mobility=w.mobility.reshape(-1), phi=w.phi.reshape(-1), nx=w.nx, ny=w.ny, dz=w.dz,
        faces=faces, iterations=rep["iterations"], relative_residual=rep["relative_residual"],
        max_speed=rep["max_speed"], pressure=pressure, diagonal=arrays["diagonal"],
        rhs=arrays["rhs"], bc=arrays["bc"])
    # Transport over the reference's field.
    S = 1
    bv = np.zeros(6 * S, np.float32)
    bv[0] = 1.0
    D = np.array([500.0], np.float32)
    vmax, K = np.float32(1.0), np.float32(0.1)
    dt = np.float32(4.0e-4)
    phi = w.phi.reshape(-1)
    levels = np.zeros(w.sites, np.float32)
    for _ in range(20):
        loss = ((vmax * phi) / (K + levels)).astype(np.float32)
        levels, _ = ref.advance_signal_grid(grid, levels, dt, 1, diffusion=D, boundary_values=bv,
                                            faces=faces, loss=loss)
    np.savez_compressed(os.path.join(OUT, "c1_adr.npz"), levels=levels, dt=dt, diffusion=D,
                        boundary_values=bv, vmax=vmax, K=K, steps=20)


def illcond():
    size = 64
    solid = np.zeros((size, size), np.uint8)
    for px in range(8, size, 16):
        for py in range(12, size - 12, 16):
            solid[px:px + 4, py:py + 4] = 1
    flat = solid.reshape(-1)
    mobility = np.zeros(size * size, np.float32)
    state = 12345
    for site in range(size * size):
        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
        exponent = np.float32(-3.0) * np.float32(state >> 8) / np.float32(16777216.0)
        mobility[site] = 0.0 if flat[site] else np.float32(10.0) ** exponent
    # The conformance case flows along y with fixed y faces.
    grid = Grid(size, size, 1, 1.0, 1.0, 1.0, solid,
                boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    ref = Reference()
    faces, rep = ref.solve_flow(grid, mobility, 1, 1.0, 1e-6, 50000)
    np.savez_compressed(os.path.join(OUT, "illcond_flow.npz"), obstacles=flat, mobility=mobility,
                        faces=faces, iterations=rep["iterations"], max_speed=rep["max_speed"])


def small():
    nx, ny = 5, 8
    solid = np.zeros((nx, ny), np.uint8)
    solid[2, 3] = 1
    x = np.arange(nx, dtype=np.float32)[:, None] * np.ones((1, ny), np.float32)
    mobility = (np.float32(0.4) + np.float32(0.1) * x).astype(np.float32).reshape(-1)
    grid = Grid(nx, ny, 1, 1.0, 1.0, 1.0, solid,
                boundary=(NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    ref = Reference()
    faces, rep = ref.solve_flow(grid, mobility, 1, 1.0, 1e-6, 50000)
    line = Grid(3, 1, 1, 1.0, 1.0, 1.0, None,
                boundary=(NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX))
    levels, _ = ref.advance_signal_grid(line, np.array([0, 1, 0], np.float32), 0.25, 1,
                                        diffusion=(1.0,))
    np.savez_compressed(os.path.join(OUT, "small_cases.npz"), obstacles=solid.reshape(-1),
                        mobility=mobility, faces=faces, iterations=rep["iterations"],
                        line_levels=levels)


def ports():
    from tests.helpers import depth_port_case, winding_channel_case
    ref = Reference()
    out = {}
    for name, builder in (("depth", depth_port_case), ("winding", winding_channel_case)):
        g, mobility, case_ports = builder()
        faces, states, rep = ref.port_flow(Grid(**g), mobility, case_ports)
        out[f"{name}_faces"] = faces
        out[f"{name}_states"] = states
        out[f"{name}_solves"] = rep["solves"]
        out[f"{name}_iterations"] = rep["iterations"]
    np.savez_compressed(os.path.join(OUT, "port_flow.npz"), **out)


def coupled():
    from tests.helpers import coupled_colony_case
    ref = Reference()
    g, levels, cells, species, plan, previous, transport = coupled_colony_case()
    out = {"levels": levels, "species": species, "previous": previous,
           "obstacles": g["obstacles"], **{f"cell_{k}": v for k, v in cells.items()},
           **{f"transport_{k}": v for k, v in transport.items()}}
    for integration in (0, 1, 2):
        for scheme in (0, 1):
            lv, sp, rep = ref.advance_coupled(Grid(**g), levels, cells, species, plan, previous,
                                              0.05, origin=(-3.0, -2.0, 0.0), scheme=scheme,
                                              integration=integration, **transport)
            key = f"i{integration}s{scheme}"
            out[f"{key}_levels"] = lv
            out[f"{key}_species"] = sp
            out[f"{key}_report"] = np.array([rep["converged"], rep["iterations"],
                                             rep["residual_rms"]], np.float64)
    np.savez_compressed(os.path.join(OUT, "coupled.npz"), **out)


def implicit():
    """Crank-Nicolson and backward Euler, upwind and van Leer, on the channel
    + trap 256x128 over the reference's own flow field: three steps of ten
    times the forward-Euler dt from the forward-Euler fixture's levels, Monod
    uptake given as the lagged loss vmax*phi/(K+c) of each step's start;
    and three forward-Euler van Leer steps at the forward-Euler dt."""
    from tests.helpers import c1_inputs
    c = c1_inputs()
    d = np.load(os.path.join(OUT, "c1_adr.npz"))
    grid = Grid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"])
    ref = Reference()
    vmax, K = np.float32(d["vmax"]), np.float32(d["K"])
    dt = np.float32(10 * d["dt"])
    out = {"dt": dt, "steps": 3}
    for integration in (1, 2):
        for scheme in (0, 1):
            levels = d["levels"].astype(np.float32)
            for _ in range(3):
                loss = ((vmax * c["phi"]) / (K + levels)).astype(np.float32)
                levels, _, rep = ref.advance_signal_grid(
                    grid, levels, dt, 1, diffusion=d["diffusion"],
                    boundary_values=d["boundary_values"], faces=c["data"]["faces"], loss=loss,
                    scheme=scheme, integration=integration, with_report=True)
            key = f"i{integration}s{scheme}"
            out[f"{key}_levels"] = levels
            out[f"{key}_report"] = np.array([rep["converged"], rep["iterations"],
                                             rep["residual_rms"]], np.float64)
    # Forward Euler under van Leer, at the forward-Euler fixture's dt.
    levels = d["levels"].astype(np.float32)
    for _ in range(3):
        loss = ((vmax * c["phi"]) / (K + levels)).astype(np.float32)
        levels, _ = ref.advance_signal_grid(
            grid, levels, np.float32(d["dt"]), 1, diffusion=d["diffusion"],
            boundary_values=d["boundary_values"], faces=c["data"]["faces"], loss=loss, scheme=1)
    out["i0s1_levels"] = levels
    np.savez_compressed(os.path.join(OUT, "c1_implicit.npz"), **out)


FIXTURES = {"c1": c1, "illcond": illcond, "small": small, "ports": ports, "coupled": coupled,
            "implicit": implicit}

if __name__ == "__main__":
    build(reference=True)
    for name in sys.argv[1:] or list(FIXTURES):
        FIXTURES[name]()
    for name in sorted(os.listdir(OUT)):
        if name.endswith(".npz"):
            print(name, os.path.getsize(os.path.join(OUT, name)))
