"""The oracle's transport on grids of several layers, pinned bit for bit
against the reference's advance_signal_grid_cpu (oracle/_ref, live) on the
reference conformance suite's own 9 x 7 x 5 grids
(tests/conformance/signal_grid_conformance_test.cpp:14-293): plain (periodic
x, fixed z, constant advection, affine reaction), masked, with a velocity
field, and with reservoir patches on three faces; forward Euler,
Crank-Nicolson and backward Euler; upwind and van Leer. The suite holds its
backends to 5e-6; the oracle is held to the reference's bits."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import VAN_LEER, UPWIND, Grid
from tests.helpers import CONFORMANCE_KINDS, CONFORMANCE_STEPS, conformance_spec


def _grid(c):
    nx, ny, nz = c["shape"]
    dx, dy, dz = c["spacing"]
    return Grid(nx, ny, nz, dx, dy, dz, c["obstacles"], boundary=c["kinds"])


@pytest.mark.parametrize("kind", CONFORMANCE_KINDS)
@pytest.mark.parametrize("integration,dt", CONFORMANCE_STEPS)
@pytest.mark.parametrize("scheme", [UPWIND, VAN_LEER])
def test_conformance_grid_matches_reference_live(oracle, reference, kind, integration, dt,
                                                 scheme):
    c = conformance_spec(kind)
    grid = _grid(c)
    kw = dict(c["kw"])
    if integration == 0:
        a = oracle.adr_step_f32(grid, c["levels"], dt, steps=2, scheme=scheme, **kw)
        b, _ = reference.advance_signal_grid(grid, c["levels"], dt, steps=2, scheme=scheme, **kw)
        assert np.array_equal(a, b)
    else:
        a, rep = oracle.adr_step_implicit(grid, c["levels"], dt, steps=2, scheme=scheme,
                                          integration=integration, **kw)
        b, _, ref = reference.advance_signal_grid(grid, c["levels"], dt, steps=2, scheme=scheme,
                                                  integration=integration, with_report=True,
                                                  **kw)
        assert np.array_equal(a, b)
        assert rep["iterations"] == ref["iterations"] and rep["converged"] == ref["converged"]
    assert not np.array_equal(a, c["levels"])
