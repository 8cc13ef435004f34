"""numpy restatement of the reference's biomass deposition and colony mobility
(python/src/microsimulator/flow.py:224-356, biomass.py:6-14). Test
infrastructure only: the product never imports it.

Pinned by the reference's own tests (python/tests/test_flow_foundations.py:75-88,
python/tests/test_flow.py:243-274) and by closed forms of the tent kernel's
voxel integrals; see tests/test_colony_oracle.py.

Grids are (nx, ny, nz) arrays in the reference's site order (x, then y, then
z fastest). Cells are arrays of positions, lengths and radii.
"""
from __future__ import annotations

import math

import numpy as np


class ColonyError(ValueError):
    """FlowError / ValueError of the reference's Python functions."""


def biomass_volume(length: float, radius: float) -> float:
    """pi r^2 (L + 2r) in double (biomass.py:6-14)."""
    if not math.isfinite(length) or length < 0 or not math.isfinite(radius) or radius <= 0:
        raise ColonyError("length must be finite and nonnegative; radius finite and positive")
    return math.pi * radius * radius * (length + 2 * radius)


def _tent_cdf(edges: np.ndarray) -> np.ndarray:
    """The tent kernel's cumulative weight at scaled edges (flow.py:245-254)."""
    return np.where(edges <= -1, 0.0,
                    np.where(edges < 0, 0.5 * (edges + 1) ** 2,
                             np.where(edges < 1, 1 - 0.5 * (1 - edges) ** 2, 1.0)))


def deposition_weights(dims, origin, spacing, position, radius, solid):
    """_deposition_weights (flow.py:224-279): the region (lower corner and
    shape) and its normalised weights, or None outside the modelled volume."""
    if any(not math.isfinite(v) for v in position):
        raise ColonyError("deposited positions must be finite")
    if any(p < o - h / 2 or p >= o + (n - 0.5) * h
           for p, o, h, n in zip(position, origin, spacing, dims)):
        return None
    lows, weights = [], []
    for p, o, h, n in zip(position, origin, spacing, dims):
        lo = max(0, math.floor((p - radius - o) / h + 0.5))
        hi = min(n, math.ceil((p + radius - o) / h + 0.5))
        edges = (o + (np.arange(lo, hi + 1) - 0.5) * h - p) / radius
        lows.append(lo)
        weights.append(np.diff(_tent_cdf(edges)))
    kernel = weights[0][:, None, None] * weights[1][None, :, None] * weights[2][None, None, :]
    if solid is not None:
        sub = solid[lows[0]:lows[0] + kernel.shape[0], lows[1]:lows[1] + kernel.shape[1],
                    lows[2]:lows[2] + kernel.shape[2]]
        kernel[sub] = 0
        connected = np.zeros(kernel.shape, dtype=bool)
        pending = [tuple(int(i) for i in np.unravel_index(int(np.argmax(kernel)), kernel.shape))]
        while pending:
            index = pending.pop()
            if connected[index] or kernel[index] <= 0:
                continue
            connected[index] = True
            for axis in range(3):
                for offset in (-1, 1):
                    adjacent = list(index)
                    adjacent[axis] += offset
                    if 0 <= adjacent[axis] < kernel.shape[axis]:
                        pending.append(tuple(adjacent))
        kernel[~connected] = 0
    total = float(kernel.sum())
    if total <= 0:
        raise ColonyError("biomass deposition has no connected fluid support")
    return tuple(lows), kernel / total


def colony_volume_fraction(dims, origin, spacing, obstacles, cells, averaging_radius=4.0):
    """colony_volume_fraction (flow.py:282-305): biomass volume per voxel
    volume, (nx, ny, nz) float64. cells: dict of position_x/y/z, lengths,
    radii arrays."""
    if not math.isfinite(averaging_radius) or averaging_radius <= 0:
        raise ColonyError("averaging radius must be finite and positive")
    dims = tuple(int(d) for d in dims)
    # The spec holds origin and spacing as float (signals.hpp:122-156), and its
    # voxel volume is their float product (signals.cpp:557-559).
    origin = tuple(float(np.float32(o)) for o in origin)
    spacing = tuple(float(np.float32(h)) for h in spacing)
    voxel = float(np.float32(np.float32(spacing[0]) * np.float32(spacing[1]))
                  * np.float32(spacing[2]))
    volume = np.zeros(dims, dtype=np.float64)
    solid = None if obstacles is None else np.asarray(obstacles, bool).reshape(dims)
    for c in range(len(cells["lengths"])):
        position = tuple(float(cells[k][c]) for k in ("position_x", "position_y", "position_z"))
        stencil = deposition_weights(dims, origin, spacing, position, averaging_radius, solid)
        if stencil is not None:
            lows, weights = stencil
            region = tuple(slice(lo, lo + w) for lo, w in zip(lows, weights.shape))
            volume[region] += biomass_volume(float(cells["lengths"][c]),
                                             float(cells["radii"][c])) * weights
    return volume / voxel


def colony_mobility(dims, origin, spacing, obstacles, cells, base=1.0, drag_coefficient=100.0,
                    max_volume_fraction=0.9, averaging_radius=4.0):
    """colony_mobility with _darcy_coefficients (flow.py:334-382): per site
    base / (1 + base * drag) in fluid, 0 in solid, with the Kozeny-Carman drag
    C f^2 / (1 - f)^3 of the column's mean volume fraction capped at the
    maximum (flow.py:91-96)."""
    density = colony_volume_fraction(dims, origin, spacing, obstacles, cells, averaging_radius)
    if not 0 < max_volume_fraction < 1:
        raise ColonyError("maximum volume fraction must lie strictly between zero and one")
    dims = tuple(int(d) for d in dims)
    if np.ndim(base) == 0:
        base = float(base)
        if not math.isfinite(base) or base <= 0.0:
            raise ColonyError("base mobility must be finite and positive")
        base_grid = np.full(dims, base)
    else:
        if len(base) != dims[0] * dims[1] * dims[2]:
            raise ColonyError("base mobility must hold one value per grid site")
        base_grid = np.asarray(base, dtype=np.float64).reshape(dims)
        if not bool(np.all(np.isfinite(base_grid))) or bool(np.any(base_grid < 0.0)):
            raise ColonyError("base mobility values must be finite and non-negative")
    solid = (np.zeros(dims, bool) if obstacles is None
             else np.asarray(obstacles, bool).reshape(dims))
    fluid = ~solid
    depth_count = np.maximum(fluid.sum(axis=2), 1)
    column_fraction = np.minimum(density.sum(axis=2) / depth_count, max_volume_fraction)
    if not math.isfinite(drag_coefficient) or drag_coefficient < 0.0:
        raise ColonyError("drag coefficient must be finite and non-negative")
    drag = drag_coefficient * column_fraction * column_fraction / (1.0 - column_fraction) ** 3
    mobility = base_grid / (1.0 + base_grid * drag[:, :, None])
    return np.where(fluid, mobility, 0.0).reshape(-1)


def colony_flow_coupling(dims, origin, spacing, obstacles, cells, velocities, base=1.0,
                         drag_coefficient=100.0, max_volume_fraction=0.9, averaging_radius=4.0):
    """colony_flow_coupling (flow.py:384-424) with _darcy_coefficients
    (359-382): the colony's mobility (as colony_mobility) and the in-plane
    body force per column, momentum * drag / mass, where the momentum is each
    cell's deposited biomass per column times its in-plane velocity and the
    mass the column's deposited biomass volume (0 where the column has none).
    velocities: (cells, 2) in-plane (x, y). Returns (mobility N, force
    (columns, 2))."""
    if not math.isfinite(averaging_radius) or averaging_radius <= 0:
        raise ColonyError("averaging radius must be finite and positive")
    dims = tuple(int(d) for d in dims)
    origin = tuple(float(np.float32(o)) for o in origin)
    spacing = tuple(float(np.float32(h)) for h in spacing)
    voxel = float(np.float32(np.float32(spacing[0]) * np.float32(spacing[1]))
                  * np.float32(spacing[2]))
    volume = np.zeros(dims, dtype=np.float64)
    momentum = np.zeros((dims[0], dims[1], 2), dtype=np.float64)
    solid = None if obstacles is None else np.asarray(obstacles, bool).reshape(dims)
    velocities = np.asarray(velocities, np.float32).reshape(-1, 2)
    for c in range(len(cells["lengths"])):
        vx, vy = float(velocities[c, 0]), float(velocities[c, 1])
        if not math.isfinite(vx) or not math.isfinite(vy):
            raise ColonyError("in-plane cell velocity must be finite")
        position = tuple(float(cells[k][c]) for k in ("position_x", "position_y", "position_z"))
        stencil = deposition_weights(dims, origin, spacing, position, averaging_radius, solid)
        if stencil is None:
            continue
        lows, weights = stencil
        region = tuple(slice(lo, lo + w) for lo, w in zip(lows, weights.shape))
        mass = biomass_volume(float(cells["lengths"][c]), float(cells["radii"][c])) * weights
        volume[region] += mass
        momentum[region[:2]] += mass.sum(axis=2)[:, :, None] * (vx, vy)
    density = volume / voxel
    if not 0 < max_volume_fraction < 1:
        raise ColonyError("maximum volume fraction must lie strictly between zero and one")
    if np.ndim(base) == 0:
        base = float(base)
        if not math.isfinite(base) or base <= 0.0:
            raise ColonyError("base mobility must be finite and positive")
        base_grid = np.full(dims, base)
    else:
        if len(base) != dims[0] * dims[1] * dims[2]:
            raise ColonyError("base mobility must hold one value per grid site")
        base_grid = np.asarray(base, dtype=np.float64).reshape(dims)
        if not bool(np.all(np.isfinite(base_grid))) or bool(np.any(base_grid < 0.0)):
            raise ColonyError("base mobility values must be finite and non-negative")
    fluid = np.ones(dims, bool) if solid is None else ~solid
    depth_count = np.maximum(fluid.sum(axis=2), 1)
    column_fraction = np.minimum(density.sum(axis=2) / depth_count, max_volume_fraction)
    if not math.isfinite(drag_coefficient) or drag_coefficient < 0.0:
        raise ColonyError("drag coefficient must be finite and non-negative")
    drag = drag_coefficient * column_fraction * column_fraction / (1.0 - column_fraction) ** 3
    mobility = np.where(fluid, base_grid / (1.0 + base_grid * drag[:, :, None]), 0.0)
    mass = volume.sum(axis=2)
    force = momentum * np.divide(drag, mass, out=np.zeros_like(drag), where=mass > 0)[:, :, None]
    return mobility.reshape(-1), force.reshape(-1, 2)
