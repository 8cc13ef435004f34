"""Shared inputs for the tests: golden fixtures and small seeded cases."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NO_FLUX, PERIODIC, FIXED = 0, 1, 2


def golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def c1_inputs():
    """Channel + trap 256x128 from the committed fixture (inputs and the
    reference's outputs)."""
    d = golden("c1_flow.npz")
    nx, ny = int(d["nx"]), int(d["ny"])
    obstacles = np.unpackbits(d["obstacles"])[: nx * ny].astype(np.uint8)
    return dict(nx=nx, ny=ny, dz=float(d["dz"]), obstacles=obstacles, mobility=d["mobility"],
                phi=d["phi"], data=d)


def illcond_inputs():
    d = golden("illcond_flow.npz")
    return dict(nx=64, ny=64, obstacles=d["obstacles"], mobility=d["mobility"], data=d)


def random_case(seed: int, nx: int, ny: int, solid_fraction: float = 0.2, decades: float = 2.0):
    """A seeded random sheet: obstacles with an open inlet column band and
    mobility 10^U(-decades, 0), flowing along x."""
    rng = np.random.default_rng(seed)
    solid = (rng.random((nx, ny)) < solid_fraction).astype(np.uint8)
    solid[:, ny // 2] = 0  # keep a through-channel
    mobility = (10.0 ** rng.uniform(-decades, 0.0, (nx, ny))).astype(np.float32)
    mobility[solid != 0] = 0.0
    return solid.reshape(-1), mobility.reshape(-1)


def split_faces(faces, nx, ny):
    nxf = (nx + 1) * ny
    nyf = nx * (ny + 1)
    return faces[:nxf], faces[nxf:nxf + nyf], faces[nxf + nyf:]


def divergence(faces, nx, ny, dx=1.0, dy=1.0):
    xf, yf, _ = split_faces(faces, nx, ny)
    xf = xf.reshape(nx + 1, ny).astype(np.float64)
    yf = yf.reshape(nx, ny + 1).astype(np.float64)
    return (xf[1:, :] - xf[:-1, :]) / dx + (yf[:, 1:] - yf[:, :-1]) / dy


def monod_loss(vmax, phi, K, levels):
    """vmax*phi/(K+c) in float32, the order the kernels use."""
    vmax = np.float32(vmax)
    K = np.float32(K)
    return ((vmax * np.asarray(phi, np.float32)) / (K + np.asarray(levels, np.float32))).astype(
        np.float32)
