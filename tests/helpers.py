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


TRANSPORT_SHAPES = [(40, 70), (130, 64), (37, 300), (5, 513), (1, 200), (90, 1)]
TRANSPORT_KINDS = [
    (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX),
    (FIXED, NO_FLUX, PERIODIC, PERIODIC, NO_FLUX, NO_FLUX),
    (PERIODIC, PERIODIC, FIXED, FIXED, NO_FLUX, NO_FLUX),
    (PERIODIC, PERIODIC, PERIODIC, PERIODIC, NO_FLUX, NO_FLUX),
]


def random_transport_case(nx, ny, S, kinds, seed):
    """Obstacles, levels in [0.5, 1], affine rates, occupancy and a random face
    field whose periodic faces agree on both sides."""
    rng = np.random.default_rng(seed)
    solid = (rng.random(nx * ny) < 0.12).astype(np.uint8)
    levels = (0.5 + 0.5 * rng.random(S * nx * ny))
    levels[np.tile(solid, S) != 0] = 0.0
    xf = rng.uniform(-1.0, 1.0, (nx + 1) * ny).astype(np.float32)
    yf = rng.uniform(-1.0, 1.0, nx * (ny + 1)).astype(np.float32)
    # Faces next to an obstacle or on a no-flux boundary carry nothing; the
    # two sides of a periodic boundary are one face.
    sm = solid.reshape(nx, ny) != 0
    X = xf.reshape(nx + 1, ny)
    X[:-1][sm] = 0.0
    X[1:][sm] = 0.0
    Y = yf.reshape(nx, ny + 1)
    Y[:, :-1][sm] = 0.0
    Y[:, 1:][sm] = 0.0
    for axis, F in ((0, X), (1, Y)):
        lo = (slice(None), 0) if axis else (0, slice(None))
        hi = (slice(None), -1) if axis else (-1, slice(None))
        if kinds[2 * axis] == PERIODIC:
            both = (F[lo] != 0) & (F[hi] != 0)
            F[lo] = np.where(both, F[lo], 0.0)
            F[hi] = F[lo]
        else:
            if kinds[2 * axis] == NO_FLUX:
                F[lo] = 0.0
            if kinds[2 * axis + 1] == NO_FLUX:
                F[hi] = 0.0
    zf = np.zeros(nx * ny * 2, np.float32)
    loss = (rng.random(S * nx * ny) * 0.5).astype(np.float32)
    source = (rng.random(S * nx * ny) * 0.2).astype(np.float32)
    loss[np.tile(solid, S) != 0] = 0.0
    source[np.tile(solid, S) != 0] = 0.0
    phi = rng.random(nx * ny).astype(np.float32)
    bv = rng.random(6 * S).astype(np.float32)
    diffusion = rng.uniform(0.05, 0.3, S).astype(np.float32)
    return dict(solid=solid, levels=levels, xf=xf, yf=yf, zf=zf, loss=loss, source=source,
                phi=phi, bv=bv, diffusion=diffusion)


def layered_case(nx, ny, nz, seed, mobility=True, axis=0):
    """A grid of several layers for shallow flow: every column fluid from the
    floor (z = 0) up to a roof of 1..nz voxels, some columns solid (pillars),
    the flow axis's two ends clear, and a mobility constant down each column
    (or none)."""
    rng = np.random.default_rng(seed)
    obstacles = np.zeros((nx, ny, nz), np.uint8)
    roof = rng.integers(1, nz + 1, (nx, ny))
    for x in range(nx):
        for y in range(ny):
            obstacles[x, y, roof[x, y]:] = 1
    pillars = rng.random((nx, ny)) < 0.1
    if axis == 0:
        pillars[0, :] = pillars[-1, :] = False
    else:
        pillars[:, 0] = pillars[:, -1] = False
    obstacles[pillars] = 1
    mob = None
    if mobility:
        column = rng.uniform(0.5, 2.0, (nx, ny)).astype(np.float32)
        mob = np.repeat(column[:, :, None], nz, axis=2).reshape(-1)
    boundary = ((FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX) if axis == 0 else
                (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX))
    return obstacles.reshape(-1), mob, boundary


LAYERED_CASES = [(12, 9, 3, 1, True), (20, 7, 4, 2, False), (15, 15, 2, 3, True),
                 (10, 12, 5, 4, False)]


def random_patches(kinds, nx, ny, S, seed):
    """Two overlapping reservoir patches on every fixed in-plane face, each
    with its own value per signal (in-plane indices: y on an x face, x on a
    y face; the second index is z, of extent one)."""
    rng = np.random.default_rng(seed)
    patches = []
    for face in range(4):
        if kinds[face] != FIXED:
            continue
        extent = ny if face < 2 else nx
        for _ in range(2):
            begin = int(rng.integers(0, max(1, extent - 2)))
            end = int(rng.integers(begin + 1, extent + 1))
            patches.append((face, (begin, 0), (end, 1), rng.random(S).astype(np.float32)))
    return patches


PATCH_KINDS = [
    (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX),
    (FIXED, FIXED, FIXED, FIXED, NO_FLUX, NO_FLUX),
    (PERIODIC, PERIODIC, FIXED, FIXED, NO_FLUX, NO_FLUX),
]



# ----------------------------------------------------------------- ports --
SPEED, PRESSURE = 0, 1
X_LOWER, X_UPPER, Y_LOWER, Y_UPPER, Z_LOWER, Z_UPPER = range(6)
SIDES_FIXED = (FIXED, FIXED, FIXED, FIXED, NO_FLUX, NO_FLUX)
DUCT = (NO_FLUX, NO_FLUX, FIXED, FIXED, NO_FLUX, NO_FLUX)


def port(face, begin, end, depth, kind, value):
    """A port over [begin, end) of a side face and the depth [0, depth)
    (tests/cpp/flow_test.cpp:42-49)."""
    return (face, (begin, 0), (end, depth), kind, value)


def depth_port_case():
    """Two inlets on one face, a third drawing fluid out through a
    perpendicular face, one outlet, around an obstacle and across varying
    mobility: 12x10x2, spacing (0.8, 1.25, 0.6)
    (tests/conformance/flow_conformance_test.cpp:167-205)."""
    nx, ny, nz = 12, 10, 2
    obstacles = np.zeros((nx, ny, nz), np.uint8)
    mobility = np.zeros((nx, ny, nz), np.float32)
    for x in range(nx):
        for y in range(ny):
            obstacles[x, y, :] = 1 if (5 <= x <= 6 and 3 <= y <= 5) else 0
            mobility[x, y, :] = np.float32(0.4) + np.float32(0.07) * np.float32((x * 3 + y * 5) % 9)
    grid = dict(nx=nx, ny=ny, nz=nz, dx=0.8, dy=1.25, dz=0.6, obstacles=obstacles.reshape(-1),
                boundary=SIDES_FIXED)
    ports = [(Y_LOWER, (0, 0), (3, 2), SPEED, 3.0), (Y_LOWER, (8, 0), (12, 2), SPEED, 1.25),
             (X_UPPER, (6, 0), (9, 2), SPEED, -0.5), (Y_UPPER, (2, 0), (10, 2), PRESSURE, 0.0)]
    return grid, mobility.reshape(-1), ports


def winding_channel_case():
    """A grid that is mostly wall with one channel two sites wide winding
    from an inlet to an outlet: 24x42x1, spacing (1, 1, 2)
    (flow_conformance_test.cpp:223-249)."""
    nx, ny = 24, 42
    obstacles = np.ones((nx, ny), np.uint8)
    for x in range(1, 23):
        for y in range(ny):
            band, run = y // 8, y % 8 < 2
            turn = x >= 21 if band % 2 == 0 else x <= 2
            if run or turn:
                obstacles[x, y] = 0
    grid = dict(nx=nx, ny=ny, nz=1, dx=1.0, dy=1.0, dz=2.0, obstacles=obstacles.reshape(-1),
                boundary=DUCT)
    ports = [(Y_LOWER, (1, 0), (3, 1), SPEED, 2.0), (Y_UPPER, (1, 0), (3, 1), PRESSURE, 0.0)]
    return grid, None, ports


def random_port_case(seed, nx, ny, nz=1, with_mobility=True):
    """Every side face fixed; two to four ports on distinct side faces (so
    none overlap), each spanning the full depth, at least one held at a
    pressure; scattered obstacles (one layer) or columns of random roof
    height with pillars (several layers); a mobility constant down each
    column, or none."""
    rng = np.random.default_rng(seed)
    if nz == 1:
        obstacles = (rng.random((nx, ny, 1)) < 0.12).astype(np.uint8)
    else:
        obstacles = np.zeros((nx, ny, nz), np.uint8)
        roof = rng.integers(1, nz + 1, (nx, ny))
        for x in range(nx):
            for y in range(ny):
                obstacles[x, y, roof[x, y]:] = 1
        obstacles[rng.random((nx, ny)) < 0.08] = 1
    mobility = None
    if with_mobility:
        column = (10.0 ** rng.uniform(-1.0, 0.0, (nx, ny))).astype(np.float32)
        mobility = np.repeat(column[:, :, None], nz, axis=2).reshape(-1)
    faces = rng.permutation(4)[: int(rng.integers(2, 5))]
    ports = []
    for k, face in enumerate(faces):
        extent = ny if face < 2 else nx
        begin = int(rng.integers(0, max(1, extent - 2)))
        end = int(rng.integers(begin + 1, extent + 1))
        if k == 0:
            kind, value = PRESSURE, float(rng.choice([0.0, rng.uniform(-1.0, 2.0)]))
        elif rng.random() < 0.7:
            kind, value = SPEED, float(rng.uniform(-2.0, 3.0))
        else:
            kind, value = PRESSURE, float(rng.uniform(-1.0, 2.0))
        ports.append((int(face), (begin, 0), (end, nz), kind, value))
    grid = dict(nx=nx, ny=ny, nz=nz, dx=float(rng.uniform(0.5, 1.5)),
                dy=float(rng.uniform(0.5, 1.5)), dz=float(rng.uniform(0.3, 1.0)),
                obstacles=obstacles.reshape(-1), boundary=SIDES_FIXED)
    return grid, mobility, ports


PORT_CASES = [(1, 20, 17, 1, True), (102, 33, 9, 1, False), (3, 16, 24, 1, True),
              (4, 14, 11, 3, True), (5, 18, 13, 2, False), (6, 9, 21, 4, True),
              (7, 40, 31, 1, True), (8, 25, 25, 3, False)]


def largest_divergence(nx, ny, nz, dx, dy, dz, obstacles, xf, yf, zf):
    """The largest divergence of the face velocities at any fluid site
    (tests/cpp/flow_test.cpp:51-75)."""
    X = np.asarray(xf, np.float64).reshape(nx + 1, ny, nz)
    Y = np.asarray(yf, np.float64).reshape(nx, ny + 1, nz)
    Z = np.asarray(zf, np.float64).reshape(nx, ny, nz + 1)
    out = ((X[1:] - X[:-1]) / dx + (Y[:, 1:] - Y[:, :-1]) / dy
           + (Z[:, :, 1:] - Z[:, :, :-1]) / dz)
    fluid = np.asarray(obstacles).reshape(nx, ny, nz) == 0
    return float(np.abs(out[fluid]).max()) if fluid.any() else 0.0


# ---------------------------------------------------------------- coupled --
# RateOp codes (cpp/include/microsimulator/rate_ops.def).
(OP_CONSTANT, OP_SPECIES, OP_POSITION_X, OP_POSITION_Y, OP_POSITION_Z, OP_LENGTH, OP_RADIUS,
 OP_GROWTH_RATE, OP_CELL_TYPE, OP_VOLUME, OP_SURFACE, OP_ADD, OP_SUBTRACT, OP_MULTIPLY,
 OP_DIVIDE, OP_POWER, OP_MINIMUM, OP_MAXIMUM, OP_NEGATE, OP_EXPONENTIAL, OP_LOGARITHM, OP_LESS,
 OP_LESS_EQUAL, OP_GREATER, OP_GREATER_EQUAL, OP_EQUAL, OP_SELECT, OP_SIGNAL,
 OP_VOLUME_CHANGE_RATE) = range(29)


def uptake_plan(S, K, transcendental=True):
    """A coupled rate plan for S signals and K species that uses every
    operation: Monod uptake of signal 0 scaled by the cell's volume (its
    negative is the grid exchange), species production and decay, and
    position-, type- and size-dependent branches. Without `transcendental`
    the exponential, logarithm and power are left out."""
    ins = []

    def op(code, a=0, b=0, c=0, v=0.0):
        ins.append((code, a, b, c, v))
        return len(ins) - 1

    c0 = op(OP_SIGNAL, 0)
    half = op(OP_CONSTANT, v=0.25)
    uptake = op(OP_DIVIDE, c0, op(OP_ADD, c0, half))
    vmax = op(OP_CONSTANT, v=0.8)
    rate = op(OP_MULTIPLY, uptake, vmax)
    volume = op(OP_VOLUME)
    taken = op(OP_MULTIPLY, op(OP_MULTIPLY, rate, volume), op(OP_CONSTANT, v=0.02))
    exchange0 = op(OP_NEGATE, taken)
    growth = op(OP_GROWTH_RATE)
    px, py = op(OP_POSITION_X), op(OP_POSITION_Y)
    side = op(OP_LESS, px, py)
    branch = op(OP_SELECT, side, rate, growth)
    kind = op(OP_CELL_TYPE)
    two = op(OP_CONSTANT, v=2.0)
    typed = op(OP_GREATER_EQUAL, kind, two)
    surface = op(OP_SURFACE)
    length, radius = op(OP_LENGTH), op(OP_RADIUS)
    slender = op(OP_MAXIMUM, op(OP_DIVIDE, length, radius), op(OP_MINIMUM, surface, two))
    change = op(OP_VOLUME_CHANGE_RATE)
    pz = op(OP_POSITION_Z)
    flags = op(OP_ADD, op(OP_LESS_EQUAL, pz, px), op(OP_EQUAL, typed, op(OP_GREATER, px, pz)))
    shape = op(OP_MULTIPLY, slender, op(OP_ADD, flags, change))
    if transcendental:
        shape = op(OP_ADD, shape, op(OP_EXPONENTIAL, op(OP_NEGATE, growth)))
        shape = op(OP_ADD, shape, op(OP_LOGARITHM, op(OP_ADD, volume, two)))
        shape = op(OP_ADD, shape, op(OP_POWER, two, half))
    species_out = []
    for k in range(K):
        level = op(OP_SPECIES, k)
        decay = op(OP_MULTIPLY, level, op(OP_CONSTANT, v=0.1 * (k + 1)))
        species_out.append(op(OP_SUBTRACT, op(OP_ADD, branch, op(OP_MULTIPLY, shape,
                                                                  op(OP_CONSTANT, v=0.01))),
                              decay))
    signal_out = [exchange0]
    for s in range(1, S):
        signal_out.append(op(OP_MULTIPLY, op(OP_SIGNAL, s), op(OP_CONSTANT, v=0.05 * s)))
    return K, ins, species_out, signal_out


def random_cells(seed, nx, ny, solid, count, dx=1.0, dy=1.0, origin=(0.0, 0.0, 0.0)):
    """Cells at random positions inside the grid whose trilinear stencils
    touch at least one fluid site."""
    rng = np.random.default_rng(seed)
    solid = np.asarray(solid).reshape(nx, ny)
    xs, ys = [], []
    while len(xs) < count:
        x = rng.uniform(0.0, (nx - 1) * dx) if nx > 1 else 0.0
        y = rng.uniform(0.0, (ny - 1) * dy) if ny > 1 else 0.0
        i0, j0 = int(x // dx) if nx > 1 else 0, int(y // dy) if ny > 1 else 0
        near = solid[i0:i0 + 2, j0:j0 + 2]
        if (near == 0).any():
            xs.append(x + origin[0])
            ys.append(y + origin[1])
    lengths = rng.uniform(1.0, 4.0, count).astype(np.float32)
    return dict(position_x=np.array(xs, np.float32), position_y=np.array(ys, np.float32),
                position_z=np.full(count, origin[2], np.float32), lengths=lengths,
                radii=rng.uniform(0.3, 0.8, count).astype(np.float32),
                growth_rates=rng.uniform(0.5, 1.5, count).astype(np.float32),
                cell_types=rng.integers(0, 4, count).astype(np.int32)), \
        (lengths * rng.uniform(0.85, 1.0, count)).astype(np.float32)


def coupled_colony_case(masked=True):
    """A one-layer version of the reference conformance suite's coupled colony
    (tests/conformance/coupled_rates_conformance_test.cpp:18-118): a 24x18
    grid of spacing (0.75, 1.25) and origin (-3, -2), periodic along x, two
    signals with affine rates, uniform advection as face velocities, 513
    cells of three species spread over the grid by its formula, and its
    13-instruction plan. Returns (grid kwargs, levels, cells, species, plan,
    previous lengths, transport kwargs)."""
    nx, ny = 24, 18
    f32 = np.float32
    solid = np.zeros(nx * ny, np.uint8)
    if masked:
        for x, y in ((11, 8), (5, 3), (17, 12), (6, 14)):
            solid[x * ny + y] = 1
    kinds = (PERIODIC, PERIODIC, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX)
    grid = dict(nx=nx, ny=ny, nz=1, dx=0.75, dy=1.25, dz=2.0, obstacles=solid, boundary=kinds)
    n = nx * ny
    index = np.arange(2 * n)
    source = (f32(0.004) * (index % 7).astype(f32)).astype(f32)
    loss = (f32(0.003) * (index % 11).astype(f32)).astype(f32)
    levels = (1.0 + f32(0.002) * (index % 113).astype(f32)).astype(f32)
    dead = np.tile(solid, 2) != 0
    source[dead] = 0.0
    loss[dead] = 0.0
    levels[dead] = 0.0
    xf = np.full((nx + 1) * ny, 0.05, f32).reshape(nx + 1, ny)
    yf = np.full(nx * (ny + 1), -0.02, f32).reshape(nx, ny + 1)
    yf[:, 0] = 0.0
    yf[:, -1] = 0.0
    sm = solid.reshape(nx, ny) != 0
    xf[:-1][sm] = 0.0
    xf[1:][sm] = 0.0
    yf[:, :-1][sm] = 0.0
    yf[:, 1:][sm] = 0.0
    both = (xf[0] != 0) & (xf[-1] != 0)
    xf[0] = np.where(both, xf[0], 0.0)
    xf[-1] = xf[0]
    faces = np.concatenate([xf.reshape(-1), yf.reshape(-1), np.zeros(2 * n, f32)])
    k = np.arange(513)
    fx = (k % 17).astype(f32) / f32(16)
    fy = ((k * 7) % 19).astype(f32) / f32(18)
    px = (f32(-3.0) + fx * f32(0.75) * f32(nx - 1)).astype(f32)
    py = (f32(-2.0) + fy * f32(1.25) * f32(ny - 1)).astype(f32)
    # Cells whose stencil would be wholly solid move to the next lattice line.
    for c in range(513):
        i0 = int(np.floor((float(px[c]) + 3.0) / 0.75))
        j0 = int(np.floor((float(py[c]) + 2.0) / 1.25))
        near = [(i, j) for i in (i0, min(i0 + 1, nx - 1)) for j in (j0, min(j0 + 1, ny - 1))]
        if all(solid[i * ny + j] for i, j in near):
            py[c] = f32(-2.0) + f32(1.25) * f32(min(j0 + 2, ny - 1))
    cells = dict(position_x=px, position_y=py, position_z=np.zeros(513, f32),
                 lengths=(f32(1.5) + f32(0.01) * (k % 31).astype(f32)).astype(f32),
                 radii=(f32(0.3) + f32(0.005) * (k % 7).astype(f32)).astype(f32),
                 growth_rates=(f32(0.01) * (k % 5).astype(f32)).astype(f32),
                 cell_types=(k % 4).astype(np.int32))
    species = np.stack([0.5 + f32(0.01) * (k % 29), 0.75 + f32(0.02) * (k % 17),
                        1.0 + f32(0.015) * (k % 13)], 1).astype(f32).reshape(-1)
    ins = [(OP_SPECIES, 0, 0, 0, 0.0), (OP_SPECIES, 1, 0, 0, 0.0), (OP_SPECIES, 2, 0, 0, 0.0),
           (OP_SIGNAL, 0, 0, 0, 0.0), (OP_SIGNAL, 1, 0, 0, 0.0), (OP_CONSTANT, 0, 0, 0, 0.02),
           (OP_CONSTANT, 0, 0, 0, -0.01), (OP_MULTIPLY, 3, 5, 0, 0.0), (OP_ADD, 0, 7, 0, 0.0),
           (OP_MULTIPLY, 1, 6, 0, 0.0), (OP_ADD, 2, 9, 0, 0.0), (OP_MULTIPLY, 0, 5, 0, 0.0),
           (OP_MULTIPLY, 4, 6, 0, 0.0)]
    plan = (3, ins, [8, 9, 10], [11, 12])
    previous = (cells["lengths"] * f32(0.97)).astype(f32)
    transport = dict(diffusion=np.array([0.08, 0.03], f32), boundary_values=np.zeros(12, f32),
                     faces=faces, source=source, loss=loss)
    return grid, levels, cells, species, plan, previous, transport


def conformance_spec(kind="plain"):
    """The reference conformance suite's 9 x 7 x 5 grid of two signals
    (tests/conformance/signal_grid_conformance_test.cpp:14-119): periodic in
    x, fixed reservoirs in z, constant advection, an affine reaction; `masked`
    adds an obstacle block and one obstacle on the periodic edge;
    `velocity_field` replaces the advection by x faces 0.05 + 0.01 y - 0.008 z;
    `patched` makes every face a fixed reservoir, with patches on x_lower,
    y_upper and z_lower (each indexed by a different pair of axes) and a
    uniform field across all six faces. Returns the grid's shape, spacing and
    kinds, and the keyword arguments of the oracle's and the reference's
    transport calls (diffusion, advection, boundary_values, faces, source,
    loss, patches, obstacles) and the levels (make_levels, :121-128)."""
    nx, ny, nz, S = 9, 7, 5, 2
    n = nx * ny * nz
    kinds = [PERIODIC, PERIODIC, NO_FLUX, NO_FLUX, FIXED, FIXED]
    bv = np.zeros(6 * S, np.float32)
    bv[0 * 6 + 4], bv[1 * 6 + 4] = 0.25, 0.5
    bv[0 * 6 + 5], bv[1 * 6 + 5] = 1.0, 0.75
    advection = np.array([0.1, -0.2, 0.05, -0.05, 0.03, -0.08], np.float32)
    index = np.arange(S * n)
    source = (np.float32(0.01) * (index % 3).astype(np.float32)).astype(np.float32)
    loss = (np.float32(0.005) * (index % 5).astype(np.float32)).astype(np.float32)
    obstacles = None
    faces = None
    patches = ()
    if kind == "masked":
        x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        solid = ((x >= 3) & (x <= 5) & (y >= 2) & (y <= 4) & (z == 2)) | \
            ((x == 0) & (y == 1) & (z == 1))
        obstacles = solid.astype(np.uint8).reshape(-1)
        mask = np.tile(obstacles, S) != 0
        source[mask] = 0.0
        loss[mask] = 0.0
    elif kind == "velocity_field":
        advection = np.zeros(3 * S, np.float32)
        fx, y, z = np.meshgrid(np.arange(nx + 1), np.arange(ny), np.arange(nz), indexing="ij")
        xf = (np.float32(0.05) + np.float32(0.01) * y.astype(np.float32)
              - np.float32(0.008) * z.astype(np.float32)).astype(np.float32).reshape(-1)
        faces = np.concatenate([xf, np.zeros(nx * (ny + 1) * nz, np.float32),
                                np.zeros(nx * ny * (nz + 1), np.float32)])
    elif kind == "patched":
        advection = np.zeros(3 * S, np.float32)
        kinds = [FIXED] * 6
        for f in range(4):
            bv[0 * 6 + f], bv[1 * 6 + f] = 0.4, 0.6
        patches = ((0, (1, 0), (4, 3), [1.5, 0.1]), (0, (3, 2), (7, 5), [0.0, 2.0]),
                   (3, (0, 1), (5, 4), [0.9, 0.2]),
                   (4, (2, 0), (9, 2), [1.2, 0.3]), (4, (6, 1), (8, 7), [0.1, 1.1]))
        faces = np.concatenate([np.full((nx + 1) * ny * nz, 0.06, np.float32),
                                np.full(nx * (ny + 1) * nz, -0.04, np.float32),
                                np.full(nx * ny * (nz + 1), 0.05, np.float32)])
    levels = (np.float32(0.5) + np.float32(0.001) *
              ((index * 37) % 211).astype(np.float32)).astype(np.float32)
    if obstacles is not None:
        levels[np.tile(obstacles, S) != 0] = 0.0
    kw = dict(diffusion=np.array([0.2, 0.05], np.float32), boundary_values=bv, faces=faces,
              source=source, loss=loss, patches=patches)
    if faces is None:
        kw["advection"] = advection
    return dict(shape=(nx, ny, nz), spacing=(1.0, 0.75, 1.25), kinds=tuple(kinds),
                obstacles=obstacles, kw=kw, levels=levels)


CONFORMANCE_KINDS = ("plain", "masked", "velocity_field", "patched")
# The suite's steps: forward Euler at 0.02, the implicit kinds at 0.5 (:267-293).
CONFORMANCE_STEPS = ((0, 0.02), (1, 0.5), (2, 0.5))
