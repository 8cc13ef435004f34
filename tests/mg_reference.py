"""CPU restatement (float64, numpy) of the device multilevel preconditioner,
for checking its kernels: 2x2 aggregation with Galerkin coarse operators,
V(2,2) damped Jacobi (omega 0.8; the sweeps on each side are a parameter),
coarse correction weighted by 1.8, the coarsest level (<= 256 cells) solved
exactly. Test infrastructure only."""
from __future__ import annotations

import numpy as np

OMEGA = 0.8
CORRECTION = 1.8
COARSE_CELLS = 256


def level0(cxl, cyl, bc, nx, ny, dx, dy):
    """Couplings to the lower x / y neighbour, boundary term (float32 products)."""
    isx = np.float32(1.0 / (float(dx) * float(dx)))
    isy = np.float32(1.0 / (float(dy) * float(dy)))
    return dict(wx=(np.asarray(cxl, np.float32) * isx).reshape(nx, ny).astype(np.float64),
                wy=(np.asarray(cyl, np.float32) * isy).reshape(nx, ny).astype(np.float64),
                bc=np.asarray(bc, np.float32).reshape(nx, ny).astype(np.float64))


def apply(L, x):
    wx, wy, bc = L["wx"], L["wy"], L["bc"]
    y = bc * x
    d = wx[1:, :] * (x[1:, :] - x[:-1, :])
    y[1:, :] += d
    y[:-1, :] -= d
    d = wy[:, 1:] * (x[:, 1:] - x[:, :-1])
    y[:, 1:] += d
    y[:, :-1] -= d
    return y


def diagonal(L):
    d = L["bc"] + L["wx"] + L["wy"]
    d[:-1, :] += L["wx"][1:, :]
    d[:, :-1] += L["wy"][:, 1:]
    return d


def coarsen(L):
    nx, ny = L["bc"].shape
    cx, cy = (nx + 1) // 2, (ny + 1) // 2

    def pad(a):
        p = np.zeros((2 * cx, 2 * cy))
        p[:nx, :ny] = a
        return p
    wx, wy, bc = pad(L["wx"]), pad(L["wy"]), pad(L["bc"])
    Wx = wx[0::2, :].reshape(cx, cy, 2).sum(axis=2)
    Wx[0, :] = 0.0
    Wy = wy[:, 0::2].reshape(cx, 2, cy).sum(axis=1)
    Wy[:, 0] = 0.0
    return dict(wx=Wx, wy=Wy, bc=bc.reshape(cx, 2, cy, 2).sum(axis=(1, 3)))


def restrict(r, shape):
    nx, ny = r.shape
    p = np.zeros((2 * shape[0], 2 * shape[1]))
    p[:nx, :ny] = r
    return p.reshape(shape[0], 2, shape[1], 2).sum(axis=(1, 3))


def prolong(e, shape):
    return np.repeat(np.repeat(e, 2, axis=0), 2, axis=1)[:shape[0], :shape[1]]


def dense(L):
    shape = L["bc"].shape
    n = L["bc"].size
    return np.column_stack([apply(L, np.eye(n)[k].reshape(shape)).ravel() for k in range(n)])


def smooth(L, x, r, sweeps):
    """`sweeps` damped-Jacobi sweeps on A x = r from x (None: from zero)."""
    for _ in range(sweeps):
        x = OMEGA * L["dinv"] * r if x is None else x + OMEGA * L["dinv"] * (r - apply(L, x))
    return x


class Multilevel:
    def __init__(self, L0, pre=2, post=2):
        self.pre, self.post = pre, post
        self.levels = [L0]
        while self.levels[-1]["bc"].size > COARSE_CELLS:
            self.levels.append(coarsen(self.levels[-1]))
        for L in self.levels:
            d = diagonal(L)
            L["dinv"] = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
        A = dense(self.levels[-1])
        self.active = np.diag(A) > 0
        self.inverse = np.linalg.inv(A[np.ix_(self.active, self.active)])

    def cycle(self, r, level=0):
        L = self.levels[level]
        if level == len(self.levels) - 1:
            e = np.zeros(r.size)
            e[self.active] = self.inverse @ r.ravel()[self.active]
            return e.reshape(r.shape)
        x = smooth(L, None, r, self.pre)
        rc = restrict(r - apply(L, x), self.levels[level + 1]["bc"].shape)
        x = x + CORRECTION * prolong(self.cycle(rc, level + 1), r.shape)
        return smooth(L, x, r, self.post)
