"""CPU restatement (float64, numpy) of the learned smoother of the multilevel
preconditioner, forward and backward, for checking the device kernels. Test
infrastructure only. The smoother has no counterpart in the reference, whose
only preconditioner is Jacobi (cpp/cpu/cpu_flow.cpp:65-69).

The smoother maps a residual r' (the residual left after the pre-smoother and
the coarse-grid correction) to a correction z. It works tile by tile, exactly
as the device kernel does. A tile is 16 x 32 output sites (x by y), and its
receptive region is the tile grown by 3.

With mask_p = (p on the grid and dinv_p > 0), for a site p of the region:
  u_p  = dinv_p r'_p
  s    = max |u| over the region (the tile's scale)
  v_p  = u_p / s
  in_p = [v, v cx, v cy, v cb], where cx = (wx_lower + wx_upper) dinv,
         cy = (wy_lower + wy_upper) dinv, cb = bc dinv
Then:
  h1 = conv3x3(in, W1)  on the tile grown by 2  (4 -> 16 channels)
  a1 = clip(h1) * mask
  h2 = conv3x3(a1, W2)  on the tile grown by 1  (16 -> 16)
  a2 = clip(h2) * mask
  n  = conv3x3(a2, W3) * mask  on the tile       (16 -> 1)
  z  = omega u + s n

clip(h) = min(max(h, -1), 1). The network has no biases and an odd
activation, so z is odd and homogeneous of degree one in r'. Weights are indexed [tap][c_in][c_out], with
tap = (dx + 1) * 3 + (dy + 1). Flat layout (3024 floats): W1 (9x4x16), then
W2 (9x16x16), then W3 (9x16).

Training loss for one pair (r', e') with A e' = r':
  d = e' - z
  L = d^T A d / e'^T A e'
A batch's loss is the mean over its pairs.

With bf16=True the forward rounds the weights and the stored activations (in,
a1, a2) to bfloat16, as the device does; sums stay exact.
"""
from __future__ import annotations

import numpy as np

TX, TY = 16, 32
HALO = 3
OMEGA = 0.8
C0, C1 = 4, 16
TAPS = [(dx, dy) for dx in (-1, 0, 1) for dy in (-1, 0, 1)]
SIZES = (9 * C0 * C1, 9 * C1 * C1, 9 * C1)
COUNT = sum(SIZES)


def bf16(a):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(flat):
    flat = np.asarray(flat, np.float64)
    assert flat.size == COUNT
    a, b = SIZES[0], SIZES[0] + SIZES[1]
    return (flat[:a].reshape(9, C0, C1), flat[a:b].reshape(9, C1, C1),
            flat[b:].reshape(9, C1, 1))


def pack(W1, W2, W3):
    return np.concatenate([np.ravel(W1), np.ravel(W2), np.ravel(W3)])


def clip(h):
    return np.clip(h, -1.0, 1.0)


def clip_slope(h):
    """Derivative of clip: 1 inside (-1, 1), 0 where the value is clipped."""
    return (np.abs(h) < 1.0).astype(np.float64)


def features(L):
    """Per-site (cx, cy, cb) from the level-0 couplings and dinv."""
    wx, wy, bc, dinv = L["wx"], L["wy"], L["bc"], L["dinv"]
    wxu = np.zeros_like(wx)
    wxu[:-1, :] = wx[1:, :]
    wyu = np.zeros_like(wy)
    wyu[:, :-1] = wy[:, 1:]
    return (wx + wxu) * dinv, (wy + wyu) * dinv, bc * dinv


def conv(act, W, out_shape):
    """act: (H + 2, W + 2, Cin) over a region one larger per side than out_shape."""
    H, Wd = out_shape
    out = np.zeros((H, Wd, W.shape[2]))
    for t, (dx, dy) in enumerate(TAPS):
        out += act[1 + dx:1 + dx + H, 1 + dy:1 + dy + Wd, :] @ W[t]
    return out


def _region(L, rp, fx0, fy0):
    nx, ny = rp.shape
    R = HALO
    xs = np.arange(fx0 - R, fx0 + TX + R)
    ys = np.arange(fy0 - R, fy0 + TY + R)
    inside = (xs[:, None] >= 0) & (xs[:, None] < nx) & (ys[None, :] >= 0) & (ys[None, :] < ny)
    xi = np.clip(xs, 0, nx - 1)[:, None]
    yi = np.clip(ys, 0, ny - 1)[None, :]
    mask = inside & (L["dinv"][xi, yi] > 0)
    return xi, yi, inside, mask


def tile_forward(L, rp, W, fx0, fy0, bf=False, feats=None):
    """Forward over one tile. Returns (n on the tile, s, cache or None)."""
    W1, W2, W3 = W
    if bf:
        W1, W2, W3 = bf16(W1), bf16(W2), bf16(W3)
    cx, cy, cb = feats if feats is not None else features(L)
    xi, yi, inside, mask = _region(L, rp, fx0, fy0)
    u = np.where(inside, L["dinv"][xi, yi] * rp[xi, yi], 0.0)
    s = np.abs(u).max()
    if s == 0.0:
        return np.zeros((TX, TY)), 0.0, None
    v = u / s
    inp = np.stack([v, v * np.where(inside, cx[xi, yi], 0), v * np.where(inside, cy[xi, yi], 0),
                    v * np.where(inside, cb[xi, yi], 0)], axis=2)
    if bf:
        inp = bf16(inp)
    m1 = mask[1:-1, 1:-1, None]
    m2 = mask[2:-2, 2:-2, None]
    m3 = mask[3:-3, 3:-3]
    h1 = conv(inp, W1, (TX + 4, TY + 4))
    a1 = np.where(m1, clip(h1), 0.0)
    if bf:
        a1 = bf16(a1)
    h2 = conv(a1, W2, (TX + 2, TY + 2))
    a2 = np.where(m2, clip(h2), 0.0)
    if bf:
        a2 = bf16(a2)
    n = np.where(m3, conv(a2, W3, (TX, TY))[..., 0], 0.0)
    cache = dict(inp=inp, h1=h1, a1=a1, h2=h2, a2=a2, m1=m1, m2=m2, m3=m3, W=(W1, W2, W3))
    return n, s, cache


def tiles(nx, ny):
    for fx0 in range(0, nx, TX):
        for fy0 in range(0, ny, TY):
            yield fx0, fy0


def correction(L, rp, W, bf=False):
    """s n over the field (the network's part of z), tile by tile."""
    nx, ny = rp.shape
    out = np.zeros((nx, ny))
    feats = features(L)
    for fx0, fy0 in tiles(nx, ny):
        n, s, _ = tile_forward(L, rp, W, fx0, fy0, bf, feats)
        hx, hy = min(TX, nx - fx0), min(TY, ny - fy0)
        out[fx0:fx0 + hx, fy0:fy0 + hy] = s * n[:hx, :hy]
    return out


def smoother(L, rp, W, bf=False):
    """z = omega dinv r' + s n."""
    return OMEGA * L["dinv"] * rp + correction(L, rp, W, bf)


def apply_operator(L, x):
    wx, wy, bc = L["wx"], L["wy"], L["bc"]
    y = bc * x
    d = wx[1:, :] * (x[1:, :] - x[:-1, :])
    y[1:, :] += d
    y[:-1, :] -= d
    d = wy[:, 1:] * (x[:, 1:] - x[:, :-1])
    y[:, 1:] += d
    y[:, :-1] -= d
    return y


def loss(L, rp, ep, W):
    d = ep - smoother(L, rp, W)
    return float(np.sum(d * apply_operator(L, d)) / np.sum(ep * apply_operator(L, ep)))


def _corr_grad(act, g, Hout, Wout):
    """dW[t][c][o] = sum_p act[p + tap][c] g[p][o]."""
    return np.stack([np.einsum("hwc,hwo->co",
                               act[1 + dx:1 + dx + Hout, 1 + dy:1 + dy + Wout, :], g)
                     for dx, dy in TAPS])


def _conv_transpose(g, W, Hin, Win):
    """dact[q][c] = sum_t sum_o W[t][c][o] g[q - tap][o], act region one larger per side."""
    out = np.zeros((Hin, Win, W.shape[1]))
    H, Wd = g.shape[:2]
    for t, (dx, dy) in enumerate(TAPS):
        out[1 + dx:1 + dx + H, 1 + dy:1 + dy + Wd, :] += g @ W[t].T
    return out


def gradients(L, rp, ep, W):
    """Loss of one pair and its gradients (dW1, dW2, dW3), by back-propagation."""
    W1, W2, W3 = W
    nx, ny = rp.shape
    d = ep - smoother(L, rp, W)
    E = float(np.sum(ep * apply_operator(L, ep)))
    Ad = apply_operator(L, d)
    value = float(np.sum(d * Ad) / E)
    g_z = -2.0 * Ad / E  # dL/dz
    G1, G2, G3 = np.zeros_like(W1), np.zeros_like(W2), np.zeros_like(W3)
    feats = features(L)
    for fx0, fy0 in tiles(nx, ny):
        n, s, c = tile_forward(L, rp, W, fx0, fy0, False, feats)
        if c is None:
            continue
        hx, hy = min(TX, nx - fx0), min(TY, ny - fy0)
        g = np.zeros((TX, TY, 1))
        g[:hx, :hy, 0] = s * g_z[fx0:fx0 + hx, fy0:fy0 + hy]
        g[..., 0] *= c["m3"]
        G3 += _corr_grad(c["a2"], g, TX, TY)
        g_a2 = _conv_transpose(g, W3, TX + 2, TY + 2)
        g_h2 = np.where(c["m2"], g_a2 * clip_slope(c["h2"]), 0.0)
        G2 += _corr_grad(c["a1"], g_h2, TX + 2, TY + 2)
        g_a1 = _conv_transpose(g_h2, W2, TX + 4, TY + 4)
        g_h1 = np.where(c["m1"], g_a1 * clip_slope(c["h1"]), 0.0)
        G1 += _corr_grad(c["inp"], g_h1, TX + 4, TY + 4)
    return value, G1, G2, G3


def initial_weights(seed=0):
    """Random first two layers (variance 1 / fan-in), zero last layer: the
    untrained smoother is the damped Jacobi post-smoother."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0.0, 1.0 / np.sqrt(9 * C0), (9, C0, C1))
    W2 = rng.normal(0.0, 1.0 / np.sqrt(9 * C1), (9, C1, C1))
    W3 = np.zeros((9, C1, 1))
    return W1, W2, W3
