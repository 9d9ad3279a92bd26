"""NumPy restatement of the learned demosaic: the network, its loss, its
backward pass and the Adam step, operation for operation as the HIP kernels
in proresraw_rs_amd/csrc/prr_net.hip compute them.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; never by
proresraw_rs_amd.

The network has no counterpart in the proresraw_rs crate. What the crate
fixes is the data it is trained on and judged against: the normalized
mosaic of `Conversion::new` (demosaic.rs:275-304) and the output of
`develop(..., Stage::Demosaiced)` (demosaic.rs:190-402).

Network, on the mosaic packed at half resolution (NHWC, 32 channels):
  x0      channels 0-3: R, G on the red row, G on the blue row, B of each
          2x2 cell of the RGGB mosaic (other patterns are flipped to RGGB);
          channel 4: the noise feature 128*sqrt(noise[0]); 5-8: the
          noise-level map 128*sigma(v) of the cell's four sensels (sigma
          the square root of develop's variance, demosaic.rs:420-425);
          9-12: the green pass's curvature ratios (horizontal, vertical)
          at the red sensel, then at the blue one, 0 where its noise tests
          leave them out (demosaic.rs:495-532); 13-16: 1 where each of the
          four sensels lies above its noise deviation at the chroma offset,
          else 0 (the test of demosaic.rs:576-587); 17-31: zero
  h0      = relu(conv0(x0) + b0)
  h_l     = h_{l-1} + relu(conv_l(h_{l-1}) + b_l),   l = 1..6
  out     = conv7(h6) + b7; channel 3*(2*dy+dx)+c is colour c of the
            sensel (2i+dy, 2j+dx) of cell (i, j)
  pred    = out (pixel-shuffled) + bilinear demosaic of the mosaic
Parameters: 8 layers of W[tap][co][ci] (tap = 3*ky + kx reads pixel
(y+ky-1, x+kx-1), zero outside) then b[co]: 9248 floats a layer.

With emulate=True the rounding points of the kernels are reproduced: bf16
weights, bf16 activations after every layer epilogue, bf16 loss and
activation gradients. Sums are taken in float64 here and in float32 by the
matrix cores, so comparisons with the kernels use a tolerance. With
emulate=False everything is float64 (finite differences, torch cross-check).
"""
import numpy as np

C = 32
TAPS = 9
LAYER_PARAMS = TAPS * C * C + C
LAYERS = 8
PARAMS = LAYERS * LAYER_PARAMS
OUT_CHANNELS = 12


def bf16(a):
    """Round float32 to the nearest bf16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x7FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def split(params):
    """Per layer: (W[9][co][ci], b[co]) views of the flat parameters."""
    p = np.asarray(params)
    out = []
    for l in range(LAYERS):
        base = l * LAYER_PARAMS
        w = p[base:base + TAPS * C * C].reshape(TAPS, C, C)
        b = p[base + TAPS * C * C:base + LAYER_PARAMS]
        out.append((w, b))
    return out


def linearize(curve8, codes):
    """Curve::linearize (curve.rs:14-24) over an array of codes."""
    c = np.array(list(curve8) + [65536], np.int64)
    code = np.asarray(codes, np.int64)
    seg, frac = code >> 13, code & 0x1FFF
    c0, c1 = c[seg], c[seg + 1]
    v = (c0 * 8192 + ((c1 - c0) & 0xFFFF) * frac + 4096) >> 13
    return np.minimum(v, 65535)


def normalized(samples, curve8, black, one_over_range):
    """one_over_range * (linearize(code) - black) in float32, the value
    develop works on (Conversion::new, demosaic.rs:293-298)."""
    lin = linearize(curve8, samples)
    return (np.float32(one_over_range) * (lin - black).astype(np.float32)).astype(np.float32)


def noise_feature(noise0):
    return np.float32(128.0) * np.sqrt(np.float32(noise0), dtype=np.float32)


def canonical(plane, flip_y, flip_x):
    """Flip a sensor-order plane to RGGB order (or back: flips are their own
    inverse)."""
    p = plane
    if flip_y:
        p = p[::-1]
    if flip_x:
        p = p[:, ::-1]
    return np.ascontiguousarray(p)


def preclip(v, clip_r, clip_b):
    """Red and blue sensels clipped at the stage's clip values."""
    v = v.copy()
    v[0::2, 0::2] = np.minimum(v[0::2, 0::2], np.float32(clip_r))
    v[1::2, 1::2] = np.minimum(v[1::2, 1::2], np.float32(clip_b))
    return v


def deviation(v, noise, off):
    """sqrt(c + max(b, a * (off + v)^2)) in float32: the noise deviation
    develop compares values with (demosaic.rs:420-425, 512, 521, 579-581)."""
    f = np.float32
    a, b, c = (f(x) for x in noise)
    t = f(off) + np.asarray(v, np.float32)
    return np.sqrt(c + np.maximum(b, a * (t * t))).astype(np.float32)


def mirror(d, pos, n, c):
    """The green pass's edge rule (mirror, demosaic.rs:547-557) on the
    neighbour arrays d[0..3] of sensels at positions `pos` of lines of n,
    centres c: the crate's if / else-if chain as masks."""
    m0 = pos == 0
    m1 = ~m0 & (pos == 1)
    ml = ~m0 & ~m1 & (pos == n - 1)
    ml2 = ~m0 & ~m1 & ~ml & (pos == n - 2)
    d0, d1, d2, d3 = d
    return [np.where(m0, d1, d0), np.where(ml, d0, d1),
            np.where(m0, d3, np.where(m1, c, d2)),
            np.where(m0, d3, np.where(ml, d2, np.where(ml2, c, d3)))]


def curvature_ratios(v, noise, gp_off, sy, sx, clip):
    """The green pass's curvature ratios (green_row_impl,
    demosaic.rs:495-532) at the sensels (sy + 2i, sx + 2j) of the clipped
    RGGB plane v: (horizontal, vertical) arrays, t[2+dir] / (t[dir] * 2)
    where the crate's noise tests apply it, else 0. Rows and columns are
    clamped (demosaic.rs:469-472, 499), greens clipped at 1 and same-colour
    sensels at `clip` (demosaic.rs:496, 502-505), then mirrored at the
    picture's edges (demosaic.rs:506-507)."""
    f = np.float32
    hgt, wid = v.shape
    ys = np.arange(sy, hgt, 2)[:, None]
    xs = np.arange(sx, wid, 2)[None, :]

    def at(dy, dx):
        return v[np.clip(ys + dy, 0, hgt - 1), np.clip(xs + dx, 0, wid - 1)]

    one, cl = f(1.0), f(clip)
    c = np.minimum(at(0, 0), cl)
    hh = mirror([np.minimum(at(0, -1), one), np.minimum(at(0, 1), one),
                 np.minimum(at(0, -2), cl), np.minimum(at(0, 2), cl)], xs, wid, c)
    vv = mirror([np.minimum(at(-1, 0), one), np.minimum(at(1, 0), one),
                 np.minimum(at(-2, 0), cl), np.minimum(at(2, 0), cl)], ys, hgt, c)
    gate = c > deviation(c, noise, gp_off)
    sums2, sums3 = hh[2] + hh[3], vv[2] + vv[3]
    t = [c * f(2.0) + sums2, c * f(2.0) + sums3, c * f(-2.0) + sums2, c * f(-2.0) + sums3]
    out = []
    for k, d in enumerate((hh, vv)):
        sd = [deviation(x, noise, gp_off) for x in d]
        p, q = d[0] - d[1], d[2] - d[3]
        any_ = ((p * q > 0) | ((d[2] - c) * (d[3] - c) > 0) | (sd[0] + sd[1] > np.abs(p)) |
                (sd[2] + sd[3] > np.abs(q)))
        all_ = (d[0] > sd[0]) & (d[1] > sd[1]) & (d[2] > sd[2]) & (d[3] > sd[3])
        with np.errstate(all="ignore"):
            r = (t[2 + k] / (t[k] * f(2.0))).astype(np.float32)
        out.append(np.where(gate & any_ & all_, r, f(0.0)).astype(np.float32))
    return out


def pack(v, feature, noise=None, off=0.0, gp_off=None, clip=(1.0, 1.0)):
    """RGGB plane (H, W), red and blue already clipped at `clip` -> x0
    (H/2, W/2, 32), bf16; with `noise` (a, b, c) and `off` (black / range),
    the noise-level map in channels 5-8; with `gp_off` (black *
    one_over_range, the green pass's offset) as well, the curvature ratios
    and above-noise tests in channels 9-16. v is the whole picture: the
    ratios read neighbours two sensels away."""
    h2, w2 = v.shape[0] // 2, v.shape[1] // 2
    x = np.zeros((h2, w2, C), np.float32)
    x[..., 0] = v[0:2 * h2:2, 0:2 * w2:2]
    x[..., 1] = v[0:2 * h2:2, 1:2 * w2:2]
    x[..., 2] = v[1:2 * h2:2, 0:2 * w2:2]
    x[..., 3] = v[1:2 * h2:2, 1:2 * w2:2]
    x[..., 4] = np.float32(feature)
    if noise is not None:
        dev = [deviation(x[..., k], noise, off) for k in range(4)]
        for k in range(4):
            x[..., 5 + k] = np.float32(128.0) * dev[k]
        if gp_off is not None:
            rr = curvature_ratios(v, noise, gp_off, 0, 0, clip[0])
            rb = curvature_ratios(v, noise, gp_off, 1, 1, clip[1])
            for k, r in enumerate(rr + rb):
                x[..., 9 + k] = r[:h2, :w2]
            for k in range(4):
                x[..., 13 + k] = (dev[k] < x[..., k]).astype(np.float32)
    return bf16(x)


def bilinear(v):
    """Bilinear demosaic of an RGGB plane, mirrored at its edges
    (index -1 -> 1, n -> n-2): (3, H, W) float32, sums in the order the
    kernel adds them."""
    v = np.asarray(v, np.float32)
    H, W = v.shape
    p = np.pad(v, 1, mode="reflect")
    f = np.float32

    def at(dy, dx):
        return p[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]

    cross = ((at(-1, 0) + at(1, 0)) + (at(0, -1) + at(0, 1))) * f(0.25)
    diag = ((at(-1, -1) + at(-1, 1)) + (at(1, -1) + at(1, 1))) * f(0.25)
    horiz = (at(0, -1) + at(0, 1)) * f(0.5)
    vert = (at(-1, 0) + at(1, 0)) * f(0.5)
    y = np.arange(H)[:, None] & 1
    x = np.arange(W)[None, :] & 1
    red, blue = (y == 0) & (x == 0), (y == 1) & (x == 1)
    g_red_row, g_blue_row = (y == 0) & (x == 1), (y == 1) & (x == 0)
    r = np.where(red, v, np.where(blue, diag, np.where(g_red_row, horiz, vert)))
    g = np.where(red | blue, cross, v)
    b = np.where(blue, v, np.where(red, diag, np.where(g_red_row, vert, horiz)))
    return np.stack([r, g, b]).astype(np.float32)


def conv(x, w, b=None):
    """3x3 convolution, zero padded: x (N, H, W, 32), w (9, co, ci) -> float64."""
    x = np.asarray(x, np.float64)
    N, H, W, _ = x.shape
    xp = np.zeros((N, H + 2, W + 2, C), np.float64)
    xp[:, 1:-1, 1:-1] = x
    out = np.zeros((N, H, W, w.shape[1]), np.float64)
    for t in range(TAPS):
        ky, kx = divmod(t, 3)
        out += xp[:, ky:ky + H, kx:kx + W, :] @ np.asarray(w[t], np.float64).T
    if b is not None:
        out += np.asarray(b, np.float64)
    return out


def conv_transpose(g, w):
    """The gradient of conv with respect to its input: a 3x3 convolution of
    g with the flipped, transposed weights."""
    wt = np.stack([np.asarray(w[8 - t], np.float64).T for t in range(TAPS)])
    return conv(g, wt)


def weight_grad(x, g):
    """dW[t][co][ci] = sum over pixels of g[p][co] * x[p + off(t)][ci] and
    db[co] = sum of g[p][co], in float64."""
    x = np.asarray(x, np.float64)
    g = np.asarray(g, np.float64)
    N, H, W, _ = x.shape
    xp = np.zeros((N, H + 2, W + 2, C), np.float64)
    xp[:, 1:-1, 1:-1] = x
    gw = np.zeros((TAPS, C, C), np.float64)
    for t in range(TAPS):
        ky, kx = divmod(t, 3)
        xs = xp[:, ky:ky + H, kx:kx + W, :].reshape(-1, C)
        gw[t] = g.reshape(-1, C).T @ xs
    return gw, g.reshape(-1, C).sum(axis=0)


def _round(a, emulate):
    return bf16(np.asarray(a, np.float32)) if emulate else np.asarray(a, np.float64)


def forward(x0, params, emulate=True):
    """x0 (N, H2, W2, 32) -> cache with h[0..6], r[1..6] and out (float)."""
    layers = split(params)
    ws = [(bf16(w) if emulate else np.asarray(w, np.float64), b) for w, b in layers]
    h, r = [], [None]
    z = conv(x0, *ws[0])
    h.append(_round(np.maximum(z, 0), emulate))
    for l in range(1, 7):
        rl = _round(np.maximum(conv(h[-1], *ws[l]), 0), emulate)
        r.append(rl)
        if emulate:
            h.append(bf16((h[-1].astype(np.float32) + rl.astype(np.float32))))
        else:
            h.append(h[-1] + rl)
    out = conv(h[6], *ws[7])
    if emulate:
        out = out.astype(np.float32)
    return {"x0": x0, "h": h, "r": r, "out": out, "w": [w for w, _ in ws]}


def shuffle(out):
    """out (N, H2, W2, >=12) -> (N, 3, 2*H2, 2*W2)."""
    N, H2, W2, _ = out.shape
    pred = np.zeros((N, 3, 2 * H2, 2 * W2), out.dtype)
    for dy in range(2):
        for dx in range(2):
            for c in range(3):
                pred[:, c, dy::2, dx::2] = out[..., 3 * (2 * dy + dx) + c]
    return pred


def unshuffle(dpred, channels=C):
    N, _, H, W = dpred.shape
    d = np.zeros((N, H // 2, W // 2, channels), dpred.dtype)
    for dy in range(2):
        for dx in range(2):
            for c in range(3):
                d[..., 3 * (2 * dy + dx) + c] = dpred[:, c, dy::2, dx::2]
    return d


def loss_and_dout(out, base, target, margin, emulate=True, borders=None, tau=0.0):
    """Mean squared error over the cells [margin, H2 - margin) x [margin,
    W2 - margin) (all 12 sub-pixel colours) and its gradient with respect to
    out; the per-element derivative 2 * diff / count is rounded to bf16.
    borders[k] = (top, bottom, left, right): the sides of crop k that lie on
    the picture's border, where the margin is 0; count stays the interior
    one. tau > 0: the mean of diff^2 + diff^4 / tau^2, derivative
    2 * diff * (1 + 2 * diff^2 / tau^2) / count."""
    N, H2, W2, _ = out.shape
    ft = np.float32 if emulate else np.float64
    pred = shuffle(np.asarray(out, ft)[..., :OUT_CHANNELS]) + np.asarray(base, ft)
    diff = pred - np.asarray(target, ft)
    interior = np.zeros((2 * H2, 2 * W2), bool)
    interior[2 * margin:2 * (H2 - margin), 2 * margin:2 * (W2 - margin)] = True
    count = N * 3 * int(interior.sum())
    mask = np.zeros((N, 1, 2 * H2, 2 * W2), bool)
    for k in range(N):
        t, b, l, r = borders[k] if borders is not None else (False,) * 4
        m0, m1 = (0 if t else margin), (0 if b else margin)
        m2, m3 = (0 if l else margin), (0 if r else margin)
        mask[k, 0, 2 * m0:2 * (H2 - m1), 2 * m2:2 * (W2 - m3)] = True
    d = np.where(mask, diff, 0)
    d64 = np.asarray(d, np.float64)
    inv = 1.0 / (tau * tau) if tau > 0 else 0.0
    loss = float(np.sum(d64 ** 2 * (1.0 + d64 ** 2 * inv)) / count)
    if emulate:
        scale = np.float32(2.0) / np.float32(count)
        d32 = d.astype(np.float32)
        inv32 = np.float32(inv)
        dpred = ((d32 * scale) * (np.float32(1.0) + np.float32(2.0) * (d32 * d32) * inv32)).astype(np.float32)
    else:
        dpred = d * (2.0 / count) * (1.0 + 2.0 * d * d * inv)
    dout = unshuffle(dpred)
    return loss, (bf16(dout) if emulate else dout)


def backward(cache, dout, emulate=True):
    """Flat parameter gradient for the loss whose gradient at `out` is dout."""
    w, h, r = cache["w"], cache["h"], cache["r"]
    grads = [None] * LAYERS
    gw, gb = weight_grad(h[6], dout)
    grads[7] = (gw, gb)
    dh = _round(conv_transpose(dout, w[7]), emulate)
    for l in range(6, 0, -1):
        dz = np.where(r[l] > 0, dh, 0)
        grads[l] = weight_grad(h[l - 1], dz)
        t = conv_transpose(dz, w[l])
        if emulate:
            dh = bf16((dh.astype(np.float64) + t).astype(np.float32))
        else:
            dh = dh + t
    dz = np.where(h[0] > 0, dh, 0)
    grads[0] = weight_grad(cache["x0"], dz)
    flat = np.zeros(PARAMS, np.float64)
    for l, (gw, gb) in enumerate(grads):
        base = l * LAYER_PARAMS
        flat[base:base + TAPS * C * C] = gw.reshape(-1)
        flat[base + TAPS * C * C:base + LAYER_PARAMS] = gb
    return flat


def adam(params, grads, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1), float32 as the kernel computes:
    m = b1*m + (1-b1)*g; v = b2*v + (1-b2)*g*g;
    p -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)."""
    f = np.float32
    g = np.asarray(grads, f)
    m = (f(b1) * m + f(1 - b1) * g).astype(f)
    v = (f(b2) * v + f(1 - b2) * (g * g)).astype(f)
    c1 = f(1.0 - b1 ** step)
    c2 = f(1.0 - b2 ** step)
    p = (params - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))).astype(f)
    return p, m, v
