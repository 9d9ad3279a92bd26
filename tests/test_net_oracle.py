"""CPU checks that pin the learned-demosaic oracle (oracle/net_oracle.py)
before it judges the HIP kernels: bf16 rounding on known values, the
bilinear base on closed forms, adjoint identities of the convolution and
its two gradients, the backward pass against central finite differences
and against torch's CPU autograd (float64, an independent formulation),
and Adam's first step against its closed form."""
import numpy as np
import pytest

from oracle import net_oracle as NO

torch = pytest.importorskip("torch")


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NO.PARAMS, np.float64)
    for l, (w, b) in enumerate(NO.split(np.zeros(NO.PARAMS))):
        base = l * NO.LAYER_PARAMS
        cin = 5 if l == 0 else NO.C
        std = scale * np.sqrt(2.0 / (9 * cin))
        wl = np.zeros((9, NO.C, NO.C))
        wl[:, :, :cin] = rng.standard_normal((9, NO.C, cin)) * std
        p[base:base + 9 * NO.C * NO.C] = wl.reshape(-1)
        p[base + 9 * NO.C * NO.C:base + NO.LAYER_PARAMS] = rng.standard_normal(NO.C) * 0.05
    return p


def rand_input(seed, n=2, h=6, w=6):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, h, w, NO.C))
    x[..., :5] = rng.uniform(0, 1, (n, h, w, 5))
    return x


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    cases = {
        1.0: 1.0,
        1.0 + 2.0 ** -8: 1.0,                    # halfway, ties to even (down)
        1.0 + 3 * 2.0 ** -8: 1.0 + 2.0 ** -6,    # halfway, ties to even (up)
        1.0 + 2.0 ** -8 + 2.0 ** -20: 1.0 + 2.0 ** -7,
        -2.5: -2.5,
        65504.0: 65536.0,                        # nearer 2^16 than 65280
    }
    for v, want in cases.items():
        assert float(NO.bf16(np.array([one * np.float32(v)]))[0]) == np.float32(want), v
    assert np.isnan(NO.bf16(np.float32(np.nan)))


def test_bilinear_per_colour_constants_are_kept():
    # Distinct constants per colour site: a neighbour rule that reads the
    # wrong sites (sides for above/below, say) picks up another colour.
    v = np.zeros((10, 12), np.float32)
    v[0::2, 0::2], v[0::2, 1::2], v[1::2, 0::2], v[1::2, 1::2] = 0.7, 0.3, 0.3, 0.1
    base = NO.bilinear(v)
    assert np.array_equal(base[0], np.full_like(v, 0.7))
    assert np.array_equal(base[1], np.full_like(v, 0.3))
    assert np.array_equal(base[2], np.full_like(v, 0.1))


def test_bilinear_reproduces_a_linear_ramp_inside():
    y, x = np.mgrid[0:12, 0:16].astype(np.float32)
    v = (0.01 * x + 0.02 * y).astype(np.float32)
    base = NO.bilinear(v)
    inner = (slice(1, -1), slice(1, -1))
    for c in range(3):
        assert np.allclose(base[c][inner], v[inner], atol=1e-6)


def test_conv_transpose_is_the_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 7, NO.C))
    g = rng.standard_normal((2, 5, 7, NO.C))
    w = rng.standard_normal((9, NO.C, NO.C))
    lhs = np.sum(NO.conv(x, w) * g)
    rhs = np.sum(x * NO.conv_transpose(g, w))
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def test_weight_grad_is_the_adjoint_in_w():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 5, 7, NO.C))
    g = rng.standard_normal((2, 5, 7, NO.C))
    dw = rng.standard_normal((9, NO.C, NO.C))
    db = rng.standard_normal(NO.C)
    gw, gb = NO.weight_grad(x, g)
    lhs = np.sum(gw * dw) + np.sum(gb * db)
    rhs = np.sum(NO.conv(x, dw, db) * g)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def loss_of(params, x0, base, target, margin):
    cache = NO.forward(x0, params, emulate=False)
    return NO.loss_and_dout(cache["out"], base, target, margin, emulate=False)


def test_backward_matches_central_differences():
    x0 = rand_input(3)
    params = rand_params(4, scale=0.8)
    rng = np.random.default_rng(5)
    n, h, w, _ = x0.shape
    base = rng.uniform(0, 1, (n, 3, 2 * h, 2 * w))
    target = rng.uniform(0, 1, (n, 3, 2 * h, 2 * w))
    cache = NO.forward(x0, params, emulate=False)
    _, dout = NO.loss_and_dout(cache["out"], base, target, 1, emulate=False)
    grad = NO.backward(cache, dout, emulate=False)
    picks = []
    for l in range(NO.LAYERS):
        base_i = l * NO.LAYER_PARAMS
        cout = NO.OUT_CHANNELS if l == NO.LAYERS - 1 else NO.C
        cin = 5 if l == 0 else NO.C
        for _ in range(4):
            t, co, ci = rng.integers(9), rng.integers(cout), rng.integers(cin)
            picks.append(base_i + t * NO.C * NO.C + co * NO.C + ci)
        picks.append(base_i + 9 * NO.C * NO.C + rng.integers(cout))
    eps = 1e-6
    for i in picks:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (loss_of(p1, x0, base, target, 1)[0] - loss_of(p2, x0, base, target, 1)[0]) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-5 * max(1e-3, abs(fd)), (i, fd, grad[i])


def torch_net(params, x0, base, target, margin):
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    x = torch.tensor(x0, dtype=torch.float64).permute(0, 3, 1, 2)

    def layer(l, t):
        wb = p[l * NO.LAYER_PARAMS:(l + 1) * NO.LAYER_PARAMS]
        w = wb[:9 * NO.C * NO.C].reshape(3, 3, NO.C, NO.C).permute(2, 3, 0, 1)
        return torch.nn.functional.conv2d(t, w, wb[9 * NO.C * NO.C:], padding=1)

    h = torch.relu(layer(0, x))
    for l in range(1, 7):
        h = h + torch.relu(layer(l, h))
    out = layer(7, h)[:, :NO.OUT_CHANNELS]
    pred = torch.nn.functional.pixel_shuffle(
        out.reshape(out.shape[0], 2, 2, 3, out.shape[2], out.shape[3])
        .permute(0, 3, 1, 2, 4, 5).reshape(out.shape[0], 12, out.shape[2], out.shape[3]), 2)
    pred = pred + torch.tensor(base)
    n, _, H, W = pred.shape
    m = 2 * margin
    diff = (pred - torch.tensor(target))[:, :, m:H - m, m:W - m]
    loss = (diff ** 2).mean()
    loss.backward()
    return loss.item(), p.grad.numpy(), out.detach().permute(0, 2, 3, 1).numpy()


def test_forward_and_backward_match_torch_autograd():
    x0 = rand_input(6, n=2, h=8, w=10)
    params = rand_params(7)
    rng = np.random.default_rng(8)
    n, h, w, _ = x0.shape
    base = rng.uniform(0, 1, (n, 3, 2 * h, 2 * w))
    target = rng.uniform(0, 1, (n, 3, 2 * h, 2 * w))
    tl, tg, tout = torch_net(params, x0, base, target, 2)
    cache = NO.forward(x0, params, emulate=False)
    assert np.allclose(cache["out"][..., :NO.OUT_CHANNELS], tout, rtol=1e-10, atol=1e-12)
    loss, dout = NO.loss_and_dout(cache["out"], base, target, 2, emulate=False)
    assert abs(loss - tl) <= 1e-12 * abs(tl)
    grad = NO.backward(cache, dout, emulate=False)
    scale = np.abs(tg).max()
    assert np.abs(grad - tg).max() <= 1e-10 * scale


def test_shuffle_round_trip():
    rng = np.random.default_rng(9)
    out = rng.standard_normal((2, 3, 4, NO.C)).astype(np.float32)
    back = NO.unshuffle(NO.shuffle(out))
    assert np.array_equal(back[..., :12], out[..., :12])
    assert not back[..., 12:].any()


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(10)
    p = rng.standard_normal(1000).astype(np.float32)
    g = rng.standard_normal(1000).astype(np.float32)
    z = np.zeros(1000, np.float32)
    p1, m1, v1 = NO.adam(p, g, z, z, 1)
    # m/(1-b1) = g and sqrt(v/(1-b2)) = |g| up to rounding: a step of lr.
    assert np.allclose(p1, p - 1e-3 * np.sign(g), rtol=0, atol=2e-9 + 1e-6 * np.abs(p).max())
    assert np.allclose(m1, 0.1 * g, rtol=1e-6)
    assert np.allclose(v1, 0.001 * g * g, rtol=1e-6)


def test_pack_and_normalize_follow_develop_units():
    curve = [0, 2000, 8000, 17000, 28000, 40000, 52000, 63000]
    codes = np.array([[0, 8192, 16000, 65535]], np.uint16)
    lin = NO.linearize(curve, codes)
    assert lin[0, 0] == 0 and lin[0, 1] == 2000 and lin[0, 3] <= 65535
    v = NO.normalized(codes, curve, 256, 1.0 / 3839.0)
    assert np.allclose(v, (lin - 256) / 3839.0, rtol=1e-6)
    mosaic = np.arange(16, dtype=np.float32).reshape(4, 4) / 16
    x = NO.pack(mosaic, 0.5)
    assert x.shape == (2, 2, NO.C)
    assert float(x[1, 0, 0]) == float(NO.bf16(mosaic[2:3, 0])[0])
    assert float(x[0, 1, 3]) == float(NO.bf16(mosaic[1:2, 3])[0])
    assert float(x[0, 0, 4]) == 0.5 and not x[..., 5:].any()


def _wgsl_params(q, clip, black=256, rng_=3839):
    """The develop uniforms of a 12-bit RGGB frame at quantizer q
    (DemosaicParams::with_white_balance's noise model, demosaic.rs:86-101)."""
    from tests import wgsl_develop as W
    F = np.float32
    base = q * q * 5.0 + 64.0 / 3.0
    noise = (F(base / 150994944.0), F(base / (rng_ * rng_ * 256.0)), F(64.0 / 3.0 / rng_ ** 2))
    return W.Params(0, black, F(1.0 / rng_), noise, clip, F(black / rng_), (1.0, 1.0))


def test_curvature_ratios_equal_the_wgsl_restatements():
    """NO.curvature_ratios (the network's input channels 9-12) equals, bit
    for bit, the curvature ratios of the second restatement of the crate's
    green pass (tests/wgsl_develop.py, written from develop.wgsl and itself
    equal bit for bit to the oracle's develop) at every red and blue sensel
    of random mosaics and constructed fields, with the ratio non-zero in
    both directions and signs, zeroed by the noise tests and by the centre's
    gate, and every edge rule taken."""
    from tests import wgsl_develop as W
    from tests.test_develop_restatement import smooth_field
    F = np.float32
    rng = np.random.default_rng(7)
    cases = []
    for i in range(24):
        w, h = 2 * int(rng.integers(3, 24)), 2 * int(rng.integers(3, 20))
        lo, hi = [(0, 4096), (200, 3000), (0, 600), (250, 4095)][i % 4]
        cases.append((rng.integers(lo, hi, size=(h, w)), [15, 45, 135][i % 3],
                      (1.0, 1.0) if i % 2 else (F(0.8), F(0.6))))
    for j, kind in enumerate(("bump", "hue_step", "ramp_cross")):
        for w, h in ((32, 24), (24, 32)):
            cases.append((smooth_field(w, h, kind, j), 15, (1.0, 1.0)))
    # Every picture has even sides, so red sensels sit at positions 0 and
    # len - 2 of their lines and blue ones at 1 and len - 1: each case takes
    # all four edge rules in both directions.
    seen = dict.fromkeys(["h+", "h-", "v+", "v-", "h0_noise", "v0_noise", "gate_closed"], 0)
    for lin, q, clip in cases:
        p = _wgsl_params(q, clip)
        lin = np.asarray(lin, np.int64)
        norm = (F(p.one_over_range) * (lin - p.black).astype(np.float32)).astype(np.float32)
        cov, want = W.Coverage(), {}
        W.green(norm, p, cov, ratios=want)
        v = NO.preclip(norm, *clip)
        off = F(p.black) * F(p.one_over_range)
        for sy, sx, cl in ((0, 0, clip[0]), (1, 1, clip[1])):
            got = NO.curvature_ratios(v, p.noise, off, sy, sx, cl)
            for k, name in enumerate("hv"):
                ref = want[name][sy::2, sx::2]
                assert np.array_equal(got[k].view(np.uint32), ref.view(np.uint32)), (name, sy)
                seen[name + "+"] += int((got[k] > 0).sum())
                seen[name + "-"] += int((got[k] < 0).sum())
        for k in ("h", "v"):
            seen[k + "0_noise"] += cov.counts.get(f"ratio_{k}_zero_by_noise", 0)
        seen["gate_closed"] += cov.counts.get("green_gate_closed", 0)
    print("ratio branches:", seen)
    assert all(n > 0 for n in seen.values()), seen


def test_above_noise_channels_follow_the_chroma_test():
    """Channels 13-16 are 1 exactly where deviation(v) < v (the chroma
    estimate's test, demosaic.rs:576-587): with a = 0 the deviation is the
    constant sqrt(b + c), so the bit turns on strictly above it."""
    F = np.float32
    noise = (F(0.0), F(1e-4), F(2.25e-4))   # deviation sqrt(3.25e-4)
    s = np.sqrt(F(3.25e-4))
    vals = [np.nextafter(s, F(0)), s, np.nextafter(s, F(1)), F(0.5)]
    v = np.array([[vals[0], vals[1]], [vals[2], vals[3]]], np.float32)
    v = np.tile(v, (3, 3))
    x = NO.pack(v, 0.0, noise, 0.0, gp_off=F(0.0))
    assert x[0, 0, 13:17].tolist() == [0.0, 0.0, 1.0, 1.0]
    assert np.all(x[..., 5:9] == NO.bf16(np.full((3, 3, 4), F(128.0) * s)))
