"""GPU checks of the learned demosaic's HIP kernels against the oracle
(oracle/net_oracle.py), through the C ABI of include/proresraw_learned.h.

Each layer is checked on the inputs the GPU itself produced (so rounding
differences do not compound): the matrix cores sum in float32 in their own
order and the oracle in float64, so activations may differ by a bf16
rounding step; packing, the loss gradient and the Adam step are exact."""
import numpy as np
import pytest

from oracle import net_oracle as NO

pytestmark = pytest.mark.gpu

CELLS, MARGIN = 24, 4


@pytest.fixture(scope="module")
def env(gpu):
    import torch
    from proresraw_rs_amd import learned as L
    from proresraw_rs_amd import synth as S
    frames = []
    for i, q in enumerate((15, 45)):
        sc = S.scene(128, 96, 0, seed=500 + i, version=0, alignment=2, qscale=(q, q),
                     white_level=4095)
        frames.append(L.frame_pair(S.encode_scene(sc)))
    return torch, L, frames


def frame_host(frame):
    """The whole normalized mosaic of a training frame, red and blue clipped
    at the training stage's clip (1)."""
    h = frame.header
    samples = frame.samples.cpu().numpy().view(np.uint16).reshape(h.coded_height(), h.coded_width())
    p = frame.params
    norm = NO.normalized(samples[:h.height, :h.width], list(h.curve), p.black, p.one_over_range)
    return NO.preclip(norm, 1.0, 1.0)


def crop_host(frame, y0, x0, cells):
    """Normalized mosaic, target and bilinear base of one crop (numpy)."""
    win = frame_host(frame)[y0:y0 + 2 * cells, x0:x0 + 2 * cells]
    target = frame.target.cpu().numpy()[:, y0:y0 + 2 * cells, x0:x0 + 2 * cells]
    return win, target, NO.bilinear(win)


def pack_host(L, frame, y0, x0, cells):
    """The oracle's packing of one crop: the whole frame packed (the
    curvature ratios read neighbours outside the crop), then cut."""
    p = frame.params
    gp_off = np.float32(p.black) * np.float32(p.one_over_range)
    x = NO.pack(frame_host(frame), L.noise_feature(frame.quantizer), p.noise,
                p.black_over_range, gp_off=gp_off)
    return x[y0 // 2:y0 // 2 + cells, x0 // 2:x0 // 2 + cells]


def run_step(L, net, frames, params, crops, lr=0.0):
    net.set_params(params)
    net.train_setup(frames, crops, CELLS, MARGIN)
    net.train_step(lr)
    return {k: net.read(k, l) for k, l in
            [(L.X0, 0), (L.OUT, 0), (L.DOUT, 0), (L.GRAD, 0), (L.LOSS, 0)]}


def bf16_close(got, want, what, steps=2):
    """Equal up to `steps` bf16 rounding steps of the larger magnitude."""
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    tol = steps * 2.0 ** -8 * np.maximum(np.abs(want), np.abs(got)) + 1e-30
    bad = np.abs(got - want) > tol
    assert bad.mean() < 1e-3, f"{what}: {bad.mean():.2e} of values off, worst {np.abs(got - want).max()}"


def test_forward_layers_loss_and_gradients_match_oracle(env):
    torch, L, frames = env
    net = L.Net()
    params = L.init_params(3)
    # Larger residual and output weights than the training start, so every
    # layer contributes measurably.
    for l in range(1, 8):
        params[l * NO.LAYER_PARAMS:l * NO.LAYER_PARAMS + 9216] *= 3.0
    params[7 * NO.LAYER_PARAMS + 9216:7 * NO.LAYER_PARAMS + 9216 + 12] = 0.01
    crops = np.array([[[0, 10, 20], [1, 40, 36]]], np.uint32)
    t = run_step(L, net, frames, params, crops)
    n = crops.shape[1]
    wins, targets, bases = zip(*[crop_host(frames[c[0]], c[1], c[2], CELLS) for c in crops[0]])

    # Packing is exact.
    x0 = np.stack([pack_host(L, frames[c[0]], c[1], c[2], CELLS) for c in crops[0]])
    assert np.abs(x0[..., 5:9]).min() > 0  # the noise-level map is there
    assert (x0[..., 9:13] > 0).any() and (x0[..., 9:13] < 0).any()  # curvature ratios
    assert (x0[..., 13:17] == 1).any() and (x0[..., 13:17] == 0).any()  # above-noise tests
    assert np.array_equal(t[L.X0], x0)

    # Each layer against the oracle on the GPU's own input.
    ws = [NO.bf16(w.astype(np.float32)) for w, _ in NO.split(params)]
    bs = [b for _, b in NO.split(params)]
    h = [net.read(L.H, l) for l in range(7)]
    r = [None] + [net.read(L.R, l) for l in range(1, 7)]
    bf16_close(h[0], NO.bf16(np.maximum(NO.conv(x0, ws[0], bs[0]), 0).astype(np.float32)), "h0")
    for l in range(1, 7):
        rr = NO.bf16(np.maximum(NO.conv(h[l - 1], ws[l], bs[l]), 0).astype(np.float32))
        bf16_close(r[l], rr, f"r{l}")
        bf16_close(h[l], NO.bf16((h[l - 1] + r[l]).astype(np.float32)), f"h{l}", steps=0)
    out = NO.conv(h[6], ws[7], bs[7])
    assert np.allclose(t[L.OUT], out, rtol=1e-4, atol=1e-5 * np.abs(out).max())

    # Loss and its gradient at out are exact given out.
    loss, dout = NO.loss_and_dout(t[L.OUT], np.stack(bases), np.stack(targets), MARGIN)
    assert np.array_equal(t[L.DOUT], dout)
    assert abs(float(t[L.LOSS][0]) - loss) <= 1e-5 * loss

    # Gradients: the oracle backward over the GPU's tensors.
    cache = {"x0": x0, "h": h, "r": r, "w": ws}
    grad = NO.backward(cache, t[L.DOUT])
    for l in range(NO.LAYERS):
        g = grad[l * NO.LAYER_PARAMS:(l + 1) * NO.LAYER_PARAMS]
        gg = t[L.GRAD][l * NO.LAYER_PARAMS:(l + 1) * NO.LAYER_PARAMS]
        scale = np.abs(g).max()
        assert scale > 0
        cos = float(np.dot(g, gg) / (np.linalg.norm(g) * np.linalg.norm(gg)))
        assert cos > 0.9999, (l, cos)
        assert np.abs(g - gg).max() <= 0.02 * scale, (l, np.abs(g - gg).max(), scale)


def test_loss_keeps_the_picture_border_cells(env):
    """Crops on the picture's corners: the sides on the border keep no
    margin in the loss (the convolution's zero padding and the mirrored
    bilinear base there are the whole-frame ones), the inner sides keep it;
    the loss gradient at out stays exact against the oracle."""
    torch, L, frames = env
    net = L.Net()
    params = L.init_params(6)
    for l in range(1, 8):
        params[l * NO.LAYER_PARAMS:l * NO.LAYER_PARAMS + 9216] *= 3.0
    h = frames[1].header
    crops = np.array([[[0, 0, 0], [1, h.height - 2 * CELLS, h.width - 2 * CELLS]]], np.uint32)
    t = run_step(L, net, frames, params, crops)
    wins, targets, bases = zip(*[crop_host(frames[c[0]], c[1], c[2], CELLS) for c in crops[0]])
    borders = [(True, False, True, False), (False, True, False, True)]
    loss, dout = NO.loss_and_dout(t[L.OUT], np.stack(bases), np.stack(targets), MARGIN,
                                  borders=borders)
    assert np.array_equal(t[L.DOUT], dout)
    assert abs(float(t[L.LOSS][0]) - loss) <= 1e-5 * loss
    # The border cells do carry gradient (and the far inner cells none).
    assert np.abs(t[L.DOUT][0, 0, 0]).max() > 0 and np.abs(t[L.DOUT][1, -1, -1]).max() > 0
    assert np.abs(t[L.DOUT][0, -1, -1]).max() == 0 and np.abs(t[L.DOUT][1, 0, 0]).max() == 0


def test_tail_weighted_loss_matches_oracle(env):
    """prr_net_set_loss(tau): the loss d^2 + d^4 / tau^2 and its bf16
    gradient at out, exact against the oracle; tau = 0 gives back the
    squared error bit for bit, and the tail term weighs large errors more
    (a planted tau in the oracle, 0.02 against 0.03, fails)."""
    torch, L, frames = env
    net = L.Net()
    params = L.init_params(8)
    for l in range(1, 8):
        params[l * NO.LAYER_PARAMS:l * NO.LAYER_PARAMS + 9216] *= 3.0
    crops = np.array([[[0, 10, 20], [1, 40, 36]]], np.uint32)
    wins, targets, bases = zip(*[crop_host(frames[c[0]], c[1], c[2], CELLS) for c in crops[0]])
    net.set_loss(0.0)
    t0 = run_step(L, net, frames, params, crops)
    net.set_loss(0.03)
    t = run_step(L, net, frames, params, crops)
    net.set_loss(0.0)
    assert np.array_equal(t[L.OUT], t0[L.OUT])
    loss, dout = NO.loss_and_dout(t[L.OUT], np.stack(bases), np.stack(targets), MARGIN, tau=0.03)
    assert np.array_equal(t[L.DOUT], dout)
    assert abs(float(t[L.LOSS][0]) - loss) <= 1e-5 * loss
    mse, dmse = NO.loss_and_dout(t[L.OUT], np.stack(bases), np.stack(targets), MARGIN)
    assert np.array_equal(t0[L.DOUT], dmse)
    assert float(t[L.LOSS][0]) > float(t0[L.LOSS][0])
    big = np.abs(dmse.astype(np.float32)) > 0
    ratio = np.abs(t[L.DOUT].astype(np.float32)[big]) / np.abs(dmse.astype(np.float32)[big])
    assert ratio.min() >= 1.0 - 2.0 ** -7 and ratio.max() > 1.5
    _, wrong = NO.loss_and_dout(t[L.OUT], np.stack(bases), np.stack(targets), MARGIN, tau=0.02)
    assert not np.array_equal(t[L.DOUT], wrong)


def test_adam_step_is_exact(env):
    torch, L, frames = env
    net = L.Net()
    params = L.init_params(4)
    crops = np.array([[[0, 0, 0], [1, 20, 16]]], np.uint32)
    t = run_step(L, net, frames, params, crops, lr=1e-3)
    want, _, _ = NO.adam(params, t[L.GRAD], np.zeros_like(params), np.zeros_like(params), 1)
    assert np.array_equal(net.params(), want)
    assert int(net.read(L.STEP)[0]) == 1


def test_graph_replay_equals_eager_and_training_is_deterministic(env):
    torch, L, frames = env
    rng = np.random.default_rng(1)
    crops = L.random_crops(rng, frames, 40, 2, CELLS)
    finals = []
    for use_graph in (False, True, True):
        net = L.Net(L.init_params(5))
        net.train_setup(frames, crops, CELLS, MARGIN)
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            if use_graph:
                net.train_step(1e-3, stream)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=stream):
                    net.train_step(1e-3, stream)
                for _ in range(39):
                    g.replay()
                    torch.cuda.synchronize()  # eager host work between replays
            else:
                for _ in range(40):
                    net.train_step(1e-3, stream)
        stream.synchronize()
        assert int(net.read(L.STEP)[0]) == 40
        finals.append(net.params())
    assert np.array_equal(finals[0], finals[1])
    assert np.array_equal(finals[1], finals[2])


def test_training_reduces_the_loss(env):
    torch, L, frames = env
    net = L.Net(L.init_params(6))
    hist = L.train(net, frames, steps=300, batch=4, cells=CELLS, margin=MARGIN, lr=1e-3, seed=2,
                   log_every=25, log=lambda *_: None)
    first, last = hist[0][1], np.mean([l for _, l in hist[-3:]])
    print("loss history", hist)
    assert last < 0.8 * first, hist


def develop_pair(torch, L, frame, stage, net, **kw):
    from proresraw_rs_amd import api as A
    h = frame.header
    n = h.width * h.height
    planes = [torch.zeros(n, dtype=torch.float32, device="cuda") for _ in range(3)]
    net.develop(h, frame.samples, frame.params, stage, planes, **kw)
    torch.cuda.synchronize()
    got = np.stack([p.cpu().numpy() for p in planes])
    fr = A.Frame(h, frame.samples.cpu().numpy().view(np.uint16), [], 0)
    plain = np.stack(A.develop_planar(fr, frame.params, stage).planes)
    return got, plain


def sample_tiles(w, h, k):
    """The sampled 64x64 tiles of the learned develop's check: k tiles
    spread over the frame's tile grid (prr_net_host.cpp sample_tiles)."""
    nbx, nby = (w + 63) // 64, (h + 63) // 64
    out = []
    for i in range(k):
        j = (i * 5 + 2) % k
        out.append((min(nbx - 1, (2 * i + 1) * nbx // (2 * k)),
                    min(nby - 1, (2 * j + 1) * nby // (2 * k))))
    return out


def tile_psnr(got, plain, w, h, bx, by):
    g = got.reshape(3, h, w)[:, 64 * by:64 * by + 64, 64 * bx:64 * bx + 64].astype(np.float64)
    p = plain.reshape(3, h, w)[:, 64 * by:64 * by + 64, 64 * bx:64 * bx + 64].astype(np.float64)
    mse = float(np.mean((g - p) ** 2))
    return 200.0 if mse == 0 else -10.0 * np.log10(mse)


def test_fallbacks_give_plain_develop_exactly(env):
    torch, L, frames = env
    from proresraw_rs_amd.api import Stage
    net = L.Net(L.init_params(7))
    # Noise feature outside the trained range.
    net.set_range(10.0, 11.0)
    got, plain = develop_pair(torch, L, frames[0], Stage.WhiteBalanced, net, check_tiles=0)
    assert net.last_check()[0]
    assert np.array_equal(got, plain)
    # A sampled-tile floor no network reaches.
    net.set_range(0.0, 10.0)
    got, plain = develop_pair(torch, L, frames[1], Stage.Rec2020, net, check_tiles=4, min_db=500.0)
    used, worst = net.last_check()
    assert used and worst < 500.0
    assert np.array_equal(got, plain)


def test_zero_network_is_the_bilinear_base(env):
    torch, L, frames = env
    from proresraw_rs_amd.api import Stage
    net = L.Net(np.zeros(NO.PARAMS, np.float32))
    net.set_range(0.0, 10.0)
    f = frames[0]
    got, plain = develop_pair(torch, L, f, Stage.Demosaiced, net, check_tiles=4, min_db=0.0)
    used, worst = net.last_check()
    assert not used
    h = f.header
    samples = f.samples.cpu().numpy().view(np.uint16).reshape(h.coded_height(), h.coded_width())
    norm = NO.normalized(samples[:h.height, :h.width], list(h.curve), f.params.black,
                         f.params.one_over_range)
    base = NO.bilinear(NO.preclip(norm, 1.0, 1.0))
    want = np.minimum(base, np.float32(1.0)).reshape(3, -1)
    assert np.array_equal(got, want)
    # The sampled-tile PSNR the kernel reports is the PSNR of those tiles.
    assert 20.0 < worst < 80.0


@pytest.mark.parametrize("pattern", [1, 2, 3])
def test_zero_network_is_the_bilinear_base_on_every_pattern(env, pattern):
    """Patterns other than RGGB go through the network flipped to RGGB
    order and come back in sensor order: with a zero network the output is
    the flipped picture's bilinear base, flipped back, exactly (the
    composition writes a cell's two sensels of a row as one pair, swapped
    by a horizontal flip)."""
    torch, L, _ = env
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.api import Stage
    sc = S.scene(128, 96, pattern, seed=600 + pattern, version=0, alignment=2, qscale=(30, 30),
                 white_level=4095)
    f = L.frame_pair(S.encode_scene(sc))
    h = f.header
    net = L.Net(np.zeros(NO.PARAMS, np.float32))
    net.set_range(0.0, 10.0)
    got, _ = develop_pair(torch, L, f, Stage.Demosaiced, net, check_tiles=0)
    samples = f.samples.cpu().numpy().view(np.uint16).reshape(h.coded_height(), h.coded_width())
    norm = NO.normalized(samples[:h.height, :h.width], list(h.curve), f.params.black,
                         f.params.one_over_range)
    red_row, red_col = int(h.bayer_pattern) >> 1, int(h.bayer_pattern) & 1
    assert (red_row, red_col) != (0, 0)
    flip = (slice(None, None, -1 if red_row else 1), slice(None, None, -1 if red_col else 1))
    base = NO.bilinear(NO.preclip(np.ascontiguousarray(norm[flip]), 1.0, 1.0))
    want = np.minimum(base, np.float32(1.0))[(slice(None),) + flip]
    assert np.array_equal(got, np.ascontiguousarray(want).reshape(3, -1))


@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_frame_packing_matches_the_oracle_on_every_pattern(env, pattern):
    """The whole-frame packing (the tiled kernel that shares each sensel's
    work through LDS) equals the oracle's packing of the frame flipped to
    RGGB order, every channel bit for bit: a size that leaves partial tiles,
    a white-balanced stage (red and blue clipped below 1) and the noise
    model of the frame's quantizer."""
    torch, L, _ = env
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.api import DemosaicParams, Stage
    sc = S.scene(150, 98, pattern, seed=700 + pattern, version=pattern & 1, alignment=2,
                 qscale=(20 + 8 * pattern, 20 + 8 * pattern), white_level=4095)
    f = L.frame_pair(S.encode_scene(sc))
    h = f.header
    prm = DemosaicParams.with_white_balance(h, f.quantizer, 2.0, 1.6)
    net = L.Net(np.zeros(NO.PARAMS, np.float32))
    net.set_range(0.0, 10.0)
    n = h.width * h.height
    planes = [torch.zeros(n, dtype=torch.float32, device="cuda") for _ in range(3)]
    net.develop(h, f.samples, prm, Stage.WhiteBalanced, planes, check_tiles=0)
    torch.cuda.synchronize()
    got = net.frame_input(h)
    samples = f.samples.cpu().numpy().view(np.uint16).reshape(h.coded_height(), h.coded_width())
    norm = NO.normalized(samples[:h.height, :h.width], list(h.curve), prm.black,
                         prm.one_over_range)
    red_row, red_col = int(h.bayer_pattern) >> 1, int(h.bayer_pattern) & 1
    v = NO.preclip(NO.canonical(norm, red_row, red_col), *prm.clip)
    gp_off = np.float32(prm.black) * np.float32(prm.one_over_range)
    want = NO.pack(v, L.noise_feature(f.quantizer), prm.noise, prm.black_over_range,
                   gp_off=gp_off, clip=prm.clip)
    assert min(prm.clip) < 1.0
    assert (want[..., 9:13] != 0).any() and (want[..., 13:17] == 1).any()
    bad = np.argwhere(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"{len(bad)} values differ, first at {bad[0]}"


def test_trained_network_passes_the_floor_and_beats_bilinear(env):
    """The committed weights on 5760x3240 frames of seeds outside the
    training set: at least 45 dB against plain develop at
    Stage::Demosaiced over the whole frame, a clear gain over the bilinear
    base alone (the method without the network); with the sampled-tile
    check on, the output is the network's exactly when every sampled tile
    reaches 45 dB, else plain develop."""
    torch, L, _ = env
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.api import Stage
    params, lo, hi = L.load_weights()
    net = L.Net(params)
    net.set_range(lo, hi)
    zero = L.Net(np.zeros(NO.PARAMS, np.float32))
    zero.set_range(0.0, 100.0)
    for q, pattern in ((15, 0), (45, 3)):
        sc = S.scene(5760, 3240, pattern, seed=987654 + q, version=0, alignment=3, qscale=(q, q),
                     white_level=4095)
        f = L.frame_pair(S.encode_scene(sc))
        got, plain = develop_pair(torch, L, f, Stage.Demosaiced, net, check_tiles=0)
        base, _ = develop_pair(torch, L, f, Stage.Demosaiced, zero, check_tiles=0)
        p_net, p_base = L.psnr(got, plain), L.psnr(base, plain)
        checked, _ = develop_pair(torch, L, f, Stage.Demosaiced, net, check_tiles=8, min_db=45.0)
        used, worst = net.last_check()
        print(f"q{q} pattern {pattern}: learned {p_net:.2f} dB, bilinear {p_base:.2f} dB, "
              f"sampled worst {worst:.2f} dB, fallback {used}")
        assert p_net >= 45.0
        assert p_net >= p_base + 4.0
        assert used == (worst < 45.0)
        assert np.array_equal(checked, plain if used else got)
        # The check runs the network on windows around the sampled tiles
        # only; its PSNRs are those of the whole-frame output on the tiles.
        want = min(tile_psnr(got, plain, 5760, 3240, bx, by)
                   for bx, by in sample_tiles(5760, 3240, 8))
        assert abs(worst - want) < 1e-3, (worst, want)


@pytest.mark.parametrize("pairs", [True, False])
@pytest.mark.parametrize("cells", [72, 100, 131])
def test_inference_convolution_equals_the_training_forward(env, cells, pairs):
    """The inference convolutions (persistent, weights held in registers,
    12 output channels; the layers two a launch or one a launch)
    do the training forward's arithmetic: on a frame that is exactly one
    training crop, the learned develop at Stage::Demosaiced equals, bit for
    bit, min(out + bilinear, 1) built from the training forward's output of
    the same crop; frame sizes that are and are not multiples of the 4 x
    64-cell tile, the paired layers' 62-column strips and 64-row segments."""
    torch, L, _ = env
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.api import Stage
    sc = S.scene(2 * cells, 2 * cells, 0, seed=4242 + cells, version=0, alignment=2,
                 qscale=(30, 30), white_level=4095)
    f = L.frame_pair(S.encode_scene(sc))
    net = L.Net(L.init_params(11))
    net.set_pairs(pairs)
    net.set_range(0.0, 100.0)
    net.train_setup([f], np.zeros((1, 1, 3), np.uint32), cells, 0)
    net.train_step(0.0)
    torch.cuda.synchronize()
    out = net.read(L.OUT)[0]
    got, _ = develop_pair(torch, L, f, Stage.Demosaiced, net, check_tiles=0)
    assert not net.last_check()[0]
    win, _, base = crop_host(f, 0, 0, cells)
    full = np.zeros_like(base)
    for dy in range(2):
        for dx in range(2):
            for c in range(3):
                full[c, dy::2, dx::2] = out[:, :, 3 * (2 * dy + dx) + c]
    want = np.minimum((full + base) * np.float32(1.0), np.float32(1.0)).astype(np.float32)
    np.testing.assert_array_equal(got, want.reshape(3, -1))


@pytest.mark.parametrize("pattern", [0, 3])
def test_paired_residual_layers_equal_single_layers(env, pattern):
    """The layers two a launch (the first with the second, residual pairs,
    the sixth with the last; the middle activations kept in LDS along a
    strip) give the planes of one layer a launch bit for bit, with the
    shipped weights, on a frame of 9 strips and 6 segments whose last ones
    are partial, and on a flipped pattern, white balanced to Rec.2020."""
    torch, L, _ = env
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.api import DemosaicParams, Stage
    sc = S.scene(1000, 700, pattern, seed=880 + pattern, version=0, alignment=3,
                 qscale=(33, 33), white_level=4095)
    f = L.frame_pair(S.encode_scene(sc))
    h = f.header
    prm = DemosaicParams.new(h, f.quantizer)
    params, lo, hi = L.load_weights()
    net = L.Net(params)
    net.set_range(lo, hi)
    out = {}
    for pairs in (True, False):
        net.set_pairs(pairs)
        planes = [torch.zeros(h.width * h.height, dtype=torch.float32, device="cuda")
                  for _ in range(3)]
        net.develop(h, f.samples, prm, Stage.Rec2020, planes, check_tiles=0)
        torch.cuda.synchronize()
        assert not net.last_check()[0]
        out[pairs] = np.stack([p.cpu().numpy() for p in planes])
    assert np.array_equal(out[True].view(np.uint32), out[False].view(np.uint32))


def test_frames_over_two_contexts_equal_one_context(env):
    """A batch of frames of two sizes spread over two contexts on their own
    streams (NetGroup) gives each frame the planes one context gives it,
    bit for bit, eagerly and in replays of a captured graph; one frame falls
    back to plain develop (its sampled-tile floor unreachable) in both."""
    torch, L, _ = env
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.api import DemosaicParams, Stage
    params, lo, hi = L.load_weights()
    frames = []
    for i, (w, h) in enumerate([(384, 256), (256, 192), (384, 256), (256, 192), (320, 224)]):
        sc = S.scene(w, h, i % 4, seed=900 + i, version=i % 2, alignment=2,
                     qscale=(20 + 5 * i, 20 + 5 * i), white_level=4095)
        frames.append(L.frame_pair(S.encode_scene(sc)))

    def planes_of(f):
        n = f.header.width * f.header.height
        return [torch.full((n,), -1.0, dtype=torch.float32, device="cuda") for _ in range(3)]

    single = L.Net(params)
    single.set_range(lo, hi)
    want = []
    for i, f in enumerate(frames):
        pl = planes_of(f)
        prm = DemosaicParams.new(f.header, f.quantizer)
        single.develop(f.header, f.samples, prm, Stage.Rec2020, pl, check_tiles=4,
                       min_db=500.0 if i == 2 else 45.0)
        torch.cuda.synchronize()
        want.append(np.stack([p.cpu().numpy() for p in pl]))

    group = L.NetGroup(params, lo, hi, n=2)
    outs = [planes_of(f) for f in frames]
    s = torch.cuda.Stream()

    def run():
        for i, f in enumerate(frames):
            prm = DemosaicParams.new(f.header, f.quantizer)
            group.develop([(f.header, f.samples, prm, outs[i], None)], Stage.Rec2020,
                          check_tiles=4, min_db=500.0 if i == 2 else 45.0, stream=s)

    def run_batch():
        # One call per floor: frame 2 alone at the unreachable floor, the
        # others together.
        jobs = [(f.header, f.samples, DemosaicParams.new(f.header, f.quantizer), outs[i], None)
                for i, f in enumerate(frames) if i != 2]
        group.develop(jobs, Stage.Rec2020, check_tiles=4, min_db=45.0, stream=s)
        f = frames[2]
        group.develop([(f.header, f.samples, DemosaicParams.new(f.header, f.quantizer), outs[2],
                        None)], Stage.Rec2020, check_tiles=4, min_db=500.0, stream=s)

    def check(what):
        for i in range(len(frames)):
            got = np.stack([p.cpu().numpy() for p in outs[i]])
            assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), (what, i)

    with torch.cuda.stream(s):
        run()
        s.synchronize()
        check("eager, one frame a call")
        run_batch()
        s.synchronize()
        check("eager, frames spread over the contexts")
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            run_batch()
        for rep in range(3):
            for o in outs:
                for p in o:
                    p.fill_(-1.0)
            s.synchronize()
            g.replay()
            s.synchronize()
            check(f"graph replay {rep}")
