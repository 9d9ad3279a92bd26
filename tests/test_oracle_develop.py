"""The oracle's develop and reduced decode against the crate's tests
(demosaic.rs:816-978, reduced.rs:285-360) and closed forms."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from proresraw_rs_amd import synth as S

RANGE = 65279 - 256


def params(pattern, wb):
    p = O.OrcParams()
    p.cfa_phase, p.black = pattern, 256
    p.one_over_range = np.float32(1.0) / np.float32(RANGE)
    p.noise[:] = [1e-6, 1e-9, 5e-9]
    p.clip[:] = [float(np.float32(1.0) / np.float32(g)) for g in wb]
    p.black_over_range = np.float32(256.0) / np.float32(RANGE)
    p.wb[:] = list(wb)
    return p


def color_at(pattern, y, x):
    rr, rc = [(0, 0), (0, 1), (1, 0), (1, 1)][pattern]
    y, x = y & 1, x & 1
    if (y, x) == (rr, rc):
        return 0
    if (y, x) == (rr ^ 1, rc ^ 1):
        return 2
    return 1


def mosaic(pattern, w, h, rgb):
    return np.array([[rgb[color_at(pattern, y, x)] for x in range(w)] for y in range(h)],
                    np.uint16).ravel()


def passes(lin, w, h, p):
    g = np.zeros(2 * w * h, np.float32)
    O.lib().orc_green_pass(lin.ctypes.data, w, h, ctypes.byref(p), g.ctypes.data)
    rgb = np.zeros(3 * w * h, np.float32)
    O.lib().orc_chroma_pass(g.ctypes.data, w, h, ctypes.byref(p), rgb.ctypes.data)
    return rgb


@pytest.mark.parametrize("pattern", range(4))
def test_flat_field_is_reproduced_exactly(pattern):
    w, h = 12, 10
    p = params(pattern, [1.0, 1.0])
    rgb = passes(mosaic(pattern, w, h, [20000] * 3), w, h, p)
    v = np.float32(p.one_over_range) * np.float32(20000 - 256)
    assert (rgb == v).all()


@pytest.mark.parametrize("pattern", range(4))
def test_colour_field_lands_on_the_right_channels(pattern):
    w, h = 16, 14
    p = params(pattern, [2.0, 1.5])
    raw = [12000, 30000, 16000]
    rgb = passes(mosaic(pattern, w, h, raw), w, h, p).reshape(-1, 3)
    norm = [np.float32(p.one_over_range) * np.float32(v - 256) for v in raw]
    want = np.array([norm[0] * 2.0, norm[1], norm[2] * 1.5], np.float32)
    assert np.abs(rgb - want).max() < 1e-6


def test_red_and_blue_clip_where_white_balance_reaches_one():
    w, h = 8, 8
    p = params(0, [2.0, 4.0])
    rgb = passes(mosaic(0, w, h, [60000, 1000, 60000]), w, h, p).reshape(-1, 3)
    assert (rgb[:, 0] == 1.0).all() and (rgb[:, 2] == 1.0).all()


def test_noise_parameters_follow_the_quantizer():
    rng = S.Rng(3)
    spec = S.spec(32, 16, 0, valid_tiles=1)
    h = O.parse_header(S.random_frame(spec, rng))
    p = O.params(h, 64, 2.0, 0.0)
    rng2 = 65023.0
    base = 5.0 * 64 * 64 + 64.0 / 3.0
    assert p.noise[0] == np.float32(base / 150994944.0)
    assert p.noise[1] == np.float32(base / (rng2 * rng2 * 256.0))
    assert p.noise[2] == np.float32(64.0 / 3.0 / (rng2 * rng2))
    assert list(p.clip) == [0.5, 1.0] and list(p.wb) == [2.0, 0.0]


def test_output_matrix_combines_gamut_matrix_and_gain():
    xyz = [1.716651187971268, -0.355670783776393, -0.25336628137366,
           -0.666684351832489, 1.616481236634939, 0.015768545813911,
           0.017639857445311, -0.042770613257809, 0.942103121235474]
    rng = S.Rng(4)
    h = O.parse_header(S.random_frame(S.spec(32, 16, 0, gain=2.0), rng))
    refl, norm = O.output_matrix(h, False), O.output_matrix(h, True)
    for i in range(9):
        assert refl[i] == np.float32(xyz[i] * 2.0)
        assert norm[i] == np.float32(xyz[i] * 2.0 * 10.0 / 9.0)


def test_quantizer_counts_every_tile_qscale_byte():
    rng = S.Rng(6)
    spec = S.spec(64, 32, 0, valid_tiles=1, qmat=[3] * 64)
    pk = bytearray(S.random_frame(spec, rng))
    h, offs, sizes, mq = O.parse_frame(bytes(pk))
    assert mq == max(pk[int(o) + 1] for o in offs)
    start = int(offs[0])
    pk[start + 1] = 255
    pk[start + 2:start + 4] = b"\0\0"
    h, s, sk, mq = O.decode_frame(bytes(pk))
    assert sk == [0] and mq == 255 and h.quant_matrix[0] * mq == 3 * 255
    assert O.decode_frame_reduced(bytes(pk), 1)[3] == 255


def develop(h, s, prm, stage):
    return np.stack(O.develop_planar(h, s, prm, stage), 1).ravel()


def test_develop_stages_compose():
    rng = S.Rng(9)
    spec = S.spec(48, 32, 0, valid_tiles=1, wb=(2.0, 1.5))
    h, s, sk, mq = O.decode_frame(S.random_frame(spec, rng))
    q = h.quant_matrix[0] * mq
    prm = O.params(h, q)
    dem = develop(h, s, prm, 0)
    plain = O.params(h, q, 1.0, 1.0)
    np.testing.assert_array_equal(dem, passes(O.linearize(h, s), 48, 32, plain))
    bal = develop(h, s, prm, 1)
    assert not np.array_equal(dem, bal)
    rec = develop(h, s, prm, 2)
    m = O.output_matrix(h, True)
    want = bal.copy()
    O.lib().orc_apply_matrix(want.ctypes.data, want.size // 3, m.ctypes.data)
    np.testing.assert_array_equal(rec, want)


def test_planar_and_interleaved_outputs_agree_across_bands():
    rng = S.Rng(10)
    spec = S.spec(80, 80, 1, valid_tiles=1, wb=(1.8, 1.3))
    h, s, sk, mq = O.decode_frame(S.random_frame(spec, rng))
    prm = O.params(h, h.quant_matrix[0] * mq)
    assert h.height % 32 != 0
    for stage in (0, 1, 2):
        inter = O.develop_interleaved(h, s, prm, stage)
        np.testing.assert_array_equal(inter, develop(h, s, prm, stage))
        # Band-parallel and single-threaded runs agree.
        np.testing.assert_array_equal(inter, O.develop_interleaved(h, s, prm, stage, threads=1))


def test_develop_runs_on_a_decoded_frame():
    rng = S.Rng(5)
    h, s, sk, mq = O.decode_frame(S.random_frame(S.spec(48, 32, 3, valid_tiles=1), rng))
    assert h.quant_matrix[0] * mq > 0
    img = develop(h, s, O.params(h, h.quant_matrix[0] * mq), 1)
    assert img.size == 48 * 32 * 3 and np.isfinite(img).all() and (img <= 1.0).all()


def test_develop_matches_a_float64_restatement_on_a_smooth_field():
    # A smooth colour ramp: the f32 chain agrees with the same formulas in
    # float64 to 1e-5 (no branch sits near its threshold on such a field).
    w, h, pattern = 24, 18, 0
    p = params(pattern, [1.0, 1.0])
    ys, xs = np.mgrid[0:h, 0:w]
    rgb_field = [8000 + 300 * xs + 100 * ys, 14000 + 200 * xs + 250 * ys, 9000 + 150 * xs + 50 * ys]
    lin = np.zeros((h, w), np.uint16)
    for y in range(h):
        for x in range(w):
            lin[y, x] = rgb_field[color_at(pattern, y, x)][y, x]
    rgb = passes(lin.ravel(), w, h, p).reshape(h, w, 3)
    norm = [(f - 256) / RANGE for f in rgb_field]
    # Interior pixels of a linear field are reproduced to within 1e-3 of
    # the true colour (bilinear-exact field, ratio-based chroma).
    inner = rgb[3:-3, 3:-3]
    for c in range(3):
        assert np.abs(inner[..., c] - norm[c][3:-3, 3:-3]).max() < 2e-3


# reduced.rs ---------------------------------------------------------------

def test_half_resolution_matches_full_decode():
    rng = S.Rng(7)
    for spec in S.specs(rng):
        spec.valid_tiles = 1
        pk = S.random_frame(spec, rng)
        h, s, sk, mq = O.decode_frame(pk)
        lin = O.linearize(h, s).reshape(h.height, h.width)
        rr, rc = [(0, 0), (0, 1), (1, 0), (1, 1)][h.bayer]
        g, b, r = (1 - rr, rc), (1 - rr, 1 - rc), (rr, rc)
        _, rgb, sk2, _ = O.decode_frame_reduced(pk, 1)
        cw = (h.width + 15) & ~15
        img = rgb.reshape(-1, cw // 2, 3)[:h.height // 2, :h.width // 2]
        at = lambda ph: lin[ph[0]::2, ph[1]::2][:h.height // 2, :h.width // 2]
        np.testing.assert_array_equal(img[..., 0], at(r))
        np.testing.assert_array_equal(img[..., 1], at(g))
        np.testing.assert_array_equal(img[..., 2], at(b))


def test_half_resolution_can_average_the_greens():
    rng = S.Rng(12)
    for spec in S.specs(rng):
        spec.valid_tiles = 1
        pk = S.random_frame(spec, rng)
        h, s, sk, mq = O.decode_frame(pk)
        lin = O.linearize(h, s).reshape(h.height, h.width).astype(np.uint32)
        rr, rc = [(0, 0), (0, 1), (1, 0), (1, 1)][h.bayer]
        g1, g2 = (1 - rr, rc), (rr, 1 - rc)
        _, rgb, _, _ = O.decode_frame_reduced(pk, 1, average=True)
        cw = (h.width + 15) & ~15
        img = rgb.reshape(-1, cw // 2, 3)[:h.height // 2, :h.width // 2]
        at = lambda ph: lin[ph[0]::2, ph[1]::2][:h.height // 2, :h.width // 2]
        np.testing.assert_array_equal(img[..., 1], (at(g1) + at(g2) + 1) // 2)


def test_reduced_sizes_and_skipped_tiles():
    rng = S.Rng(8)
    for spec in S.specs(rng):
        pk = S.random_frame(spec, rng)
        h, s, sk, mq = O.decode_frame(pk)
        cw, ch = O.coded(h)
        for shift in (1, 2, 3):
            for avg in (False, True):
                _, rgb, sk2, mq2 = O.decode_frame_reduced(pk, shift, avg)
                assert rgb.size == 3 * (cw >> shift) * (ch >> shift)
                assert sk2 == sk and mq2 == mq
