"""Pins the oracle's develop (oracle/prr_oracle.c, restated from
src/demosaic.rs:413-720) against a second restatement written from the
crate's WGSL develop (src/gpu/develop.wgsl:95-303) in NumPy float32
(tests/wgsl_develop.py), bit for bit, on random mosaics and on fields built
to drive the chain's hard terms:

- the green pass's curvature ratio (demosaic.rs:512-531) non-zero in both
  directions and of both signs, and zeroed by the noise test;
- the chroma estimate's weight a / (b + a) strictly between 0 and 1 with
  mean_mid != mean_x (demosaic.rs:589-603), and both of its noise
  fallbacks (centre within its noise; a neighbour within its noise);
- all four mirror cases of the green pass at each picture side
  (demosaic.rs:537-557);

each asserted from the restatement's branch counters. The exact fused
multiply-add the restatement uses is itself checked against rational
arithmetic, including the halfway cases where a float64 sum rounded again
to float32 goes wrong."""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from proresraw_rs_amd import synth as S

from . import wgsl_develop as W

F = np.float32


def exact_fma(a, b, c):
    """a * b + c rounded once to float32, by rational arithmetic: the
    nearest float32 to the exact value, ties to even."""
    x = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    r = F(float(x))  # float64 rounding first, then float32: may be off by one
    best = None
    for cand in (np.nextafter(r, F(-np.inf)), r, np.nextafter(r, F(np.inf))):
        d = abs(Fraction(float(cand)) - x)
        if best is None or d < best[0] or (d == best[0] and
                                           int(np.asarray(cand).view(np.uint32)) % 2 == 0):
            best = (d, cand)
    return best[1]


def test_exact_fma_against_rational_arithmetic():
    rng = np.random.default_rng(7)
    a = (rng.standard_normal(4000) * 10.0 ** rng.integers(-6, 6, 4000)).astype(np.float32)
    b = (rng.standard_normal(4000) * 10.0 ** rng.integers(-6, 6, 4000)).astype(np.float32)
    c = (rng.standard_normal(4000) * 10.0 ** rng.integers(-12, 6, 4000)).astype(np.float32)
    got = W.fma32(a, b, c)
    for i in range(0, 4000, 7):
        assert got[i] == exact_fma(a[i], b[i], c[i]), i
    # Halfway cases: a * b = 1 + 2^-11 + 2^-24 lies exactly between two
    # float32 values; a tiny c decides the rounding, which rounding the
    # float64 sum again to float32 gets wrong.
    one = F(1.0) + F(2.0 ** -12)
    for c_val, want in ((2.0 ** -60, 1.0 + 2.0 ** -11 + 2.0 ** -23),
                        (-(2.0 ** -60), 1.0 + 2.0 ** -11),
                        (0.0, 1.0 + 2.0 ** -11)):
        cc = F(c_val)
        got = W.fma32(one, one, cc)
        assert got == F(want) == exact_fma(one, one, cc), c_val
    naive = F(float(one) * float(one) + float(F(2.0 ** -60)))
    assert naive != F(1.0 + 2.0 ** -11 + 2.0 ** -23)  # the case a naive fma misses


def frame_header(w, h, pattern, version=0, curve=None, white_level=4095):
    """An oracle header for a w x h picture (from an encoded packet)."""
    pk = S.encode_scene(S.scene(w, h, pattern, seed=5, version=version, alignment=0,
                                qscale=(15, 15), curve=curve, white_level=white_level,
                                n_discs=2, n_gradients=1))
    return O.parse_frame(pk)[0]


def oracle_planes(h, codes, prm, stage):
    cw, ch = O.coded(h)
    samples = np.zeros(cw * ch, np.uint16)
    samples.reshape(ch, cw)[:h.height, :h.width] = codes
    return [p.reshape(h.height, h.width) for p in O.develop_planar(h, samples, prm, stage,
                                                                  threads=1)]


def params(h, quantizer, wb=(1.7, 1.3), noise=None):
    prm = O.params(h, quantizer, *wb)
    if noise is not None:
        for i in range(3):
            prm.noise[i] = noise[i]
    return prm


def compare(h, codes, prm, stage, cov, what, plant=None):
    want = oracle_planes(h, codes, prm, stage)
    got, _ = W.develop(codes, list(h.curve), W.Params.of(prm, stage, O.output_matrix(h)), cov,
                       plant)
    for k in range(3):
        g, r = np.asarray(got[k], np.float32), np.asarray(want[k], np.float32)
        if not np.array_equal(g.view(np.uint32), r.view(np.uint32)):
            bad = np.argwhere(g.view(np.uint32) != r.view(np.uint32))
            y, x = bad[0]
            raise AssertionError(f"{what} stage {stage} plane {k}: {len(bad)} values differ, "
                                 f"first at ({x}, {y}): wgsl {g[y, x]!r} oracle {r[y, x]!r}")


def random_mosaic(rng, w, h, lo, hi):
    return rng.integers(lo, hi, size=(h, w)).astype(np.uint16)


def smooth_field(w, h, kind, seed):
    """Code values of constructed fields (default curve: code = linear for
    values below 512): a curvature bump, a hue step, a ramp with a cross."""
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    if kind == "bump":
        v = 600 + 1800 * np.exp(-((xs - w / 2) ** 2 + (ys - h / 2) ** 2) / (0.08 * w * h))
    elif kind == "hue_step":
        v = np.where(xs < w / 2, 900.0, 2600.0) * np.where((xs + ys) % 2 == 0, 1.0, 0.55)
    else:
        v = 400 + 40 * xs + 25 * ys + 800 * ((np.abs(xs - w / 3) < 1.5) | (np.abs(ys - h / 2) < 1.5))
    v = v + rng.normal(0, 4, v.shape)
    return np.clip(np.round(v), 0, 4095).astype(np.uint16)


COVERAGE_NEEDED = [
    "ratio_h_pos", "ratio_h_neg", "ratio_v_pos", "ratio_v_neg", "ratio_h_zero_by_noise",
    "ratio_v_zero_by_noise", "green_gate_closed", "estimate_weight_interior",
    "estimate_fallback_centre", "estimate_fallback_neighbour",
    "mirror_h_0", "mirror_h_1", "mirror_h_last", "mirror_h_last2",
    "mirror_v_0", "mirror_v_1", "mirror_v_last", "mirror_v_last2",
]


def test_oracle_develop_equals_the_wgsl_restatement():
    rng = np.random.default_rng(2024)
    cov = W.Coverage()
    cases = 0
    # Random mosaics: every Bayer pattern, odd and even sizes, value ranges
    # from below black to the top of the curve, custom curves (v1), the
    # noise model of several quantizers and one with b = c = 0.
    for i in range(24):
        pattern = i % 4
        w, h = 2 * int(rng.integers(3, 20)), 2 * int(rng.integers(3, 17))
        version = i % 3 == 2
        curve = [0, 300, 700, 1400, 2600, 4100, 6000, 9000] if version else None
        hdr = frame_header(w, h, pattern, version=int(version), curve=curve)
        lo, hi = [(0, 65536), (200, 3000), (0, 600), (250, 40000)][i % 4]
        codes = random_mosaic(rng, w, h, lo, hi)
        q = [15, 45, 135, 1][i % 4]
        noise = (0.0, 0.0, 0.0) if i == 13 else None
        prm = params(hdr, q, noise=noise)
        for stage in (0, 1, 2):
            compare(hdr, codes, prm, stage, cov, f"random {i} {w}x{h} p{pattern}")
            cases += 1
    # Constructed fields (and their transposes for the vertical direction).
    for j, kind in enumerate(("bump", "hue_step", "ramp_cross")):
        for pattern in (0, 3):
            for w, h in ((32, 24), (24, 32)):
                hdr = frame_header(w, h, pattern)
                codes = smooth_field(w, h, kind, j)
                prm = params(hdr, 15)
                for stage in (0, 2):
                    compare(hdr, codes, prm, stage, cov, f"{kind} {w}x{h} p{pattern}")
                    cases += 1
    print(f"{cases} comparisons; branch counts:")
    for k in COVERAGE_NEEDED:
        print(f"  {k}: {cov.counts.get(k, 0)}")
    missing = [k for k in COVERAGE_NEEDED if cov.counts.get(k, 0) == 0]
    assert missing == [], f"branches never taken: {missing}"


def test_green_pass_alone_equals_the_wgsl_restatement():
    """orc_green_pass (green_pass, demosaic.rs:453-464, on linear values)
    against the restatement's pass 1, both outputs of every sensel."""
    rng = np.random.default_rng(99)
    lib = O.lib()
    cov = W.Coverage()
    for i in range(8):
        w, h = 2 * int(rng.integers(3, 15)), 2 * int(rng.integers(3, 15))
        hdr = frame_header(w, h, i % 4)
        lin = rng.integers(0, 5000, size=(h, w)).astype(np.uint16)
        prm = params(hdr, [15, 45][i % 2])
        out = np.zeros(2 * w * h, np.float32)
        lib.orc_green_pass(lin.ctypes.data, w, h, ctypes.byref(prm), out.ctypes.data)
        (rb, gr), _ = W.green_from_linear(lin, W.Params.of(prm, 1), cov)
        out = out.reshape(h, w, 2)
        assert np.array_equal(out[..., 0].view(np.uint32), rb.view(np.uint32))
        assert np.array_equal(out[..., 1].view(np.uint32), gr.view(np.uint32))


@pytest.mark.parametrize("mistake", ["ratio_denominator", "weight_swap"])
def test_a_planted_mistake_in_the_restatement_is_seen(mistake):
    """The comparison is sharp: a factor of 2 dropped from the ratio's
    denominator, or a and b swapped in the weight, makes outputs differ from
    the oracle's."""
    rng = np.random.default_rng(5)
    hdr = frame_header(24, 20, 0)
    # Codes whose linear values sit well above black, so the ratio and the
    # weight are in play (checked on the faithful restatement first).
    codes = random_mosaic(rng, 24, 20, 9000, 30000)
    prm = params(hdr, 15)
    cov = W.Coverage()
    compare(hdr, codes, prm, 1, cov, "faithful")
    assert cov.counts["ratio_h_pos"] + cov.counts["ratio_h_neg"] > 0
    assert cov.counts["estimate_weight_interior"] > 0
    with pytest.raises(AssertionError):
        compare(hdr, codes, prm, 1, W.Coverage(), "planted", plant=mistake)
