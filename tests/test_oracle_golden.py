"""The oracle's transforms pinned by the crate's golden vectors, its
2^20-block hash and closed forms (idct.rs:462-595)."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def block(fn, c, q):
    c = np.ascontiguousarray(c, np.int16)
    q = np.ascontiguousarray(q, np.int16)
    out = np.zeros(64, np.uint16)
    fn(c.ctypes.data, q.ctypes.data, out.ctypes.data)
    return out


def apple(c, q=None):
    return block(O.lib().orc_apple_block, c, np.ones(64, np.int16) if q is None else q)


def ffmpeg(c, q=None):
    return block(O.lib().orc_ffmpeg_block, c, np.ones(64, np.int16) if q is None else q)


def test_apple_matches_reference_vectors():
    data = np.fromfile(os.path.join(GOLDEN, "idct_vectors.bin"), dtype="<u2")
    assert data.size % 192 == 0
    recs = data.reshape(-1, 192)
    assert len(recs) == 512
    assert sum(1 for r in recs if (r[64:128].astype(np.int16) <= 0).any()) == 85
    for i, r in enumerate(recs):
        got = apple(r[:64].astype(np.int16), r[64:128].astype(np.int16))
        np.testing.assert_array_equal(got, r[128:], err_msg=f"record {i}")


def test_apple_matches_reference_hash():
    want = int(open(os.path.join(GOLDEN, "idct_hash.txt")).read().strip(), 16)
    assert O.lib().orc_testgen_hash(0x4a54, 1 << 20, 0) == want


def test_testgen_hash_detects_a_changed_block():
    # The hash pins the generator and transform together: one block
    # fewer or a different seed changes it.
    want = int(open(os.path.join(GOLDEN, "idct_hash.txt")).read().strip(), 16)
    assert O.lib().orc_testgen_hash(0x4a55, 1 << 10, 0) != want
    assert O.lib().orc_testgen_hash(0x4a54, (1 << 10) - 1, 0) != O.lib().orc_testgen_hash(0x4a54, 1 << 10, 0)


@pytest.mark.parametrize("fn", [apple, ffmpeg])
def test_dc_only_block_is_flat(fn):
    for dc in [-32768, -5000, -1, 0, 1, 800, 32767]:
        b = np.zeros(64, np.int16)
        b[0] = dc
        assert (fn(b).astype(np.int64) == dc + 32768).all(), dc


def float_reference(c):
    """Orthonormal 2-D IDCT scaled like the integer ones (idct.rs:469-493)."""
    out = np.zeros((8, 8))
    for y in range(8):
        for x in range(8):
            s = 0.0
            for v in range(8):
                for u in range(8):
                    cu = 1 / math.sqrt(2) if u == 0 else 1.0
                    cv = 1 / math.sqrt(2) if v == 0 else 1.0
                    s += cu * cv * c[8 * v + u] * math.cos((2 * x + 1) * u * math.pi / 16) * \
                        math.cos((2 * y + 1) * v * math.pi / 16)
            out[y, x] = s * 2 + 32768
    return out.ravel()


def random_blocks(n):
    seed = 12345
    out = []
    for _ in range(n):
        b = []
        for _ in range(64):
            seed = (seed * 1664525 + 1013904223) & 0xFFFFFFFF
            b.append(((seed >> 16) % 2001) - 1000)
        out.append(np.array(b, np.int16))
    return out


def test_both_transforms_match_the_float_reference_closely():
    for b in random_blocks(50):
        want = float_reference(b.astype(np.float64))
        assert np.abs(apple(b).astype(np.float64) - want).max() <= 3.0
        assert np.abs(ffmpeg(b).astype(np.float64) - want).max() <= 2.0


def test_a_dropped_term_fails_the_float_bound():
    # The float bound is tight enough to catch a transposed or missing
    # coefficient: transposing the block moves outputs far past it.
    b = random_blocks(1)[0]
    want = float_reference(b.astype(np.float64))
    t = b.reshape(8, 8).T.ravel().copy()
    assert np.abs(apple(t).astype(np.float64) - want).max() > 3.0
    assert np.abs(ffmpeg(t).astype(np.float64) - want).max() > 2.0


def test_dequantizes_with_saturation():
    b = np.zeros(64, np.int16)
    b[0] = 3000
    q = np.ones(64, np.int16)
    q[0] = 20
    sat = np.zeros(64, np.int16)
    sat[0] = 32767
    np.testing.assert_array_equal(apple(b, q), apple(sat))
    np.testing.assert_array_equal(ffmpeg(b, q), ffmpeg(sat))
    assert O.lib().orc_dequantize(-32768, -1) == 32767
    assert O.lib().orc_dequantize(-300, 300) == -32768


def test_code_values_are_centered_and_clamped():
    L = O.lib()
    assert L.orc_code_value(0) == 32768
    assert L.orc_code_value(-32768) == 0
    assert L.orc_code_value(-40000) == 0
    assert L.orc_code_value(32767) == 65535
    assert L.orc_code_value(40000) == 65535
    assert L.orc_output_code(-0.9) == 0
    assert L.orc_output_code(32768.5) == 32768
    assert L.orc_output_code(70000.0) == 65535
    assert L.orc_output_code(float("nan")) == 65535


def reduced_block(fn, c, q, n):
    c = np.ascontiguousarray(c, np.int16)
    q = np.ascontiguousarray(q, np.int16)
    out = np.zeros(n, np.uint16)
    fn(c.ctypes.data, q.ctypes.data, out.ctypes.data)
    return out


def test_reduced_transforms_dc_only_are_flat():
    L = O.lib()
    q = np.ones(64, np.int16)
    for dc in [-32768, -777, 0, 5, 32767]:
        c4 = np.zeros(16, np.int16)
        c4[0] = dc
        assert (reduced_block(L.orc_apple_block_4x4, c4, q, 16).astype(np.int64) == dc + 32768).all()
        c2 = np.zeros(4, np.int16)
        c2[0] = dc
        assert (reduced_block(L.orc_apple_block_2x2, c2, q, 4).astype(np.int64) == dc + 32768).all()


def float_reference_4x4(d):
    """The textbook 4-point 2-D IDCT of dequantized corner coefficients,
    scaled like the 8-point one (a DC-only block decodes to its DC) and
    clamped to the code range."""
    out = np.zeros((4, 4))
    for y in range(4):
        for x in range(4):
            s = 0.0
            for v in range(4):
                for u in range(4):
                    cu = 1 / math.sqrt(2) if u == 0 else 1.0
                    cv = 1 / math.sqrt(2) if v == 0 else 1.0
                    s += cu * cv * d[4 * v + u] * math.cos((2 * x + 1) * u * math.pi / 8) * \
                        math.cos((2 * y + 1) * v * math.pi / 8)
            out[y, x] = 2 * s + 32768
    return np.clip(out.ravel(), 0, 65535)


def test_quarter_transform_matches_the_float_4_point_idct():
    L = O.lib()
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(300):
        c4 = rng.integers(-1000, 1001, 16).astype(np.int16)
        q = rng.integers(1, 8, 64).astype(np.int16)
        d = [int(c4[k]) * int(q[8 * (k // 4) + k % 4]) for k in range(16)]
        got = reduced_block(L.orc_apple_block_4x4, c4, q, 16).astype(np.float64)
        worst = max(worst, float(np.abs(got - float_reference_4x4(d)).max()))
    assert worst <= 2.0, worst
    # A transposed corner is far outside the bound.
    c4 = rng.integers(-1000, 1001, 16).astype(np.int16)
    t = c4.reshape(4, 4).T.ravel().copy()
    got = reduced_block(L.orc_apple_block_4x4, t, np.ones(64, np.int16), 16).astype(np.float64)
    assert np.abs(got - float_reference_4x4(c4.astype(np.float64))).max() > 100


def test_eighth_transform_is_sums_and_differences():
    L = O.lib()
    q = np.arange(1, 65, dtype=np.int16)
    c = np.array([3, -2, 5, 7], np.int16)
    d = [3 * 1, -2 * 2, 5 * 9, 7 * 10]
    want = [d[0] + d[1] + d[2] + d[3], d[0] - d[1] + d[2] - d[3],
            d[0] + d[1] - d[2] - d[3], d[0] - d[1] - d[2] + d[3]]
    got = reduced_block(L.orc_apple_block_2x2, c, q, 4).astype(np.int64) - 32768
    assert list(got) == want


def test_apple_and_ffmpeg_differ_by_a_few_codes():
    # Both approximate the same transform: on dense unsaturated blocks they
    # differ by a few code values (bitstream.md section 7: +-1-3).
    worst = 0
    for b in random_blocks(200):
        d = apple(b).astype(np.int64) - ffmpeg(b).astype(np.int64)
        worst = max(worst, int(np.abs(d).max()))
    assert 0 < worst <= 3, worst
