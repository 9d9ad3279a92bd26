"""The oracle's lossless JPEG tiles (oracle/lj92_oracle.c) pinned by the
crate's own tests (examples/prr2dng/lj92.rs:346-399: round trips of random
and extreme samples, a flat tile, a skewed histogram), by an independent
T.81 decoder (tests/lj92_decoder.py), by Huffman optimality against a
textbook heap construction, and by the CFA-phase rule of the tile's edge
padding (dng.rs:115-128)."""
import heapq

import numpy as np
import pytest

from oracle import oracle as O

from .lj92_decoder import decode


def round_trip(samples, width, height):
    data = O.lj92_encode(np.array(samples, dtype=np.uint16).reshape(height, width), 2)
    out, w, h, c = decode(data)
    assert (w, h, c) == (width, height, 2)
    assert out == list(samples)
    return data


def xorshift(state):
    def next_():
        nonlocal state
        m = (1 << 64) - 1
        state ^= (state << 13) & m
        state ^= state >> 7
        state ^= (state << 17) & m
        return state
    return next_


def test_random_and_extreme_samples_survive():
    nxt = xorshift(0x1234_5678_9abc_def0)
    round_trip([nxt() & 0xFFFF for _ in range(64 * 16)], 64, 16)
    round_trip([0 if (i // 2) % 2 == 0 else 0x8000 for i in range(32 * 4)], 32, 4)
    round_trip([0xFFFF] * 32, 16, 2)


def test_a_flat_tile_still_gets_a_valid_table():
    data = round_trip([1000] * 256, 16, 16)
    assert len(data) < 16 * 16 // 8 + 64


def test_skewed_histograms_need_long_codes():
    counts, values, code, length = O.lj92_table([1 << (2 * i) for i in range(17)])
    assert all(1 <= int(l) <= 16 for l in length)
    kraft = sum(0.5 ** int(l) for l in length)
    assert kraft < 1.0
    assert not any(int(c) == (1 << int(l)) - 1 for c, l in zip(code, length))


def heap_huffman_cost(freqs):
    """Total bits of an optimal prefix code (textbook Huffman)."""
    heap = [f for f in freqs if f > 0]
    heapq.heapify(heap)
    cost = 0
    while len(heap) > 1:
        a, b = heapq.heappop(heap), heapq.heappop(heap)
        cost += a + b
        heapq.heappush(heap, a + b)
    return cost


def test_tables_are_optimal_canonical_codes():
    rng = np.random.default_rng(7)
    for trial in range(300):
        hist = [int(x) for x in rng.integers(0, 5000, 17)]
        for z in rng.integers(0, 17, rng.integers(0, 12)):
            hist[z] = 0
        if trial % 7 == 0:
            hist = [0] * 17
            hist[int(rng.integers(0, 17))] = int(rng.integers(1, 100000))
        counts, values, code, length = O.lj92_table(hist)
        used = [v for v in range(17) if length[v]]
        assert sorted(used) == sorted(int(v) for v in values)
        assert sorted(used) == [v for v in range(17) if hist[v]]
        assert sum(int(c) for c in counts) == len(values)
        # Canonical: codes follow BITS / HUFFVAL order, prefix-free.
        words = sorted((int(length[v]), int(code[v])) for v in used)
        for (l1, c1), (l2, c2) in zip(words, words[1:]):
            assert (c1 + 1) << (l2 - l1) <= c2
        # The used codes leave exactly one more code, of length L (the
        # reserved all-ones code, the longest); when L < 16 no length was
        # limited and the code with the reserved symbol is an optimal
        # Huffman code.
        lmax = max(int(length[v]) for v in used)
        kraft = sum(2.0 ** -int(length[v]) for v in used)
        L = int(round(-np.log2(1.0 - kraft)))
        assert 2.0 ** -L == pytest.approx(1.0 - kraft, abs=0, rel=1e-12)
        assert lmax <= L <= 16
        if L < 16:
            cost = sum(hist[v] * int(length[v]) for v in used)
            assert cost + L == heap_huffman_cost(hist + [1])


def test_a_transposed_prediction_would_fail_the_round_trip():
    """The decoder reads the left neighbour two sensels back: a stream
    coded from the transposed tile decodes to the transposed samples only,
    so a transposed encoder would be caught."""
    rng = np.random.default_rng(3)
    a = rng.integers(0, 4096, (16, 16), dtype=np.uint16)
    data = O.lj92_encode(a.T, 2)
    out, *_ = decode(data)
    assert out == a.T.reshape(-1).tolist() and out != a.reshape(-1).tolist()


def test_tile_edges_repeat_the_same_cfa_phase():
    rng = np.random.default_rng(11)
    h, w = 270, 300
    v = rng.integers(0, 65536, (h, w), dtype=np.uint16)
    t = O.dng_tile(v, 256, 256)
    assert np.array_equal(t[:h - 256, :w - 256], v[256:, 256:])
    for y in range(256):
        for x in (w - 256, w - 255, 100, 255):
            sy, sx = 256 + y, 256 + x
            ry = sy if sy < h else (h - 2 if (sy - h) % 2 == 0 else h - 1)
            rx = sx if sx < w else (w - 2 if (sx - w) % 2 == 0 else w - 1)
            # Same CFA phase as the sensel it stands for, and on the last two rows/columns.
            assert (sy - ry) % 2 == 0 and (sx - rx) % 2 == 0
            assert t[y, x] == v[ry, rx]


def test_frame_tiles_decode_to_their_tiles():
    rng = np.random.default_rng(5)
    h, w = 300, 520
    base = rng.integers(200, 3800, (h, w)).astype(np.int64)
    v = (base + rng.integers(-30, 30, (h, w))).clip(0, 65535).astype(np.uint16)
    data, offsets, lengths = O.dng_lj92_tiles(v, threads=4)
    assert len(offsets) == 3 * 2
    assert int(offsets[0]) == 0
    assert all(int(offsets[i]) + int(lengths[i]) == int(offsets[i + 1]) for i in range(5))
    assert int(offsets[-1]) + int(lengths[-1]) == len(data)
    for t in (0, 2, 5):
        stream = data[int(offsets[t]):int(offsets[t]) + int(lengths[t])]
        out, *_ = decode(stream)
        want = O.dng_tile(v, (t % 3) * 256, (t // 3) * 256)
        assert out == want.reshape(-1).tolist()


def test_padding_repeats_the_edge_of_the_same_colour():
    """dng.rs:241-252: a 4x2 picture whose value is its column plus 10 x its
    row; past the right edge the last two columns alternate, past the
    bottom the last two rows."""
    values = np.array([[i % 4 + 10 * (i // 4) for i in range(8)]], dtype=np.uint16).reshape(2, 4)
    tile = O.dng_tile(values, 0, 0)
    assert tile[0, :8].tolist() == [0, 1, 2, 3, 2, 3, 2, 3]
    assert tile[1, :4].tolist() == [10, 11, 12, 13]
    assert tile[2, 0] == 0
    assert tile[3, 0] == 10
