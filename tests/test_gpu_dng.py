"""DNG raw tiles coded on the device (csrc/prr_dng.hip through
prr_dng_lj92_tiles[_device]) against the oracle's restatement of the
crate's DNG writer and lossless JPEG encoder: every byte, offset and length
equal, on pictures with partial edge tiles, odd sizes, flat, noisy,
0xFF-heavy and length-limited content, and on a decoded frame's code values
in device memory at the coded stride."""
import numpy as np
import pytest

from oracle import oracle as O
from proresraw_rs_amd import dng as D
from proresraw_rs_amd import synth as S
from proresraw_rs_amd.workloads import config1

from .lj92_decoder import decode

pytestmark = pytest.mark.gpu


def check(values, width=None, height=None):
    v = np.ascontiguousarray(values, dtype=np.uint16)
    w = v.shape[1] if width is None else width
    h = v.shape[0] if height is None else height
    print(f"dng tiles {w}x{h} (stride {v.shape[1]})", flush=True)
    data, offsets, lengths = D.lj92_tiles(v, w, h)
    want, w_off, w_len = O.dng_lj92_tiles(v[:h, :w], threads=8)
    assert np.array_equal(offsets, w_off)
    assert np.array_equal(lengths, w_len)
    assert data == want
    return data, offsets, lengths


def skewed_tile(rng):
    """A tile whose difference category counts make the optimal code a
    chain (each count exceeds the sum of all smaller ones but the last, with
    the reserved symbol's 1 among them), 17 deep, so lengths are limited to
    16 (Annex K.3)."""
    seq = [1, 1, 1]
    while len(seq) < 18:
        seq.append(sum(seq[:-1]) + 1)
    counts = seq[1:]
    counts[-1] += 65536 - sum(counts)
    cats = np.concatenate([np.full(n, c) for c, n in enumerate(counts)])
    rng.shuffle(cats)
    mags = np.where(cats == 0, 0, (1 << np.maximum(cats - 1, 0)).astype(np.int64))
    d = np.where(rng.random(65536) < 0.5, mags, -mags).astype(np.int64)
    d[cats == 16] = -32768
    d = d.reshape(256, 256)
    t = np.zeros((256, 256), dtype=np.int64)
    for c in range(256):
        if c >= 2:
            t[:, c] = t[:, c - 2] + d[:, c]
        else:
            t[0, c] = 32768 + d[0, c]
            for r in range(1, 256):
                t[r, c] = t[r - 1, c] + d[r, c]
    return (t & 0xFFFF).astype(np.uint16)


def category_counts(t):
    """Categories of lj92's prediction differences over one tile."""
    x = t.astype(np.int64)
    pred = np.empty_like(x)
    pred[:, 2:] = x[:, :-2]
    pred[1:, :2] = x[:-1, :2]
    pred[0, :2] = 32768
    d = ((x - pred) & 0xFFFF)
    d = np.where(d >= 32768, d - 65536, d)
    cats = np.array([int(abs(int(v))).bit_length() for v in d.reshape(-1)])
    return [int(c) for c in np.bincount(cats, minlength=17)]


def huffman_depth(freqs):
    """Depth of the deepest leaf of a textbook Huffman tree."""
    import heapq
    heap = [(f, 0, i) for i, f in enumerate(freqs) if f > 0]
    heapq.heapify(heap)
    k = len(freqs)
    while len(heap) > 1:
        a, b = heapq.heappop(heap), heapq.heappop(heap)
        heapq.heappush(heap, (a[0] + b[0], max(a[1], b[1]) + 1, k))
        k += 1
    return heap[0][1]


def test_pictures_match_the_oracle(gpu):
    rng = np.random.default_rng(21)
    check(np.full((256, 256), 1000))
    check(rng.integers(0, 65536, (256, 256)))
    check(np.where(rng.random((300, 520)) < 0.5, 0xFFFF, 0))
    smooth = np.add.outer(np.arange(270), np.arange(777)) * 3 % 4096
    check(smooth + rng.integers(0, 12, smooth.shape))
    check(rng.integers(250, 4096, (17, 9)))            # one small tile, odd width
    check(rng.integers(0, 4096, (513, 2)))             # two columns
    # dng.rs:241-252's 4x2 picture: its one tile decodes to the padded tile.
    small = np.array([[i % 4 + 10 * (i // 4) for i in range(8)]], dtype=np.uint16).reshape(2, 4)
    data, _, _ = check(small)
    out, *_ = decode(data)
    assert out == O.dng_tile(small, 0, 0).reshape(-1).tolist()
    t = skewed_tile(rng)
    assert huffman_depth(category_counts(t) + [1]) > 16  # the limit is exercised
    data, offsets, lengths = check(t)
    out, *_ = decode(data)
    assert out == t.reshape(-1).tolist()


def test_stride_and_padding_of_a_frame(gpu):
    rng = np.random.default_rng(4)
    v = rng.integers(0, 4096, (300, 640)).astype(np.uint16)
    check(v, 600, 290)       # rows wider than the picture, partial tiles
    check(v, 599, 299)       # odd sizes


def test_decoded_code_values_on_the_device(gpu):
    import torch
    P = gpu
    packet = S.encode_scene(config1(1000, 520, seed=9))
    h, samples, skipped, mq = O.decode_frame(packet)
    coded_w, coded_h = O.coded(h)
    samples = samples.reshape(coded_h, coded_w)
    want = O.dng_lj92_tiles(samples[:h.height, :h.width], threads=8)
    slot_samples = torch.from_numpy(samples.view(np.int16).reshape(-1)).cuda()
    n = D.tile_count(h.width, h.height)
    out = torch.empty(D.capacity(h.width, h.height), dtype=torch.uint8, device="cuda")
    off = torch.empty(n, dtype=torch.int32, device="cuda")
    ln = torch.empty(n, dtype=torch.int32, device="cuda")
    D.lj92_tiles_device(slot_samples, h.width, h.height, coded_w, out, off, ln)
    torch.cuda.synchronize()
    offsets = off.cpu().numpy().view(np.uint32)
    lengths = ln.cpu().numpy().view(np.uint32)
    total = int(offsets[-1]) + int(lengths[-1])
    assert np.array_equal(offsets, want[1]) and np.array_equal(lengths, want[2])
    assert out[:total].cpu().numpy().tobytes() == want[0]
    # Twice in a row on the same context and buffers: same bytes.
    D.lj92_tiles_device(slot_samples, h.width, h.height, coded_w, out, off, ln)
    torch.cuda.synchronize()
    assert out[:total].cpu().numpy().tobytes() == want[0]
