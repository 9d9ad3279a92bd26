"""The device decoder's per-stream symbol budget is never reached where the
crate's coefficient loop (coefficients.rs:37-146) ends.

The device loop counts AC symbols from `blocks` and stops a stream at
budget(bytes, codes) = 16 (8 bytes + codes) + 2048 (then the frame fails
with PRR_E_DECODE_OVERRUN). The bound it must stay above, derived in the
design notes from the reader's rules (bits.rs:91-105, 116-134):

- while unread words remain, a consume either refills (loads a word) or
  lowers `extra` by its length, at least 1, and `extra` only grows by 32 a
  refill: at most 8 bytes + bytes/4 + 96 symbols;
- once the stream is read, the window only shifts; a consume moves it
  unless its length is exactly 64 (codewords are at most 69 bits), which
  after the first cycle only a level or a zero run can have, so each later
  run symbol shifts the 64-bit window and at most 65 cycles remain; a
  cycle's run of levels stops when n reaches the coefficient count, and a
  zero run can lower n by at most 14 (golomb values of the RN codebooks
  wrap by at most 2^4), so those cycles hold at most codes + 65 * 17
  symbols;

so AC symbols <= 8.25 bytes + codes + 1250, below the budget for every
stream (bytes >= 1, codes >= 4). The oracle's instrumented copy of the loop
(orc_count_symbols) counts the symbols of adversarial streams: random
bytes, sparse bits (long leading-zero views: over-long and 64-bit
codewords, wrapped zero runs), all four alignments, 1 to 16 blocks and
every coefficient count."""
import numpy as np

from oracle import oracle as O


def derived_bound(nbytes, codes):
    return 8.25 * nbytes + codes + 1250


def device_budget(nbytes, codes, blocks):
    # The device counter starts at `blocks`.
    return 16 * (8 * nbytes + codes) + 2048 - blocks


def test_the_derived_bound_is_below_the_device_budget():
    for log2_blocks in range(5):
        for log2_coefs in (2, 4, 6):
            codes = 1 << (log2_blocks + log2_coefs)
            for nbytes in list(range(1, 64)) + [100, 1000, 65535]:
                assert derived_bound(nbytes, codes) < device_budget(nbytes, codes,
                                                                    1 << log2_blocks)


def streams(rng):
    # Random bytes of every length class.
    for n in list(range(1, 40)) + [64, 100, 257, 1000, 4000]:
        yield "random", rng.integers(0, 256, n, dtype=np.uint8).tobytes()
    # Sparse bits: views with 20-31 leading zeros give codewords of 33 to 69
    # bits (over-long ones wrap the window's shifts).
    for n in (4, 8, 16, 33, 64, 200, 1000):
        for density in (0.002, 0.01, 0.03, 0.1):
            bits = (rng.random(8 * n) < density).astype(np.uint8)
            yield f"sparse{density}", np.packbits(bits).tobytes()
    # Single set bits at every offset of a short stream: every leading-zero
    # count, including the 29-31 that make 64-bit codewords.
    for n in (4, 5, 8, 12):
        for k in range(8 * n):
            b = bytearray(n)
            b[k // 8] = 0x80 >> (k % 8)
            yield "onebit", bytes(b)
    # Repeated 0x00000001 words and near-empty tails.
    for reps in (1, 2, 7, 50):
        yield "words", b"\x00\x00\x00\x01" * reps
        yield "words_ff", (b"\x00\x00\x00\x01" * reps) + b"\xff"


def test_symbol_counts_stay_below_the_bound():
    rng = np.random.default_rng(11)
    worst = 0.0
    worst_what = None
    total = 0
    long_views = 0
    for kind, data in streams(rng):
        for align in range(4):
            for log2_blocks in range(5):
                for log2_coefs in (6, 4, 2):
                    codes = 1 << (log2_blocks + log2_coefs)
                    n, bits, capped = O.count_symbols(data, align, log2_blocks, log2_coefs,
                                                      cap=1 << 24)
                    assert not capped, (kind, len(data), align, log2_blocks, log2_coefs)
                    bound = derived_bound(len(data), codes)
                    assert n <= bound, (kind, len(data), align, log2_blocks, log2_coefs, n)
                    assert n < device_budget(len(data), codes, 1 << log2_blocks)
                    total += 1
                    long_views += bits > 8 * len(data) + 64
                    if n / bound > worst:
                        worst, worst_what = n / bound, (kind, len(data), align, log2_blocks,
                                                        log2_coefs, n)
    print(f"{total} streams; largest count / bound {worst:.3f} at {worst_what}; "
          f"{long_views} consumed more bits than the stream holds (over-long codewords)")
    assert long_views > 0  # the adversarial streams did reach the over-long regime


def test_real_frame_streams():
    """Every stream of a seeded 5760x3240 q45 frame (36,540 streams) stays
    far below the budget; the census also gives the longest chain."""
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.workloads import stream_frame
    pk = S.encode_scene(stream_frame(1920, 1080, seed=9))
    nbytes, syms = O.stream_census(pk)
    ok = nbytes > 0
    # symbols include the DC values (blocks per stream, 8 at alignment 3
    # for full tiles)
    assert (syms[ok] < 16 * (8 * nbytes[ok].astype(np.int64) + 512) + 2048).all()
    assert syms.max() > 0
