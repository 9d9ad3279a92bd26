"""The oracle against the crate's known-answer tests: bits.rs:202-269,
coefficients.rs:153-229, curve.rs:36-70, tile.rs:156-250,
header.rs:333-489, decoder.rs:313-427, synth.rs:228-274 and the wrapped
zero run of gpu.rs:1410-1443."""
import ctypes
import struct

import numpy as np
import pytest

from oracle import oracle as O
from proresraw_rs_amd import synth as S

from .bitwriter import (AC_CB, AC_INITIAL, DC_CB, DC_RESET, FIRST_DC, LN_CB, LN_INITIAL,
                        RN_INITIAL, SCAN, BitWriter, wrapped_zero_run_stream)


def read(data, align, cbs):
    buf = np.frombuffer(bytes(data) or b"\0", np.uint8)
    cb = np.array(cbs, np.uint16)
    out = np.zeros(len(cbs), np.uint32)
    n = O.lib().orc_read_symbols(buf.ctypes.data, len(data), align, cb.ctypes.data, len(cbs),
                                 out.ctypes.data)
    return [int(v) for v in out[:n]]


def codeword(p, cb):
    v, l = ctypes.c_uint32(), ctypes.c_uint32()
    O.lib().orc_codeword(p, cb, ctypes.byref(v), ctypes.byref(l))
    return v.value, l.value


def decode_component(data, align, log2_blocks, log2_coefs=6):
    n = 1 << (log2_blocks + log2_coefs)
    buf = np.frombuffer(bytes(data) or b"\0", np.uint8)
    c = np.full(n, 7, np.int16)
    O.lib().orc_decode_component_reduced(buf.ctypes.data, len(data), align, log2_blocks,
                                         log2_coefs, c.ctypes.data)
    return c


# bits.rs ------------------------------------------------------------------

def test_rice_codewords():
    assert read(bytes([0b1110_1100]), 0, [0x132, 0x132]) == [0b11, 4 + 0b10]


def test_exp_golomb_codewords():
    assert read(bytes([0b0110_1000]), 0, [0x031]) == [0b01101 + 2 - 8]


def test_round_trips_through_the_test_encoder_at_every_alignment():
    cbs = [0x2BC, 0x064, 0x031, 0x000, 0x433]
    values = [0, 1, 2, 3, 7, 8, 15, 16, 100, 1000, 4095, 65535]
    w = BitWriter()
    expected = []
    for cb in cbs:
        for v in values:
            if BitWriter().symbol(cb, v) <= 32:
                w.symbol(cb, v)
                expected.append((cb, v))
    w.put(1, 1)
    data = w.finish()
    for align in range(4):
        got = read(data, align, [cb for cb, _ in expected])
        assert got == [v for _, v in expected], align


def test_ends_on_zero_window_and_past_the_end():
    assert read(bytes([0, 0, 0, 0, 0x80]), 0, [0]) == []
    assert read(bytes([0b1000_0000]), 0, [0, 0]) == [0]
    assert read(b"", 3, [0]) == []


def test_overlong_codeword_shifts_wrap():
    p = 1 << 11
    assert codeword(p, 0) == (1 + (p >> 24) - 1, 40)


# coefficients.rs -----------------------------------------------------------

def test_empty_stream_is_all_zero():
    assert (decode_component(b"", 0, 2) == 0).all()


def test_decodes_dc_values_with_running_sign():
    w = BitWriter()
    w.symbol(FIRST_DC, 5)
    w.symbol(DC_RESET, 3)
    c = decode_component(w.finish(), 0, 1)
    assert c[0] == -3
    assert c[64] == -5


def test_dc_sign_persists_until_a_zero_delta():
    w = BitWriter()
    w.symbol(FIRST_DC, 0)
    w.symbol(DC_RESET, 1)
    w.symbol(DC_CB[1], 2)
    w.symbol(DC_CB[1], 0)
    w.symbol(DC_CB[0], 2)
    c = decode_component(w.finish(), 0, 3)
    assert [int(c[b * 64]) for b in range(5)] == [0, -1, -2, -2, -1]


def test_decodes_ac_runs():
    w = BitWriter()
    w.symbol(FIRST_DC, 0)
    w.symbol(LN_INITIAL, 2)
    w.symbol(AC_INITIAL, 0)
    w.put(0, 1)
    w.symbol(AC_CB[0], 2)
    w.put(1, 1)
    w.symbol(RN_INITIAL, 0)
    w.symbol(AC_CB[2], 1)
    w.put(0, 1)
    c = decode_component(w.finish(), 0, 0)
    assert c[SCAN[1]] == 1
    assert c[SCAN[2]] == -3
    assert c[SCAN[3]] == 0
    assert c[SCAN[4]] == 2
    assert int((c != 0).sum()) == 3


def test_interleaves_blocks_by_frequency():
    w = BitWriter()
    w.symbol(FIRST_DC, 0)
    w.symbol(DC_RESET, 0)
    w.symbol(LN_INITIAL, 2)
    w.symbol(AC_INITIAL, 4)
    w.put(0, 1)
    w.symbol(AC_CB[4], 6)
    w.put(0, 1)
    c = decode_component(w.finish(), 0, 1)
    assert c[SCAN[1]] == 5
    assert c[64 + SCAN[1]] == 7


def test_a_wrapped_zero_run_overwrites():
    c = decode_component(wrapped_zero_run_stream(), 0, 0)
    assert c[SCAN[13]] == 2


def random_levels(rng, log2_blocks, density):
    n = 64 << log2_blocks
    lv = np.zeros(n, np.int16)
    mask = rng.random(n) < density
    mag = np.minimum(rng.geometric(0.15, n), 2000)
    lv[mask] = (mag * rng.choice([-1, 1], n))[mask]
    lv[::64] = rng.integers(-3000, 3000, 1 << log2_blocks)
    return lv


def test_valid_streams_round_trip():
    # Streams the encoder writes (a restatement of the grammar's inverse)
    # decode back to the coefficients, at every alignment: pins the decoder
    # on valid input, independent of the reference's hand-made streams.
    rng = np.random.default_rng(7)
    for trial in range(120):
        log2 = trial % 5
        lv = random_levels(rng, log2, [0.02, 0.2, 0.7][trial % 3])
        data = S.encode_component(lv, log2)
        for align in range(4):
            np.testing.assert_array_equal(decode_component(data, align, log2), lv)


def test_reduced_decodes_keep_the_leading_scan_positions():
    # Quarter and eighth decodes read the same stream and stop after scan
    # position 16 / 4 of every block (coefficients.rs:29-33, codebook.rs:104-110).
    rng = np.random.default_rng(8)
    for trial in range(60):
        log2 = trial % 5
        lv = random_levels(rng, log2, 0.4)
        data = S.encode_component(lv, log2)
        full = lv.reshape(-1, 8, 8)
        q = decode_component(data, 0, log2, 4).reshape(-1, 4, 4)
        e = decode_component(data, 0, log2, 2).reshape(-1, 2, 2)
        np.testing.assert_array_equal(q, full[:, :4, :4])
        np.testing.assert_array_equal(e, full[:, :2, :2])


# curve.rs -------------------------------------------------------------------

def lin(curve, code):
    c = np.array(curve, np.uint16)
    return O.lib().orc_linearize(c.ctypes.data, code)


DEFAULT = [0, 512, 1024, 2048, 4096, 8192, 16384, 32768]


def test_default_curve_hits_control_points():
    for i, p in enumerate(DEFAULT):
        assert lin(DEFAULT, i * 8192) == p
    assert lin(DEFAULT, 65535) == 65532


def test_interpolates_linearly_within_a_segment():
    c = [0, 100, 200, 300, 400, 500, 600, 700]
    assert lin(c, 4096) == 50
    assert lin(c, 8192 + 2048) == 125
    assert lin(c, 57344 + 4096) == (700 + 65536) // 2


def test_rounds_to_nearest():
    c = [0, 1, 2, 3, 4, 5, 6, 7]
    assert lin(c, 4095) == 0
    assert lin(c, 4096) == 1


def test_curve_matches_a_rational_interpolation():
    # Round-to-nearest of the exact interpolation, with the 16-bit wrap of
    # a falling segment (curve.rs:18-22).
    c = [0, 224, 288, 468, 1162, 3339, 9431, 25377]
    for code in range(0, 65536, 97):
        seg, frac = code >> 13, code & 0x1FFF
        c0, c1 = c[seg], (c[seg + 1] if seg < 7 else 65536)
        want = min((c0 * 8192 + ((c1 - c0) & 0xFFFF) * frac + 4096) >> 13, 65535)
        assert lin(c, code) == want


# tile.rs ---------------------------------------------------------------------

def widths(w, h, a):
    n = O.lib().orc_tile_count(w, h, a)
    x, y = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    l2 = np.zeros(n, np.uint8)
    O.lib().orc_tile_geometry(w, h, a, x.ctypes.data, y.ctypes.data, l2.ctypes.data)
    return x, y, l2


def test_full_rows_use_the_largest_tiles():
    assert O.lib().orc_tiles_per_row(5888, 3) == 46
    assert O.lib().orc_tile_count(5888, 3312, 3) == 46 * 207
    x, y, l2 = widths(5888, 3312, 3)
    assert (16 << l2[:46] == 128).all()


def test_remainder_splits_into_halving_tiles():
    assert O.lib().orc_tiles_per_row(4264, 3) == 35
    x, y, l2 = widths(4264, 2408, 3)
    w = list(16 << l2[:35].astype(int))
    assert w[:33] == [128] * 33 and w[33:] == [32, 16]
    assert x[34] == 4256 and l2[34] == 0
    assert y[150 * 35] == 2400


def test_alignment_zero_and_four():
    x, y, l2 = widths(48, 16, 0)
    assert list(16 << l2.astype(int)) == [16, 16, 16]
    assert O.lib().orc_tiles_per_row(4224, 4) == 17
    x, y, l2 = widths(4224, 3024, 4)
    assert list(16 << l2[:17].astype(int)) == [256] * 16 + [128]


def tile_parse(data):
    buf = np.frombuffer(bytes(data) + b"\0", np.uint8)
    qs = ctypes.c_uint8()
    offs, lens = np.zeros(4, np.uint32), np.zeros(4, np.uint32)
    rc = O.lib().orc_tile_parse(buf.ctypes.data, len(data), ctypes.byref(qs), offs.ctypes.data,
                                lens.ctypes.data)
    return rc, qs.value, list(offs), list(lens)


def test_tile_header_splits_streams():
    data = bytes([8 << 3, 15, 0, 2, 0, 1, 0, 3, 10, 11, 20, 30, 31, 32, 40, 41])
    rc, qs, offs, lens = tile_parse(data)
    assert rc == 0 and qs == 15
    assert [data[o:o + l] for o, l in zip(offs, lens)] == [bytes([10, 11]), bytes([20]),
                                                           bytes([30, 31, 32]), bytes([40, 41])]


def test_tile_header_rejects_what_apple_rejects():
    def tile(hdr, qscale, sizes, length):
        d = bytes([hdr << 3, qscale]) + b"".join(struct.pack(">H", s) for s in sizes)
        d = d + b"\xAA" * (length - len(d))
        return tile_parse(d[:length])[0] == 0
    assert tile(8, 1, [1, 1, 1], 12)
    assert not tile(7, 1, [1, 1, 1], 12)
    assert not tile(8, 0, [1, 1, 1], 12)
    assert not tile(8, 1, [0, 1, 1], 12)
    assert not tile(8, 1, [1, 0, 1], 12)
    assert not tile(8, 1, [1, 1, 0], 12)
    assert not tile(8, 1, [1, 1, 1], 11)
    data = bytes([8 << 3, 1, 0, 9, 0, 0, 0, 0, 1, 2])
    assert tile_parse(data)[0] != 0
    assert tile_parse(data[:5])[0] != 0
    assert tile_parse(bytes([16 << 3, 1, 0, 0, 0, 0, 0, 0, 1, 2]))[0] != 0


# header.rs / decoder.rs -----------------------------------------------------

def body(flags, tail=b""):
    b = bytearray(70)
    b[2:6] = b"test"
    b[6:8] = struct.pack(">H", 64)
    b[8:10] = struct.pack(">H", 32)
    b[10:14] = bytes([1, 2, 3, 4])
    b[14:16] = struct.pack(">H", 3)
    b[16:18] = struct.pack(">H", 0xFDFF)
    b[18:22] = struct.pack(">f", 2.0)
    b[22:26] = struct.pack(">f", 1.5)
    for i in range(9):
        b[26 + 4 * i:30 + 4 * i] = struct.pack(">f", float(i))
    b[62:66] = struct.pack(">f", 11.8)
    b[66:68] = struct.pack(">H", 5600)
    b[68:70] = struct.pack(">H", flags)
    return bytes(b) + bytes(tail)


def packet_of(b, size=None):
    p = struct.pack(">I", 0) + b"prrf" + struct.pack(">H", len(b) + 2) + b
    return struct.pack(">I", size if size is not None else len(p)) + p[4:]


def parse_err(pk):
    try:
        O.parse_header(pk)
    except O.OracleError as e:
        return e.code, e.a, e.b
    return None


def test_parses_fixed_fields_and_defaults():
    h = O.parse_header(packet_of(body(0b0110)))
    assert h.version == 0 and bytes(h.vendor) == b"test"
    assert (h.width, h.height) == (64, 32)
    assert list(h.crop) == [1, 2, 3, 4]
    assert h.bayer == 3 and h.white_level == 65279
    assert (h.wb_red, h.wb_blue) == (2.0, 1.5)
    assert h.color_matrix[5] == 5.0
    assert h.gain == np.float32(11.8)
    assert h.wb_cct == 5600 and h.alignment == 3
    assert list(h.quant_matrix) == [1] * 64 and list(h.curve) == DEFAULT


def test_parses_quant_matrix_and_curve():
    pts = [0, 224, 288, 468, 1162, 3339, 9431, 25377]
    tail = bytes([7] * 64) + b"".join(struct.pack(">H", p) for p in pts)
    b = bytearray(body(0b11_1001, tail))
    b[1] = 1
    h = O.parse_header(packet_of(bytes(b)))
    assert h.alignment == 4 and list(h.quant_matrix) == [7] * 64 and list(h.curve) == pts


def test_rejects_bad_input():
    short = struct.pack(">I", 20) + b"prrf" + struct.pack(">H", 72) + bytes(10)
    assert parse_err(short)[:3] == (1, 80, 20)
    b = bytearray(body(0))
    b[1] = 2
    assert parse_err(packet_of(bytes(b))) == (8, 2, 0)
    b = bytearray(body(0))
    b[7] = 65
    assert parse_err(packet_of(bytes(b)))[0] == 9
    assert parse_err(packet_of(body(0b1010))) == (10, 5, 0)
    assert parse_err(packet_of(body(1, b"\0" * 10)))[0] == 1


def test_rejects_what_apple_rejects():
    qmat = bytearray([63] * 64)
    qmat[5] = 64
    assert parse_err(packet_of(body(1, qmat))) == (3, 5, 64)
    qmat[5] = 0
    assert parse_err(packet_of(body(1, qmat))) == (3, 5, 0)

    def curve(points, version):
        b = bytearray(body(0x10, b"".join(struct.pack(">H", p) for p in points)))
        b[1] = version
        return parse_err(packet_of(bytes(b)))
    rising = [0, 512, 1024, 2048, 4096, 8192, 16384, 32768]
    assert curve(rising, 1) is None
    assert curve(rising, 0)[0] == 5
    assert curve([0, 512, 512, 2048, 4096, 8192, 16384, 32768], 1)[0] == 4
    assert curve([0, 512, 1024, 2048, 4096, 8192, 16384, 100], 1)[0] == 4


def test_coded_dimensions_round_up_to_tiles():
    b = bytearray(body(0))
    b[6:8] = struct.pack(">H", 4264)
    b[8:10] = struct.pack(">H", 2408)
    h = O.parse_header(packet_of(bytes(b)))
    assert O.coded(h) == (4272, 2416)


def decode_err(pk):
    try:
        O.decode_frame(pk)
    except O.OracleError as e:
        return e.code, e.a, e.b, e.tag
    return None


def test_rejects_bad_prefixes():
    assert decode_err(bytes(4))[0] == 1
    p = bytearray(80)
    p[:4] = struct.pack(">I", 80)
    p[4:8] = b"xxxx"
    assert decode_err(bytes(p))[0] == 6 and decode_err(bytes(p))[3] == b"xxxx"
    p[4:8] = b"prrf"
    p[8:10] = struct.pack(">H", 61)
    assert decode_err(bytes(p))[:2] == (7, 61)
    p[:4] = struct.pack(">I", 81)
    assert decode_err(bytes(p))[0] == 2


def empty_tile_packet():
    b = bytearray(70)
    b[6:8] = struct.pack(">H", 16)
    b[8:10] = struct.pack(">H", 16)
    p = bytearray(struct.pack(">I", 0) + b"prrf" + struct.pack(">H", 72) + bytes(b))
    p += struct.pack(">H", 12) + bytes([8 << 3, 1, 0, 1, 0, 1, 0, 1, 0, 0, 0, 0])
    p[:4] = struct.pack(">I", len(p))
    return p


def test_empty_tile_decodes_to_mid_grey():
    p = empty_tile_packet()
    h, s, sk, mq = O.decode_frame(bytes(p))
    assert (h.width, h.height) == (16, 16) and sk == []
    assert (s == 32768).all()
    assert (O.linearize(h, s) == 4096).all()
    at = len(p) - 12
    p[at - 2:at] = struct.pack(">H", 8)
    del p[at + 8:]
    p[at + 2:at + 8] = bytes(6)
    p[:4] = struct.pack(">I", len(p))
    h, s, sk, mq = O.decode_frame(bytes(p))
    assert sk == [0] and (s == 0).all()


def test_frame_size_may_be_below_the_packet_length():
    p = bytearray(80)
    p[4:8] = b"prrf"
    p[8:10] = struct.pack(">H", 61)
    for size in (0, 81):
        p[:4] = struct.pack(">I", size)
        assert decode_err(bytes(p))[0] == 2
    p[:4] = struct.pack(">I", 20)
    assert decode_err(bytes(p))[:2] == (7, 61)


def test_missing_tile_is_an_error():
    b = bytearray(70)
    b[6:8] = struct.pack(">H", 16)
    b[8:10] = struct.pack(">H", 16)
    p = bytearray(struct.pack(">I", 0) + b"prrf" + struct.pack(">H", 72) + bytes(b))
    p += struct.pack(">H", 40) + bytes(8)
    p[:4] = struct.pack(">I", len(p))
    assert decode_err(bytes(p))[:2] == (11, 0)


# synth.rs ----------------------------------------------------------------------

def test_random_frames_decode_without_panicking():
    rng = S.Rng(1)
    for spec in S.specs(rng):
        pk = S.random_frame(spec, rng)
        h, s, sk, mq = O.decode_frame(pk)
        assert h.width == spec.width and h.bayer == spec.pattern
        assert int((s == 32768).sum()) < s.size // 2


def test_the_header_carries_the_spec_and_valid_tiles_are_never_skipped():
    rng = S.Rng(2)
    spec = S.spec(48, 32, 3, vendor=b"atm0", crop=[16, 8, 4, 2], white_level=4095 + 256,
                  wb=(1.5, 0.75), color_matrix=[[0.7, 0.2, 0.1], [0.3, 1.0, -0.3], [0.0, -0.4, 1.5]],
                  gain=11.8, wb_cct=5235, valid_tiles=1)
    h, s, sk, mq = O.decode_frame(S.random_frame(spec, rng))
    assert bytes(h.vendor) == b"atm0" and list(h.crop) == [16, 8, 4, 2]
    assert h.white_level == 4095 + 256 and (h.wb_red, h.wb_blue) == (1.5, 0.75)
    assert np.allclose(list(h.color_matrix), [0.7, 0.2, 0.1, 0.3, 1.0, -0.3, 0.0, -0.4, 1.5])
    assert h.gain == np.float32(11.8) and h.wb_cct == 5235 and sk == []


def test_scene_encoder_round_trips_through_the_oracle():
    # The encoder is not a reference: this checks it produces frames whose
    # decode is close to the source mosaic (quantization error only).
    sc = S.scene(160, 96, 0, seed=5, version=0, alignment=1, qscale=(4, 4), threads=2)
    pk = S.encode_scene(sc)
    h, s, sk, mq = O.decode_frame(pk)
    assert sk == [] and mq == 4
    src = S.scene_mosaic(sc).ravel().astype(np.float64)
    err = O.linearize(h, s).astype(np.float64) - src
    assert np.sqrt((err ** 2).mean()) < 3.0


def test_corrupted_tiles_are_exactly_the_chosen_ones():
    sc = S.scene(256, 64, 3, seed=9, version=1, alignment=3, qscale=(45, 45),
                 curve=[0, 100, 300, 700, 1300, 2100, 3100, 4400], corrupt_fraction=0.25,
                 threads=2)
    pk = S.encode_scene(sc)
    h, s, sk, mq = O.decode_frame(pk)
    n = O.lib().orc_tile_count(256, 64, 3)
    assert len(sk) == round(0.25 * n)
