"""CPU-only checks of the C ABI: the library loads and exports every
symbol include/*.h declares, and its host-side frame structure functions
(no device work) agree with the oracle, including every error variant."""
import ctypes
import os
import re
import struct

import numpy as np
import pytest

import proresraw_rs_amd as P
from oracle import oracle as O
from proresraw_rs_amd import _native as N
from proresraw_rs_amd import synth as S

from .test_oracle_known_answers import body, empty_tile_packet, packet_of

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if fn.endswith(".h"):
            text = open(os.path.join(INCLUDE, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(prr_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert missing == []
    # The Python binding declares exactly the header's functions.
    assert set(N.SIGNATURES) == names


def test_version_string():
    assert b"gfx950" in N.lib().prr_version()


def err_of(fn, *args):
    try:
        fn(*args)
    except P.Error as e:
        return e
    return None


def oracle_err(pk):
    try:
        O.parse_frame(pk)
    except O.OracleError as e:
        return e
    return None


def frame_info_err(pk):
    return err_of(P.parse_frame_info, pk)


def test_header_fields_match_the_oracle():
    rng = S.Rng(21)
    for spec in S.specs(rng):
        pk = S.random_frame(spec, rng)
        h, info = P.parse_frame_info(pk)
        o = O.parse_header(pk)
        assert (h.width, h.height, h.alignment, int(h.bayer_pattern)) == (o.width, o.height, o.alignment, o.bayer)
        assert h.quant_matrix == list(o.quant_matrix) and h.curve == list(o.curve)
        assert h.white_level == o.white_level and h.flags == o.flags
        assert info.nb_tiles == O.lib().orc_tile_count(o.width, o.height, o.alignment)
        assert info.tiles_per_row == O.lib().orc_tiles_per_row(o.width, o.alignment)
        assert P.header_len(pk) == 10 + struct.unpack(">H", pk[8:10])[0] - 2


def bad_packets():
    """One packet per frame-level error the crate reports."""
    out = []
    out.append(bytes(4))                                        # Truncated prefix
    p = bytearray(80)
    p[:4] = struct.pack(">I", 80)
    p[4:8] = b"xxxx"
    out.append(bytes(p))                                        # BadTag
    p[4:8] = b"prrf"
    p[8:10] = struct.pack(">H", 61)
    out.append(bytes(p))                                        # HeaderTooShort
    p[:4] = struct.pack(">I", 81)
    out.append(bytes(p))                                        # FrameSizeMismatch
    p[:4] = struct.pack(">I", 0)
    out.append(bytes(p))                                        # FrameSizeMismatch (0)
    b = bytearray(body(0))
    b[1] = 2
    out.append(packet_of(bytes(b)))                             # UnsupportedVersion
    b = bytearray(body(0))
    b[7] = 65
    out.append(packet_of(bytes(b)))                             # BadDimensions
    out.append(packet_of(body(0b1010)))                         # BadAlignment
    out.append(packet_of(body(1, b"\0" * 10)))                  # Truncated qmat
    q = bytearray([63] * 64)
    q[5] = 64
    out.append(packet_of(body(1, q)))                           # BadQuantMatrix
    rising = b"".join(struct.pack(">H", v) for v in [0, 512, 1024, 2048, 4096, 8192, 16384, 32768])
    out.append(packet_of(body(0x10, rising)))                   # UnexpectedCurve (v0)
    flat = b"".join(struct.pack(">H", v) for v in [0, 512, 512, 2048, 4096, 8192, 16384, 32768])
    b = bytearray(body(0x10, flat))
    b[1] = 1
    out.append(packet_of(bytes(b)))                             # BadCurve
    short = struct.pack(">I", 20) + b"prrf" + struct.pack(">H", 72) + bytes(10)
    out.append(short)                                           # Truncated header
    pk = bytearray(empty_tile_packet())
    out.append(bytes(pk[:83]))                                  # Truncated size table
    return out


def test_frame_level_errors_match_the_oracle():
    for i, pk in enumerate(bad_packets()):
        o = oracle_err(pk)
        e = frame_info_err(pk)
        assert o is not None and e is not None, i
        assert (e.code, e.a, e.b) == (o.code, o.a, o.b), (i, e, o)
        if e.kind == "BadTag":
            assert e.tag == o.tag
        if e.kind == "BadCurve":
            assert e.points == o.points


def test_parse_frame_matches_the_oracle():
    """prr_parse_frame (parse_frame, decoder.rs:76-123, on the host): tile
    offsets and sizes, max_qscale (counting invalid and short tiles) and
    frame_quantizer (demosaic.rs:50-52) equal the oracle's on random frames,
    frames with corrupt tiles, and frames whose size table runs tiles past
    the packet (TileOutOfBounds naming the same tile); every frame-level
    error as parse_frame_info's."""
    rng = S.Rng(33)
    pr = np.random.default_rng(33)
    checked = oob = 0
    for i, spec in enumerate(S.specs(rng)):
        spec.valid_tiles = i % 2
        pk = S.random_frame(spec, rng)
        variants = [pk]
        h, info = P.parse_frame_info(pk)
        if info.nb_tiles:
            # A tile entry grown past the packet, and a packet cut short.
            k = int(pr.integers(0, info.nb_tiles))
            bad = bytearray(pk)
            bad[info.header_end + 2 * k:info.header_end + 2 * k + 2] = b"\xff\xff"
            variants.append(bytes(bad))
            variants.append(pk[:max(info.header_end + 2 * info.nb_tiles,
                                    len(pk) - int(pr.integers(1, 200)))])
        for v in variants:
            try:
                oh, offs, sizes, mq = O.parse_frame(v)
                want = None
            except O.OracleError as e:
                want = e
            try:
                got = P.parse_frame(v)
                err = None
            except P.Error as e:
                err = e
            if want is not None:
                assert err is not None and (err.code, err.a, err.b) == (want.code, want.a, want.b)
                oob += want.code == 11
                continue
            assert err is None, err
            assert np.array_equal(got.tile_offsets, offs) and np.array_equal(got.tile_sizes, sizes)
            assert got.max_qscale == mq
            assert P.frame_quantizer(got) == oh.quant_matrix[0] * mq
            tiles = got.tiles(v)
            assert all(len(t) == n for t, n in zip(tiles, got.tile_sizes))
            checked += 1
    print(f"{checked} frames parsed, {oob} TileOutOfBounds")
    assert checked >= 15 and oob >= 5
    for pk in bad_packets():
        o = oracle_err(pk)
        e = err_of(P.parse_frame, pk)
        assert (e.code, e.a, e.b) == (o.code, o.a, o.b)
    # A short tile table capacity is refused with the count to allocate.
    pk = S.encode_scene(S.scene(64, 32, 0, seed=2, version=0, alignment=1, threads=1))
    out = N.PrrParsedFrame()
    offs = np.zeros(2, np.uint64)
    sizes = np.zeros(2, np.uint32)
    err = N.PrrError()
    buf = np.frombuffer(pk, np.uint8)
    rc = N.lib().prr_parse_frame(buf.ctypes.data, buf.size, ctypes.byref(out), offs.ctypes.data,
                                 sizes.ctypes.data, 2, ctypes.byref(err))
    assert rc == 100 and err.a == out.info.nb_tiles > 2


def test_error_payloads_carry_the_crate_fields():
    e = frame_info_err(bad_packets()[0])
    assert e.kind == "Truncated" and (e.needed, e.available) == (10, 4) and e.what == "frame prefix"
    e = frame_info_err(bad_packets()[3])
    assert e.kind == "FrameSizeMismatch" and (e.declared, e.actual) == (81, 80)
    e = frame_info_err(bad_packets()[9])
    assert e.kind == "BadQuantMatrix" and (e.index, e.value) == (5, 64)


def test_demosaic_params_and_matrix_match_the_oracle():
    rng = S.Rng(23)
    for spec in S.specs(rng)[:8]:
        spec.white_level = 4095 + 256 * (spec.pattern % 2)
        spec.gain = 1.5
        pk = S.random_frame(spec, rng)
        h, _ = P.parse_frame_info(pk)
        o = O.parse_header(pk)
        for q, wb in [(15, (2.0, 1.6)), (45 * 63, (0.0, 3.0)), (0, (1.0, 1.0))]:
            p = P.DemosaicParams.with_white_balance(h, q, *wb)
            po = O.params(o, q, *wb)
            assert (p.cfa_phase, p.black) == (po.cfa_phase, po.black)
            assert np.float32(p.one_over_range) == np.float32(po.one_over_range)
            assert list(np.float32(p.noise)) == list(np.float32(list(po.noise)))
            assert list(np.float32(p.clip)) == list(np.float32(list(po.clip)))
            assert np.float32(p.black_over_range) == np.float32(po.black_over_range)
        for exp in (P.Exposure.Normalized, P.Exposure.Reflection):
            m = np.array(P.output_matrix(h, exp), np.float32)
            assert np.array_equal(m, O.output_matrix(o, exp == P.Exposure.Normalized))


def test_output_size_is_checked_before_any_device_work():
    # A wrong caller buffer is reported like the crate's "output size" panic.
    pk = S.encode_scene(S.scene(32, 32, 0, seed=1, version=0, alignment=1, threads=1))
    h, info = P.parse_frame_info(pk)
    lib = N.lib()
    err = N.PrrError()
    out = np.zeros(10, np.uint16)
    hdr = N.PrrHeader()
    buf = np.frombuffer(pk, np.uint8)
    rc = lib.prr_decode_frame(None, buf.ctypes.data, buf.size, 0, ctypes.byref(hdr),
                              out.ctypes.data, out.size, None, 0, None, None, ctypes.byref(err))
    assert rc == 101  # no context: invalid argument, nothing touched


def test_fourccs():
    assert P.is_prores_raw_fourcc(b"aprn") and P.is_prores_raw_fourcc(b"APRH")
    assert not P.is_prores_raw_fourcc(b"apch")


def test_new_entry_points_reject_bad_arguments_before_device_work():
    """Argument checks of the pipeline, DNG-tile and orientation entry points
    come before any HIP call, so they hold without a device."""
    lib = N.lib()
    err = N.PrrError()
    h = ctypes.c_void_p()
    assert lib.prr_pipeline_create(None, 3, ctypes.byref(h), ctypes.byref(err)) == 101
    assert lib.prr_pipeline_submit(None, None, None, None, 0, 0, 0, 0.0, 0.0, 2, None, None, None,
                                   None, ctypes.byref(err)) == 101
    assert lib.prr_pipeline_sync(None, ctypes.byref(err)) == 101
    assert lib.prr_pipeline_wait(None, None, ctypes.byref(err)) == 101
    lib.prr_pipeline_destroy(None)
    # A submit whose planes, packet or transform are refused queues nothing:
    # every check comes before the pipeline itself is touched (the handle
    # below is never dereferenced).
    fake_pipe = ctypes.c_void_p(8)
    hdr, info = N.PrrHeader(), N.PrrFrameInfo()
    p8, p4 = ctypes.c_void_p(64), ctypes.c_void_p(68)
    for packet, idct, r, g, b, what in ((p8, 0, p4, p8, p8, b"plane alignment"),
                                        (p8, 0, p8, p8, p4, b"plane alignment"),
                                        (ctypes.c_void_p(66), 0, p8, p8, p8, b"packet alignment"),
                                        (p8, 2, p8, p8, p8, b"samples or idct")):
        err = N.PrrError()
        assert lib.prr_pipeline_submit(fake_pipe, ctypes.byref(hdr), ctypes.byref(info), packet,
                                       16, idct, 0, 0.0, 0.0, 2, r, g, b, None,
                                       ctypes.byref(err)) == 101
        assert err.what == what
    # The allocation-free decode checks its arguments before its context.
    assert lib.prr_decode_frame_device_noalloc(None, None, None, None, 0, 0, None, None, None,
                                               None, None, ctypes.byref(err)) == 101
    # A device decode stages each tile's coefficients in its sample buffer,
    # which must be 16-byte aligned (the linear picture 4-byte aligned):
    # refused before the context is used (the handle is never dereferenced).
    for samples, linear in ((ctypes.c_void_p(72), None), (p8, ctypes.c_void_p(66))):
        for fn in (lib.prr_decode_frame_device, lib.prr_decode_frame_device_noalloc):
            err = N.PrrError()
            assert fn(ctypes.c_void_p(8), ctypes.byref(hdr), ctypes.byref(info), p8, 16, 0,
                      samples, linear, None, p8, None, ctypes.byref(err)) == 101
            assert err.what == b"samples or linear alignment"
    # DNG tiles: a picture narrower than two sensels, a stride below the width.
    buf = (ctypes.c_uint16 * 16)()
    out = (ctypes.c_uint8 * 16)()
    offs = (ctypes.c_uint32 * 1)()
    total = ctypes.c_size_t()
    fake_ctx = ctypes.c_void_p(1)  # never dereferenced: the size check comes first
    for w, h_, stride in ((1, 4, 1), (4, 1, 4), (4, 4, 3)):
        rc = lib.prr_dng_lj92_tiles(fake_ctx, buf, w, h_, stride, out, 16, offs, offs, 1,
                                    ctypes.byref(total), ctypes.byref(err))
        assert rc == 101, (w, h_, stride)
    assert lib.prr_dng_lj92_capacity(5760, 3240) == 299 * lib.prr_dng_lj92_capacity(256, 256)
    # Orientation: no context, an orientation outside 1..8.
    crop = (ctypes.c_uint8 * 4)()
    p = ctypes.c_void_p(16)
    assert lib.prr_orient_planes_device(None, p, p, p, 8, 8, 8, crop, 1, p, p, p, 8, None,
                                        ctypes.byref(err)) == 101
    ow, oh = ctypes.c_uint32(), ctypes.c_uint32()
    assert lib.prr_oriented_size(8, 8, crop, 0, ctypes.byref(ow), ctypes.byref(oh)) == 101
    assert lib.prr_oriented_size(8, 8, crop, 9, ctypes.byref(ow), ctypes.byref(oh)) == 101
    assert lib.prr_oriented_size(8, 6, crop, 6, ctypes.byref(ow), ctypes.byref(oh)) == 0
    assert (ow.value, oh.value) == (6, 8)
