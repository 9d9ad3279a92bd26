"""ctypes wrapper of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker; never by proresraw_rs_amd.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcError(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("what", ctypes.c_char * 28), ("a", ctypes.c_uint64),
                ("b", ctypes.c_uint64), ("points", ctypes.c_uint16 * 8), ("tag", ctypes.c_uint8 * 4)]


class OrcHeader(ctypes.Structure):
    _fields_ = [
        ("version", ctypes.c_uint8), ("vendor", ctypes.c_uint8 * 4), ("bayer", ctypes.c_uint8),
        ("width", ctypes.c_uint16), ("height", ctypes.c_uint16), ("crop", ctypes.c_uint8 * 4),
        ("white_level", ctypes.c_uint16), ("wb_cct", ctypes.c_uint16), ("flags", ctypes.c_uint16),
        ("alignment", ctypes.c_uint8), ("quant_matrix", ctypes.c_uint8 * 64),
        ("curve", ctypes.c_uint16 * 8), ("wb_red", ctypes.c_float), ("wb_blue", ctypes.c_float),
        ("color_matrix", ctypes.c_float * 9), ("gain", ctypes.c_float)]


class OrcParams(ctypes.Structure):
    _fields_ = [("cfa_phase", ctypes.c_int32), ("black", ctypes.c_int32),
                ("one_over_range", ctypes.c_float), ("noise", ctypes.c_float * 3),
                ("clip", ctypes.c_float * 2), ("black_over_range", ctypes.c_float),
                ("wb", ctypes.c_float * 2)]


class OracleError(Exception):
    def __init__(self, e: OrcError):
        self.code, self.a, self.b = e.code, int(e.a), int(e.b)
        self.what = e.what.decode()
        self.points, self.tag = list(e.points), bytes(e.tag)
        super().__init__(f"oracle error {self.code} {self.what} a={self.a} b={self.b}")


vp, sz, i32, u32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64
E = ctypes.POINTER(OrcError)
HP = ctypes.POINTER(OrcHeader)
PP = ctypes.POINTER(OrcParams)

_SIG = {
    "orc_codeword": (None, [u32, u32, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    "orc_read_symbols": (sz, [vp, sz, ctypes.c_uint, vp, sz, vp]),
    "orc_decode_component_reduced": (None, [vp, sz, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, vp]),
    "orc_count_symbols": (i32, [vp, sz, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, u64,
                                ctypes.POINTER(u64), ctypes.POINTER(u64)]),
    "orc_frame_stream_census": (sz, [vp, sz, vp, vp, sz]),
    "orc_apple_block": (None, [vp, vp, vp]),
    "orc_ffmpeg_block": (None, [vp, vp, vp]),
    "orc_apple_block_4x4": (None, [vp, vp, vp]),
    "orc_apple_block_2x2": (None, [vp, vp, vp]),
    "orc_output_code": (ctypes.c_uint16, [ctypes.c_float]),
    "orc_code_value": (ctypes.c_uint16, [ctypes.c_int32]),
    "orc_dequantize": (ctypes.c_int16, [ctypes.c_int16, ctypes.c_int16]),
    "orc_testgen_hash": (u64, [u64, u64, i32]),
    "orc_testgen_block": (None, [u64, u64, vp, vp]),
    "orc_linearize": (ctypes.c_uint16, [vp, ctypes.c_uint16]),
    "orc_header_len": (i32, [vp, sz, ctypes.POINTER(sz), E]),
    "orc_parse_header": (i32, [vp, sz, HP, E]),
    "orc_tiles_per_row": (sz, [ctypes.c_uint, ctypes.c_uint]),
    "orc_tile_count": (sz, [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]),
    "orc_tile_geometry": (None, [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, vp, vp, vp]),
    "orc_tile_parse": (i32, [vp, sz, vp, vp, vp]),
    "orc_parse_frame": (i32, [vp, sz, HP, vp, vp, sz, ctypes.POINTER(sz), vp, E]),
    "orc_decode_frame": (i32, [vp, sz, i32, i32, vp, sz, vp, sz, ctypes.POINTER(sz), vp, E]),
    "orc_linearize_frame": (None, [vp, vp, sz, sz, sz, vp, i32]),
    "orc_decode_frame_reduced": (i32, [vp, sz, ctypes.c_uint, i32, i32, vp, sz, vp, sz,
                                       ctypes.POINTER(sz), vp, E]),
    "orc_params_with_white_balance": (None, [HP, u32, ctypes.c_float, ctypes.c_float, PP]),
    "orc_output_matrix": (None, [HP, i32, vp]),
    "orc_develop": (None, [HP, vp, sz, PP, i32, i32, vp, vp, vp]),
    "orc_green_pass": (None, [vp, sz, sz, PP, vp]),
    "orc_chroma_pass": (None, [vp, sz, sz, PP, vp]),
    "orc_apply_matrix": (None, [vp, sz, vp]),
    "orc_dng_tile": (None, [vp, sz, sz, sz, sz, vp]),
    "orc_lj92_table": (i32, [vp, vp, vp, vp, vp]),
    "orc_lj92_encode": (sz, [vp, sz, sz, sz, vp, sz]),
    "orc_lj92_tile_capacity": (sz, []),
    "orc_dng_lj92_tiles": (sz, [vp, sz, sz, vp, sz, vp, vp, i32]),
}

_lib = None
THREADS = min(os.cpu_count() or 1, 16)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _arr(packet):
    a = np.frombuffer(bytes(packet), dtype=np.uint8)
    return a, (a.ctypes.data if a.size else None)


def parse_header(packet):
    a, p = _arr(packet)
    h, e = OrcHeader(), OrcError()
    rc = lib().orc_parse_header(p, a.size, ctypes.byref(h), ctypes.byref(e))
    if rc:
        raise OracleError(e)
    return h


def parse_frame(packet):
    """(header, offsets, sizes, max_qscale)."""
    a, p = _arr(packet)
    h, e = OrcHeader(), OrcError()
    n, mq = ctypes.c_size_t(), ctypes.c_uint8()
    rc = lib().orc_parse_frame(p, a.size, ctypes.byref(h), None, None, 0, ctypes.byref(n),
                               ctypes.byref(mq), ctypes.byref(e))
    if rc:
        raise OracleError(e)
    offs = np.zeros(n.value + 1, np.uint64)
    sizes = np.zeros(n.value + 1, np.uint32)
    lib().orc_parse_frame(p, a.size, ctypes.byref(h), offs.ctypes.data, sizes.ctypes.data,
                          n.value, ctypes.byref(n), ctypes.byref(mq), ctypes.byref(e))
    return h, offs[:n.value], sizes[:n.value], mq.value


def coded(h):
    return (h.width + 15) & ~15, (h.height + 15) & ~15


def decode_frame(packet, ffmpeg=False, threads=None):
    """(header, samples[coded], skipped list, max_qscale)."""
    a, p = _arr(packet)
    h = parse_frame(packet)[0]
    cw, ch = coded(h)
    n = lib().orc_tile_count(h.width, h.height, h.alignment)
    samples = np.zeros(cw * ch, np.uint16)
    skipped = np.zeros(max(n, 1), np.uint32)
    nsk, mq, e = ctypes.c_size_t(), ctypes.c_uint8(), OrcError()
    rc = lib().orc_decode_frame(p, a.size, int(ffmpeg), threads or THREADS, samples.ctypes.data,
                                samples.size, skipped.ctypes.data, skipped.size,
                                ctypes.byref(nsk), ctypes.byref(mq), ctypes.byref(e))
    if rc:
        raise OracleError(e)
    return h, samples, [int(x) for x in skipped[:nsk.value]], mq.value


def count_symbols(data, align, log2_blocks, log2_coefs=6, cap=1 << 26):
    """(AC symbols, bits consumed, gave up) of decode_component_reduced on
    one stream (an instrumented copy of its loop)."""
    a, p = _arr(data)
    n, bits = ctypes.c_uint64(), ctypes.c_uint64()
    capped = lib().orc_count_symbols(p, a.size, align, log2_blocks, log2_coefs, cap,
                                     ctypes.byref(n), ctypes.byref(bits))
    return n.value, bits.value, bool(capped)


def stream_census(packet):
    """Per stream of a frame (4 per tile, decode order): (bytes, symbols)."""
    a, p = _arr(packet)
    h = parse_frame(packet)[0]
    n = lib().orc_tile_count(h.width, h.height, h.alignment)
    nbytes = np.zeros(4 * n, np.uint32)
    syms = np.zeros(4 * n, np.uint32)
    got = lib().orc_frame_stream_census(p, a.size, nbytes.ctypes.data, syms.ctypes.data, n)
    assert got == n
    return nbytes, syms


def linearize(h, samples, threads=None):
    curve = np.array(list(h.curve), np.uint16)
    out = np.zeros(h.width * h.height, np.uint16)
    cw, _ = coded(h)
    lib().orc_linearize_frame(curve.ctypes.data, samples.ctypes.data, cw, h.width, h.height,
                              out.ctypes.data, threads or THREADS)
    return out


def decode_frame_reduced(packet, shift, average=False, threads=None):
    a, p = _arr(packet)
    h = parse_frame(packet)[0]
    cw, ch = coded(h)
    n = lib().orc_tile_count(h.width, h.height, h.alignment)
    rgb = np.zeros(3 * (cw >> shift) * (ch >> shift), np.uint16)
    skipped = np.zeros(max(n, 1), np.uint32)
    nsk, mq, e = ctypes.c_size_t(), ctypes.c_uint8(), OrcError()
    rc = lib().orc_decode_frame_reduced(p, a.size, shift, int(average), threads or THREADS,
                                        rgb.ctypes.data, rgb.size, skipped.ctypes.data,
                                        skipped.size, ctypes.byref(nsk), ctypes.byref(mq),
                                        ctypes.byref(e))
    if rc:
        raise OracleError(e)
    return h, rgb, [int(x) for x in skipped[:nsk.value]], mq.value


def params(h, quantizer, red=None, blue=None):
    pr = OrcParams()
    lib().orc_params_with_white_balance(ctypes.byref(h), quantizer,
                                        h.wb_red if red is None else red,
                                        h.wb_blue if blue is None else blue, ctypes.byref(pr))
    return pr


def output_matrix(h, normalized=True):
    m = np.zeros(9, np.float32)
    lib().orc_output_matrix(ctypes.byref(h), int(normalized), m.ctypes.data)
    return m


def develop_planar(h, samples, prm, stage, threads=None):
    n = h.width * h.height
    r, g, b = (np.zeros(n, np.float32) for _ in range(3))
    cw, _ = coded(h)
    lib().orc_develop(ctypes.byref(h), samples.ctypes.data, cw, ctypes.byref(prm), int(stage),
                      threads or THREADS, r.ctypes.data, g.ctypes.data, b.ctypes.data)
    return r, g, b


def develop_interleaved(h, samples, prm, stage, threads=None):
    out = np.zeros(3 * h.width * h.height, np.float32)
    cw, _ = coded(h)
    lib().orc_develop(ctypes.byref(h), samples.ctypes.data, cw, ctypes.byref(prm), int(stage),
                      threads or THREADS, out.ctypes.data, None, None)
    return out


# ---- lossless JPEG tiles of a DNG (oracle/lj92_oracle.c) -------------------

def dng_tile(values, left, top):
    """dng.rs:115-128: the 256x256 tile at (left, top) of a height x width
    picture (uint16, row-major)."""
    v = np.ascontiguousarray(values, dtype=np.uint16)
    out = np.zeros((256, 256), dtype=np.uint16)
    lib().orc_dng_tile(v.ctypes.data, v.shape[1], v.shape[0], left, top, out.ctypes.data)
    return out


def lj92_table(histogram):
    """HuffmanTable::optimal (lj92.rs:105-206): (counts[16], values,
    codes[17], lengths[17])."""
    h = np.ascontiguousarray(histogram, dtype=np.uint64)
    counts = np.zeros(16, dtype=np.uint8)
    values = np.zeros(17, dtype=np.uint8)
    code = np.zeros(17, dtype=np.uint16)
    length = np.zeros(17, dtype=np.uint8)
    n = lib().orc_lj92_table(h.ctypes.data, counts.ctypes.data, values.ctypes.data,
                             code.ctypes.data, length.ctypes.data)
    return counts, values[:n].copy(), code, length


def lj92_encode(samples, components=2):
    """encode (lj92.rs:14-88) of a height x width uint16 array."""
    s = np.ascontiguousarray(samples, dtype=np.uint16)
    h, w = s.shape
    cap = 128 + 4 * s.size + 16
    out = np.zeros(cap, dtype=np.uint8)
    n = lib().orc_lj92_encode(s.ctypes.data, w, h, components, out.ctypes.data, cap)
    assert n <= cap
    return out[:n].tobytes()


def dng_lj92_tiles(values, threads=None):
    """dng.rs:92-113: every 256x256 tile's stream in tile order:
    (bytes, offsets, lengths)."""
    v = np.ascontiguousarray(values, dtype=np.uint16)
    h, w = v.shape
    n = ((w + 255) // 256) * ((h + 255) // 256)
    cap = n * lib().orc_lj92_tile_capacity()
    out = np.zeros(cap, dtype=np.uint8)
    offsets = np.zeros(n, dtype=np.uint32)
    lengths = np.zeros(n, dtype=np.uint32)
    total = lib().orc_dng_lj92_tiles(v.ctypes.data, w, h, out.ctypes.data, cap,
                                     offsets.ctypes.data, lengths.ctypes.data, threads or THREADS)
    return out[:total].tobytes(), offsets, lengths
