"""Python mirror of proresraw_rs's public API (src/lib.rs:50-66) over the
C ABI: the same names, argument meanings and error behaviour, with every
decode, linearize and develop running in the HIP kernels.

Errors are raised as :class:`Error` carrying the crate's variant name
(``Error.kind``) and payload; malformed tiles are not errors but are listed
in ``skipped_tiles``. Where the crate panics on a wrong output size
(``*_into``), :class:`ValueError` is raised.
"""
import ctypes
import enum
import threading
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import _native as N

BLACK_LEVEL = 256  # header.rs:7
PREFIX_LEN = 10  # decoder.rs:13

_KINDS = {
    1: "Truncated", 2: "FrameSizeMismatch", 3: "BadQuantMatrix", 4: "BadCurve",
    5: "UnexpectedCurve", 6: "BadTag", 7: "HeaderTooShort", 8: "UnsupportedVersion",
    9: "BadDimensions", 10: "BadAlignment", 11: "TileOutOfBounds", 12: "BadTileHeader",
    100: "OutputSize", 101: "InvalidArgument", 102: "Device", 103: "DecodeOverrun",
}


class Error(Exception):
    """Error (error.rs:7-63): ``kind`` is the variant name; payload fields
    are attributes named as in the crate (needed/available, declared/actual,
    index/value, width/height, tile, points, tag, ...)."""

    def __init__(self, e: N.PrrError):
        self.code = e.code
        self.kind = _KINDS.get(e.code, f"Code{e.code}")
        self.what = e.what.decode(errors="replace")
        self.a, self.b = int(e.a), int(e.b)
        k = self.kind
        if k == "Truncated":
            self.needed, self.available = self.a, self.b
        elif k == "FrameSizeMismatch":
            self.declared, self.actual = self.a, self.b
        elif k == "BadQuantMatrix":
            self.index, self.value = self.a, self.b
        elif k == "BadCurve":
            self.points = list(e.points)
        elif k == "BadTag":
            self.tag = bytes(e.tag)
        elif k == "HeaderTooShort":
            self.value = self.a
        elif k == "UnsupportedVersion":
            self.value = self.a
        elif k == "BadDimensions":
            self.width, self.height = self.a, self.b
        elif k == "BadAlignment":
            self.value = self.a
        elif k in ("TileOutOfBounds", "BadTileHeader"):
            self.tile = self.a
        super().__init__(f"{self.kind}({self.what or ''} a={self.a} b={self.b})")

    def __eq__(self, other):
        return isinstance(other, Error) and (self.code, self.a, self.b) == (other.code, other.a, other.b)

    def __hash__(self):
        return hash((self.code, self.a, self.b))


def _check(rc: int, err: N.PrrError):
    if rc:
        if rc == 100:
            raise ValueError(err.what.decode(errors="replace") or "output size")
        raise Error(err)


class Idct(enum.IntEnum):
    """idct.rs:19-26."""
    Apple = 0
    Ffmpeg = 1


class Scale(enum.IntEnum):
    """reduced.rs:25-61; the value is log2 of the factor."""
    Half = 1
    Quarter = 2
    Eighth = 3

    def factor(self) -> int:
        return 1 << int(self)


class Greens(enum.IntEnum):
    """reduced.rs:64-72."""
    First = 0
    Average = 1


class Stage(enum.IntEnum):
    """demosaic.rs:112-123."""
    Demosaiced = 0
    WhiteBalanced = 1
    Rec2020 = 2


class Exposure(enum.IntEnum):
    """demosaic.rs:127-133."""
    Normalized = 0
    Reflection = 1


class BayerPattern(enum.IntEnum):
    """header.rs:21-30."""
    Rggb = 0
    Grbg = 1
    Gbrg = 2
    Bggr = 3

    def red_position(self) -> Tuple[int, int]:
        return [(0, 0), (0, 1), (1, 0), (1, 1)][int(self)]

    def stream_phases(self):
        r, c = self.red_position()
        return [(1 - r, c), (r, 1 - c), (1 - r, 1 - c), (r, c)]


@dataclass
class FrameHeader:
    """FrameHeader (header.rs:112-146)."""
    version: int
    vendor: bytes
    width: int
    height: int
    crop: Tuple[int, int, int, int]
    bayer_pattern: BayerPattern
    white_level: int
    wb_red: float
    wb_blue: float
    color_matrix: List[List[float]]
    gain: float
    wb_cct: int
    flags: int
    alignment: int
    quant_matrix: List[int]
    curve: List[int]
    _raw: N.PrrHeader = field(repr=False, compare=False, default=None)

    @classmethod
    def _from_raw(cls, h: N.PrrHeader):
        raw = N.PrrHeader()
        ctypes.memmove(ctypes.byref(raw), ctypes.byref(h), ctypes.sizeof(N.PrrHeader))
        cm = list(h.color_matrix)
        return cls(
            version=h.version, vendor=bytes(h.vendor), width=h.width, height=h.height,
            crop=tuple(h.crop), bayer_pattern=BayerPattern(h.bayer), white_level=h.white_level,
            wb_red=h.wb_red, wb_blue=h.wb_blue, color_matrix=[cm[0:3], cm[3:6], cm[6:9]],
            gain=h.gain, wb_cct=h.wb_cct, flags=h.flags, alignment=h.alignment,
            quant_matrix=list(h.quant_matrix), curve=list(h.curve), _raw=raw)

    def coded_width(self) -> int:
        return (self.width + 15) & ~15

    def coded_height(self) -> int:
        return (self.height + 15) & ~15


_tls = threading.local()


def context():
    """The calling thread's device context (one HIP stream + scratch)."""
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        lib = N.lib()
        handle = ctypes.c_void_p()
        err = N.PrrError()
        _check(lib.prr_context_create(0, ctypes.byref(handle), ctypes.byref(err)), err)
        ctx = _Context(handle)
        _tls.ctx = ctx
    return ctx


class _Context:
    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        try:
            N.lib().prr_context_destroy(self.handle)
        except Exception:
            pass


def _buf(packet) -> Tuple[ctypes.c_void_p, int, object]:
    arr = np.frombuffer(bytes(packet), dtype=np.uint8) if not isinstance(packet, np.ndarray) \
        else np.ascontiguousarray(packet, dtype=np.uint8)
    return ctypes.c_void_p(arr.ctypes.data if arr.size else 0), int(arr.size), arr


def header_len(packet) -> int:
    """header_len (decoder.rs:58-73)."""
    p, n, keep = _buf(packet)
    out = ctypes.c_size_t()
    err = N.PrrError()
    _check(N.lib().prr_header_len(p, n, ctypes.byref(out), ctypes.byref(err)), err)
    return out.value


def parse_header(packet) -> FrameHeader:
    """parse_header (decoder.rs:45-53)."""
    p, n, keep = _buf(packet)
    h = N.PrrHeader()
    err = N.PrrError()
    _check(N.lib().prr_parse_header(p, n, ctypes.byref(h), ctypes.byref(err)), err)
    return FrameHeader._from_raw(h)


def parse_frame_info(packet):
    """The prefix, header and size-table checks of parse_frame
    (decoder.rs:76-103): (FrameHeader, PrrFrameInfo)."""
    p, n, keep = _buf(packet)
    h = N.PrrHeader()
    info = N.PrrFrameInfo()
    err = N.PrrError()
    _check(N.lib().prr_parse_frame_info(p, n, ctypes.byref(h), ctypes.byref(info),
                                        ctypes.byref(err)), err)
    return FrameHeader._from_raw(h), info


@dataclass
class ParsedFrame:
    """ParsedFrame (decoder.rs:15-29): the header, the tile grid's sizes,
    each tile's payload as (offset, size) in the packet, and the largest
    tile qscale byte."""
    header: FrameHeader
    info: N.PrrFrameInfo = field(repr=False)
    tile_offsets: np.ndarray
    tile_sizes: np.ndarray
    max_qscale: int
    _raw: N.PrrParsedFrame = field(repr=False, compare=False, default=None)

    def tiles(self, packet) -> List[bytes]:
        """The tile payloads (ParsedFrame::tiles) as slices of `packet`."""
        b = bytes(packet)
        return [b[int(o):int(o) + int(n)] for o, n in zip(self.tile_offsets, self.tile_sizes)]


def parse_frame(packet) -> ParsedFrame:
    """parse_frame (decoder.rs:76-123): the size table walked on the host
    (no device work)."""
    p, n, keep = _buf(packet)
    out = N.PrrParsedFrame()
    err = N.PrrError()
    lib = N.lib()
    _check(lib.prr_parse_frame(p, n, ctypes.byref(out), None, None, 0, ctypes.byref(err)), err)
    count = out.info.nb_tiles
    offsets = np.zeros(max(count, 1), dtype=np.uint64)
    sizes = np.zeros(max(count, 1), dtype=np.uint32)
    _check(lib.prr_parse_frame(p, n, ctypes.byref(out), ctypes.c_void_p(offsets.ctypes.data),
                               ctypes.c_void_p(sizes.ctypes.data), offsets.size,
                               ctypes.byref(err)), err)
    return ParsedFrame(FrameHeader._from_raw(out.header), out.info, offsets[:count],
                       sizes[:count], int(out.max_qscale), out)


def frame_quantizer(parsed: ParsedFrame) -> int:
    """frame_quantizer (demosaic.rs:50-52)."""
    return int(N.lib().prr_frame_quantizer(ctypes.byref(parsed._raw)))


class Frame:
    """Frame (decoder.rs:131-217): code values over the coded size,
    skipped tile indices and the largest tile qscale."""

    def __init__(self, header: FrameHeader, samples: np.ndarray, skipped_tiles: List[int],
                 max_qscale: int):
        self.header = header
        self.samples = samples
        self.skipped_tiles = skipped_tiles
        self.max_qscale = max_qscale

    def quantizer(self) -> int:
        return N.lib().prr_quantizer(ctypes.byref(self.header._raw), self.max_qscale)

    def width(self) -> int:
        return self.header.width

    def height(self) -> int:
        return self.header.height

    def stride(self) -> int:
        return self.header.coded_width()

    def row(self, y: int) -> np.ndarray:
        s = self.stride()
        return self.samples[y * s:y * s + self.width()]

    def rows(self):
        return (self.row(y) for y in range(self.height()))

    def linearize(self) -> np.ndarray:
        out = np.zeros(self.width() * self.height(), dtype=np.uint16)
        self.linearize_into(out)
        return out

    def linearize_into(self, out: np.ndarray):
        if out.dtype != np.uint16 or not out.flags.c_contiguous:
            raise ValueError("output must be contiguous uint16")
        if out.size != self.width() * self.height():
            raise ValueError("output size")
        if self.width() == 0:
            return
        err = N.PrrError()
        s = np.ascontiguousarray(self.samples, dtype=np.uint16)
        _check(N.lib().prr_linearize_into(context().handle, ctypes.byref(self.header._raw),
                                          s.ctypes.data, s.size, out.ctypes.data, out.size,
                                          ctypes.byref(err)), err)


def decode_frame_with(packet, idct: Idct = Idct.Apple) -> Frame:
    """decode_frame_with (decoder.rs:226-267) on the GPU."""
    header, info = parse_frame_info(packet)
    p, n, keep = _buf(packet)
    samples = np.zeros(info.coded_width * info.coded_height, dtype=np.uint16)
    skipped = np.zeros(max(info.nb_tiles, 1), dtype=np.uint32)
    nsk = ctypes.c_size_t()
    mq = ctypes.c_uint8()
    h = N.PrrHeader()
    err = N.PrrError()
    _check(N.lib().prr_decode_frame(context().handle, p, n, int(idct), ctypes.byref(h),
                                    samples.ctypes.data, samples.size, skipped.ctypes.data,
                                    skipped.size, ctypes.byref(nsk), ctypes.byref(mq),
                                    ctypes.byref(err)), err)
    return Frame(header, samples, [int(t) for t in skipped[:nsk.value]], mq.value)


def decode_frame(packet) -> Frame:
    """decode_frame (decoder.rs:221-223): the Apple-compatible transform."""
    return decode_frame_with(packet, Idct.Apple)


class RgbFrame:
    """RgbFrame (reduced.rs:75-132)."""

    def __init__(self, header, scale, greens, samples, skipped_tiles, max_qscale):
        self.header = header
        self.scale = scale
        self.greens = greens
        self.samples = samples
        self.skipped_tiles = skipped_tiles
        self.max_qscale = max_qscale

    def width(self) -> int:
        return self.header.width >> int(self.scale)

    def height(self) -> int:
        return self.header.height >> int(self.scale)

    def stride(self) -> int:
        return self.header.coded_width() >> int(self.scale)

    def row(self, y: int) -> np.ndarray:
        s = 3 * y * self.stride()
        return self.samples[s:s + 3 * self.width()]

    def quantizer(self) -> int:
        return N.lib().prr_quantizer(ctypes.byref(self.header._raw), self.max_qscale)

    def to_packed(self) -> np.ndarray:
        if self.height() == 0:
            return np.zeros(0, np.uint16)
        return np.concatenate([self.row(y) for y in range(self.height())])


def decode_frame_reduced_with(packet, scale: Scale, greens: Greens = Greens.First) -> RgbFrame:
    """decode_frame_reduced_with (reduced.rs:147-201) on the GPU."""
    header, info = parse_frame_info(packet)
    p, n, keep = _buf(packet)
    s = int(scale)
    rgb = np.zeros(3 * (info.coded_width >> s) * (info.coded_height >> s), dtype=np.uint16)
    skipped = np.zeros(max(info.nb_tiles, 1), dtype=np.uint32)
    nsk = ctypes.c_size_t()
    mq = ctypes.c_uint8()
    h = N.PrrHeader()
    err = N.PrrError()
    _check(N.lib().prr_decode_frame_reduced(context().handle, p, n, s, int(greens), ctypes.byref(h),
                                            rgb.ctypes.data, rgb.size, skipped.ctypes.data,
                                            skipped.size, ctypes.byref(nsk), ctypes.byref(mq),
                                            ctypes.byref(err)), err)
    return RgbFrame(header, Scale(s), Greens(int(greens)), rgb,
                    [int(t) for t in skipped[:nsk.value]], mq.value)


def decode_frame_reduced(packet, scale: Scale) -> RgbFrame:
    """decode_frame_reduced (reduced.rs:142-144)."""
    return decode_frame_reduced_with(packet, scale, Greens.First)


@dataclass
class DemosaicParams:
    """DemosaicParams (demosaic.rs:57-107)."""
    cfa_phase: int
    black: int
    one_over_range: float
    noise: Tuple[float, float, float]
    clip: Tuple[float, float]
    black_over_range: float
    wb: Tuple[float, float]

    @classmethod
    def _from_raw(cls, p: N.PrrDemosaicParams):
        return cls(p.cfa_phase, p.black, p.one_over_range, tuple(p.noise), tuple(p.clip),
                   p.black_over_range, tuple(p.wb))

    def _raw(self) -> N.PrrDemosaicParams:
        p = N.PrrDemosaicParams()
        p.cfa_phase, p.black, p.one_over_range = self.cfa_phase, self.black, self.one_over_range
        p.noise[:] = list(self.noise)
        p.clip[:] = list(self.clip)
        p.black_over_range = self.black_over_range
        p.wb[:] = list(self.wb)
        return p

    @classmethod
    def new(cls, header: FrameHeader, quantizer: int):
        p = N.PrrDemosaicParams()
        N.lib().prr_demosaic_params_new(ctypes.byref(header._raw), quantizer, ctypes.byref(p))
        return cls._from_raw(p)

    @classmethod
    def with_white_balance(cls, header: FrameHeader, quantizer: int, red: float, blue: float):
        p = N.PrrDemosaicParams()
        N.lib().prr_demosaic_params_with_white_balance(ctypes.byref(header._raw), quantizer,
                                                       red, blue, ctypes.byref(p))
        return cls._from_raw(p)


def output_matrix(header: FrameHeader, exposure: Exposure = Exposure.Normalized) -> List[float]:
    """output_matrix (demosaic.rs:138-153)."""
    m = (ctypes.c_float * 9)()
    N.lib().prr_output_matrix(ctypes.byref(header._raw), int(exposure), m)
    return list(m)


@dataclass
class RgbImage:
    """RgbImage (demosaic.rs:157-165)."""
    width: int
    height: int
    data: np.ndarray


@dataclass
class RgbPlanes:
    """RgbPlanes (demosaic.rs:169-178)."""
    width: int
    height: int
    planes: List[np.ndarray]


def _samples(frame: Frame) -> np.ndarray:
    return np.ascontiguousarray(frame.samples, dtype=np.uint16)


def develop_into(frame: Frame, params: DemosaicParams, stage: Stage, out: np.ndarray):
    """develop_into (demosaic.rs:223-227)."""
    if out.dtype != np.float32 or not out.flags.c_contiguous:
        raise ValueError("output must be contiguous float32")
    s = _samples(frame)
    raw = params._raw()
    err = N.PrrError()
    _check(N.lib().prr_develop_into(context().handle, ctypes.byref(frame.header._raw),
                                    s.ctypes.data, s.size, ctypes.byref(raw), int(stage),
                                    out.ctypes.data, out.size, ctypes.byref(err)), err)


def develop_planar_into(frame: Frame, params: DemosaicParams, stage: Stage, out):
    """develop_planar_into (demosaic.rs:250-261): out = [r, g, b]."""
    r, g, b = out
    for plane in (r, g, b):
        if plane.dtype != np.float32 or not plane.flags.c_contiguous:
            raise ValueError("planes must be contiguous float32")
        if plane.size != frame.width() * frame.height():
            raise ValueError("plane size")
    s = _samples(frame)
    raw = params._raw()
    err = N.PrrError()
    _check(N.lib().prr_develop_planar_into(context().handle, ctypes.byref(frame.header._raw),
                                           s.ctypes.data, s.size, ctypes.byref(raw), int(stage),
                                           r.ctypes.data, g.ctypes.data, b.ctypes.data, r.size,
                                           ctypes.byref(err)), err)


def develop_with(frame: Frame, params: DemosaicParams, stage: Stage) -> RgbImage:
    """develop_with (demosaic.rs:206-215)."""
    w, h = frame.width(), frame.height()
    data = np.zeros(3 * w * h, dtype=np.float32)
    develop_into(frame, params, stage, data)
    return RgbImage(w, h, data)


def develop(frame: Frame, stage: Stage = Stage.WhiteBalanced) -> RgbImage:
    """develop (demosaic.rs:190-193): the header's white balance."""
    return develop_with(frame, DemosaicParams.new(frame.header, frame.quantizer()), stage)


def develop_planar(frame: Frame, params: DemosaicParams, stage: Stage) -> RgbPlanes:
    """develop_planar (demosaic.rs:232-242)."""
    w, h = frame.width(), frame.height()
    planes = [np.zeros(w * h, dtype=np.float32) for _ in range(3)]
    develop_planar_into(frame, params, stage, planes)
    return RgbPlanes(w, h, planes)


FOURCCS = (b"aprn", b"aprh")  # lib.rs:69


def is_prores_raw_fourcc(fourcc: bytes) -> bool:
    """is_prores_raw_fourcc (lib.rs:71-75)."""
    return bytes(fourcc).lower() in FOURCCS
