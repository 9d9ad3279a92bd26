"""The sample table of one video track of a QuickTime movie, mirroring the
crate's `mov` module (src/mov.rs): read_video_track, read_at, Track, Sample
and Orientation, through the C ABI of include/proresraw_mov.h (C++ host code
reading the `moov` atom and positional reads of samples).

Also here: `read_samples` (many samples into one buffer on host threads,
the step that feeds a batch of frames to the device) and `build_movie`, a
writer of small QuickTime movies in the layout the crate's tests build by
hand (mov.rs:800-1010), used for synthetic clips.
"""
import ctypes
import enum
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N

PRR_E_MOV = 104


class MovError(Exception):
    """mov::Error (mov.rs:22-45): why a movie could not be read."""


@dataclass(frozen=True)
class Sample:
    """Sample (mov.rs:91-98)."""
    offset: int
    size: int


class Orientation(enum.IntEnum):
    """Orientation (mov.rs:121-157), valued as its EXIF number."""
    Identity = 1
    MirrorHorizontal = 2
    Rotate180 = 3
    MirrorVertical = 4
    Transpose = 5
    Rotate90 = 6
    Transverse = 7
    Rotate270 = 8

    def exif(self) -> int:
        return int(self)

    def swaps_dimensions(self) -> bool:
        return int(self) >= 5


@dataclass
class Track:
    """Track (mov.rs:100-119)."""
    fourcc: bytes
    fps: Tuple[int, int]
    variable_durations: bool
    samples: List[Sample]
    creation_time: Optional[int]
    orientation: Orientation
    warnings: List[str] = field(default_factory=list)
    metadata: List[Tuple[str, str]] = field(default_factory=list)

    def text(self, keys: Sequence[str]) -> Optional[str]:
        """Track::text (mov.rs:159-169)."""
        for key in keys:
            for k, v in self.metadata:
                if k == key:
                    return v
        return None


def _fd(file):
    """(fd, close?) for a path, an open file or a descriptor."""
    if isinstance(file, int):
        return file, False
    if hasattr(file, "fileno"):
        return file.fileno(), False
    return os.open(os.fspath(file), os.O_RDONLY), True


def _msg(buf):
    return buf.value.decode("utf-8", errors="replace")


def is_prores_raw_fourcc(fourcc) -> bool:
    """is_prores_raw_fourcc (lib.rs:72-75), in the C library."""
    b = (ctypes.c_uint8 * 4)(*bytes(fourcc)[:4])
    return bool(N.lib().prr_is_prores_raw_fourcc(b))


def read_video_track(file, accept=None) -> Track:
    """read_video_track (mov.rs:171-176): the first enabled video track of
    the movie whose sample description `accept(fourcc: bytes)` takes (None:
    the ProRes RAW fourccs). `file` is a path, an open file or a
    descriptor."""
    lib = N.lib()
    fd, close = _fd(file)
    try:
        out = ctypes.c_void_p()
        msg = ctypes.create_string_buffer(1024)
        held = []
        cb = ctypes.cast(None, N.MOV_ACCEPT)
        if accept is not None:
            def trampoline(p, _user):
                try:
                    return 1 if accept(bytes(p[i] for i in range(4))) else 0
                except Exception:  # an exception cannot cross the C frame
                    return 0
            cb = N.MOV_ACCEPT(trampoline)
            held.append(cb)
        rc = lib.prr_mov_read_video_track(fd, cb, None, ctypes.byref(out), msg, len(msg))
        if rc == PRR_E_MOV:
            raise MovError(_msg(msg))
        if rc:
            raise MovError(f"prr_mov_read_video_track returned {rc}")
        try:
            return _track(lib, out)
        finally:
            lib.prr_mov_track_free(out)
    finally:
        if close:
            os.close(fd)


def _track(lib, h) -> Track:
    fourcc = (ctypes.c_uint8 * 4)()
    lib.prr_mov_track_fourcc(h, fourcc)
    num, den = ctypes.c_uint32(), ctypes.c_uint32()
    lib.prr_mov_track_fps(h, ctypes.byref(num), ctypes.byref(den))
    ptr = ctypes.POINTER(N.PrrMovSample)()
    n = lib.prr_mov_track_samples(h, ctypes.byref(ptr))
    if n:
        arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint64)), (2 * n,))
        offsets = arr[0::2].tolist()
        sizes = (arr[1::2] & 0xFFFFFFFF).tolist()
        samples = [Sample(o, s) for o, s in zip(offsets, sizes)]
    else:
        samples = []
    ct = ctypes.c_uint64()
    creation = int(ct.value) if lib.prr_mov_track_creation_time(h, ctypes.byref(ct)) else None
    def text(fn, i):
        n = ctypes.c_size_t()
        p = fn(h, i, ctypes.byref(n))
        return ctypes.string_at(p, n.value).decode("utf-8")

    warnings = [text(lib.prr_mov_track_warning, i) for i in range(lib.prr_mov_track_n_warnings(h))]
    metadata = [(text(lib.prr_mov_track_metadata_key, i), text(lib.prr_mov_track_metadata_value, i))
                for i in range(lib.prr_mov_track_n_metadata(h))]
    return Track(fourcc=bytes(fourcc), fps=(num.value, den.value),
                 variable_durations=bool(lib.prr_mov_track_variable_durations(h)),
                 samples=samples, creation_time=creation,
                 orientation=Orientation(lib.prr_mov_track_orientation(h)),
                 warnings=warnings, metadata=metadata)


def read_at(file, offset: int, size: int) -> bytes:
    """read_at (mov.rs:178-182): `size` bytes at byte `offset`."""
    fd, close = _fd(file)
    try:
        buf = (ctypes.c_uint8 * max(size, 1))()
        msg = ctypes.create_string_buffer(512)
        rc = N.lib().prr_mov_read_at(fd, offset, buf, size, msg, len(msg))
        if rc == PRR_E_MOV:
            raise MovError(_msg(msg))
        if rc:
            raise MovError(f"prr_mov_read_at returned {rc}")
        return bytes(buf)[:size]
    finally:
        if close:
            os.close(fd)


def packet_layout(samples: Sequence[Sample], align=4, pad=16):
    """Byte offsets of each sample in one buffer where every packet starts
    `align`-aligned and is followed by `pad` readable bytes (the decode's
    packet rule), and the buffer's total size."""
    offsets = np.zeros(len(samples), dtype=np.uint64)
    at = 0
    for i, s in enumerate(samples):
        offsets[i] = at
        at += ((s.size + align - 1) // align) * align + pad
    return offsets, at


def read_samples(file, samples: Sequence[Sample], dst, offsets, threads=0):
    """Read every sample into `dst` (a writable buffer: a numpy array or a
    pinned torch tensor's data pointer as int) at `offsets`, on `threads`
    host threads (0: up to 16)."""
    fd, close = _fd(file)
    try:
        n = len(samples)
        table = (N.PrrMovSample * max(n, 1))()
        for i, s in enumerate(samples):
            table[i].offset, table[i].size = s.offset, s.size
        offs = np.ascontiguousarray(offsets, dtype=np.uint64)
        ptr = dst if isinstance(dst, int) else dst.ctypes.data
        msg = ctypes.create_string_buffer(512)
        rc = N.lib().prr_mov_read_samples(
            fd, table, n, ctypes.c_void_p(ptr),
            offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), threads, msg, len(msg))
        if rc == PRR_E_MOV:
            raise MovError(_msg(msg))
        if rc:
            raise MovError(f"prr_mov_read_samples returned {rc}")
    finally:
        if close:
            os.close(fd)


# ---- writing small movies --------------------------------------------------

def _atom(kind: bytes, body: bytes) -> bytes:
    return (len(body) + 8).to_bytes(4, "big") + kind + body


def _full(kind: bytes, version: int, flags: int, body: bytes) -> bytes:
    return _atom(kind, ((version << 24) | flags).to_bytes(4, "big") + body)


def _words(values) -> bytes:
    return b"".join((v & 0xFFFFFFFF).to_bytes(4, "big") for v in values)


@dataclass
class TrackSpec:
    """One track of a movie for build_movie; the defaults describe the
    five-sample track the crate's tests start from (mov.rs:838-855)."""
    handler: bytes = b"vide"
    fourcc: bytes = b"aprh"
    enabled: bool = True
    timescale: int = 600
    durations: List[int] = field(default_factory=lambda: [20] * 5)
    sizes: List[int] = field(default_factory=lambda: [3, 5, 7, 2, 4])
    size_table: str = "stsz"            # "stsz", "stz2-4", "stz2-8", "stz2-16"
    chunks: List[int] = field(default_factory=lambda: [2, 2, 1])  # samples per chunk
    co64: bool = False
    edits: List[tuple] = field(default_factory=list)  # (duration, media time[, rate 16.16])
    tkhd_version: int = 0
    matrix: Tuple[int, int, int, int] = (0x10000, 0, 0, 0x10000)  # a b c d, 16.16
    payloads: Optional[List[bytes]] = None  # sample bytes; default: size copies of the index


def _trak(spec: TrackSpec, data_at: int) -> bytes:
    at = 40 if spec.tkhd_version == 0 else 52
    tkhd = bytearray(at + 44)
    tkhd[0:4] = (int(spec.enabled) | (spec.tkhd_version << 24)).to_bytes(4, "big")
    a, b, c, d = spec.matrix
    for i, v in enumerate((a, b, 0, c, d, 0, 0, 0, 0x40000000)):
        tkhd[at + 4 * i:at + 4 * i + 4] = (v & 0xFFFFFFFF).to_bytes(4, "big")
    body = _atom(b"tkhd", bytes(tkhd))
    if spec.edits:
        elst = _words([0, len(spec.edits)])
        for e in spec.edits:
            duration, time, rate = (tuple(e) + (0x10000,))[:3]
            elst += _words([duration, time, rate])
        body += _atom(b"edts", _atom(b"elst", elst))
    entry = bytearray(78)
    entry[7] = 1
    stbl = _full(b"stsd", 0, 0, _words([1]) + _atom(spec.fourcc, bytes(entry)))
    runs = []
    for dur in spec.durations:
        if runs and runs[-1][1] == dur:
            runs[-1][0] += 1
        else:
            runs.append([1, dur])
    stbl += _full(b"stts", 0, 0, _words([len(runs)] + [x for r in runs for x in r]))
    stsc, count = b"", 0
    for i, n in enumerate(spec.chunks):
        if i == 0 or spec.chunks[i - 1] != n:
            stsc += _words([i + 1, n, 1])
            count += 1
    stbl += _full(b"stsc", 0, 0, _words([count]) + stsc)
    n = len(spec.sizes)
    if spec.size_table == "stsz":
        stbl += _full(b"stsz", 0, 0, _words([0, n]) + _words(spec.sizes))
    else:
        bits = int(spec.size_table.split("-")[1])
        if bits == 4:
            table = bytes(((spec.sizes[i] << 4) | (spec.sizes[i + 1] if i + 1 < n else 0)) & 0xFF
                          for i in range(0, n, 2))
        elif bits == 8:
            table = bytes(s & 0xFF for s in spec.sizes)
        else:
            table = b"".join((s & 0xFFFF).to_bytes(2, "big") for s in spec.sizes)
        stbl += _full(b"stz2", 0, 0, bytes([0, 0, 0, bits]) + _words([n]) + table)
    offsets, pos, it = [], data_at, iter(spec.sizes)
    for k in spec.chunks:
        offsets.append(pos)
        pos += sum(next(it) for _ in range(k))
    if spec.co64:
        stbl += _full(b"co64", 0, 0, _words([len(offsets)]) +
                      b"".join(o.to_bytes(8, "big") for o in offsets))
    else:
        stbl += _full(b"stco", 0, 0, _words([len(offsets)] + offsets))
    minf = _full(b"vmhd", 0, 1, bytes(8)) + _atom(b"stbl", stbl)
    mdia = _full(b"mdhd", 0, 0, _words([0, 0, spec.timescale, 0, 0]))
    mdia += _full(b"hdlr", 0, 0, b"mhlr" + spec.handler + bytes(13))
    mdia += _atom(b"minf", minf)
    body += _atom(b"mdia", mdia)
    return _atom(b"trak", body)


def _payload(spec: TrackSpec, i: int) -> bytes:
    if spec.payloads is not None:
        return spec.payloads[i]
    return bytes([i & 0xFF]) * spec.sizes[i]


def build_movie(tracks: Sequence[TrackSpec], creation=0, meta: Optional[bytes] = None) -> bytes:
    """A movie of `ftyp`, `wide`, one `mdat` holding every track's samples
    in turn, then `moov` (mvhd at timescale 600, the tracks, and `meta` as
    the movie's metadata atom body when given)."""
    out = _atom(b"ftyp", b"qt  \0\0\x02\0qt  ") + _atom(b"wide", b"")
    data, starts = b"", []
    for spec in tracks:
        starts.append(len(out) + 8 + len(data))
        data += b"".join(_payload(spec, i) for i in range(len(spec.sizes)))
    out += _atom(b"mdat", data)
    moov = _full(b"mvhd", 0, 0, _words([creation, 0, 600, 0]))
    for spec, start in zip(tracks, starts):
        moov += _trak(spec, start)
    if meta is not None:
        moov += _atom(b"meta", meta)
    return out + _atom(b"moov", moov)


def clip_movie(packets: Sequence[bytes], fps=(24000, 1001), fourcc=b"aprh",
               frames_per_chunk=4) -> bytes:
    """A single-track movie of ProRes RAW packets at `fps` (timescale
    fps[0], each sample fps[1] long), with chunk offsets in `co64`."""
    sizes = [len(p) for p in packets]
    chunks = [frames_per_chunk] * (len(packets) // frames_per_chunk)
    if len(packets) % frames_per_chunk:
        chunks.append(len(packets) % frames_per_chunk)
    spec = TrackSpec(fourcc=fourcc, timescale=fps[0], durations=[fps[1]] * len(packets),
                     sizes=sizes, chunks=chunks, co64=True, payloads=list(packets))
    return build_movie([spec])
