"""Oracle (test infrastructure only): a pure-Python restatement of the
crate's QuickTime sample-table reader, src/mov.rs.

Only tests/ may import this module; the product reader is the C++ code
behind include/proresraw_mov.h. Each function cites the mov.rs lines it
follows. It is pinned by the crate's own movie tests (mov.rs:1013-1303,
restated in tests/test_mov.py against hand-built movies), which cover every
sample-size table form, both chunk-offset forms, edit lists, durations,
display matrices, metadata forms and the broken-file messages.

`read_video_track(data, accept)` takes the whole file as bytes and returns a
dict with the Track's fields, or raises MovError with the crate's message.
"""
import math

MAX_MOOV = 64 << 20                        # mov.rs:83
MAX_SAMPLES = 10_000_000                   # mov.rs:85
FIRST_ATOMS = (b"ftyp", b"moov", b"mdat", b"wide", b"free", b"skip", b"pnot")  # mov.rs:87-89


class MovError(Exception):
    """mov::Error (mov.rs:22-45): one message."""


def bail(msg):
    raise MovError(msg)


def name(kind):
    """`escape_ascii` of a fourcc (mov.rs:772-775)."""
    out = []
    for c in kind:
        if c == 9:
            out.append("\\t")
        elif c == 13:
            out.append("\\r")
        elif c == 10:
            out.append("\\n")
        elif c == 0x5C:
            out.append("\\\\")
        elif c == 0x27:
            out.append("\\'")
        elif c == 0x22:
            out.append('\\"')
        elif 0x20 <= c < 0x7F:
            out.append(chr(c))
        else:
            out.append(f"\\x{c:02x}")
    return "".join(out)


def fmt_f64(v):
    """Rust's `{}` of an f64: the shortest digits that read back, no exponent."""
    if v == 0:
        return "-0" if math.copysign(1, v) < 0 else "0"
    r = repr(float(v))  # Python's repr is also the shortest round-trip form
    if "e" in r or "E" in r:
        mant, exp = r.lower().split("e")
        exp = int(exp)
        neg = mant.startswith("-")
        mant = mant.lstrip("-")
        digits = mant.replace(".", "")
        point = (mant.index(".") if "." in mant else len(mant)) + exp
        if point <= 0:
            s = "0." + "0" * (-point) + digits
        elif point >= len(digits):
            s = digits + "0" * (point - len(digits))
        else:
            s = digits[:point] + "." + digits[point:]
        s = s.rstrip("0").rstrip(".") if "." in s else s
        return ("-" if neg else "") + s
    if r.endswith(".0"):
        r = r[:-2]
    return r


def be32(b, at=0):
    return int.from_bytes(b[at:at + 4], "big")


def be64(b, at=0):
    return int.from_bytes(b[at:at + 8], "big")


def entries(body, at, count, width, kind):
    """mov.rs:729-747."""
    n = count * width
    if at + n <= len(body):
        return body[at:at + n]
    bail(f"truncated `{name(kind)}` atom: {count} entries do not fit in {len(body)} bytes")


def u32_at(body, at, kind):
    return be32(entries(body, at, 1, 4, kind))


def u64_at(body, at, kind):
    return be64(entries(body, at, 1, 8, kind))


def children(data):
    """mov.rs:693-719."""
    atoms = []
    rest = data
    while rest:
        if len(rest) < 8:
            if any(rest):
                bail("truncated atom header")
            break
        kind = bytes(rest[4:8])
        size32 = be32(rest)
        if size32 == 1:
            header, size = 16, u64_at(rest, 8, kind)
        elif size32 == 0:
            header, size = 8, len(rest)
        else:
            header, size = 8, size32
        if not (header <= size <= len(rest)):
            bail(f"truncated `{name(kind)}` atom: it claims {size} bytes and {len(rest)} are left")
        atoms.append((kind, rest[header:size]))
        rest = rest[size:]
    return atoms


def find(atoms, kind):
    for k, body in atoms:
        if k == kind:
            return body
    return None


def require(atoms, kind, parent):
    """mov.rs:721-727."""
    body = find(atoms, kind)
    if body is None:
        bail(f"{parent} has no `{name(kind)}` atom")
    return body


def times(body, kind):
    """mov.rs:641-649."""
    if not body:
        bail(f"truncated `{name(kind)}` atom")
    if body[0] == 0:
        return u32_at(body, 4, kind), u32_at(body, 12, kind)
    if body[0] == 1:
        return u64_at(body, 4, kind), u32_at(body, 20, kind)
    bail(f"`{name(kind)}` version {body[0]} is not supported")


def sample_count(count):
    """mov.rs:749-755."""
    if count > MAX_SAMPLES:
        bail(f"the video track lists {count} samples, more than the {MAX_SAMPLES} this reader "
             "accepts")
    return count


ORIENTATIONS = {  # sign pattern of a b c d -> EXIF value (mov.rs:672-680)
    (1, 0, 0, 1): 1, (-1, 0, 0, 1): 2, (-1, 0, 0, -1): 3, (1, 0, 0, -1): 4,
    (0, 1, 1, 0): 5, (0, 1, -1, 0): 6, (0, -1, -1, 0): 7, (0, -1, 1, 0): 8,
}


def orientation(tkhd, warnings):
    """mov.rs:651-692."""
    if not tkhd:
        bail("truncated `tkhd` atom")
    if tkhd[0] == 0:
        at = 40
    elif tkhd[0] == 1:
        at = 52
    else:
        bail(f"`tkhd` version {tkhd[0]} is not supported")

    def entry(i):
        v = u32_at(tkhd, at + 4 * i, b"tkhd")
        return (v - (1 << 32) if v >= 1 << 31 else v) / 65536.0

    a, b, c, d = entry(0), entry(1), entry(3), entry(4)
    sign = lambda v: (v > 0) - (v < 0)  # noqa: E731
    key = (sign(a), sign(b), sign(c), sign(d))
    if key in ORIENTATIONS:
        return ORIENTATIONS[key]
    warnings.append(f"the video track's display matrix [{fmt_f64(a)} {fmt_f64(b)}; {fmt_f64(c)} "
                    f"{fmt_f64(d)}] is not a quarter turn or mirror, so the frames are shown as "
                    "stored")
    return 1


def common_duration(durations):
    """mov.rs:507-521."""
    body = durations[:max(len(durations) - 1, 1)]
    counts = {}
    for d in body:
        counts[d] = counts.get(d, 0) + 1
    if not counts:
        return 1, False
    best = max(counts.values())
    mode = min(d for d, c in counts.items() if c == best)
    return mode, len(counts) > 1


def to_i32(v):
    return v - (1 << 32) if v >= 1 << 31 else v


def to_i64(v):
    return v - (1 << 64) if v >= 1 << 63 else v


def edit(elst, durations, timescale, movie_timescale, warnings):
    """mov.rs:522-596: the samples the first edit that shows media presents."""
    if not elst:
        bail("truncated `elst` atom")
    version = elst[0]
    count = u32_at(elst, 4, b"elst")
    if version == 0:
        t = entries(elst, 8, count, 12, b"elst")
        edits = [(be32(t, 12 * i), to_i32(be32(t, 12 * i + 4)), to_i32(be32(t, 12 * i + 8)))
                 for i in range(count)]
    elif version == 1:
        t = entries(elst, 8, count, 20, b"elst")
        edits = [(be64(t, 20 * i), to_i64(be64(t, 20 * i + 8)), to_i32(be32(t, 20 * i + 16)))
                 for i in range(count)]
    else:
        bail(f"`elst` version {version} is not supported")
    if len(edits) > 1:
        warnings.append(f"the video track's edit list has {len(edits)} edits; only the first that "
                        "shows media is honoured")
    if edits and edits[0][1] < 0:
        warnings.append("the video track's edit list starts with an empty edit (a delay), which "
                        "is ignored")
    shown = next((e for e in edits if e[1] >= 0), None)
    if shown is None:
        if edits:
            warnings.append("the video track's edit list shows no media; every sample is "
                            "presented")
        return 0, len(durations)
    duration, start, rate = shown
    if rate != 0x10000:
        warnings.append(f"the video track's edit plays at rate {fmt_f64(rate / 65536.0)}; frames "
                        "are read at normal speed")
    start = start * movie_timescale
    if duration == 0:
        end = 1 << 128
    elif movie_timescale == 0:
        bail("the movie's timescale is 0")
    else:
        end = start + duration * timescale
    time = 0
    first = last = len(durations)
    for i, d in enumerate(durations):
        if time >= end:
            last = i
            break
        if time >= start and first == len(durations):
            first = i
        time += d * movie_timescale
    return first, max(last, first)


def utf8_lossy(b):
    return bytes(b).decode("utf-8", errors="replace")


def text_metadata(meta):
    """mov.rs:597-639."""
    if len(meta) >= 8 and meta[4:8] == b"hdlr":
        body = meta
    elif len(meta) >= 4:
        body = meta[4:]
    else:
        bail("truncated `meta` atom")
    atoms = children(body)
    keys, ilst = find(atoms, b"keys"), find(atoms, b"ilst")
    if keys is None or ilst is None:
        return []
    count = u32_at(keys, 4, b"keys")
    names = []
    rest = keys[8:]
    for _ in range(count):
        if len(rest) < 4:
            bail("truncated `keys` atom")
        size = be32(rest)
        stop = max(size, 8)
        if stop > len(rest):
            bail("truncated `keys` atom")
        names.append(utf8_lossy(rest[8:stop]))
        rest = rest[stop:]
    items = []
    for index, item in children(ilst):
        i = be32(index) - 1
        if i < 0 or i >= len(names):
            continue
        for kind, data in children(item):
            if kind == b"data" and len(data) >= 8 and be32(data) == 1:
                items.append((names[i], utf8_lossy(data[8:]).rstrip("\0")))
    return items


def video_track(trak):
    """mov.rs:303-341."""
    atoms = children(trak)
    tkhd = require(atoms, b"tkhd", "a track")
    if u32_at(tkhd, 0, b"tkhd") & 1 == 0:
        return None
    mdia = children(require(atoms, b"mdia", "a track"))
    handler = u32_at(require(mdia, b"hdlr", "a track's media"), 8, b"hdlr")
    if handler.to_bytes(4, "big") != b"vide":
        return None
    _, timescale = times(require(mdia, b"mdhd", "the video track"), b"mdhd")
    if timescale == 0:
        bail("the video track's timescale is 0")
    minf = children(require(mdia, b"minf", "the video track"))
    stbl = children(require(minf, b"stbl", "the video track"))
    stsd = require(stbl, b"stsd", "the video track")
    if u32_at(stsd, 4, b"stsd") == 0:
        bail("the video track has no sample description")
    fourcc = u32_at(stsd, 12, b"stsd").to_bytes(4, "big")
    elst = None
    edts = find(atoms, b"edts")
    if edts is not None:
        elst = find(children(edts), b"elst")
    return {"fourcc": fourcc, "timescale": timescale, "tkhd": tkhd, "elst": elst, "stbl": stbl}


def sizes_of(stbl):
    """mov.rs:385-419."""
    stsz = find(stbl, b"stsz")
    if stsz is not None:
        size = u32_at(stsz, 4, b"stsz")
        count = sample_count(u32_at(stsz, 8, b"stsz"))
        if size != 0:
            return [size] * count
        t = entries(stsz, 12, count, 4, b"stsz")
        return [be32(t, 4 * i) for i in range(count)]
    stz2 = find(stbl, b"stz2")
    if stz2 is None:
        bail("the video track has no sample sizes (`stsz` or `stz2`)")
    if len(stz2) < 8:
        bail("truncated `stz2` atom")
    field = stz2[7]
    count = sample_count(u32_at(stz2, 8, b"stz2"))
    if field == 4:
        t = entries(stz2, 12, (count + 1) // 2, 1, b"stz2")
        return [(t[i // 2] & 15) if i & 1 else (t[i // 2] >> 4) for i in range(count)]
    if field == 8:
        return list(entries(stz2, 12, count, 1, b"stz2"))
    if field == 16:
        t = entries(stz2, 12, count, 2, b"stz2")
        return [int.from_bytes(t[2 * i:2 * i + 2], "big") for i in range(count)]
    bail(f"`stz2` field size {field} is not 4, 8 or 16 bits")


def durations_of(stbl, samples):
    """mov.rs:420-438."""
    stts = require(stbl, b"stts", "the video track")
    count = u32_at(stts, 4, b"stts")
    t = entries(stts, 8, count, 8, b"stts")
    total = sum(be32(t, 8 * i) for i in range(count))
    if total != samples:
        bail(f"the video track's sample table disagrees: `stts` times {total} samples, the sizes "
             f"list {samples}")
    out = []
    for i in range(count):
        duration = be32(t, 8 * i + 4)
        if duration == 0:
            bail("the video track has a sample of duration 0")
        out.extend([duration] * be32(t, 8 * i))
    return out


def locate(stbl, sizes, length):
    """mov.rs:439-505."""
    stco = find(stbl, b"stco")
    if stco is not None:
        count = u32_at(stco, 4, b"stco")
        t = entries(stco, 8, count, 4, b"stco")
        chunks = [be32(t, 4 * i) for i in range(count)]
    else:
        co64 = find(stbl, b"co64")
        if co64 is None:
            bail("the video track has no chunk offsets (`stco` or `co64`)")
        count = u32_at(co64, 4, b"co64")
        t = entries(co64, 8, count, 8, b"co64")
        chunks = [be64(t, 8 * i) for i in range(count)]
    stsc = require(stbl, b"stsc", "the video track")
    count = u32_at(stsc, 4, b"stsc")
    t = entries(stsc, 8, count, 12, b"stsc")
    runs = [(be32(t, 12 * i), be32(t, 12 * i + 4)) for i in range(count)]
    ok = (bool(runs) and runs[0][0] == 1 and all(a[0] < b[0] for a, b in zip(runs, runs[1:]))
          and runs[-1][0] <= len(chunks))
    if not ok:
        bail(f"the video track's `stsc` does not list runs of its {len(chunks)} chunks from chunk "
             "1 in order")
    samples = []
    it = iter(sizes)
    for i, (first, per_chunk) in enumerate(runs):
        end = runs[i + 1][0] - 1 if i + 1 < len(runs) else len(chunks)
        for chunk in chunks[first - 1:end]:
            offset = chunk
            for _ in range(per_chunk):
                size = next(it, None)
                if size is None:
                    bail("the video track's sample table disagrees: its chunks hold more than "
                         f"the {len(sizes)} samples it sizes")
                stop = min(offset + size, (1 << 64) - 1)
                if stop > length:
                    bail(f"sample {len(samples)} (bytes {offset}..{stop}) lies beyond the end of "
                         f"the file at {length} bytes (a truncated file?)")
                samples.append((offset, size))
                offset = stop
    if len(samples) != len(sizes):
        bail(f"the video track's sample table disagrees: its chunks hold {len(samples)} of the "
             f"{len(sizes)} samples it sizes")
    return samples


def read_track(video, length, movie_timescale):
    """Video::read (mov.rs:344-378)."""
    sizes = sizes_of(video["stbl"])
    if not sizes:
        bail("the video track has no samples")
    durations = durations_of(video["stbl"], len(sizes))
    samples = locate(video["stbl"], sizes, length)
    mode, variable = common_duration(durations)
    divisor = math.gcd(video["timescale"], mode)
    warnings = []
    orient = orientation(video["tkhd"], warnings)
    if video["elst"] is not None:
        first, last = edit(video["elst"], durations, video["timescale"], movie_timescale,
                           warnings)
    else:
        first, last = 0, len(samples)
    if first >= last:
        bail(f"the video track's edit list presents none of its {len(samples)} samples")
    return {"fourcc": video["fourcc"], "fps": (video["timescale"] // divisor, mode // divisor),
            "variable_durations": variable, "samples": samples[first:last],
            "creation_time": None, "orientation": orient, "warnings": warnings, "metadata": []}


def parse_movie(moov, length, accept):
    """mov.rs:259-292."""
    atoms = children(moov)
    mvhd = require(atoms, b"mvhd", "the movie")
    creation_time, movie_timescale = times(mvhd, b"mvhd")
    rejected = None
    for kind, trak in atoms:
        if kind != b"trak":
            continue
        video = video_track(trak)
        if video is None:
            continue
        if not accept(video["fourcc"]):
            if rejected is None:
                rejected = video["fourcc"]
            continue
        track = read_track(video, length, movie_timescale)
        track["creation_time"] = creation_time or None
        meta = find(atoms, b"meta")
        if meta is not None:
            try:
                track["metadata"] = text_metadata(meta)
            except MovError as e:
                track["warnings"].append(f"the movie's metadata is unreadable: {e}")
        return track
    if rejected is not None:
        bail(f"the video track is `{name(rejected)}`, not a format this reader accepts")
    bail("the movie has no enabled video track")


def read_moov(data):
    """mov.rs:205-255 on an in-memory file."""
    length = len(data)
    position = 0
    while position < length:
        available = min(length - position, 16)
        if available < 8:
            bail(f"truncated atom header at byte {position}")
        header = data[position:position + available]
        kind = bytes(header[4:8])
        if position == 0 and kind not in FIRST_ATOMS:
            bail(f"not a QuickTime movie: it starts with `{name(kind)}`, not `ftyp` or `moov`")
        size32 = be32(header)
        if size32 == 1:
            if available != 16:
                bail(f"truncated `{name(kind)}` atom header")
            header_size, size = 16, be64(header, 8)
        elif size32 == 0:
            header_size, size = 8, length - position
        else:
            header_size, size = 8, size32
        if size < header_size:
            bail(f"invalid `{name(kind)}` atom size {size} at byte {position}")
        if size > length - position:
            bail(f"truncated `{name(kind)}` atom at byte {position}: it claims {size} bytes and "
                 f"the file has {length - position} left (an interrupted recording?)")
        if kind == b"moov":
            body = size - header_size
            if body > MAX_MOOV:
                bail(f"the `moov` atom is {body} bytes, more than the {MAX_MOOV} this reader "
                     "accepts")
            return data[position + header_size:position + size]
        position += size
    if length == 0:
        bail("not a QuickTime movie: the file is empty")
    bail("the movie has no `moov` atom, so its frames cannot be found (an interrupted "
         "recording?)")


def is_prores_raw_fourcc(fourcc):
    """lib.rs:72-75."""
    return bytes(fourcc).lower() in (b"aprn", b"aprh")


def read_video_track(data, accept=is_prores_raw_fourcc):
    """read_video_track (mov.rs:171-176) over the file's bytes."""
    data = bytes(data)
    return parse_movie(read_moov(data), len(data), accept)
