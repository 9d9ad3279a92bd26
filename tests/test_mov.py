"""The movie sample-table reader (include/proresraw_mov.h) against the crate's
own movie tests (src/mov.rs:1013-1303, restated here on the same hand-built
movies) and against the oracle restatement (oracle/mov_oracle.py) on random
and damaged movies. Host code only: no GPU is needed."""
import os
import random

import numpy as np
import pytest

from oracle import mov_oracle as O
from proresraw_rs_amd import mov as M
from proresraw_rs_amd.mov import Orientation, TrackSpec

ONE = 0x10000


def as_dict(t: M.Track):
    return {"fourcc": t.fourcc, "fps": t.fps, "variable_durations": t.variable_durations,
            "samples": [(s.offset, s.size) for s in t.samples],
            "creation_time": t.creation_time, "orientation": int(t.orientation),
            "warnings": t.warnings, "metadata": t.metadata}


def read_both(tmp_path, data, accept=None, name="hand.mov"):
    """The track (or error message) from the library and from the oracle,
    which must agree exactly."""
    path = tmp_path / name
    path.write_bytes(data)
    try:
        got = as_dict(M.read_video_track(str(path), accept))
    except M.MovError as e:
        got = ("error", str(e))
    try:
        want = O.read_video_track(data, accept or O.is_prores_raw_fourcc)
    except O.MovError as e:
        want = ("error", str(e))
    assert got == want
    return got, path


def track(tmp_path, specs, creation=0, accept=None, meta=None):
    got, path = read_both(tmp_path, M.build_movie(specs, creation, meta), accept)
    assert not isinstance(got, tuple), got
    return got, path


def message(tmp_path, data, accept=None):
    got, _ = read_both(tmp_path, data, accept)
    assert isinstance(got, tuple), "the movie should be rejected"
    return got[1]


def assert_samples(path, t, first, sizes):
    """Sample i's bytes are `size` copies of its index (mov.rs:791-806)."""
    assert [s for _, s in t["samples"]] == list(sizes)
    for i, (offset, size) in enumerate(t["samples"]):
        assert M.read_at(str(path), offset, size) == bytes([first + i]) * size, f"sample {i}"


def test_chunks_of_several_samples_resolve_after_wide(tmp_path):
    spec = TrackSpec()
    t, path = track(tmp_path, [spec], creation=3_000_000_000)
    assert t["fourcc"] == b"aprh"
    assert t["fps"] == (30, 1)
    assert not t["variable_durations"]
    assert t["creation_time"] == 3_000_000_000
    assert_samples(path, t, 0, spec.sizes)


def test_stco_and_co64_agree(tmp_path):
    a, _ = track(tmp_path, [TrackSpec()])
    b, _ = track(tmp_path, [TrackSpec(co64=True)])
    assert a == b


def test_compact_sizes_of_every_field_width(tmp_path):
    for table in ("stz2-4", "stz2-8", "stz2-16"):
        spec = TrackSpec(size_table=table)
        t, path = track(tmp_path, [spec])
        assert_samples(path, t, 0, spec.sizes)


def test_edit_list_skips_samples_before_the_first_edit(tmp_path):
    spec = TrackSpec(edits=[(80, 15)])
    t, path = track(tmp_path, [spec])
    assert t["warnings"] == []
    assert_samples(path, t, 1, spec.sizes[1:])
    spec = TrackSpec(edits=[(20, 20)])
    t, path = track(tmp_path, [spec])
    assert_samples(path, t, 1, spec.sizes[1:2])


def test_odd_edit_lists_use_the_first_non_empty_edit_and_warn(tmp_path):
    spec = TrackSpec(edits=[(40, -1), (40, 40), (20, 0)])
    t, path = track(tmp_path, [spec])
    assert_samples(path, t, 2, spec.sizes[2:4])
    assert len(t["warnings"]) == 2, t["warnings"]
    assert "3 edits" in t["warnings"][0]
    assert "empty edit" in t["warnings"][1]
    assert "presents none" in message(tmp_path, M.build_movie([TrackSpec(edits=[(20, 1000)])]))


def test_edit_rates_and_edits_showing_no_media_warn(tmp_path):
    t, _ = track(tmp_path, [TrackSpec(edits=[(0, 20, 0x8000)])])
    assert t["warnings"] == ["the video track's edit plays at rate 0.5; frames are read at "
                             "normal speed"]
    assert len(t["samples"]) == 4
    t, _ = track(tmp_path, [TrackSpec(edits=[(40, -1)])])
    assert len(t["samples"]) == 5
    assert t["warnings"][-1] == ("the video track's edit list shows no media; every sample is "
                                 "presented")


def test_the_most_common_duration_sets_the_rate(tmp_path):
    t, _ = track(tmp_path, [TrackSpec(durations=[20, 20, 21, 20, 7])])
    assert t["fps"] == (30, 1)
    assert t["variable_durations"]
    t, _ = track(tmp_path, [TrackSpec(durations=[20, 20, 20, 20, 7])])
    assert not t["variable_durations"]
    # Ties go to the shorter duration (mov.rs:507-521).
    t, _ = track(tmp_path, [TrackSpec(durations=[25, 24, 25, 24, 9])])
    assert t["fps"] == (25, 1)


def test_other_tracks_and_disabled_video_are_skipped(tmp_path):
    sound = TrackSpec(handler=b"soun", fourcc=b"lpcm")
    timecode = TrackSpec(handler=b"tmcd", fourcc=b"tmcd")
    disabled = TrackSpec(enabled=False, fourcc=b"aprn")
    video = TrackSpec(sizes=[9, 8, 7, 6, 5])
    t, _ = track(tmp_path, [sound, timecode, disabled, video])
    assert t["fourcc"] == b"aprh"
    assert [s for _, s in t["samples"]] == video.sizes
    assert message(tmp_path, M.build_movie([sound])) == "the movie has no enabled video track"


def test_display_matrices_set_the_orientation(tmp_path):
    cases = [
        ((ONE, 0, 0, ONE), Orientation.Identity),
        ((-ONE, 0, 0, ONE), Orientation.MirrorHorizontal),
        ((-ONE, 0, 0, -ONE), Orientation.Rotate180),
        ((ONE, 0, 0, -ONE), Orientation.MirrorVertical),
        ((0, ONE, ONE, 0), Orientation.Transpose),
        ((0, ONE, -ONE, 0), Orientation.Rotate90),
        ((0, -ONE, -ONE, 0), Orientation.Transverse),
        ((0, -ONE, ONE, 0), Orientation.Rotate270),
        ((-2 * ONE, 0, 0, -ONE // 2), Orientation.Rotate180),  # scale is ignored
    ]
    for matrix, orientation in cases:
        for version in (0, 1):
            t, _ = track(tmp_path, [TrackSpec(matrix=matrix, tkhd_version=version)])
            assert t["orientation"] == int(orientation), (matrix, version)
            assert t["warnings"] == []


def test_orientations_have_their_exif_values():
    for i, o in enumerate(Orientation):
        assert o.exif() == i + 1
        assert o.swaps_dimensions() == (i >= 4)


def test_other_display_matrices_show_frames_as_stored(tmp_path):
    t, _ = track(tmp_path, [TrackSpec(matrix=(46341, 46341, -46341, 46341))])
    assert t["orientation"] == int(Orientation.Identity)
    assert len(t["warnings"]) == 1
    assert t["warnings"][0] == ("the video track's display matrix [0.7071075439453125 "
                                "0.7071075439453125; -0.7071075439453125 0.7071075439453125] is "
                                "not a quarter turn or mirror, so the frames are shown as stored")


def test_other_codecs_are_named(tmp_path):
    text = message(tmp_path, M.build_movie([TrackSpec(fourcc=b"apch")]))
    assert "the video track is `apch`" in text
    # A caller's own acceptance function takes it.
    t, _ = track(tmp_path, [TrackSpec(fourcc=b"apch")], accept=lambda f: f == b"apch")
    assert t["fourcc"] == b"apch"


def test_broken_files_are_described(tmp_path):
    data = M.build_movie([TrackSpec()])
    assert "truncated `moov`" in message(tmp_path, data[:-10])
    inner = bytearray(data)
    stsz = inner.find(b"stsz") - 4
    inner[stsz:stsz + 4] = (1000).to_bytes(4, "big")
    assert "truncated `stsz`" in message(tmp_path, bytes(inner))
    moov = data.rfind(b"moov") - 4
    assert "no `moov`" in message(tmp_path, data[:moov])
    beyond = bytearray(data)
    stco = beyond.find(b"stco") + 12
    beyond[stco:stco + 4] = len(data).to_bytes(4, "big")
    assert "beyond the end of the file" in message(tmp_path, bytes(beyond))
    assert "not a QuickTime movie" in message(tmp_path, b"\x89PNG\r\n\x1a\n\0\0\0\rIHDR")
    assert message(tmp_path, b"\x89PNG\r\n\x1a\n\0\0\0\rIHDR") == (
        "not a QuickTime movie: it starts with `\\r\\n\\x1a\\n`, not `ftyp` or `moov`")
    assert "not a QuickTime movie" in message(tmp_path, b"")


def meta_children(values=((1, b"Apple"), (0, bytes([1, 2, 3])))):
    keys = (0).to_bytes(4, "big") + (2).to_bytes(4, "big")
    for key in (b"com.apple.quicktime.make", b"com.example.table"):
        keys += (len(key) + 8).to_bytes(4, "big") + b"mdta" + key
    ilst = b""
    for index, (kind, value) in enumerate(values, start=1):
        data = M._atom(b"data", kind.to_bytes(4, "big") + bytes(4) + value)
        ilst += M._atom(index.to_bytes(4, "big"), data)
    return (M._full(b"hdlr", 0, 0, b"\0\0\0\0mdta" + bytes(12)) + M._atom(b"keys", keys)
            + M._atom(b"ilst", ilst))


def test_text_metadata_in_both_meta_forms(tmp_path):
    expected = [("com.apple.quicktime.make", "Apple")]
    children = meta_children()
    t, _ = track(tmp_path, [TrackSpec()], meta=children)
    assert t["metadata"] == expected
    t, _ = track(tmp_path, [TrackSpec()], meta=bytes(4) + children)
    assert t["metadata"] == expected
    path = tmp_path / "meta.mov"
    path.write_bytes(M.build_movie([TrackSpec()], meta=children))
    assert M.read_video_track(str(path)).text(["com.apple.proapps.manufacturer",
                                "com.apple.quicktime.make"]) == "Apple"
    # Invalid UTF-8 becomes U+FFFD per ill-formed subpart; trailing NULs go.
    odd = meta_children(((1, b"A\xff\xe2\x82B\xf0\x9f\x98\x80\0\0"), (1, b"\xed\xa0\x80z")))
    t, _ = track(tmp_path, [TrackSpec()], meta=odd)
    assert t["metadata"][0][1] == "A��B\U0001F600"
    # A damaged metadata atom costs the tags, not the frames.
    t, _ = track(tmp_path, [TrackSpec()], meta=bytes(4) + b"\0\0\0\x30keys")
    assert t["metadata"] == [] and t["warnings"][0].startswith(
        "the movie's metadata is unreadable: truncated `keys` atom")


def test_fourccs():
    for f, want in ((b"aprn", True), (b"APRH", True), (b"aPrN", True), (b"apch", False)):
        assert M.is_prores_raw_fourcc(f) == want == O.is_prores_raw_fourcc(f)


def test_read_samples_gathers_every_sample(tmp_path):
    spec = TrackSpec(sizes=[11, 1, 300, 7, 64, 2], durations=[1001] * 6, timescale=24000,
                     chunks=[3, 3])
    t, path = track(tmp_path, [spec])
    samples = [M.Sample(o, s) for o, s in t["samples"]]
    offsets, total = M.packet_layout(samples)
    assert all(int(o) % 4 == 0 for o in offsets)
    buf = np.zeros(total, dtype=np.uint8)
    M.read_samples(str(path), samples, buf, offsets, threads=3)
    for i, s in enumerate(samples):
        got = bytes(buf[int(offsets[i]):int(offsets[i]) + s.size])
        assert got == bytes([i]) * s.size
    # The failing sample with the lowest index names the error.
    bad = samples[:2] + [M.Sample(10 ** 6, 8), M.Sample(2 * 10 ** 6, 8)]
    with pytest.raises(M.MovError, match=r"^reading 8 bytes at byte 1000000: "):
        M.read_samples(str(path), bad, buf, offsets[:4], threads=4)


def random_spec(rng):
    n = rng.randint(1, 12)
    sizes = [rng.randint(0, 40) for _ in range(n)]
    chunks, left = [], n
    while left:
        k = rng.randint(1, left)
        chunks.append(k)
        left -= k
    base = rng.choice([20, 1001, 512])
    durations = [base if rng.random() < 0.8 else rng.randint(1, 2 * base) for _ in range(n)]
    table = rng.choice(["stsz", "stz2-8", "stz2-16"] + (["stz2-4"] if max(sizes) < 16 else []))
    edits = []
    for _ in range(rng.choice([0, 0, 1, 2, 3])):
        edits.append((rng.choice([0, 20, 40, 100]), rng.choice([-1, 0, 15, 20, 60, 5000]),
                      rng.choice([0x10000, 0x10000, 0x8000, 0x20000])))
    m = rng.choice([(ONE, 0, 0, ONE), (0, ONE, -ONE, 0), (-ONE, 0, 0, -ONE),
                    (rng.randint(-ONE, ONE), rng.randint(-ONE, ONE), rng.randint(-ONE, ONE),
                     rng.randint(-ONE, ONE))])
    return TrackSpec(fourcc=rng.choice([b"aprh", b"aprn", b"APRN", b"apch"]),
                     enabled=rng.random() < 0.9,
                     handler=b"vide" if rng.random() < 0.9 else b"soun",
                     timescale=rng.choice([600, 24000, 30000, 1]), durations=durations,
                     sizes=sizes, size_table=table, chunks=chunks, co64=rng.random() < 0.5,
                     edits=edits, tkhd_version=rng.randint(0, 1), matrix=m)


def test_random_movies_match_the_oracle(tmp_path):
    rng = random.Random(0x6d6f76)
    outcomes = set()
    for _ in range(300):
        specs = [random_spec(rng) for _ in range(rng.randint(1, 3))]
        meta = meta_children() if rng.random() < 0.3 else None
        got, _ = read_both(tmp_path, M.build_movie(specs, rng.choice([0, 7]), meta))
        outcomes.add("error" if isinstance(got, tuple) else "track")
    assert outcomes == {"error", "track"}


def test_damaged_movies_match_the_oracle(tmp_path):
    """Every damage to the `moov` body (a changed byte, a changed size word,
    a cut file) reads the same track or the same message as the oracle."""
    rng = random.Random(0x626164)
    base = [M.build_movie([random_spec(rng)], 5, meta_children() if k % 2 else None)
            for k in range(8)]
    messages = set()
    for i in range(1500):
        data = bytearray(rng.choice(base))
        moov = data.rfind(b"moov") - 4
        kind = i % 4
        if kind == 0:
            at = rng.randrange(moov, len(data))
            data[at] = rng.randrange(256)
        elif kind == 1:
            at = rng.randrange(moov, len(data) - 4)
            word = int.from_bytes(data[at:at + 4], "big")
            value = rng.choice([0, 1, 7, 8, 0xFFFFFFFF, rng.randrange(1 << 32), word + 1, word - 1])
            data[at:at + 4] = (value & 0xFFFFFFFF).to_bytes(4, "big")
        elif kind == 2:
            data = data[:rng.randrange(0, len(data))]
        else:
            for _ in range(4):
                at = rng.randrange(moov, len(data))
                data[at] ^= 1 << rng.randrange(8)
        got, _ = read_both(tmp_path, bytes(data))
        if isinstance(got, tuple):
            messages.add(got[1].split(":")[0][:40])
    assert len(messages) > 15, messages


def test_clip_movie_of_packets_reads_back(tmp_path):
    packets = [bytes([i]) * (100 + 37 * i) for i in range(10)]
    data = M.clip_movie(packets, fps=(24000, 1001), frames_per_chunk=3)
    got, path = read_both(tmp_path, data)
    assert got["fps"] == (24000, 1001) and not got["variable_durations"]
    for (offset, size), p in zip(got["samples"], packets):
        assert M.read_at(str(path), offset, size) == p
