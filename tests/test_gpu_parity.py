"""GPU parity: the HIP path against the oracle on the same packets.

Code values, linear values, reduced RGB and skipped tiles must be
bit-exact; develop output within 1e-5 relative max-abs (max |gpu - oracle|
over max |oracle|), the bar for floating point, and also bit-exact, which
the develop kernel's operation-for-operation arithmetic gives (asserted, so
that a change to it that moves any bit is seen)."""
import numpy as np
import pytest

import proresraw_rs_amd as P
from oracle import oracle as O
from proresraw_rs_amd import synth as S

from .helpers import config0, config1

pytestmark = pytest.mark.gpu

DEVELOP_RTOL = 1e-5


def say(*a):
    print(*a, flush=True)


def random_frames(seed, valid=False):
    rng = S.Rng(seed)
    for spec in S.specs(rng):
        spec.valid_tiles = 1 if valid else 0
        yield spec, S.random_frame(spec, rng)


@pytest.mark.parametrize("idct", [P.Idct.Apple, P.Idct.Ffmpeg])
def test_random_frames_decode_bit_exact(gpu, idct):
    for i, (spec, pk) in enumerate(random_frames(11)):
        say(f"random frame {i} {spec.width}x{spec.height} a{spec.alignment} p{spec.pattern} {idct.name}")
        f = P.decode_frame_with(pk, idct)
        h, s, sk, mq = O.decode_frame(pk, ffmpeg=idct == P.Idct.Ffmpeg)
        assert f.skipped_tiles == sk
        assert f.max_qscale == mq
        np.testing.assert_array_equal(f.samples, s)
        np.testing.assert_array_equal(f.linearize(), O.linearize(h, s))


@pytest.mark.parametrize("shift", [1, 2, 3])
@pytest.mark.parametrize("greens", [P.Greens.First, P.Greens.Average])
def test_random_frames_reduced_bit_exact(gpu, shift, greens):
    for i, (spec, pk) in enumerate(random_frames(8)):
        say(f"reduced {i} {spec.width}x{spec.height} a{spec.alignment} shift {shift} {greens.name}")
        r = P.decode_frame_reduced_with(pk, P.Scale(shift), greens)
        h, rgb, sk, mq = O.decode_frame_reduced(pk, shift, greens == P.Greens.Average)
        assert r.skipped_tiles == sk
        assert r.max_qscale == mq
        np.testing.assert_array_equal(r.samples, rgb)


def check_develop(frame, h, samples, prm, oprm, stage):
    r, g, b = O.develop_planar(h, samples, oprm, stage)
    planes = P.develop_planar(frame, prm, P.Stage(stage))
    worst = 0.0
    exact = True
    for got, want in zip(planes.planes, (r, g, b)):
        scale = max(float(np.abs(want).max()), 1e-30)
        worst = max(worst, float(np.abs(got - want).max()) / scale)
        exact &= bool(np.array_equal(got, want))
    say(f"  stage {stage}: relative max-abs {worst:.3g}, bit-exact {exact}")
    assert worst <= DEVELOP_RTOL
    assert exact
    return exact


def test_config0_decode_and_develop(gpu):
    say("config0 1920x1080 RGGB v0 a2 q15")
    pk = S.encode_scene(config0())
    f = P.decode_frame(pk)
    h, s, sk, mq = O.decode_frame(pk)
    assert f.skipped_tiles == sk == []
    np.testing.assert_array_equal(f.samples, s)
    prm = P.DemosaicParams.with_white_balance(f.header, f.quantizer(), 2.0, 1.6)
    oprm = O.params(h, h.quant_matrix[0] * mq, 2.0, 1.6)
    for stage in (0, 1, 2):
        check_develop(f, h, s, prm, oprm, stage)


@pytest.mark.parametrize("size", [(4096, 2160), (4104, 2168)])
def test_config1_both_idcts(gpu, size):
    w, hh = size
    say(f"config1 {w}x{hh} BGGR v1 a3 q45, 2% corrupt")
    pk = S.encode_scene(config1(w, hh))
    for idct in (P.Idct.Apple, P.Idct.Ffmpeg):
        f = P.decode_frame_with(pk, idct)
        h, s, sk, mq = O.decode_frame(pk, ffmpeg=idct == P.Idct.Ffmpeg)
        assert len(sk) > 0
        assert f.skipped_tiles == sk
        np.testing.assert_array_equal(f.samples, s)
        np.testing.assert_array_equal(f.linearize(), O.linearize(h, s))
    prm = P.DemosaicParams.new(f.header, f.quantizer())
    oprm = O.params(h, h.quant_matrix[0] * mq)
    check_develop(f, h, s, prm, oprm, 2)


def test_develop_random_frames_all_stages(gpu):
    for i, (spec, pk) in enumerate(random_frames(5, valid=True)):
        if i % 3:
            continue
        say(f"develop random {i} {spec.width}x{spec.height} p{spec.pattern}")
        f = P.decode_frame(pk)
        h, s, sk, mq = O.decode_frame(pk)
        q = h.quant_matrix[0] * mq
        prm = P.DemosaicParams.with_white_balance(f.header, q, 1.8, 1.3)
        oprm = O.params(h, q, 1.8, 1.3)
        for stage in (0, 1, 2):
            check_develop(f, h, s, prm, oprm, stage)


def test_develop_outside_the_in_range_division_bounds(gpu):
    """Frames whose chroma quotients may fall outside div_in_range's bounds
    take the general division: a white level of 400 (one_over_range 1/144)
    and a noise model with n1 = n2 = 0; both bit-exact against the oracle,
    as the frames inside the bounds are."""
    import dataclasses
    sc = S.scene(640, 360, 0, seed=5, version=0, alignment=1, qscale=(15, 45), white_level=400)
    pk = S.encode_scene(sc)
    f = P.decode_frame(pk)
    h, s, sk, mq = O.decode_frame(pk)
    np.testing.assert_array_equal(f.samples, s)
    assert f.header.white_level == 400
    q = h.quant_matrix[0] * mq
    prm = P.DemosaicParams.with_white_balance(f.header, q, 2.0, 1.6)
    oprm = O.params(h, q, 2.0, 1.6)
    assert prm.one_over_range > 2.0 ** -8
    for stage in (0, 1, 2):
        check_develop(f, h, s, prm, oprm, stage)
    pk0 = S.encode_scene(config0())
    f0 = P.decode_frame(pk0)
    h0, s0, sk0, mq0 = O.decode_frame(pk0)
    q0 = h0.quant_matrix[0] * mq0
    prm0 = P.DemosaicParams.new(f0.header, q0)
    prm0 = dataclasses.replace(prm0, noise=(prm0.noise[0], 0.0, 0.0))
    oprm0 = O.params(h0, q0)
    oprm0.noise[1] = 0.0
    oprm0.noise[2] = 0.0
    check_develop(f0, h0, s0, prm0, oprm0, 2)


def test_frames_beyond_the_register_scan(gpu):
    """More than 8192 tiles: the single-frame size-table scan takes more
    than one pass and the batch scan several chunks (frame decode, reduced
    decode, a batch of three frames, a tile past the packet)."""
    from proresraw_rs_amd.batch import Batch
    import torch
    sc = S.scene(2048, 1088, 1, seed=21, version=1, alignment=0, qscale=(15, 45),
                 corrupt_fraction=0.02, threads=4)
    pk = S.encode_scene(sc)
    hdr, info = P.parse_frame_info(pk)
    say(f"2048x1088 a0: {info.nb_tiles} tiles")
    assert info.nb_tiles > 8192
    f = P.decode_frame(pk)
    h, s, sk, mq = O.decode_frame(pk)
    assert f.skipped_tiles == sk and len(sk) > 0 and f.max_qscale == mq
    np.testing.assert_array_equal(f.samples, s)
    r = P.decode_frame_reduced_with(pk, P.Scale.Half, P.Greens.Average)
    h, rgb, sk2, mq2 = O.decode_frame_reduced(pk, 1, True)
    assert r.skipped_tiles == sk2 and r.max_qscale == mq2
    np.testing.assert_array_equal(r.samples, rgb)
    # A tile that runs past the packet in the scan's second pass of 8192
    # tiles (and the batch scan's fifth chunk of 2048): TileOutOfBounds
    # names it, on both paths.
    k = 8500
    bad = bytearray(pk)
    bad[info.header_end + 2 * k:info.header_end + 2 * k + 2] = b"\xff\xff"
    bad = bytes(bad)
    with pytest.raises(O.OracleError) as oe:
        O.decode_frame(bad)
    with pytest.raises(P.Error) as ge:
        P.decode_frame(bad)
    assert oe.value.code == 11 and oe.value.a == k
    assert ge.value.kind == "TileOutOfBounds" and ge.value.tile == k
    # One batch: the frame decoded whole, reduced at Half, and with the bad
    # tile, which affects no other frame.
    b = Batch()
    slot = b.add_full(pk, P.Idct.Apple)
    slot_r = b.add_reduced(pk, P.Scale.Half, P.Greens.Average)
    slot_bad = b.add_full(bad, P.Idct.Apple)
    st = torch.cuda.Stream()
    b.run(st)
    st.synchronize()
    assert slot.skipped_tiles() == sk and slot.status_host()[1] == mq
    np.testing.assert_array_equal(slot.samples.cpu().numpy().view(np.uint16), s)
    assert slot_r.skipped_tiles() == sk2 and slot_r.status_host()[1] == mq2
    np.testing.assert_array_equal(slot_r.rgb.cpu().numpy().view(np.uint16), rgb)
    assert slot.status_host()[0] == slot_r.status_host()[0] == 0xFFFFFFFF
    assert slot_bad.status_host()[0] == k


def device_decode_develop(pk, stage, wb=None, linear_input=False):
    """decode (with the fused linearize) then develop on the device through
    the C ABI; develop reads the code values or, with linear_input, the
    linearized picture."""
    import ctypes
    import torch
    from proresraw_rs_amd import _native as N
    lib, ctx = N.lib(), P.context()
    hdr, info = P.parse_frame_info(pk)
    dev = torch.device("cuda")
    padded = ((len(pk) + 3) & ~3) + 16
    packet = torch.zeros(padded, dtype=torch.uint8, device=dev)
    packet[:len(pk)] = torch.frombuffer(bytearray(pk), dtype=torch.uint8).to(dev)
    samples = torch.zeros(info.coded_width * info.coded_height, dtype=torch.int16, device=dev)
    n = hdr.width * hdr.height
    linear = torch.zeros(n, dtype=torch.int16, device=dev)
    flags = torch.zeros(max(info.nb_tiles, 1), dtype=torch.uint8, device=dev)
    status = torch.zeros(4, dtype=torch.int32, device=dev)
    planes = [torch.full((n,), -1.0, dtype=torch.float32, device=dev) for _ in range(3)]
    s = torch.cuda.Stream()
    hs = ctypes.c_void_p(s.cuda_stream)
    err = N.PrrError()
    assert lib.prr_decode_frame_device(ctx.handle, ctypes.byref(hdr._raw), ctypes.byref(info),
                                       packet.data_ptr(), len(pk), 0, samples.data_ptr(),
                                       linear.data_ptr(), flags.data_ptr(), status.data_ptr(), hs,
                                       ctypes.byref(err)) == 0
    fn = lib.prr_develop_linear_device if linear_input else lib.prr_develop_frame_device
    src = linear if linear_input else samples
    assert fn(ctx.handle, ctypes.byref(hdr._raw), src.data_ptr(), status.data_ptr(),
              1 if wb else 0, ctypes.c_float(wb[0] if wb else 0.0),
              ctypes.c_float(wb[1] if wb else 0.0), int(stage), planes[0].data_ptr(),
              planes[1].data_ptr(), planes[2].data_ptr(), hs, ctypes.byref(err)) == 0
    s.synchronize()
    return [p.cpu().numpy() for p in planes]


@pytest.mark.parametrize("size", [(4096, 2160), (4104, 2168)])
def test_develop_from_the_linearized_picture(gpu, size):
    """prr_develop_linear_device on the decode's linear output equals
    prr_develop_frame_device on the code values, bit for bit."""
    w, hh = size
    pk = S.encode_scene(config1(w, hh))
    for stage, wb in ((P.Stage.Rec2020, None), (P.Stage.WhiteBalanced, (2.0, 1.6))):
        say(f"{w}x{hh} {stage.name} wb={wb}")
        a = device_decode_develop(pk, stage, wb)
        b = device_decode_develop(pk, stage, wb, linear_input=True)
        for p, q in zip(a, b):
            np.testing.assert_array_equal(p.view(np.uint32), q.view(np.uint32))
    rng = S.Rng(9)
    for i, spec in enumerate(S.specs(rng)):
        spec.valid_tiles = 1
        rpk = S.random_frame(spec, rng)
        if i % 4:
            continue
        a = device_decode_develop(rpk, P.Stage.Demosaiced)
        b = device_decode_develop(rpk, P.Stage.Demosaiced, linear_input=True)
        for p, q in zip(a, b):
            np.testing.assert_array_equal(p.view(np.uint32), q.view(np.uint32))


@pytest.mark.parametrize("spw", [8, 16, 32, 64])
def test_every_decode_launch_shape_is_bit_exact(gpu, spw):
    """The single-frame decode gives the oracle's code values, linear values
    and skipped tiles at every streams-per-workgroup setting, on the
    4104x2168 partial-tile frame (full decode, both transforms) and on
    reduced decodes."""
    import ctypes
    from proresraw_rs_amd import _native as N
    lib, ctx = N.lib(), P.context()
    err = N.PrrError()
    assert lib.prr_set_decode_config(ctx.handle, spw, ctypes.byref(err)) == 0
    try:
        pk = S.encode_scene(config1(4104, 2168))
        for idct in (P.Idct.Apple, P.Idct.Ffmpeg):
            f = P.decode_frame_with(pk, idct)
            h, s, sk, mq = O.decode_frame(pk, ffmpeg=idct == P.Idct.Ffmpeg)
            assert f.skipped_tiles == sk and f.max_qscale == mq
            np.testing.assert_array_equal(f.samples, s)
        for shift in (1, 2, 3):
            r = P.decode_frame_reduced_with(pk, P.Scale(shift), P.Greens.Average)
            h, rgb, sk, mq = O.decode_frame_reduced(pk, shift, True)
            assert r.skipped_tiles == sk
            np.testing.assert_array_equal(r.samples, rgb)
    finally:
        assert lib.prr_set_decode_config(ctx.handle, 0, ctypes.byref(err)) == 0
