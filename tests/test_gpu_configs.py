"""The benchmark configurations at their own sizes, on the GPU, against the
oracle, and the interleaved develop layout:

- interleaved develop (develop_into / develop / develop_with,
  demosaic.rs:190-227) through the host entry point prr_develop_into, the
  device entry point with d_g == d_b == NULL, and a batch develop job with
  one interleaved buffer, for every stage;
- the end-to-end stream's two frame kinds in one batch: 8192x4320 v1
  alignment 4 qscale 15 and 5760x3240 v0 alignment 3 qscale 45, full
  decode + linearize + develop(Rec2020) (decoder.rs:221-307,
  demosaic.rs:250-261);
- the mixed batch's four sizes (1920x1080, 4096x2160, 5760x3240,
  4032x3024) each in every mode (full decode + develop(WhiteBalanced),
  Half, Quarter, Eighth, Half with Greens::Average; reduced.rs:147-201)
  in one batch, with the launch shape the autotuner picks for it.

Integer outputs bit-exact; develop within 1e-5 relative max-abs and, as
the kernel's arithmetic gives, bit-exact (both asserted)."""
import ctypes

import numpy as np
import pytest

import proresraw_rs_amd as P
from oracle import oracle as O
from proresraw_rs_amd import synth as S
from proresraw_rs_amd.batch import Batch
from proresraw_rs_amd.workloads import config0, seeded_curve, seeded_qmat

pytestmark = pytest.mark.gpu

DEVELOP_RTOL = 1e-5


def say(*a):
    print(*a, flush=True)


def same_floats(got, want, what):
    got = np.asarray(got, dtype=np.float32).reshape(-1)
    want = np.asarray(want, dtype=np.float32).reshape(-1)
    rel = float(np.abs(got - want).max()) / max(float(np.abs(want).max()), 1e-30)
    exact = bool(np.array_equal(got.view(np.uint32), want.view(np.uint32)))
    say(f"  {what}: relative max-abs {rel:.3g}, bit-exact {exact}")
    assert rel <= DEVELOP_RTOL, what
    assert exact, what


def interleaved_frames():
    yield "config0 1920x1080", S.encode_scene(config0())
    rng = S.Rng(17)
    for i, spec in enumerate(S.specs(rng)):
        spec.valid_tiles = 1
        pk = S.random_frame(spec, rng)
        if i % 5 == 0:
            yield f"random {i} {spec.width}x{spec.height} p{spec.pattern}", pk


def test_interleaved_develop_host_entry_point(gpu):
    """P.develop / develop_with / develop_into (prr_develop_into) against
    the oracle's interleaved develop."""
    for name, pk in interleaved_frames():
        f = P.decode_frame(pk)
        h, s, sk, mq = O.decode_frame(pk)
        np.testing.assert_array_equal(f.samples, s)
        q = h.quant_matrix[0] * mq
        for stage in (P.Stage.Demosaiced, P.Stage.WhiteBalanced, P.Stage.Rec2020):
            img = P.develop(f, stage)
            same_floats(img.data, O.develop_interleaved(h, s, O.params(h, q), int(stage)),
                        f"{name} develop {stage.name}")
            prm = P.DemosaicParams.with_white_balance(f.header, f.quantizer(), 1.7, 1.25)
            out = np.full(3 * f.width() * f.height(), -3.0, dtype=np.float32)
            P.develop_into(f, prm, stage, out)
            same_floats(out, O.develop_interleaved(h, s, O.params(h, q, 1.7, 1.25), int(stage)),
                        f"{name} develop_into {stage.name}")


def test_interleaved_develop_device_entry_point(gpu):
    """prr_develop_device with d_g == d_b == NULL writes develop_into's
    layout into d_r."""
    import torch
    from proresraw_rs_amd import _native as N
    lib, ctx = N.lib(), P.context()
    for name, pk in interleaved_frames():
        f = P.decode_frame(pk)
        h, s, sk, mq = O.decode_frame(pk)
        q = h.quant_matrix[0] * mq
        d_samples = torch.from_numpy(np.ascontiguousarray(f.samples).view(np.int16)).cuda()
        prm = P.DemosaicParams.new(f.header, f.quantizer())
        raw = prm._raw()
        out = torch.full((3 * f.width() * f.height(),), -1.0, device="cuda")
        st = torch.cuda.Stream()
        err = N.PrrError()
        for stage in (0, 1, 2):
            rc = lib.prr_develop_device(ctx.handle, ctypes.byref(f.header._raw),
                                        d_samples.data_ptr(), ctypes.byref(raw), stage,
                                        out.data_ptr(), None, None,
                                        ctypes.c_void_p(st.cuda_stream), ctypes.byref(err))
            assert rc == 0
            st.synchronize()
            same_floats(out.cpu().numpy(), O.develop_interleaved(h, s, O.params(h, q), stage),
                        f"{name} device interleaved stage {stage}")


def test_interleaved_develop_in_a_batch(gpu):
    import torch
    packets = list(interleaved_frames())
    b = Batch()
    plan = []
    for i, (name, pk) in enumerate(packets):
        stage = P.Stage(i % 3)
        plan.append((name, pk, stage,
                     b.add_full(pk, P.Idct.Apple, develop=stage, wb=(2.0, 1.6),
                                interleaved=True)))
    s = torch.cuda.Stream()
    b.run(s)
    s.synchronize()
    for name, pk, stage, slot in plan:
        h, smp, sk, mq = O.decode_frame(pk)
        q = h.quant_matrix[0] * mq
        want = O.develop_interleaved(h, smp, O.params(h, q, 2.0, 1.6), int(stage))
        same_floats(slot.planes[0].cpu().numpy(), want, f"batch {name} {stage.name}")


def check_full(name, pk, slot, stage, wb=None):
    st = slot.status_host()
    assert st[0] == 0xFFFFFFFF and st[3] == 0, (name, st)
    h, s, sk, mq = O.decode_frame(pk)
    assert slot.skipped_tiles() == sk and st[1] == mq and st[2] == len(sk), name
    np.testing.assert_array_equal(slot.samples.cpu().numpy().view(np.uint16), s)
    if slot.linear is not None:
        np.testing.assert_array_equal(slot.linear.cpu().numpy().view(np.uint16),
                                      O.linearize(h, s))
    q = h.quant_matrix[0] * mq
    prm = O.params(h, q) if wb is None else O.params(h, q, *wb)
    for k, (got, want) in enumerate(zip(slot.planes, O.develop_planar(h, s, prm, int(stage)))):
        same_floats(got.cpu().numpy(), want, f"{name} plane {k}")


def test_end_to_end_stream_frame_kinds_in_one_batch(gpu):
    """One 8192x4320 v1 a4 q15 frame (custom curve) and one 5760x3240 v0 a3
    q45 frame, RGGB, decoded together in one persistent launch with the
    linearize fused, developed to Rec.2020."""
    import torch
    kinds = [("8192x4320 v1 a4 q15", S.scene(8192, 4320, 0, seed=3000, version=1, alignment=4,
                                             qscale=(15, 15), curve=seeded_curve(40))),
             ("5760x3240 v0 a3 q45", S.scene(5760, 3240, 0, seed=3001, version=0, alignment=3,
                                             qscale=(45, 45)))]
    b = Batch()
    plan = []
    for name, sc in kinds:
        say(f"encode {name}")
        pk = S.encode_scene(sc)
        plan.append((name, pk, b.add_full(pk, P.Idct.Apple, develop=P.Stage.Rec2020,
                                          linear=True)))
    s = torch.cuda.Stream()
    b.run(s)
    s.synchronize()
    say("batch", b.info())
    for name, pk, slot in plan:
        say(f"check {name}")
        check_full(name, pk, slot, P.Stage.Rec2020)


def test_mixed_batch_sizes_in_every_mode_autotuned(gpu, tmp_path):
    """The mixed batch's four sizes, each as a full decode +
    develop(WhiteBalanced), Half, Quarter, Eighth and Half with
    Greens::Average, in one batch run with the shape the autotuner picks."""
    import torch
    sizes = [(1920, 1080), (4096, 2160), (5760, 3240), (4032, 3024)]
    packets = {}
    for i, (w, h) in enumerate(sizes):
        for j in range(2):
            version = (i + j) % 2
            k = 2 * i + j
            sc = S.scene(w, h, 3 * (k % 2), seed=1200 + k, version=version, alignment=k % 5,
                         qscale=(15 + (k * 7) % 31, 15 + (k * 7) % 31),
                         qmat=seeded_qmat(k) if k % 3 == 0 else None,
                         curve=seeded_curve(k) if version == 1 else None)
            say(f"encode {w}x{h} v{version} a{k % 5}")
            packets[(i, j)] = S.encode_scene(sc)
    modes = [("develop", None), ("reduced", (P.Scale.Half, P.Greens.First)),
             ("reduced", (P.Scale.Quarter, P.Greens.First)),
             ("reduced", (P.Scale.Eighth, P.Greens.First)),
             ("reduced", (P.Scale.Half, P.Greens.Average))]
    b = Batch()
    plan = []
    for i, (w, h) in enumerate(sizes):
        for m, (kind, arg) in enumerate(modes):
            pk = packets[(i, m % 2)]
            if kind == "develop":
                slot = b.add_full(pk, P.Idct.Apple, develop=P.Stage.WhiteBalanced, wb=(2.0, 1.6))
            else:
                slot = b.add_reduced(pk, *arg)
            plan.append((f"{w}x{h} {kind} {arg}", pk, kind, arg, slot))
    s = torch.cuda.Stream()
    tuned = b.autotune(s, cache=str(tmp_path / "tuning.json"))
    say(f"autotuned spw {tuned['spw']} chunk {tuned['chunk']} waves {tuned['waves']}")
    for _, _, _, _, slot in plan:
        for t in (slot.samples, slot.rgb):
            if t is not None:
                t.fill_(0x5A5A)
        for p in slot.planes or ():
            p.fill_(-1.0)
    b.run(s)
    s.synchronize()
    for name, pk, kind, arg, slot in plan:
        say(f"check {name}")
        if kind == "develop":
            check_full(name, pk, slot, P.Stage.WhiteBalanced, wb=(2.0, 1.6))
            continue
        scale, greens = arg
        st = slot.status_host()
        assert st[0] == 0xFFFFFFFF and st[3] == 0
        h, rgb, sk, mq = O.decode_frame_reduced(pk, int(scale), greens == P.Greens.Average)
        assert slot.skipped_tiles() == sk and st[1] == mq
        np.testing.assert_array_equal(slot.rgb.cpu().numpy().view(np.uint16), rgb)
