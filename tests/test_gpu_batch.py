"""Batched decode: every frame of a mixed batch (sizes, versions,
alignments, qscales, patterns, IDCTs, reduced scales, develop stages)
against the oracle, bit-exact for integer outputs and within 1e-5 for
develop; then again through a captured graph replayed several times."""
import numpy as np
import pytest

import proresraw_rs_amd as P
from oracle import oracle as O
from proresraw_rs_amd import synth as S
from proresraw_rs_amd.batch import Batch
from proresraw_rs_amd.workloads import seeded_curve, seeded_qmat

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def mixed_packets():
    out = []
    rng = S.Rng(31)
    for i, spec in enumerate(S.specs(rng)):
        out.append(("random%d" % i, S.random_frame(spec, rng)))
    sizes = [(192, 108), (256, 128), (336, 200), (400, 150)]
    for i in range(12):
        w, h = sizes[i % 4]
        version = i % 2
        sc = S.scene(w, h, 3 * (i % 2), seed=100 + i, version=version, alignment=i % 5,
                     qscale=(15, 45), qmat=seeded_qmat(i) if i % 3 == 0 else None,
                     curve=seeded_curve(i) if version == 1 else None,
                     corrupt_fraction=0.05 if i % 4 == 1 else 0.0, threads=4)
        out.append(("scene%d" % i, S.encode_scene(sc)))
    return out


def build(packets, **cfg):
    b = Batch(**cfg)
    plan = []
    for i, (name, pk) in enumerate(packets):
        k = i % 6
        if k == 0:
            plan.append((name, pk, "full", P.Idct.Apple, None, b.add_full(pk, P.Idct.Apple, linear=True)))
        elif k == 1:
            plan.append((name, pk, "full", P.Idct.Ffmpeg, None, b.add_full(pk, P.Idct.Ffmpeg)))
        elif k == 2:
            st = P.Stage(i % 3)
            plan.append((name, pk, "develop", P.Idct.Apple, st,
                         b.add_full(pk, P.Idct.Apple, develop=st, wb=(2.0, 1.6) if i % 2 else None)))
        else:
            scale = P.Scale(1 + (i % 3))
            greens = P.Greens.Average if i % 4 == 0 else P.Greens.First
            plan.append((name, pk, "reduced", scale, greens, b.add_reduced(pk, scale, greens)))
    return b, plan


def verify(plan):
    for name, pk, kind, a, c, slot in plan:
        say(f"  check {name} {kind} {a} {c}")
        st = slot.status_host()
        assert st[0] == 0xFFFFFFFF and st[3] == 0
        if kind == "reduced":
            h, rgb, sk, mq = O.decode_frame_reduced(pk, int(a), c == P.Greens.Average)
            assert slot.skipped_tiles() == sk and st[1] == mq
            np.testing.assert_array_equal(slot.rgb.cpu().numpy().view(np.uint16), rgb)
            continue
        h, s, sk, mq = O.decode_frame(pk, ffmpeg=a == P.Idct.Ffmpeg)
        assert slot.skipped_tiles() == sk and st[1] == mq
        np.testing.assert_array_equal(slot.samples.cpu().numpy().view(np.uint16), s)
        if slot.linear is not None:
            np.testing.assert_array_equal(slot.linear.cpu().numpy().view(np.uint16), O.linearize(h, s))
        if kind == "develop":
            q = h.quant_matrix[0] * mq
            prm = O.params(h, q) if slot.wb is None else O.params(h, q, *slot.wb)
            want = O.develop_planar(h, s, prm, int(c))
            for got, ref in zip(slot.planes, want):
                g = got.cpu().numpy()
                rel = float(np.abs(g - ref).max()) / max(float(np.abs(ref).max()), 1e-30)
                assert rel <= 1e-5, (name, rel)


def test_mixed_batch_matches_the_oracle():
    import torch
    packets = mixed_packets()
    b, plan = build(packets)
    for i, (name, pk, kind, a, c, slot) in enumerate(plan):
        slot.wb = (2.0, 1.6) if kind == "develop" and i % 2 else None
    s = torch.cuda.Stream()
    say("batch info before run", b.info())
    b.run(s)
    s.synchronize()
    say("batch info", b.info())
    verify(plan)


def test_batch_graph_replays_match_the_oracle():
    import torch
    packets = mixed_packets()[::2]
    b, plan = build(packets, chunk=2)
    for i, (name, pk, kind, a, c, slot) in enumerate(plan):
        slot.wb = (2.0, 1.6) if kind == "develop" and i % 2 else None
    s = torch.cuda.Stream()
    b.run(s)  # uploads the frame table
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            b.run(s)
    for rep in range(3):
        for name, pk, kind, a, c, slot in plan:
            for t in (slot.samples, slot.rgb, slot.linear):
                if t is not None:
                    t.fill_(0x5A5A)
            if slot.planes:
                for p in slot.planes:
                    p.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        say(f"replay {rep}")
        verify(plan)


def test_a_refused_frame_leaves_the_batch_as_it_was():
    """prr_batch_add checks every argument before it changes the batch: a
    develop frame with a stage out of range is refused, and the batch still
    holds and decodes exactly the frames it had."""
    import torch
    packets = mixed_packets()[:6]
    b, plan = build(packets)
    for name, pk, kind, a, c, slot in plan:
        slot.wb = None
    before = b.info()["tile_groups"]
    with pytest.raises(P.Error):
        b.add_full(packets[0][1], P.Idct.Apple, develop=7)
    assert b.info()["tile_groups"] == before and len(b.slots) == len(plan)
    s = torch.cuda.Stream()
    b.run(s)
    s.synchronize()
    verify(plan)
