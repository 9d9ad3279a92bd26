"""The frame pipeline (prr_pipeline_*): frames of different sizes,
patterns, qscales and inverse transforms submitted back to back, decoded on one stream and
developed on another through a ring of buffer sets, give planes bit-equal to
the oracle's develop of the oracle's decode (within the 1e-5 relative bar;
exact in practice) and each frame's status."""
import numpy as np
import pytest

from oracle import oracle as O
from proresraw_rs_amd import synth as S
from proresraw_rs_amd.workloads import config1

pytestmark = pytest.mark.gpu


def test_pipelined_frames_match_the_oracle(gpu):
    import torch
    P = gpu
    from proresraw_rs_amd.pipeline import Pipeline
    scenes = [S.scene(256, 128, 0, seed=61, version=0, alignment=2, qscale=(15, 15)),
              config1(200, 120, seed=62),
              S.scene(384, 160, 3, seed=63, version=1, alignment=4, qscale=(40, 40)),  # larger
              config1(256, 136, seed=64)]
    packets = [S.encode_scene(sc) for sc in scenes]
    frames = []
    for pk in packets:
        hdr, info = P.parse_frame_info(pk)
        pad = ((len(pk) + 3) & ~3) + 16
        d = torch.zeros(pad, dtype=torch.uint8, device="cuda")
        d[:len(pk)] = torch.frombuffer(bytearray(pk), dtype=torch.uint8).cuda()
        frames.append((pk, hdr, info, d))
    pipe = Pipeline(depth=2)
    order = [0, 1, 0, 2, 3, 1, 2, 2, 0, 3]
    outs = []
    for i, f in enumerate(order):
        pk, hdr, info, d = frames[f]
        planes = [torch.full((hdr.width * hdr.height,), -7.0, device="cuda") for _ in range(3)]
        status = torch.zeros(4, dtype=torch.int32, device="cuda")
        ffmpeg = i % 3 == 1
        print(f"pipeline frame {i}: scene {f} {hdr.width}x{hdr.height} ffmpeg={ffmpeg}", flush=True)
        pipe.submit(hdr, info, d, len(pk), planes, stage=P.Stage.Rec2020, status=status,
                    idct=P.Idct.Ffmpeg if ffmpeg else P.Idct.Apple)
        outs.append((f, ffmpeg, planes, status))
    pipe.sync()
    want = {}
    for f, ffmpeg, planes, status in outs:
        pk = frames[f][0]
        if (f, ffmpeg) not in want:
            h, samples, skipped, mq = O.decode_frame(pk, ffmpeg=ffmpeg)
            want[(f, ffmpeg)] = (O.develop_planar(h, samples, O.params(h, h.quant_matrix[0] * mq),
                                                  2), mq)
        ref, mq = want[(f, ffmpeg)]
        st = status.cpu().numpy().view(np.uint32)
        assert st[0] == 0xFFFFFFFF and st[1] == mq and st[3] == 0
        for p, r in zip(planes, ref):
            got = p.cpu().numpy()
            rel = float(np.abs(got - r.reshape(-1)).max()) / max(float(np.abs(r).max()), 1e-30)
            assert rel <= 1e-5
    pipe.close()


def test_pipeline_waits_for_packets_uploaded_on_a_side_stream(gpu):
    """Packets copied from pinned memory with non_blocking=True on a torch
    side stream, into ONE device buffer that every frame reuses, with no
    host synchronization between frames: prr_pipeline_wait orders each
    submit after its upload, and prr_pipeline_join orders the next upload
    after the frames that read the buffer. Every frame's planes equal the
    oracle's."""
    import torch
    P = gpu
    from proresraw_rs_amd.pipeline import Pipeline
    scenes = [config1(512, 272, seed=81),
              S.scene(448, 256, 0, seed=82, version=0, alignment=2, qscale=(15, 15)),
              S.scene(512, 272, 3, seed=83, version=1, alignment=4, qscale=(30, 30))]
    packets = [S.encode_scene(sc) for sc in scenes]
    pad = max(((len(pk) + 3) & ~3) + 16 for pk in packets)
    hosts = []
    for pk in packets:
        t = torch.zeros(pad, dtype=torch.uint8).pin_memory()
        t[:len(pk)] = torch.frombuffer(bytearray(pk), dtype=torch.uint8)
        hosts.append(t)
    d_packet = torch.zeros(pad, dtype=torch.uint8, device="cuda")
    side = torch.cuda.Stream()
    pipe = Pipeline(depth=2)
    order = [0, 1, 2, 0, 2, 1, 1, 0]
    outs = []
    torch.cuda.synchronize()
    for i, f in enumerate(order):
        pk = packets[f]
        hdr, info = P.parse_frame_info(pk)
        planes = [torch.full((hdr.width * hdr.height,), -7.0, device="cuda") for _ in range(3)]
        status = torch.zeros(4, dtype=torch.int32, device="cuda")
        print(f"side-stream frame {i}: scene {f}", flush=True)
        pipe.join(side)  # the upload waits for the frames still reading d_packet
        with torch.cuda.stream(side):
            d_packet.copy_(hosts[f], non_blocking=True)
        pipe.wait(side)  # the submit waits for the upload
        pipe.submit(hdr, info, d_packet, len(pk), planes, stage=P.Stage.Rec2020, status=status)
        outs.append((f, planes, status))
    pipe.sync()
    want = {}
    for f, planes, status in outs:
        if f not in want:
            h, samples, skipped, mq = O.decode_frame(packets[f])
            want[f] = (O.develop_planar(h, samples, O.params(h, h.quant_matrix[0] * mq), 2), mq)
        ref, mq = want[f]
        st = status.cpu().numpy().view(np.uint32)
        assert st[0] == 0xFFFFFFFF and st[1] == mq and st[3] == 0
        for p, r in zip(planes, ref):
            np.testing.assert_array_equal(p.cpu().numpy().view(np.uint32),
                                          r.reshape(-1).view(np.uint32))
    pipe.close()
