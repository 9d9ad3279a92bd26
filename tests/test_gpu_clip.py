"""A movie decoded window by window on the device (proresraw_rs_amd.clip):
every frame's code values, linear values and skipped tiles are bit-exact
against the oracle's decode_frame of the same sample, and its planes equal
the oracle's develop (within 1e-5 relative max-abs; bit-exact in practice)."""
import numpy as np
import pytest

from oracle import oracle as O
from proresraw_rs_amd import mov as M
from proresraw_rs_amd import synth as S
from proresraw_rs_amd.workloads import config1

pytestmark = pytest.mark.gpu


def clip_packets():
    packets = []
    for i in range(7):
        if i % 3 == 2:
            sc = config1(192 + 8 * i, 96, seed=40 + i)        # v1, corrupt tiles, custom curve
        else:
            sc = S.scene(160 if i < 4 else 224, 112, 3 * (i % 2), seed=30 + i, version=i % 2,
                         alignment=i % 5, qscale=(15 + 5 * i, 15 + 5 * i))
        packets.append(S.encode_scene(sc))
    return packets


@pytest.mark.parametrize("develop", [None, 1])
def test_clip_windows_match_the_oracle(gpu, tmp_path, develop):
    P = gpu
    from proresraw_rs_amd.clip import ClipDecoder
    packets = clip_packets()
    path = tmp_path / "clip.mov"
    path.write_bytes(M.clip_movie(packets, frames_per_chunk=2))
    stage = P.Stage(develop) if develop is not None else None
    dec = ClipDecoder(path, window=3, develop=stage, linear=True, threads=4)
    seen = []
    for window in dec.frames():
        for f in window:
            print(f"clip frame {f.index} ({f.header.width}x{f.header.height})", flush=True)
            h, samples, skipped, mq = O.decode_frame(packets[f.index])
            assert f.skipped_tiles() == skipped
            got = f.samples.cpu().numpy().view(np.uint16)
            assert np.array_equal(got, samples.reshape(-1))
            lin = f.linear.cpu().numpy().view(np.uint16)
            assert np.array_equal(lin, O.linearize(h, samples).reshape(-1))
            if develop is not None:
                want = O.develop_planar(h, samples, O.params(h, h.quant_matrix[0] * mq), develop)
                for plane, ref in zip(f.planes, want):
                    p = plane.cpu().numpy()
                    rel = float(np.abs(p - ref.reshape(-1)).max()) / max(
                        float(np.abs(ref).max()), 1e-30)
                    assert rel <= 1e-5
            seen.append(f.index)
    assert seen == list(range(len(packets)))
    # A second pass over part of the clip reuses every buffer.
    done, _ = dec.run(first=2, count=4)
    assert done == 4
    dec.close()
