"""Crop and orientation on the device (csrc/prr_orient.hip) against the
oracle: exact copies of every plane for all eight orientations, several
crops and sizes (partial and single tiles), float planes and uint16 CFA."""
import itertools

import numpy as np
import pytest

from oracle import orient_oracle as O
from proresraw_rs_amd import orient as R

pytestmark = pytest.mark.gpu


def test_planes_and_cfa_match_the_oracle(gpu):
    import torch
    rng = np.random.default_rng(8)
    cases = [((w, h), m, w + 5) for (w, h), m in itertools.product(
        [(130, 70), (64, 64), (1, 3), (257, 129)], [(0, 0, 0, 0), (3, 2, 1, 4)])]
    # 16-byte aligned rows and a crop keeping the width a multiple of 4: the
    # float4 path for full tiles, the element path for the partial ones.
    cases += [((256, 192), (4, 4, 2, 2), 256), ((200, 136), (0, 0, 0, 0), 204),
              ((328, 260), (8, 0, 3, 1), 332)]
    for (w, h), margins, stride in cases:
        planes = [rng.standard_normal((h, stride)).astype(np.float32) for _ in range(3)]
        cfa = rng.integers(0, 65536, (h, stride)).astype(np.uint16)
        d_planes = [torch.from_numpy(p.reshape(-1)).cuda() for p in planes]
        d_cfa = torch.from_numpy(cfa.view(np.int16).reshape(-1)).cuda()
        for o in range(1, 9):
            print(f"orient {w}x{h} crop {margins} orientation {o}", flush=True)
            ow, oh = R.oriented_size(w, h, margins, o)
            ostride = (ow + 3) // 4 * 4 if stride % 4 == 0 else ow + 3
            outs = [torch.full((oh * ostride,), -1.0, device="cuda") for _ in range(3)]
            R.orient_planes_device(d_planes, w, h, stride, margins, o, outs, ostride)
            out_cfa = torch.zeros(oh * ostride, dtype=torch.int16, device="cuda")
            R.orient_cfa_device(d_cfa, w, h, stride, margins, o, out_cfa, ostride)
            torch.cuda.synchronize()
            for p, q in zip(planes, outs):
                want = O.crop_orient(p[:, :w], margins, o)
                got = q.cpu().numpy().reshape(oh, ostride)
                assert np.array_equal(got[:, :ow], want)
                assert np.all(got[:, ow:] == -1.0)  # nothing past the row (when padded)
            want = O.crop_orient(cfa[:, :w], margins, o)
            got = out_cfa.cpu().numpy().view(np.uint16).reshape(oh, ostride)[:, :ow]
            assert np.array_equal(got, want)
