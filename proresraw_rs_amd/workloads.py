"""Seeded synthetic frame configurations shared by the benchmark and the
tests (the frame descriptions of the benchmark configurations)."""
import numpy as np

from . import synth as S


def seeded_qmat(seed):
    rng = S.Rng(seed)
    return [1 + rng.below(63) for _ in range(64)]


def seeded_curve(seed, top=4400):
    """A strictly rising 8-point curve covering 12-bit data."""
    rng = S.Rng(seed)
    pts, prev = [], -1
    gamma = 1.5 + rng.below(1000) / 1000.0
    for i in range(8):
        v = int(round(top * (i / 8.0) ** gamma)) + rng.below(16)
        v = max(v, prev + 1)
        pts.append(v)
        prev = v
    return pts


def seeded_matrix(seed):
    rng = S.Rng(seed)
    base = np.array([[0.6, 0.3, 0.1], [0.25, 0.7, 0.05], [0.05, 0.1, 0.85]])
    jitter = np.array([[rng.below(2001) - 1000 for _ in range(3)] for _ in range(3)]) / 20000.0
    return (base + jitter).astype(np.float32)


def config0(width=1920, height=1080, seed=1):
    """1920x1080 RGGB, v0, alignment 2, qscale 15, default qmat and curve,
    white 4095; developed WhiteBalanced with gains (2.0, 1.6)."""
    return S.scene(width, height, 0, seed=seed, version=0, alignment=2, qscale=(15, 15),
                   white_level=4095)


def config1(width=4096, height=2160, seed=2):
    """BGGR, v1 (a custom curve needs v1), alignment 3, qscale 45, seeded qmat
    and 8-point curve, 2% of tiles corrupted, seeded camera matrix, gain 1."""
    return S.scene(width, height, 3, seed=seed, version=1, alignment=3, qscale=(45, 45),
                   qmat=seeded_qmat(seed + 100), curve=seeded_curve(seed + 200),
                   white_level=4095, corrupt_fraction=0.02, color_matrix=seeded_matrix(seed + 300),
                   gain=1.0, wb=(1.9, 1.4))


def stream_frame(width=5760, height=3240, seed=3001):
    """The 5760x3240 frame kind of the end-to-end stream: RGGB, v0 (default
    curve), alignment 3, qscale 45, default quant matrix, white 4095."""
    return S.scene(width, height, 0, seed=seed, version=0, alignment=3, qscale=(45, 45),
                   white_level=4095)


def stream_frame_8k(width=8192, height=4320, seed=3000):
    """The 8192x4320 frame kind of the end-to-end stream: RGGB, v1 with a
    seeded 8-point curve, alignment 4, qscale 15."""
    return S.scene(width, height, 0, seed=seed, version=1, alignment=4, qscale=(15, 15),
                   curve=seeded_curve(40), white_level=4095)
