"""Synthetic frames: the crate's synth module (src/synth.rs) and the seeded
scene encoder, from libprr_synth.so (plain C, no device)."""
import ctypes
import os

import numpy as np

from . import _native as N


class Spec(ctypes.Structure):
    """synth::Spec (synth.rs:34-89) plus the bitstream version."""
    _fields_ = [
        ("width", ctypes.c_uint16),
        ("height", ctypes.c_uint16),
        ("alignment", ctypes.c_uint8),
        ("pattern", ctypes.c_uint8),
        ("version", ctypes.c_uint8),
        ("valid_tiles", ctypes.c_uint8),
        ("has_qmat", ctypes.c_uint8),
        ("has_curve", ctypes.c_uint8),
        ("qmat", ctypes.c_uint8 * 64),
        ("curve", ctypes.c_uint16 * 8),
        ("vendor", ctypes.c_uint8 * 4),
        ("crop", ctypes.c_uint8 * 4),
        ("white_level", ctypes.c_uint16),
        ("wb_cct", ctypes.c_uint16),
        ("wb", ctypes.c_float * 2),
        ("color_matrix", ctypes.c_float * 9),
        ("gain", ctypes.c_float),
    ]


class Scene(ctypes.Structure):
    _fields_ = [
        ("spec", Spec),
        ("seed", ctypes.c_uint64),
        ("qscale_min", ctypes.c_uint8),
        ("qscale_max", ctypes.c_uint8),
        ("n_gradients", ctypes.c_uint8),
        ("reserved", ctypes.c_uint8),
        ("n_discs", ctypes.c_uint32),
        ("corrupt_fraction", ctypes.c_float),
        ("dn_per_electron", ctypes.c_float),
        ("read_sigma", ctypes.c_float),
        ("threads", ctypes.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(N.SYNTH_PATH):
            raise OSError(f"{N.SYNTH_PATH} is missing: build with `python -m proresraw_rs_amd.build`")
        h = ctypes.CDLL(N.SYNTH_PATH)
        SP = ctypes.POINTER(Spec)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        h.prs_spec_default.argtypes = [SP, ctypes.c_uint16, ctypes.c_uint16, ctypes.c_uint8]
        h.prs_rng_new.restype = ctypes.c_uint64
        h.prs_rng_new.argtypes = [ctypes.c_uint64]
        h.prs_rng_next.restype = ctypes.c_uint32
        h.prs_rng_next.argtypes = [u64p]
        h.prs_specs.argtypes = [u64p, SP]
        h.prs_random_frame.restype = ctypes.c_size_t
        h.prs_random_frame.argtypes = [SP, u64p, ctypes.c_void_p, ctypes.c_size_t]
        h.prs_frame_streams.restype = ctypes.c_size_t
        h.prs_frame_streams.argtypes = [SP, ctypes.c_uint8, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_size_t]
        h.prs_encode_scene.restype = ctypes.c_size_t
        h.prs_encode_scene.argtypes = [ctypes.POINTER(Scene), ctypes.c_void_p, ctypes.c_size_t]
        h.prs_scene_mosaic.argtypes = [ctypes.POINTER(Scene), ctypes.c_void_p]
        h.prs_encode_component.restype = ctypes.c_size_t
        h.prs_encode_component.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p,
                                           ctypes.c_size_t]
        _lib = h
    return _lib


class Rng:
    """synth::Rng (synth.rs:9-29), xorshift64*."""

    def __init__(self, seed: int):
        self.state = ctypes.c_uint64(lib().prs_rng_new(seed))

    def next_u32(self) -> int:
        return lib().prs_rng_next(ctypes.byref(self.state))

    def below(self, bound: int) -> int:
        return self.next_u32() % bound


def spec(width, height, pattern=0, **kw) -> Spec:
    """Spec::new (synth.rs:71-88) with keyword overrides; `qmat` and `curve`
    set the optional header fields."""
    s = Spec()
    lib().prs_spec_default(ctypes.byref(s), width, height, pattern)
    for k, v in kw.items():
        if k == "qmat":
            if v is None:
                s.has_qmat = 0
            else:
                s.has_qmat = 1
                s.qmat[:] = list(v)
        elif k == "curve":
            if v is None:
                s.has_curve = 0
            else:
                s.has_curve = 1
                s.curve[:] = list(v)
        elif k == "wb":
            s.wb[:] = list(v)
        elif k in ("color_matrix",):
            s.color_matrix[:] = [float(x) for x in np.asarray(v, dtype=np.float32).ravel()]
        elif k in ("vendor", "crop"):
            getattr(s, k)[:] = list(v)
        else:
            setattr(s, k, v)
    return s


def specs(rng: Rng):
    """synth::specs (synth.rs:188-221): 20 specs."""
    out = (Spec * 20)()
    lib().prs_specs(ctypes.byref(rng.state), out)
    return list(out)


def random_frame(s: Spec, rng: Rng) -> bytes:
    """synth::random_frame (synth.rs:92-94)."""
    saved = ctypes.c_uint64(rng.state.value)
    n = lib().prs_random_frame(ctypes.byref(s), ctypes.byref(saved), None, 0)
    buf = (ctypes.c_uint8 * n)()
    got = lib().prs_random_frame(ctypes.byref(s), ctypes.byref(rng.state), buf, n)
    assert got == n
    return bytes(buf)



def frame_streams(s: Spec, qscale: int, streams) -> bytes:
    """synth::frame with every tile synth::tile(qscale, streams)."""
    bufs = [ctypes.create_string_buffer(bytes(x), max(len(x), 1)) for x in streams]
    ptrs = (ctypes.c_void_p * 4)(*[ctypes.addressof(b) for b in bufs])
    lens = (ctypes.c_size_t * 4)(*[len(x) for x in streams])
    n = lib().prs_frame_streams(ctypes.byref(s), qscale, ptrs, lens, None, 0)
    out = (ctypes.c_uint8 * n)()
    lib().prs_frame_streams(ctypes.byref(s), qscale, ptrs, lens, out, n)
    return bytes(out)


def scene(width, height, pattern=0, *, seed=1, version=0, alignment=2, qscale=(15, 15),
          qmat=None, curve=None, white_level=4095, corrupt_fraction=0.0, n_gradients=8,
          n_discs=200, dn_per_electron=0.5, read_sigma=3.0, wb=(1.0, 1.0),
          color_matrix=None, gain=1.0, threads=None) -> Scene:
    """Settings of a seeded scene frame (8 gradients + 200 hard-edged discs,
    Poisson-Gaussian noise of gain 0.5 DN/e- and read sigma 3 DN)."""
    sc = Scene()
    sc.spec = spec(width, height, pattern, version=version, alignment=alignment, qmat=qmat,
                   curve=curve, white_level=white_level, wb=wb, gain=gain,
                   **({"color_matrix": color_matrix} if color_matrix is not None else {}))
    sc.seed = seed
    sc.qscale_min, sc.qscale_max = qscale
    sc.n_gradients = n_gradients
    sc.n_discs = n_discs
    sc.corrupt_fraction = corrupt_fraction
    sc.dn_per_electron = dn_per_electron
    sc.read_sigma = read_sigma
    sc.threads = threads if threads is not None else min(os.cpu_count() or 1, 16)
    return sc


def encode_scene(sc: Scene) -> bytes:
    """The scene encoded into a frame packet."""
    n = lib().prs_encode_scene(ctypes.byref(sc), None, 0)
    if n == 0:
        raise ValueError("empty scene")
    buf = (ctypes.c_uint8 * n)()
    got = lib().prs_encode_scene(ctypes.byref(sc), buf, n)
    if got != n:
        raise RuntimeError("scene encoder size mismatch")
    return bytes(buf)


def scene_mosaic(sc: Scene) -> np.ndarray:
    """The noisy linear mosaic the encoder started from (height, width)."""
    out = np.zeros(sc.spec.width * sc.spec.height, dtype=np.uint16)
    lib().prs_scene_mosaic(ctypes.byref(sc), out.ctypes.data)
    return out.reshape(sc.spec.height, sc.spec.width)


def encode_component(levels, log2_blocks: int) -> bytes:
    """Entropy code 64 << log2_blocks quantized coefficients (block after
    block, raster order) into one component stream."""
    lv = np.ascontiguousarray(levels, dtype=np.int16)
    assert lv.size == 64 << log2_blocks
    n = lib().prs_encode_component(lv.ctypes.data, log2_blocks, None, 0)
    buf = (ctypes.c_uint8 * max(n, 1))()
    lib().prs_encode_component(lv.ctypes.data, log2_blocks, buf, n)
    return bytes(buf)[:n]
