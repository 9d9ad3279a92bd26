"""DNG raw tiles on the device: the LosslessJpeg branch of the DNG writer
(examples/prr2dng/dng.rs:92-128) with lj92::encode
(examples/prr2dng/lj92.rs:14-93) for every 256x256 tile, through
prr_dng_lj92_tiles[_device] of include/proresraw_amd.h (HIP kernels in
csrc/prr_dng.hip)."""
import ctypes

import numpy as np

from . import _native as N
from .api import Error, context

TILE = 256


def tile_count(width: int, height: int) -> int:
    return ((width + TILE - 1) // TILE) * ((height + TILE - 1) // TILE)


def capacity(width: int, height: int) -> int:
    """Worst-case bytes of a picture's tile streams."""
    return N.lib().prr_dng_lj92_capacity(width, height)


def lj92_tiles(values: np.ndarray, width: int = None, height: int = None):
    """Every tile's lossless JPEG stream of a height x width uint16 picture
    (rows may be wider than `width`: the visible code values of a frame at
    its coded stride): (bytes, offsets, lengths), streams back to back in
    tile order."""
    v = np.ascontiguousarray(values, dtype=np.uint16)
    h, stride = v.shape
    width = stride if width is None else width
    height = h if height is None else height
    n = tile_count(width, height)
    cap = capacity(width, height)
    out = np.empty(cap, dtype=np.uint8)
    offsets = np.zeros(n, dtype=np.uint32)
    lengths = np.zeros(n, dtype=np.uint32)
    total = ctypes.c_size_t()
    err = N.PrrError()
    rc = N.lib().prr_dng_lj92_tiles(context().handle, v.ctypes.data, width, height, stride,
                                    out.ctypes.data, cap, offsets.ctypes.data,
                                    lengths.ctypes.data, n, ctypes.byref(total), ctypes.byref(err))
    if rc:
        raise Error(err)
    return out[:total.value].tobytes(), offsets, lengths


def lj92_tiles_device(d_values, width: int, height: int, stride: int, d_out, d_offsets,
                      d_lengths, stream=None):
    """The same on torch device tensors (uint16 picture as int16, uint8
    output of at least capacity(width, height) bytes, int32 offsets and
    lengths of tile_count entries), enqueued on `stream`."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    err = N.PrrError()
    rc = N.lib().prr_dng_lj92_tiles_device(
        context().handle, ctypes.c_void_p(d_values.data_ptr()), width, height, stride,
        ctypes.c_void_p(d_out.data_ptr()), d_out.numel(), ctypes.c_void_p(d_offsets.data_ptr()),
        ctypes.c_void_p(d_lengths.data_ptr()), ctypes.c_void_p(s.cuda_stream), ctypes.byref(err))
    if rc:
        raise Error(err)
