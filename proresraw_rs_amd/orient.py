"""Crop and display orientation on the device (prr_oriented_size,
prr_oriented_bayer, prr_orient_planes_device, prr_orient_cfa_device of
include/proresraw_amd.h; kernel in csrc/prr_orient.hip). The crate records
mov::Orientation (src/mov.rs:121-157) and header::Crop (src/header.rs:98-109)
for readers to apply; these apply them to developed planes or CFA data."""
import ctypes

from . import _native as N
from .api import Error, context


def _crop(margins):
    return (ctypes.c_uint8 * 4)(*[int(m) for m in (margins or (0, 0, 0, 0))])


def oriented_size(width, height, margins, orientation):
    """(width, height) after the crop (left, right, top, bottom) and EXIF
    orientation 1..8."""
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    rc = N.lib().prr_oriented_size(width, height, _crop(margins), int(orientation),
                                   ctypes.byref(w), ctypes.byref(h))
    if rc:
        raise ValueError(f"bad size or orientation: {width}x{height}, {orientation}")
    return w.value, h.value


def oriented_bayer(bayer, margins, width, height, orientation):
    """The BayerPattern code of the cropped, turned CFA picture."""
    b = N.lib().prr_oriented_bayer(int(bayer), _crop(margins), width, height, int(orientation))
    if b < 0:
        raise ValueError("bad pattern, size or orientation")
    return b


def orient_planes_device(planes, width, height, in_stride, margins, orientation, out_planes,
                         out_stride, stream=None):
    """Three float32 device planes to three oriented ones (torch tensors)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    err = N.PrrError()
    p = [ctypes.c_void_p(t.data_ptr()) for t in planes]
    q = [ctypes.c_void_p(t.data_ptr()) for t in out_planes]
    rc = N.lib().prr_orient_planes_device(context().handle, *p, width, height, in_stride,
                                          _crop(margins), int(orientation), *q, out_stride,
                                          ctypes.c_void_p(s.cuda_stream), ctypes.byref(err))
    if rc:
        raise Error(err)


def orient_cfa_device(values, width, height, in_stride, margins, orientation, out, out_stride,
                      stream=None):
    """A uint16 CFA picture (torch int16 on the device) to its oriented copy."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    err = N.PrrError()
    rc = N.lib().prr_orient_cfa_device(context().handle, ctypes.c_void_p(values.data_ptr()), width,
                                       height, in_stride, _crop(margins), int(orientation),
                                       ctypes.c_void_p(out.data_ptr()), out_stride,
                                       ctypes.c_void_p(s.cuda_stream), ctypes.byref(err))
    if rc:
        raise Error(err)
