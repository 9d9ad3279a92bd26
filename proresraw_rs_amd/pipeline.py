"""A stream of frames through the device pipeline (prr_pipeline_* of
include/proresraw_amd.h): frame i + 1's decode on one HIP stream beside
frame i's develop on another, intermediate buffers in a ring owned by the
native pipeline. Packets and planes are torch device tensors."""
import ctypes

from . import _native as N
from .api import Error, Idct, Stage, context


class Pipeline:
    def __init__(self, depth=3):
        self.lib = N.lib()
        self.ctx = context()
        h = ctypes.c_void_p()
        err = N.PrrError()
        if self.lib.prr_pipeline_create(self.ctx.handle, depth, ctypes.byref(h),
                                        ctypes.byref(err)):
            raise Error(err)
        self.handle = h

    def submit(self, header, info, d_packet, length, planes, stage=Stage.Rec2020,
               idct=Idct.Apple, wb=None, status=None):
        """Enqueue decode + develop(&frame, stage) of one frame (`d_packet`
        a 4-byte aligned device tensor padded past `length`); `planes` three
        float32 device tensors of width * height; `status`, optional, an
        int32 x 4 device tensor for the frame's prr_device_status."""
        err = N.PrrError()
        r, g, b = (ctypes.c_void_p(p.data_ptr()) for p in planes)
        rc = self.lib.prr_pipeline_submit(
            self.handle, ctypes.byref(header._raw), ctypes.byref(info),
            ctypes.c_void_p(d_packet.data_ptr()), length, int(idct), 1 if wb else 0,
            wb[0] if wb else 0.0, wb[1] if wb else 0.0, int(stage), r, g, b,
            ctypes.c_void_p(status.data_ptr()) if status is not None else None,
            ctypes.byref(err))
        if rc:
            raise Error(err)

    def wait(self, stream):
        """Frames submitted from now on start after the work queued on
        `stream` so far (a packet uploaded there, planes read there)."""
        err = N.PrrError()
        if self.lib.prr_pipeline_wait(self.handle, ctypes.c_void_p(stream.cuda_stream),
                                      ctypes.byref(err)):
            raise Error(err)

    def join(self, stream):
        """`stream` waits for every frame submitted so far."""
        err = N.PrrError()
        if self.lib.prr_pipeline_join(self.handle, ctypes.c_void_p(stream.cuda_stream),
                                      ctypes.byref(err)):
            raise Error(err)

    def sync(self):
        err = N.PrrError()
        if self.lib.prr_pipeline_sync(self.handle, ctypes.byref(err)):
            raise Error(err)

    def timing(self, enable):
        """(mean decode ms, mean develop ms, frames) of the frames timed
        since the last call; times the frames from now on if `enable`."""
        a, b, n = ctypes.c_float(), ctypes.c_float(), ctypes.c_uint64()
        err = N.PrrError()
        if self.lib.prr_pipeline_timing(self.handle, int(enable), ctypes.byref(a), ctypes.byref(b),
                                        ctypes.byref(n), ctypes.byref(err)):
            raise Error(err)
        return a.value, b.value, n.value

    def close(self):
        if self.handle:
            self.lib.prr_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
