"""Decoding a ProRes RAW movie on the device, window by window.

The crate's callers read a clip with `mov::read_video_track`, then loop over
its samples with `mov::read_at` and `decode_frame` (README, examples/
decode.rs, examples/prr2dng/main.rs:237). Here a clip is decoded in windows
of frames: host threads read a window's samples with positional reads into
pinned host memory (prr_mov_read_samples) while the device decodes the
previous window; one copy on a copy stream brings the window's packets to
HBM; one persistent batched launch (prr_batch_*) decodes, and optionally
develops, every frame of the window. Results are those of decode_frame and
develop on each frame (the batch's own parity); the windows only change when
work happens.
"""
import ctypes
import os
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _native as N
from . import mov as M
from .api import Error, FrameHeader, Idct, Stage, context
from .batch import KIND_DECODE, KIND_DEVELOP


def _torch():
    import torch
    return torch


@dataclass
class ClipFrame:
    """One decoded frame of a window: its index in the clip and its device
    buffers (valid until the window after next is decoded)."""
    index: int
    header: FrameHeader
    info: N.PrrFrameInfo
    samples: object          # torch int16, coded_width * coded_height
    linear: Optional[object]  # torch int16, width * height
    planes: Optional[list]   # three torch float32 planes, width * height
    flags: object            # torch uint8, one per tile
    status: object           # torch int32 x 4 (prr_device_status)

    def skipped_tiles(self) -> List[int]:
        f = self.flags[:self.info.nb_tiles].cpu().numpy()
        return [int(i) for i in np.nonzero(f)[0]]

    def check(self):
        st = self.status.cpu().numpy().view(np.uint32)
        if st[0] != 0xFFFFFFFF:
            e = N.PrrError()
            e.code, e.a = 11, int(st[0])
            raise Error(e)
        if st[3]:
            e = N.PrrError()
            e.code = 103
            raise Error(e)


class _Outputs:
    """Device output buffers of one window position, reused while the frame
    size stays the same."""

    def __init__(self):
        self.key = None

    def ensure(self, info, header, develop, linear):
        torch = _torch()
        key = (info.coded_width, info.coded_height, info.nb_tiles, header.width, header.height,
               develop is not None, linear)
        if key == self.key:
            return
        dev = torch.device("cuda")
        npix = header.width * header.height
        self.samples = torch.empty(info.coded_width * info.coded_height, dtype=torch.int16,
                                   device=dev)
        self.linear = torch.empty(npix, dtype=torch.int16, device=dev) if linear else None
        self.planes = ([torch.empty(npix, dtype=torch.float32, device=dev) for _ in range(3)]
                       if develop is not None else None)
        self.flags = torch.empty(max(info.nb_tiles, 1), dtype=torch.uint8, device=dev)
        self.status = torch.empty(4, dtype=torch.int32, device=dev)
        self.key = key


class ClipDecoder:
    """Decode (and develop) a movie's ProRes RAW track in windows of
    `window` frames. `develop` is a Stage (develop(&frame, stage) with the
    header's white balance) or None (decode only); `linear` also writes
    Frame::linearize's output."""

    def __init__(self, path, window=16, idct=Idct.Apple, develop: Optional[Stage] = None,
                 linear=False, threads=16, streams_per_wg=64, chunk=1):
        torch = _torch()
        self.path = os.fspath(path)
        self.track = M.read_video_track(self.path)
        self.fd = os.open(self.path, os.O_RDONLY)
        self.window = window
        self.idct, self.develop, self.linear = idct, develop, linear
        self.threads = threads
        self.lib = N.lib()
        self.ctx = context()
        self.stream = torch.cuda.Stream()
        self.copy_stream = torch.cuda.Stream()
        # Two of everything: window w uses set w % 2 while set (w - 1) % 2 is
        # still decoding.
        self.pinned = [None, None]
        self.dev_packets = [None, None]
        self.outputs = [[_Outputs() for _ in range(window)] for _ in range(2)]
        self.batches = []
        for _ in range(2):
            h = ctypes.c_void_p()
            err = N.PrrError()
            if self.lib.prr_batch_create(self.ctx.handle, ctypes.byref(h), ctypes.byref(err)):
                raise Error(err)
            if self.lib.prr_batch_set_config(h, streams_per_wg, chunk, 0, ctypes.byref(err)):
                raise Error(err)
            self.batches.append(h)
        self.copied = [torch.cuda.Event(), torch.cuda.Event()]
        self.decoded = [torch.cuda.Event(), torch.cuda.Event()]
        self.read_seconds = 0.0

    def close(self):
        if self.fd is not None:
            self.stream.synchronize()
            for h in self.batches:
                self.lib.prr_batch_destroy(h)
            self.batches = []
            os.close(self.fd)
            self.fd = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _buffers(self, k, total):
        torch = _torch()
        if self.pinned[k] is None or self.pinned[k].numel() < total:
            cap = int(total * 1.25) + 4096
            self.pinned[k] = torch.empty(cap, dtype=torch.uint8, pin_memory=True)
            self.dev_packets[k] = torch.empty(cap, dtype=torch.uint8, device="cuda")
        return self.pinned[k], self.dev_packets[k]

    def _enqueue(self, w, first, count):
        """Read window w's samples, copy them to the device and enqueue their
        batched decode; returns the window's ClipFrames."""
        k = w % 2
        samples = self.track.samples[first:first + count]
        offsets, total = M.packet_layout(samples)
        # The pinned buffer of set k was last read by the copy of window w - 2.
        self.copied[k].synchronize()
        pinned, dev = self._buffers(k, total)
        t0 = time.perf_counter()
        M.read_samples(self.fd, samples, pinned.data_ptr(), offsets, self.threads)
        self.read_seconds += time.perf_counter() - t0
        # The device packets of set k were last read by the decode of window w - 2.
        self.copy_stream.wait_event(self.decoded[k])
        with _torch().cuda.stream(self.copy_stream):
            dev[:total].copy_(pinned[:total], non_blocking=True)
        self.copied[k].record(self.copy_stream)
        batch = self.batches[k]
        err = N.PrrError()
        if self.lib.prr_batch_clear(batch, ctypes.byref(err)):
            raise Error(err)
        frames = []
        for j, s in enumerate(samples):
            o = int(offsets[j])
            hdr, info = N.PrrHeader(), N.PrrFrameInfo()
            if self.lib.prr_parse_frame_info(ctypes.c_void_p(pinned.data_ptr() + o), s.size,
                                             ctypes.byref(hdr), ctypes.byref(info),
                                             ctypes.byref(err)):
                raise Error(err)
            header = FrameHeader._from_raw(hdr)
            out = self.outputs[k][j]
            out.ensure(info, header, self.develop, self.linear)
            bf = N.PrrBatchFrame()
            bf.d_packet, bf.len = dev.data_ptr() + o, s.size
            bf.kind = KIND_DEVELOP if self.develop is not None else KIND_DECODE
            bf.idct = int(self.idct)
            bf.d_samples = out.samples.data_ptr()
            bf.d_linear = out.linear.data_ptr() if out.linear is not None else None
            if self.develop is not None:
                bf.stage = int(self.develop)
                bf.d_r, bf.d_g, bf.d_b = [p.data_ptr() for p in out.planes]
            bf.d_tile_flags, bf.d_status = out.flags.data_ptr(), out.status.data_ptr()
            if self.lib.prr_batch_add(batch, ctypes.byref(hdr), ctypes.byref(info),
                                      ctypes.byref(bf), ctypes.byref(err)):
                raise Error(err)
            frames.append(ClipFrame(first + j, header, info, out.samples, out.linear,
                                    out.planes, out.flags, out.status))
        self.stream.wait_event(self.copied[k])
        if self.lib.prr_batch_run(batch, ctypes.c_void_p(self.stream.cuda_stream),
                                  ctypes.byref(err)):
            raise Error(err)
        self.decoded[k].record(self.stream)
        return frames

    def frames(self, first=0, count=None):
        """Yield each window's ClipFrames once its decode has finished; the
        next window's reads and copy are already under way by then."""
        n = len(self.track.samples) if count is None else min(count,
                                                               len(self.track.samples) - first)
        starts = list(range(first, first + n, self.window))
        pending = None
        for w, s in enumerate(starts):
            cur = self._enqueue(w, s, min(self.window, first + n - s))
            if pending is not None:
                self.decoded[(w - 1) % 2].synchronize()
                yield pending
            pending = cur
        if pending is not None:
            self.decoded[(len(starts) - 1) % 2].synchronize()
            yield pending

    def run(self, first=0, count=None):
        """Decode the frames and return (frames decoded, seconds)."""
        t0 = time.perf_counter()
        done = 0
        for window in self.frames(first, count):
            done += len(window)
        return done, time.perf_counter() - t0
