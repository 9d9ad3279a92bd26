"""Batched decode over torch device buffers: many frames of any size,
version, alignment and qscale in one persistent launch (prr_batch_* of the
C ABI). torch only allocates device memory and provides the stream."""
import ctypes
import hashlib
import json
import os

import numpy as np

from . import _native as N
from .api import Error, Greens, Idct, Scale, Stage, context, parse_frame_info

KIND_DECODE, KIND_DEVELOP, KIND_REDUCED = 0, 1, 2


def _torch():
    import torch
    return torch


class Slot:
    """The device buffers of one frame of a batch."""

    def __init__(self, header, info, packet, kind):
        torch = _torch()
        dev = torch.device("cuda")
        self.header, self.info, self.kind = header, info, kind
        padded = ((len(packet) + 3) & ~3) + 16
        self.packet = torch.zeros(padded, dtype=torch.uint8, device=dev)
        self.packet[:len(packet)] = torch.frombuffer(bytearray(packet), dtype=torch.uint8).to(dev)
        self.len = len(packet)
        self.flags = torch.zeros(max(info.nb_tiles, 1), dtype=torch.uint8, device=dev)
        self.status = torch.zeros(4, dtype=torch.int32, device=dev)
        self.samples = self.linear = self.rgb = None
        self.planes = None
        self.scale = None

    def status_host(self):
        return self.status.cpu().numpy().view(np.uint32)

    def skipped_tiles(self):
        f = self.flags[:self.info.nb_tiles].cpu().numpy()
        return [int(i) for i in np.nonzero(f)[0]]

    def check(self):
        """Raise the crate's error for this frame, if any."""
        st = self.status_host()
        if st[0] != 0xFFFFFFFF:
            e = N.PrrError()
            e.code, e.a = 11, int(st[0])
            raise Error(e)
        if st[3]:
            e = N.PrrError()
            e.code = 103
            raise Error(e)


class Batch:
    def __init__(self, streams_per_wg=0, chunk=0, waves_per_cu=0):
        self.lib = N.lib()
        self.ctx = context()
        h = ctypes.c_void_p()
        err = N.PrrError()
        self._check(self.lib.prr_batch_create(self.ctx.handle, ctypes.byref(h), ctypes.byref(err)), err)
        self.handle = h
        self.slots = []
        if streams_per_wg or chunk or waves_per_cu:
            self.set_config(streams_per_wg, chunk, waves_per_cu)

    @staticmethod
    def _check(rc, err):
        if rc:
            raise Error(err)

    def set_config(self, streams_per_wg=0, chunk=0, waves_per_cu=0):
        err = N.PrrError()
        self._check(self.lib.prr_batch_set_config(self.handle, streams_per_wg, chunk, waves_per_cu,
                                                  ctypes.byref(err)), err)

    def _add(self, slot, bf):
        bf.d_packet, bf.len = slot.packet.data_ptr(), slot.len
        bf.d_tile_flags, bf.d_status = slot.flags.data_ptr(), slot.status.data_ptr()
        err = N.PrrError()
        self._check(self.lib.prr_batch_add(self.handle, ctypes.byref(slot.header._raw),
                                           ctypes.byref(slot.info), ctypes.byref(bf),
                                           ctypes.byref(err)), err)
        self.slots.append(slot)
        return slot

    def add_full(self, packet, idct=Idct.Apple, develop=None, wb=None, linear=False,
                 interleaved=False):
        """A full decode (decode_frame_with); with develop=Stage, also
        develop(&frame, stage) (or with_white_balance(wb) when wb is given)
        into three planes, or with interleaved=True into one interleaved RGB
        buffer (develop_into's layout)."""
        torch = _torch()
        header, info = parse_frame_info(packet)
        slot = Slot(header, info, packet, KIND_DEVELOP if develop is not None else KIND_DECODE)
        dev = torch.device("cuda")
        slot.samples = torch.zeros(info.coded_width * info.coded_height, dtype=torch.int16, device=dev)
        npix = header.width * header.height
        if linear:
            slot.linear = torch.zeros(npix, dtype=torch.int16, device=dev)
        bf = N.PrrBatchFrame()
        bf.kind = slot.kind
        bf.idct = int(idct)
        bf.d_samples = slot.samples.data_ptr()
        bf.d_linear = slot.linear.data_ptr() if linear else None
        if develop is not None:
            bf.stage = int(develop)
            bf.explicit_wb = 1 if wb is not None else 0
            if wb is not None:
                bf.wb_red, bf.wb_blue = wb
            if interleaved:
                slot.planes = [torch.zeros(3 * npix, dtype=torch.float32, device=dev)]
                bf.d_r, bf.d_g, bf.d_b = slot.planes[0].data_ptr(), None, None
            else:
                slot.planes = [torch.zeros(npix, dtype=torch.float32, device=dev)
                               for _ in range(3)]
                bf.d_r, bf.d_g, bf.d_b = [p.data_ptr() for p in slot.planes]
        return self._add(slot, bf)

    def add_reduced(self, packet, scale: Scale, greens: Greens = Greens.First):
        """decode_frame_reduced_with(packet, scale, greens)."""
        torch = _torch()
        header, info = parse_frame_info(packet)
        slot = Slot(header, info, packet, KIND_REDUCED)
        s = int(scale)
        slot.scale = s
        slot.rgb = torch.zeros(3 * (info.coded_width >> s) * (info.coded_height >> s),
                               dtype=torch.int16, device=torch.device("cuda"))
        bf = N.PrrBatchFrame()
        bf.kind = KIND_REDUCED
        bf.scale, bf.greens = s, int(greens)
        bf.d_rgb = slot.rgb.data_ptr()
        return self._add(slot, bf)

    def run(self, stream=None):
        """Enqueue the batch on `stream` (a torch stream; default: current)."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        err = N.PrrError()
        self._check(self.lib.prr_batch_run(self.handle, ctypes.c_void_p(s.cuda_stream),
                                           ctypes.byref(err)), err)

    def signature(self) -> str:
        """What the best launch shape depends on: every frame's size,
        alignment, mode and tile count."""
        parts = []
        for s in self.slots:
            h = s.header
            parts.append(f"{h.width}x{h.height}a{h.alignment}k{s.kind}s{s.scale or 0}"
                         f"t{s.info.nb_tiles}")
        return hashlib.sha1(";".join(parts).encode()).hexdigest()[:16]

    def time_run(self, stream=None, reps=3):
        """Mean milliseconds of `reps` runs after one untimed run (which also
        uploads the frame table after a configuration change)."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        self.run(s)
        s.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            self.run(s)
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def autotune(self, stream=None, spws=(16, 32, 64), chunks=(1, 2, 4, 8), waves=(0,),
                 reps=3, cache=None):
        """Pick streams per workgroup, queue chunk and waves per CU by timing
        each candidate on this batch; the choice is kept in a JSON cache
        (PRR_AUTOTUNE_CACHE, else ~/.cache/proresraw_amd/batch_tuning.json)
        under the batch's signature and reused without timing. Returns
        {"spw", "chunk", "waves", "ms", "cached", "trials"}."""
        path = cache or os.environ.get("PRR_AUTOTUNE_CACHE") or os.path.join(
            os.path.expanduser("~"), ".cache", "proresraw_amd", "batch_tuning.json")
        key = self.signature()
        table = {}
        if os.path.exists(path):
            try:
                with open(path) as f:
                    table = json.load(f)
            except (OSError, ValueError):
                table = {}
        if key in table:
            best = dict(table[key])
            self.set_config(best["spw"], best["chunk"], best["waves"])
            best["cached"] = True
            return best
        trials = []
        for spw in spws:
            for chunk in chunks:
                for w in waves:
                    self.set_config(spw, chunk, w)
                    trials.append({"spw": spw, "chunk": chunk, "waves": w,
                                   "ms": self.time_run(stream, reps)})
        best = dict(min(trials, key=lambda t: t["ms"]))
        self.set_config(best["spw"], best["chunk"], best["waves"])
        table[key] = best
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(table, f, indent=1)
            os.replace(tmp, path)
        except OSError:
            pass
        best.update(cached=False, trials=trials)
        return best

    def info(self):
        n, g, l = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        self.lib.prr_batch_info(self.handle, ctypes.byref(n), ctypes.byref(g), ctypes.byref(l))
        return {"tile_groups": n.value, "grid": g.value, "lds_bytes": l.value}

    def __del__(self):
        try:
            self.lib.prr_batch_destroy(self.handle)
        except Exception:
            pass
