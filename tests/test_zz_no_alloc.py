"""The allocation-free decode (prr_decode_frame_device_noalloc, static
device scratch) against prr_decode_frame_device (context scratch) through
the C ABI: code values, linear values, tile flags and the device status
bit-exact on two calls in a row, the second on a smaller frame (so the
scratch keeps the first call's offsets and order past the second frame's
end), in both inverse transforms; and the refusal of a frame with more
tiles than the scratch holds."""
import ctypes

import numpy as np
import pytest

import proresraw_rs_amd as P
from proresraw_rs_amd import _native as N
from proresraw_rs_amd import synth as S
from proresraw_rs_amd.workloads import config1

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _decode(fn, pk, idct, stream):
    import torch
    lib, ctx = N.lib(), P.context()
    hdr, info = P.parse_frame_info(pk)
    dev = torch.device("cuda")
    padded = ((len(pk) + 3) & ~3) + 16
    packet = torch.zeros(padded, dtype=torch.uint8, device=dev)
    packet[:len(pk)] = torch.frombuffer(bytearray(pk), dtype=torch.uint8).to(dev)
    # Garbage in every output, so a value the decode does not write shows.
    samples = torch.full((info.coded_width * info.coded_height,), 0x5A5A, dtype=torch.int16,
                         device=dev)
    linear = torch.full((hdr.width * hdr.height,), 0x3C3C, dtype=torch.int16, device=dev)
    flags = torch.full((max(info.nb_tiles, 1),), 7, dtype=torch.uint8, device=dev)
    status = torch.full((4,), 0x11111111, dtype=torch.int32, device=dev)
    err = N.PrrError()
    rc = getattr(lib, fn)(ctx.handle, ctypes.byref(hdr._raw), ctypes.byref(info),
                          packet.data_ptr(), len(pk), int(idct), samples.data_ptr(),
                          linear.data_ptr(), flags.data_ptr(), status.data_ptr(),
                          ctypes.c_void_p(stream.cuda_stream), ctypes.byref(err))
    assert rc == 0, (fn, rc, err.what)
    stream.synchronize()
    return [t.cpu().numpy() for t in (samples, linear, flags, status)]


def test_noalloc_decode_equals_the_allocating_decode(gpu):
    import torch
    s = torch.cuda.Stream()
    big = S.encode_scene(config1(1024, 544, seed=71))
    small = S.encode_scene(S.scene(336, 200, 0, seed=72, version=0, alignment=2,
                                   qscale=(15, 45)))
    for idct in (P.Idct.Apple, P.Idct.Ffmpeg):
        for name, pk in (("1024x544 a3 corrupt", big), ("336x200 a2", small)):
            say(f"noalloc {name} {idct.name}")
            want = _decode("prr_decode_frame_device", pk, idct, s)
            got = _decode("prr_decode_frame_device_noalloc", pk, idct, s)
            for w, g in zip(want, got):
                np.testing.assert_array_equal(g, w)
            st = got[3].view(np.uint32)
            assert st[0] == 0xFFFFFFFF and st[3] == 0


def test_noalloc_refuses_frames_beyond_its_scratch(gpu):
    lib, ctx = N.lib(), P.context()
    pk = S.encode_scene(S.scene(256, 128, 0, seed=73, version=0, alignment=2, qscale=(15, 15)))
    hdr, info = P.parse_frame_info(pk)
    info.nb_tiles = 262145  # one tile past the scratch: refused before any launch
    err = N.PrrError()
    rc = lib.prr_decode_frame_device_noalloc(ctx.handle, ctypes.byref(hdr._raw),
                                             ctypes.byref(info), ctypes.c_void_p(256), len(pk), 0,
                                             ctypes.c_void_p(256), None, None,
                                             ctypes.c_void_p(256), None, ctypes.byref(err))
    assert rc == 101 and err.a == 262145 and err.b == 262144
