"""ctypes declarations of the C ABI in include/proresraw_amd.h,
include/proresraw_learned.h and include/proresraw_mov.h.

Loading fails loudly when the native library has not been built: there is
no fallback implementation behind this module.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libproresraw_amd.so")
SYNTH_PATH = os.path.join(LIB_DIR, "libprr_synth.so")

c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_vp = ctypes.c_void_p


class PrrError(ctypes.Structure):
    _fields_ = [
        ("code", ctypes.c_int32),
        ("what", ctypes.c_char * 28),
        ("a", ctypes.c_uint64),
        ("b", ctypes.c_uint64),
        ("points", ctypes.c_uint16 * 8),
        ("tag", ctypes.c_uint8 * 4),
    ]


class PrrHeader(ctypes.Structure):
    _fields_ = [
        ("version", ctypes.c_uint8),
        ("vendor", ctypes.c_uint8 * 4),
        ("bayer", ctypes.c_uint8),
        ("width", ctypes.c_uint16),
        ("height", ctypes.c_uint16),
        ("crop", ctypes.c_uint8 * 4),
        ("white_level", ctypes.c_uint16),
        ("wb_cct", ctypes.c_uint16),
        ("flags", ctypes.c_uint16),
        ("alignment", ctypes.c_uint8),
        ("quant_matrix", ctypes.c_uint8 * 64),
        ("curve", ctypes.c_uint16 * 8),
        ("wb_red", ctypes.c_float),
        ("wb_blue", ctypes.c_float),
        ("color_matrix", ctypes.c_float * 9),
        ("gain", ctypes.c_float),
    ]


class PrrFrameInfo(ctypes.Structure):
    _fields_ = [
        ("coded_width", ctypes.c_uint32),
        ("coded_height", ctypes.c_uint32),
        ("tiles_per_row", ctypes.c_uint32),
        ("tile_rows", ctypes.c_uint32),
        ("nb_tiles", ctypes.c_uint32),
        ("header_end", ctypes.c_size_t),
    ]


class PrrParsedFrame(ctypes.Structure):
    _fields_ = [
        ("header", PrrHeader),
        ("info", PrrFrameInfo),
        ("max_qscale", ctypes.c_uint8),
    ]


class PrrDemosaicParams(ctypes.Structure):
    _fields_ = [
        ("cfa_phase", ctypes.c_int32),
        ("black", ctypes.c_int32),
        ("one_over_range", ctypes.c_float),
        ("noise", ctypes.c_float * 3),
        ("clip", ctypes.c_float * 2),
        ("black_over_range", ctypes.c_float),
        ("wb", ctypes.c_float * 2),
    ]


class PrrDeviceStatus(ctypes.Structure):
    _fields_ = [
        ("oob_tile", ctypes.c_uint32),
        ("max_qscale", ctypes.c_uint32),
        ("n_skipped", ctypes.c_uint32),
        ("overrun", ctypes.c_uint32),
    ]


class PrrBatchFrame(ctypes.Structure):
    _fields_ = [
        ("d_packet", ctypes.c_void_p),
        ("len", ctypes.c_size_t),
        ("kind", ctypes.c_int32),
        ("idct", ctypes.c_int32),
        ("scale", ctypes.c_int32),
        ("greens", ctypes.c_int32),
        ("stage", ctypes.c_int32),
        ("explicit_wb", ctypes.c_int32),
        ("wb_red", ctypes.c_float),
        ("wb_blue", ctypes.c_float),
        ("d_samples", ctypes.c_void_p),
        ("d_linear", ctypes.c_void_p),
        ("d_rgb", ctypes.c_void_p),
        ("d_r", ctypes.c_void_p),
        ("d_g", ctypes.c_void_p),
        ("d_b", ctypes.c_void_p),
        ("d_tile_flags", ctypes.c_void_p),
        ("d_status", ctypes.c_void_p),
    ]


class PrrNetFrame(ctypes.Structure):
    _fields_ = [
        ("d_samples", ctypes.c_void_p),
        ("stride", ctypes.c_uint32),
        ("header", PrrHeader),
        ("params", PrrDemosaicParams),
        ("d_target", ctypes.c_void_p),
    ]


class PrrMovSample(ctypes.Structure):
    _fields_ = [
        ("offset", ctypes.c_uint64),
        ("size", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


MOV_ACCEPT = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_uint8), ctypes.c_void_p)

NET_PARAMS = 73984
NET_LAYER_PARAMS = 9248

E = ctypes.POINTER(PrrError)
H = ctypes.POINTER(PrrHeader)
FI = ctypes.POINTER(PrrFrameInfo)
DP = ctypes.POINTER(PrrDemosaicParams)
sz = ctypes.c_size_t
i32 = ctypes.c_int
u32 = ctypes.c_uint32

# name -> (restype, argtypes); every symbol include/proresraw_amd.h declares.
SIGNATURES = {
    "prr_context_create": (i32, [i32, ctypes.POINTER(c_vp), E]),
    "prr_context_destroy": (None, [c_vp]),
    "prr_context_stream": (c_vp, [c_vp]),
    "prr_version": (ctypes.c_char_p, []),
    "prr_set_decode_config": (i32, [c_vp, u32, E]),
    "prr_header_len": (i32, [c_vp, sz, ctypes.POINTER(sz), E]),
    "prr_parse_header": (i32, [c_vp, sz, H, E]),
    "prr_parse_frame_info": (i32, [c_vp, sz, H, FI, E]),
    "prr_parse_frame": (i32, [c_vp, sz, ctypes.POINTER(PrrParsedFrame), c_vp, c_vp, sz, E]),
    "prr_quantizer": (u32, [H, ctypes.c_uint8]),
    "prr_frame_quantizer": (u32, [ctypes.POINTER(PrrParsedFrame)]),
    "prr_demosaic_params_new": (None, [H, u32, DP]),
    "prr_demosaic_params_with_white_balance": (None, [H, u32, ctypes.c_float, ctypes.c_float, DP]),
    "prr_output_matrix": (None, [H, i32, c_vp]),
    "prr_decode_frame": (i32, [c_vp, c_vp, sz, i32, H, c_vp, sz, c_vp, sz,
                               ctypes.POINTER(sz), c_u8p, E]),
    "prr_linearize_into": (i32, [c_vp, H, c_vp, sz, c_vp, sz, E]),
    "prr_decode_frame_reduced": (i32, [c_vp, c_vp, sz, i32, i32, H, c_vp, sz, c_vp, sz,
                                       ctypes.POINTER(sz), c_u8p, E]),
    "prr_develop_planar_into": (i32, [c_vp, H, c_vp, sz, DP, i32, c_vp, c_vp, c_vp, sz, E]),
    "prr_develop_into": (i32, [c_vp, H, c_vp, sz, DP, i32, c_vp, sz, E]),
    "prr_decode_frame_device": (i32, [c_vp, H, FI, c_vp, sz, i32, c_vp, c_vp, c_vp, c_vp, c_vp, E]),
    "prr_decode_frame_device_noalloc": (i32, [c_vp, H, FI, c_vp, sz, i32, c_vp, c_vp, c_vp, c_vp,
                                              c_vp, E]),
    "prr_linearize_device": (i32, [c_vp, H, c_vp, c_vp, c_vp, E]),
    "prr_decode_frame_reduced_device": (i32, [c_vp, H, FI, c_vp, sz, i32, i32, c_vp, c_vp, c_vp,
                                              c_vp, E]),
    "prr_develop_device": (i32, [c_vp, H, c_vp, DP, i32, c_vp, c_vp, c_vp, c_vp, E]),
    "prr_develop_frame_device": (i32, [c_vp, H, c_vp, c_vp, i32, ctypes.c_float, ctypes.c_float,
                                       i32, c_vp, c_vp, c_vp, c_vp, E]),
    "prr_develop_linear_device": (i32, [c_vp, H, c_vp, c_vp, i32, ctypes.c_float, ctypes.c_float,
                                        i32, c_vp, c_vp, c_vp, c_vp, E]),
    "prr_pipeline_create": (i32, [c_vp, u32, ctypes.POINTER(c_vp), E]),
    "prr_pipeline_submit": (i32, [c_vp, H, FI, c_vp, sz, i32, i32, ctypes.c_float, ctypes.c_float,
                                  i32, c_vp, c_vp, c_vp, c_vp, E]),
    "prr_pipeline_wait": (i32, [c_vp, c_vp, E]),
    "prr_pipeline_join": (i32, [c_vp, c_vp, E]),
    "prr_pipeline_sync": (i32, [c_vp, E]),
    "prr_pipeline_timing": (i32, [c_vp, i32, ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint64),
                                  E]),
    "prr_pipeline_destroy": (None, [c_vp]),
    "prr_dng_lj92_capacity": (sz, [u32, u32]),
    "prr_dng_lj92_tiles_device": (i32, [c_vp, c_vp, u32, u32, u32, c_vp, sz, c_vp, c_vp, c_vp, E]),
    "prr_dng_lj92_tiles": (i32, [c_vp, c_vp, u32, u32, u32, c_vp, sz, c_vp, c_vp, sz,
                                 ctypes.POINTER(sz), E]),
    "prr_oriented_size": (i32, [u32, u32, c_vp, i32, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    "prr_oriented_bayer": (i32, [i32, c_vp, u32, u32, i32]),
    "prr_orient_planes_device": (i32, [c_vp, c_vp, c_vp, c_vp, u32, u32, u32, c_vp, i32, c_vp,
                                       c_vp, c_vp, u32, c_vp, E]),
    "prr_orient_cfa_device": (i32, [c_vp, c_vp, u32, u32, u32, c_vp, i32, c_vp, u32, c_vp, E]),
    "prr_batch_create": (i32, [c_vp, ctypes.POINTER(c_vp), E]),
    "prr_batch_add": (i32, [c_vp, H, FI, ctypes.POINTER(PrrBatchFrame), E]),
    "prr_batch_set_config": (i32, [c_vp, u32, u32, u32, E]),
    "prr_batch_clear": (i32, [c_vp, E]),
    "prr_batch_run": (i32, [c_vp, c_vp, E]),
    "prr_batch_info": (i32, [c_vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    "prr_batch_destroy": (None, [c_vp]),
    "prr_net_create": (i32, [c_vp, ctypes.POINTER(c_vp), E]),
    "prr_net_destroy": (None, [c_vp]),
    "prr_net_set_params": (i32, [c_vp, c_vp, E]),
    "prr_net_get_params": (i32, [c_vp, c_vp, E]),
    "prr_net_set_range": (i32, [c_vp, ctypes.c_float, ctypes.c_float, E]),
    "prr_net_train_setup": (i32, [c_vp, ctypes.POINTER(PrrNetFrame), u32, c_vp, u32, u32, u32,
                                  u32, E]),
    "prr_net_set_schedule": (i32, [c_vp, ctypes.c_float, u32, E]),
    "prr_net_set_loss": (i32, [c_vp, ctypes.c_float, E]),
    "prr_net_train_step": (i32, [c_vp, ctypes.c_float, c_vp, E]),
    "prr_net_read": (i32, [c_vp, i32, u32, c_vp, sz, E]),
    "prr_net_develop_device": (i32, [c_vp, H, c_vp, DP, c_vp, i32, u32, ctypes.c_float, c_vp,
                                     c_vp, c_vp, c_vp, E]),
    "prr_net_last_check": (i32, [c_vp, ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_float), E]),
    "prr_net_set_pairs": (i32, [c_vp, i32, E]),
    "prr_is_prores_raw_fourcc": (i32, [c_vp]),
    "prr_mov_read_video_track": (i32, [i32, MOV_ACCEPT, c_vp, ctypes.POINTER(c_vp),
                                       ctypes.c_char_p, sz]),
    "prr_mov_track_free": (None, [c_vp]),
    "prr_mov_track_fourcc": (None, [c_vp, c_vp]),
    "prr_mov_track_fps": (None, [c_vp, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    "prr_mov_track_variable_durations": (i32, [c_vp]),
    "prr_mov_track_samples": (sz, [c_vp, ctypes.POINTER(ctypes.POINTER(PrrMovSample))]),
    "prr_mov_track_creation_time": (i32, [c_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "prr_mov_track_orientation": (i32, [c_vp]),
    "prr_mov_track_n_warnings": (sz, [c_vp]),
    "prr_mov_track_warning": (c_vp, [c_vp, sz, ctypes.POINTER(sz)]),
    "prr_mov_track_n_metadata": (sz, [c_vp]),
    "prr_mov_track_metadata_key": (c_vp, [c_vp, sz, ctypes.POINTER(sz)]),
    "prr_mov_track_metadata_value": (c_vp, [c_vp, sz, ctypes.POINTER(sz)]),
    "prr_mov_track_text": (c_vp, [c_vp, ctypes.POINTER(ctypes.c_char_p), sz, ctypes.POINTER(sz)]),
    "prr_mov_read_at": (i32, [i32, ctypes.c_uint64, c_vp, sz, ctypes.c_char_p, sz]),
    "prr_mov_read_samples": (i32, [i32, ctypes.POINTER(PrrMovSample), sz, c_vp,
                                   ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint, ctypes.c_char_p,
                                   sz]),
}

_lib = None


def _share_torch_runtime():
    """torch ships its own HIP runtime under the same soname
    (libamdhip64.so.7) as the one the library links. Loading torch first
    makes the library bind to that runtime, so a process that also uses
    torch has one HIP runtime (device memory, streams and graphs shared);
    loading the library first would give the process two runtimes, and
    torch would then find no device."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """The loaded C-ABI library (raises OSError when it is not built)."""
    global _lib
    if _lib is None:
        _share_torch_runtime()
        if not os.path.exists(LIB_PATH):
            raise OSError(
                f"{LIB_PATH} is missing: build it with `python -m proresraw_rs_amd.build` "
                "(hipcc for gfx950); there is no CPU fallback")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib
