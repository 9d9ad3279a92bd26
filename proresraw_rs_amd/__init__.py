"""proresraw_rs_amd: ProRes RAW decode and develop on AMD MI355X (gfx950).

The public names mirror the Rust crate proresraw_rs (src/lib.rs:50-66);
every decode, linearize and develop runs in the hand-written HIP kernels of
``csrc/prr_kernels.hip`` through the C ABI of ``include/proresraw_amd.h``.
"""
from .api import (  # noqa: F401
    BLACK_LEVEL,
    FOURCCS,
    PREFIX_LEN,
    BayerPattern,
    DemosaicParams,
    Error,
    Exposure,
    Frame,
    FrameHeader,
    Greens,
    ParsedFrame,
    Idct,
    RgbFrame,
    RgbImage,
    RgbPlanes,
    Scale,
    Stage,
    context,
    decode_frame,
    decode_frame_reduced,
    decode_frame_reduced_with,
    decode_frame_with,
    develop,
    develop_into,
    develop_planar,
    develop_planar_into,
    develop_with,
    frame_quantizer,
    header_len,
    is_prores_raw_fourcc,
    output_matrix,
    parse_frame,
    parse_frame_info,
    parse_header,
)

__version__ = "0.1.0"
