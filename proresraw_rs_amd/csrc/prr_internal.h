/* Internal interface between the host layer (prr_host.cpp) and the HIP
 * kernels (prr_kernels.hip). */
#ifndef PRR_INTERNAL_H
#define PRR_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/proresraw_amd.h"

namespace prr {

/* Everything a decode kernel needs to know about one frame. */
struct FrameDesc {
    const uint8_t *packet;   /* device, 4-byte aligned, padded */
    uint64_t len;            /* packet length in bytes */
    uint64_t *offsets;       /* device scratch, nb_tiles + 1 tile starts */
    uint32_t table_off;      /* byte offset of the tile size table */
    uint32_t nb_tiles;
    uint32_t tiles_per_row;
    uint32_t columns;        /* 16-sensel columns */
    uint32_t alignment;
    uint32_t coded_width;
    uint32_t coded_height;
    uint32_t width;
    uint32_t height;
    uint32_t bayer;
    uint8_t qmat[64];
    uint16_t curve[8];
};

/* Develop parameters with the stage applied (demosaic.rs:286-304). */
struct DevelopDesc {
    const uint16_t *samples; /* code values at stride `stride` */
    uint32_t stride;
    uint32_t width;
    uint32_t height;
    int32_t cfa_phase;
    int32_t black;
    float one_over_range;
    float noise[3];
    float clip[2];
    float black_over_range;
    float wb[2];
    int32_t has_matrix;
    float m[9];
    uint16_t curve[8];
    /* When set, the noise model comes from the quantizer qmat0 * the
     * frame's max_qscale, read on the device (develop(&frame, stage)). */
    const prr_device_status *qstatus;
    uint32_t qmat0;
    uint32_t range;
    /* When set, the develop kernel runs only if *gate is non-zero (the
     * learned path's fallback decided on the device). */
    const uint32_t *gate;
    /* When set, `samples` holds linearized values (Frame::linearize) and the
     * curve is not applied again. */
    uint32_t linear_input;
    /* When set, the first workgroup copies *qstatus here (a pipeline
     * frame's status handed to its caller without a dispatch of its own). */
    prr_device_status *status_out;
};

enum BatchMode : uint32_t {
    BATCH_FULL_APPLE = 0,
    BATCH_FULL_FFMPEG = 1,
    BATCH_HALF = 2,
    BATCH_QUARTER = 3,
    BATCH_EIGHTH = 4,
};

/* Tiles of one workgroup of the batch's size-table scan (a multiple of
 * every group size, 1 to 16 tiles). */
enum { SCAN_CHUNK = 2048 };

/* One frame of a batch as the persistent kernel sees it. */
struct BatchFrame {
    FrameDesc f;
    uint32_t mode;
    uint32_t average;
    uint32_t spw;
    uint16_t *samples;       /* full decode: also the stage of its coefficients */
    uint16_t *linear;        /* full decode, optional */
    uint16_t *rgb;           /* reduced decode */
    uint8_t *flags;
    prr_device_status *status;
};

/* Develop tiles (pixels): every develop launch covers its frame with
 * tiles of this size. */
constexpr uint32_t DEVELOP_TILE_W = 64, DEVELOP_TILE_H = 16;

/* A develop of a batch: its frame's tiles are tiles first_tile ..
 * first_tile + tiles_x * ceil(height / DEVELOP_TILE_H) - 1 of the batch's
 * develop launch, row by row. */
struct BatchDevelop {
    DevelopDesc d;
    float *r, *g, *b;
    uint32_t first_tile, tiles_x;
};

/* Launch configuration of the full and reduced decode kernels. */
struct DecodeLaunch {
    uint32_t streams_per_wg; /* 16, 32 or 64 lanes decoding streams */
    uint32_t lds_limit;      /* bytes of LDS a workgroup may use */
    uint32_t solo_tiles;     /* longest tiles decoded one per workgroup */
};

/* Status reset + size-table scan + tile order (k_prepare), then the
 * decode; `queue` is device scratch of nb_tiles uint32. */
/* Crop + orientation of planes or a CFA picture (prr_orient.hip). */
struct OrientDesc {
    const void *in[3];
    void *out[3];
    uint32_t planes;      /* 3 planes of float, or 1 plane of uint16 */
    uint32_t orient;      /* EXIF 1..8 */
    uint32_t cw, ch;      /* cropped size */
    uint32_t ow, oh;      /* output size (cw x ch, swapped for 5..8) */
    uint32_t left, top;   /* crop origin in the input */
    uint32_t in_stride, out_stride;
};
hipError_t launch_orient(const OrientDesc &d, int element_bytes, hipStream_t s);

/* Lossless JPEG tiles of a DNG (prr_dng.hip). */
enum { LJ92_TILE = 256 };
/* Raw (unstuffed) entropy-coded words a tile may need: 31 bits a sample at
 * most (a 16-bit code and 15 extra bits), rounded up. */
enum { LJ92_RAW_WORDS = 63504 };
/* Bytes a tile's complete stream may need: headers, every raw byte
 * stuffed, EOI. */
enum { LJ92_TILE_BYTES = 128 + 2 * 4 * LJ92_RAW_WORDS };
struct Lj92Record {
    uint32_t raw_bits;   /* entropy-coded bits before the padding */
    uint32_t ff;         /* 0xFF bytes among the raw bytes */
    uint32_t header_len; /* SOI, SOF3, DHT, SOS */
    uint32_t total;      /* the tile's complete stream in bytes */
    uint8_t header[96];
};
hipError_t launch_lj92(const uint16_t *values, uint32_t stride, uint32_t width, uint32_t height,
                       uint32_t *raw, Lj92Record *records, uint8_t *out, uint32_t *offsets,
                       uint32_t *lengths, hipStream_t s);

hipError_t launch_decode(const FrameDesc &f, int ffmpeg, uint16_t *samples, uint16_t *linear,
                         uint8_t *flags, prr_device_status *status, uint32_t *queue,
                         const DecodeLaunch &cfg, hipStream_t s);
hipError_t launch_decode_reduced(const FrameDesc &f, int shift, int average, uint16_t *rgb,
                                 uint8_t *flags, prr_device_status *status, uint32_t *queue,
                                 const DecodeLaunch &cfg, hipStream_t s);
hipError_t launch_linearize(const uint16_t *samples, uint32_t stride, uint32_t width,
                            uint32_t height, const uint16_t *curve, uint16_t *out,
                            hipStream_t s);
hipError_t launch_develop(const DevelopDesc &d, float *r, float *g, float *b, hipStream_t s);
hipError_t launch_batch(const BatchFrame *d_frames, uint32_t n_frames, uint32_t n_items,
                        uint32_t *d_queue, uint32_t *d_head, uint32_t *d_hist,
                        const uint32_t *d_chunks, uint32_t n_chunks, uint32_t *d_chunk_bytes,
                        const BatchDevelop *d_jobs, uint32_t n_jobs, uint32_t dev_tiles, uint32_t grid,
                        uint32_t chunk, size_t lds, hipStream_t s);
size_t decode_lds_bytes(uint32_t spw, uint32_t per_block, uint32_t amax);
/* LDS of a full decode workgroup (any spw). */
size_t staged_decode_lds_bytes(uint32_t amax);

/* Layout of the allocation-free decode's static device scratch: tile
 * offsets (PRR_NOALLOC_MAX_TILES + 1 uint64, padded to 256 bytes), then the
 * tile order (PRR_NOALLOC_MAX_TILES uint32). */
enum : size_t {
    NOALLOC_OFFSETS_BYTES = ((size_t)8 * (PRR_NOALLOC_MAX_TILES + 1) + 255) & ~(size_t)255,
    NOALLOC_SCRATCH_BYTES = NOALLOC_OFFSETS_BYTES + (size_t)4 * PRR_NOALLOC_MAX_TILES,
};
enum { NOALLOC_MAX_DEVICES = 64 };
/* The device address of `device`'s copy of the static scratch (looked up
 * once per device; `device` must be current). */
hipError_t noalloc_scratch(int device, void **out);

/* Host side (prr_host.cpp). */
/* Conversion::new (demosaic.rs:286-304): the stage folded into the params. */
DevelopDesc develop_desc(const prr_header *hdr, const uint16_t *d_samples,
                         const prr_demosaic_params *p, int stage);
/* develop(&frame, stage) with the noise model read on the device. */
DevelopDesc develop_desc_device(const prr_header *hdr, const uint16_t *d_samples,
                                const prr_device_status *d_status, int explicit_wb, float red,
                                float blue, int stage);
int context_device(const prr_context *ctx);
const void *batch_decode_kernel();

} // namespace prr

#endif
