/* proresraw_amd: ProRes RAW decode and develop on AMD Instinct MI355X
 * (gfx950) with hand-written HIP kernels, behind a C ABI.
 *
 * The entry points replace the public API of the Rust crate proresraw_rs
 * (re-exported at src/lib.rs:50-66); each declaration names the reference
 * function it stands in for. Conventions:
 *   - every function returns 0 (PRR_OK) or a PRR_E_* status and, when `err`
 *     is not NULL, fills it with the matching error payload;
 *   - callers own every buffer; sizes are passed explicitly, and a size that
 *     does not match returns PRR_E_OUTPUT_SIZE where the crate panics
 *     ("output size", "plane size");
 *   - malformed tiles are not errors: they are reported as skipped tiles;
 *   - a prr_context owns one HIP stream and the device scratch of one host
 *     thread; use one context per thread. Host-buffer entry points are
 *     synchronous on the context's stream; *_device entry points take
 *     device pointers and enqueue on the given hipStream_t as is (NULL is
 *     the HIP null stream; prr_context_stream gives the context's).
 */
#ifndef PRORESRAW_AMD_H
#define PRORESRAW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 1..12 mirror the variants of the crate's Error enum
 * (src/error.rs:7-63) in declaration order. */
enum prr_status {
    PRR_OK = 0,
    PRR_E_TRUNCATED = 1,            /* Error::Truncated {what, needed=a, available=b} */
    PRR_E_FRAME_SIZE_MISMATCH = 2,  /* Error::FrameSizeMismatch {declared=a, actual=b} */
    PRR_E_BAD_QUANT_MATRIX = 3,     /* Error::BadQuantMatrix {index=a, value=b} */
    PRR_E_BAD_CURVE = 4,            /* Error::BadCurve(points) */
    PRR_E_UNEXPECTED_CURVE = 5,     /* Error::UnexpectedCurve */
    PRR_E_BAD_TAG = 6,              /* Error::BadTag(tag) */
    PRR_E_HEADER_TOO_SHORT = 7,     /* Error::HeaderTooShort(a) */
    PRR_E_UNSUPPORTED_VERSION = 8,  /* Error::UnsupportedVersion(a) */
    PRR_E_BAD_DIMENSIONS = 9,       /* Error::BadDimensions {width=a, height=b} */
    PRR_E_BAD_ALIGNMENT = 10,       /* Error::BadAlignment(a) */
    PRR_E_TILE_OUT_OF_BOUNDS = 11,  /* Error::TileOutOfBounds {tile=a} */
    PRR_E_BAD_TILE_HEADER = 12,     /* Error::BadTileHeader {tile=a} */
    PRR_E_OUTPUT_SIZE = 100,        /* a caller buffer has the wrong size (crate: panic) */
    PRR_E_INVALID_ARGUMENT = 101,   /* NULL pointer, bad enum value, misaligned buffer */
    PRR_E_DEVICE = 102,             /* a HIP runtime call failed; `what` names it */
    PRR_E_DECODE_OVERRUN = 103,     /* a stream exceeded the per-stream symbol bound */
};

typedef struct {
    int32_t code;
    char what[28];
    uint64_t a;
    uint64_t b;
    uint16_t points[8];
    uint8_t tag[4];
} prr_error;

/* FrameHeader (src/header.rs:112-146). */
typedef struct {
    uint8_t version;
    uint8_t vendor[4];
    uint8_t bayer;            /* BayerPattern code: 0 RGGB, 1 GRBG, 2 GBRG, 3 BGGR */
    uint16_t width;
    uint16_t height;
    uint8_t crop[4];          /* left, right, top, bottom */
    uint16_t white_level;
    uint16_t wb_cct;
    uint16_t flags;
    uint8_t alignment;
    uint8_t quant_matrix[64];
    uint16_t curve[8];
    float wb_red;
    float wb_blue;
    float color_matrix[9];    /* row-major camera RGB to XYZ */
    float gain;
} prr_header;

/* Sizes derived from a header (header.rs:249-256, tile.rs:47-72,
 * reduced.rs:99-111). */
typedef struct {
    uint32_t coded_width;
    uint32_t coded_height;
    uint32_t tiles_per_row;
    uint32_t tile_rows;
    uint32_t nb_tiles;
    size_t header_end;        /* bytes from the packet start to the size table */
} prr_frame_info;

/* DemosaicParams (src/demosaic.rs:57-74). */
typedef struct {
    int32_t cfa_phase;
    int32_t black;
    float one_over_range;
    float noise[3];
    float clip[2];
    float black_over_range;
    float wb[2];
} prr_demosaic_params;

enum prr_idct { PRR_IDCT_APPLE = 0, PRR_IDCT_FFMPEG = 1 };               /* idct.rs:19-26 */
enum prr_scale { PRR_SCALE_HALF = 1, PRR_SCALE_QUARTER = 2, PRR_SCALE_EIGHTH = 3 }; /* reduced.rs:25-32 */
enum prr_greens { PRR_GREENS_FIRST = 0, PRR_GREENS_AVERAGE = 1 };          /* reduced.rs:64-72 */
enum prr_stage { PRR_STAGE_DEMOSAICED = 0, PRR_STAGE_WHITE_BALANCED = 1,
                 PRR_STAGE_REC2020 = 2 };                                    /* demosaic.rs:112-123 */
enum prr_exposure { PRR_EXPOSURE_NORMALIZED = 0, PRR_EXPOSURE_REFLECTION = 1 }; /* demosaic.rs:127-133 */

/* Per-frame result the device path writes: the first out-of-bounds tile
 * (UINT32_MAX when none; parse_frame's TileOutOfBounds), the largest tile
 * qscale byte (ParsedFrame::max_qscale, decoder.rs:104-116), the number of
 * skipped tiles and a non-zero `overrun` if a stream hit the symbol bound. */
typedef struct {
    uint32_t oob_tile;
    uint32_t max_qscale;
    uint32_t n_skipped;
    uint32_t overrun;
} prr_device_status;

typedef struct prr_context prr_context;

/* ---- context ---------------------------------------------------------- */
int prr_context_create(int device, prr_context **out, prr_error *err);
void prr_context_destroy(prr_context *ctx);
/* The context's HIP stream (hipStream_t). */
void *prr_context_stream(prr_context *ctx);
/* Version string of this library. */
const char *prr_version(void);
/* Decode launch shape (no counterpart in the crate, which has no launch to
 * shape): lanes decoding streams per workgroup, 8, 16, 32 or 64, or 0 for
 * the automatic width (the default): the narrowest that keeps the decode at
 * 1.25 waves a SIMD or fewer, counting every frame decoding at once (a
 * 4096x2160 frame decodes 16 wide, a 5760x3240 one 32 wide, two 5760x3240
 * frames of a pipeline 64 wide). */
int prr_set_decode_config(prr_context *ctx, uint32_t streams_per_wg, prr_error *err);

/* ---- host-side frame structure (no device work) ----------------------- */
/* header_len (decoder.rs:58-73). */
int prr_header_len(const uint8_t *packet, size_t len, size_t *out, prr_error *err);
/* parse_header (decoder.rs:45-53). */
int prr_parse_header(const uint8_t *packet, size_t len, prr_header *out, prr_error *err);
/* Prefix checks of parse_frame (decoder.rs:76-103) plus the grid sizes. */
int prr_parse_frame_info(const uint8_t *packet, size_t len, prr_header *hdr, prr_frame_info *info,
                         prr_error *err);
/* ParsedFrame (decoder.rs:15-29): the header, the tile grid's sizes and
 * the largest tile qscale byte; the tile payloads are the
 * tile_offsets / tile_sizes prr_parse_frame writes. */
typedef struct {
    prr_header header;
    prr_frame_info info;
    uint8_t max_qscale;
} prr_parsed_frame;
/* parse_frame (decoder.rs:76-123) on the host, for callers that only parse:
 * the prefix and header checks of prr_parse_frame_info, then the size-table
 * walk: tile i's payload is bytes [tile_offsets[i], tile_offsets[i] +
 * tile_sizes[i]) of the packet, max_qscale is the largest byte after each
 * tile's start (read for invalid and short tiles too, 0 past the packet),
 * and the first tile running past the packet is PRR_E_TILE_OUT_OF_BOUNDS
 * {tile = a}. tile_offsets and tile_sizes may both be NULL; otherwise they
 * hold `cap` entries, and cap < nb_tiles returns PRR_E_OUTPUT_SIZE (a =
 * nb_tiles) with out->header and out->info filled, so a caller can size
 * them. */
int prr_parse_frame(const uint8_t *packet, size_t len, prr_parsed_frame *out,
                    uint64_t *tile_offsets, uint32_t *tile_sizes, size_t cap, prr_error *err);
/* quantizer / Frame::quantizer (decoder.rs:33-35, 150-152). */
uint32_t prr_quantizer(const prr_header *hdr, uint8_t max_qscale);
/* frame_quantizer (demosaic.rs:50-52): the quantizer of a parsed frame. */
uint32_t prr_frame_quantizer(const prr_parsed_frame *parsed);
/* DemosaicParams::new / with_white_balance (demosaic.rs:80-106). */
void prr_demosaic_params_new(const prr_header *hdr, uint32_t quantizer,
                             prr_demosaic_params *out);
void prr_demosaic_params_with_white_balance(const prr_header *hdr, uint32_t quantizer, float red,
                                            float blue, prr_demosaic_params *out);
/* output_matrix (demosaic.rs:138-153). */
void prr_output_matrix(const prr_header *hdr, int exposure, float m[9]);

/* ---- host-buffer entry points (drop-in; synchronous) ------------------ */
/* decode_frame / decode_frame_with (decoder.rs:221-267): `samples` holds
 * coded_width*coded_height code values; skipped tile indices ascend. */
int prr_decode_frame(prr_context *ctx, const uint8_t *packet, size_t len, int idct,
                     prr_header *hdr, uint16_t *samples, size_t samples_len, uint32_t *skipped,
                     size_t skipped_cap, size_t *n_skipped, uint8_t *max_qscale,
                     prr_error *err);
/* Frame::linearize_into (decoder.rs:197-216): `samples` at stride
 * coded_width, `out` width*height linear values. */
int prr_linearize_into(prr_context *ctx, const prr_header *hdr, const uint16_t *samples,
                       size_t samples_len, uint16_t *out, size_t out_len, prr_error *err);
/* decode_frame_reduced / decode_frame_reduced_with (reduced.rs:142-201):
 * `rgb` holds 3*(coded_width>>scale)*(coded_height>>scale) values. */
int prr_decode_frame_reduced(prr_context *ctx, const uint8_t *packet, size_t len, int scale,
                             int greens, prr_header *hdr, uint16_t *rgb, size_t rgb_len,
                             uint32_t *skipped, size_t skipped_cap, size_t *n_skipped,
                             uint8_t *max_qscale, prr_error *err);
/* develop_planar_into (demosaic.rs:250-261): planes of width*height. */
int prr_develop_planar_into(prr_context *ctx, const prr_header *hdr, const uint16_t *samples,
                            size_t samples_len, const prr_demosaic_params *params, int stage,
                            float *r, float *g, float *b, size_t plane_len, prr_error *err);
/* develop_into (demosaic.rs:223-227): interleaved RGB, 3*width*height. */
int prr_develop_into(prr_context *ctx, const prr_header *hdr, const uint16_t *samples,
                     size_t samples_len, const prr_demosaic_params *params, int stage,
                     float *rgb, size_t rgb_len, prr_error *err);

/* ---- device-resident twins (enqueue only) ------------------------------ */
/* Decode a packet already in device memory. `hdr` and `info` come from
 * prr_parse_frame_info on the packet's first bytes. `d_packet` must be 4-byte
 * aligned and readable for ((len + 3) & ~3) + 8 bytes. `d_samples` (16-byte
 * aligned) holds coded_width*coded_height values; the decode also uses it
 * to stage each tile's coefficients before they are transformed, so its
 * contents are undefined until the decode completes. `d_linear` (4-byte
 * aligned), when not NULL, receives the linearized visible picture
 * (width*height) in the same pass;
 * `d_tile_flags`, when not NULL, gets one byte per tile (1 = skipped);
 * `d_status` receives the frame's prr_device_status. */
int prr_decode_frame_device(prr_context *ctx, const prr_header *hdr, const prr_frame_info *info,
                            const uint8_t *d_packet, size_t len, int idct, uint16_t *d_samples,
                            uint16_t *d_linear, uint8_t *d_tile_flags,
                            prr_device_status *d_status, void *stream, prr_error *err);
/* No counterpart in the crate: prr_decode_frame_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments, rules
 * and results; it calls no hipMalloc, hipFree or allocating routine. Its
 * scratch (tile offsets and tile order) is a static device array of
 * 3 MiB per GPU, set up by prr_context_create and held until the process
 * ends, so frames of more than PRR_NOALLOC_MAX_TILES tiles are refused
 * with PRR_E_INVALID_ARGUMENT (a 8192x4320 frame has at most 138,240).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (on different streams, or from several threads) race, and the caller
 * serializes them. */
#define PRR_NOALLOC_MAX_TILES 262144u
int prr_decode_frame_device_noalloc(prr_context *ctx, const prr_header *hdr,
                                    const prr_frame_info *info, const uint8_t *d_packet,
                                    size_t len, int idct, uint16_t *d_samples,
                                    uint16_t *d_linear, uint8_t *d_tile_flags,
                                    prr_device_status *d_status, void *stream, prr_error *err);
/* Frame::linearize_into on device buffers. */
int prr_linearize_device(prr_context *ctx, const prr_header *hdr, const uint16_t *d_samples,
                         uint16_t *d_out, void *stream, prr_error *err);
/* Reduced decode of a packet in device memory (same packet rules). */
int prr_decode_frame_reduced_device(prr_context *ctx, const prr_header *hdr,
                                    const prr_frame_info *info, const uint8_t *d_packet,
                                    size_t len, int scale, int greens, uint16_t *d_rgb,
                                    uint8_t *d_tile_flags, prr_device_status *d_status,
                                    void *stream, prr_error *err);
/* develop_planar_into on device buffers; d_g == d_b == NULL selects the
 * interleaved layout of develop_into with 3*width*height values in d_r. */
int prr_develop_device(prr_context *ctx, const prr_header *hdr, const uint16_t *d_samples,
                       const prr_demosaic_params *params, int stage, float *d_r, float *d_g,
                       float *d_b, void *stream, prr_error *err);
/* develop / develop_with (demosaic.rs:190-215) on device buffers with the
 * frame's quantizer read on the device from `d_status` (its max_qscale),
 * so decode and develop enqueue back to back with no readback:
 * DemosaicParams::new (explicit_wb = 0) or ::with_white_balance(wb_red,
 * wb_blue) (explicit_wb = 1). Planes as in prr_develop_device. */
int prr_develop_frame_device(prr_context *ctx, const prr_header *hdr, const uint16_t *d_samples,
                             const prr_device_status *d_status, int explicit_wb, float wb_red,
                             float wb_blue, int stage, float *d_r, float *d_g, float *d_b,
                             void *stream, prr_error *err);
/* prr_develop_frame_device reading the linearized picture instead of the
 * code values: d_linear is Frame::linearize's output (width*height values,
 * e.g. the d_linear of prr_decode_frame_device), so the curve is not
 * applied a second time. Same planes, bit for bit, as develop(&frame) on
 * the code values those came from (demosaic.rs:190-215; Conversion::new
 * linearizes each code value first, demosaic.rs:293-298). */
int prr_develop_linear_device(prr_context *ctx, const prr_header *hdr, const uint16_t *d_linear,
                              const prr_device_status *d_status, int explicit_wb, float wb_red,
                              float wb_blue, int stage, float *d_r, float *d_g, float *d_b,
                              void *stream, prr_error *err);

/* ---- frame pipeline (no counterpart in the crate, whose callers decode
 * and develop one frame per call): a stream of frames whose packets are in
 * device memory, decoded (with the linearize fused) on one HIP stream and
 * developed from the linear picture on a second, so frame i + 1's decode
 * (a latency-bound entropy chain) runs beside frame i's develop (VALU
 * bound). Consecutive frames decode on two streams of their own, so two
 * frames' decodes overlap (one frame has too few streams to fill the chip
 * with wide decode waves). The pipeline owns `depth` sets of intermediate
 * buffers (code values, linear values, tile flags, status) used in turn; a
 * decode waits on the device for the develop that last used its set. Each
 * frame's planes equal prr_develop_frame_device's after
 * prr_decode_frame_device. ---- */
typedef struct prr_pipeline prr_pipeline;
/* depth: buffer sets, 2..8 (0: 3). */
int prr_pipeline_create(prr_context *ctx, uint32_t depth, prr_pipeline **out, prr_error *err);
/* Enqueue one frame (packet rules of prr_decode_frame_device; the packet
 * must stay valid until the pipeline is joined or synchronized): decode with
 * `idct`, then develop(&frame, stage) (explicit_wb as
 * prr_develop_frame_device) into the planes. d_status_out, when not NULL,
 * receives the frame's prr_device_status after its develop. Returns at once;
 * an argument it refuses queues no work.
 * Ordering: the decode runs on a stream of the pipeline's own and the
 * develop on another, so neither waits for work the caller queued on its
 * streams. A caller that writes the packet, or still reads earlier planes,
 * on a stream of its own calls prr_pipeline_wait(p, that stream) first;
 * then prr_pipeline_join (or _sync) before reading the planes. */
int prr_pipeline_submit(prr_pipeline *p, const prr_header *hdr, const prr_frame_info *info,
                        const uint8_t *d_packet, size_t len, int idct, int explicit_wb,
                        float wb_red, float wb_blue, int stage, float *d_r, float *d_g,
                        float *d_b, prr_device_status *d_status_out, prr_error *err);
/* Make the pipeline wait, on the device, for the work queued on `stream`
 * so far: frames submitted after this call start their decode and their
 * develop only after it (a packet uploaded or planes read on the caller's
 * stream). */
int prr_pipeline_wait(prr_pipeline *p, void *stream, prr_error *err);
/* Make `stream` wait, on the device, for every frame submitted so far. */
int prr_pipeline_join(prr_pipeline *p, void *stream, prr_error *err);
/* Wait on the host for every frame submitted so far. */
int prr_pipeline_sync(prr_pipeline *p, prr_error *err);
/* Per-kernel timing with HIP events on the pipeline's own streams: returns
 * the mean decode (prepare + decode) and develop milliseconds of the frames
 * submitted since the last call (waiting for them), resets, and times the
 * frames submitted from now on when `enable` is set. */
int prr_pipeline_timing(prr_pipeline *p, int enable, float *decode_ms, float *develop_ms,
                        uint64_t *frames, prr_error *err);
void prr_pipeline_destroy(prr_pipeline *p);

/* ---- DNG raw tiles: the LosslessJpeg branch of the DNG writer
 * (examples/prr2dng/dng.rs:92-113): every 256x256 tile of a width x height
 * picture, padded as dng.rs `tile` pads it (sensels past the right or
 * bottom edge repeat the last column or row of the same CFA phase,
 * dng.rs:115-128), coded as lj92::encode(tile, 256, 256, 2)
 * (examples/prr2dng/lj92.rs:14-93: two interleaved components, predictor
 * 1, one optimal Huffman table a tile). The streams are written back to
 * back in tile order (row-major over tiles) with each one's byte offset and
 * length (TileOffsets / TileByteCounts relative to the first). The picture
 * is the visible code values (Frame::rows, stride coded_width) or the
 * linearized values (stride width). width and height are at least 2. ---- */
/* Worst-case bytes of the streams (every tile at 31 bits a sample, every
 * byte stuffed). */
size_t prr_dng_lj92_capacity(uint32_t width, uint32_t height);
/* Device buffers, enqueued on `stream`: d_out holds at least
 * prr_dng_lj92_capacity bytes, d_offsets and d_lengths one uint32 per tile.
 * Uses the context's scratch: calls on one context are serialized by the
 * caller. */
int prr_dng_lj92_tiles_device(prr_context *ctx, const uint16_t *d_values, uint32_t width,
                              uint32_t height, uint32_t stride, uint8_t *d_out, size_t out_cap,
                              uint32_t *d_offsets, uint32_t *d_lengths, void *stream,
                              prr_error *err);
/* Host buffers, synchronous: `n_tiles` must be the picture's tile count;
 * *total receives the streams' size, and PRR_E_OUTPUT_SIZE is returned
 * (with *total set) when out_cap is smaller. */
int prr_dng_lj92_tiles(prr_context *ctx, const uint16_t *values, uint32_t width, uint32_t height,
                       uint32_t stride, uint8_t *out, size_t out_cap, uint32_t *offsets,
                       uint32_t *lengths, size_t n_tiles, size_t *total, prr_error *err);

/* ---- crop and orientation for display. The crate records them and
 * leaves them to readers: mov::Orientation (src/mov.rs:121-157, the eight
 * EXIF symmetries of a track's display matrix) and header::Crop
 * (src/header.rs:98-109), written as the DNG Orientation and DefaultCrop
 * tags (examples/prr2dng/dng.rs:155-158, 190-196). These entry points apply
 * them on the device: the picture is cropped by crop = {left, right, top,
 * bottom} (not at all when the margins leave no sensel, as dng.rs writes no
 * DefaultCrop then), then turned as EXIF orientation 1..8 says to show it
 * upright (6 = turned 90 degrees clockwise). ---- */
/* Size after crop and orientation (PRR_E_INVALID_ARGUMENT for an
 * orientation outside 1..8 or an empty picture). */
int prr_oriented_size(uint32_t width, uint32_t height, const uint8_t crop[4], int orientation,
                      uint32_t *out_width, uint32_t *out_height);
/* BayerPattern code (0 RGGB, 1 GRBG, 2 GBRG, 3 BGGR) of a CFA picture with
 * pattern `bayer` after prr_orient_cfa_device; -1 when invalid. */
int prr_oriented_bayer(int bayer, const uint8_t crop[4], uint32_t width, uint32_t height,
                       int orientation);
/* The three develop planes (width x height at in_stride floats a row) to
 * three output planes of prr_oriented_size at out_stride; enqueued on
 * `stream`. Inputs and outputs must not overlap. */
int prr_orient_planes_device(prr_context *ctx, const float *d_r, const float *d_g,
                             const float *d_b, uint32_t width, uint32_t height,
                             uint32_t in_stride, const uint8_t crop[4], int orientation,
                             float *d_out_r, float *d_out_g, float *d_out_b,
                             uint32_t out_stride, void *stream, prr_error *err);
/* A CFA picture (code or linear values) likewise. */
int prr_orient_cfa_device(prr_context *ctx, const uint16_t *d_in, uint32_t width,
                          uint32_t height, uint32_t in_stride, const uint8_t crop[4],
                          int orientation, uint16_t *d_out, uint32_t out_stride, void *stream,
                          prr_error *err);

/* ---- batches (no counterpart in the crate, which decodes one frame per
 * call): many frames of any size, version, alignment and qscale decoded
 * in one persistent launch whose waves pull tile groups from a device
 * queue ordered by coded byte length; each frame is a full decode (with
 * develop or without) or a reduced decode. Results equal the per-frame
 * entry points'. ------------------------------------------------------- */
enum prr_batch_kind { PRR_BATCH_DECODE = 0, PRR_BATCH_DEVELOP = 1, PRR_BATCH_REDUCED = 2 };

typedef struct {
    const uint8_t *d_packet;  /* prr_decode_frame_device's packet rules */
    size_t len;
    int32_t kind;             /* prr_batch_kind */
    int32_t idct;             /* DECODE / DEVELOP */
    int32_t scale;            /* REDUCED: prr_scale */
    int32_t greens;           /* REDUCED: prr_greens */
    int32_t stage;            /* DEVELOP: prr_stage */
    int32_t explicit_wb;      /* DEVELOP: 0 header gains, 1 wb_red / wb_blue */
    float wb_red;
    float wb_blue;
    uint16_t *d_samples;      /* DECODE / DEVELOP: coded_width * coded_height, 16-byte aligned */
    uint16_t *d_linear;       /* DECODE / DEVELOP, optional: width * height, 4-byte aligned */
    uint16_t *d_rgb;          /* REDUCED */
    float *d_r, *d_g, *d_b;   /* DEVELOP: planes (or interleaved in d_r) */
    uint8_t *d_tile_flags;    /* optional, one byte per tile */
    prr_device_status *d_status;
} prr_batch_frame;

typedef struct prr_batch prr_batch;
int prr_batch_create(prr_context *ctx, prr_batch **out, prr_error *err);
/* Add a frame (at most 256 per batch); `hdr`/`info` from prr_parse_frame_info. */
int prr_batch_add(prr_batch *batch, const prr_header *hdr, const prr_frame_info *info,
                  const prr_batch_frame *frame, prr_error *err);
/* Remove every frame, keeping the batch's device buffers for the next
 * frames (a clip decoded window by window); the launch shape stays. */
int prr_batch_clear(prr_batch *batch, prr_error *err);
/* Launch shape: streams per workgroup (16/32/64, 64 by default; a frame's
 * is halved until its group fits the batch's 32 KiB LDS budget; frames
 * already added are laid out again), queue entries taken per atomic grab,
 * waves per CU of the persistent launch (0 = as many as fit); 0 keeps the
 * current spw / chunk.
 * Takes effect at the next prr_batch_run, which then uploads the frame
 * table again (capture a graph only after that run). */
int prr_batch_set_config(prr_batch *batch, uint32_t streams_per_wg, uint32_t chunk,
                         uint32_t waves_per_cu, prr_error *err);
/* Enqueue the whole batch on `stream`. The first call after adding frames
 * also uploads the frame table and synchronizes; later calls only launch
 * kernels and may be captured in a HIP graph. */
int prr_batch_run(prr_batch *batch, void *stream, prr_error *err);
/* Tile groups, persistent grid size and LDS bytes per workgroup of the
 * last run. */
int prr_batch_info(prr_batch *batch, uint32_t *n_items, uint32_t *grid, uint32_t *lds_bytes);
void prr_batch_destroy(prr_batch *batch);

#ifdef __cplusplus
}
#endif
#endif
