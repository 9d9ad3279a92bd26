/* Oracle: a CPU restatement of proresraw_rs's decode and develop path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or the timed CPU baseline). The product path never calls it.
 *
 * Every function cites the reference file:line it restates. The reference
 * is the Rust crate proresraw_rs (its src/ files); it cannot be compiled here (no
 * Rust toolchain), so this C restatement is pinned instead by the crate's
 * own golden vectors and known-answer tests (see tests/test_oracle_*.py).
 */
#ifndef PRR_ORACLE_H
#define PRR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes, one per variant of error.rs:7-63. */
enum {
    ORC_OK = 0,
    ORC_E_TRUNCATED = 1,
    ORC_E_FRAME_SIZE_MISMATCH = 2,
    ORC_E_BAD_QUANT_MATRIX = 3,
    ORC_E_BAD_CURVE = 4,
    ORC_E_UNEXPECTED_CURVE = 5,
    ORC_E_BAD_TAG = 6,
    ORC_E_HEADER_TOO_SHORT = 7,
    ORC_E_UNSUPPORTED_VERSION = 8,
    ORC_E_BAD_DIMENSIONS = 9,
    ORC_E_BAD_ALIGNMENT = 10,
    ORC_E_TILE_OUT_OF_BOUNDS = 11,
    ORC_E_BAD_TILE_HEADER = 12,
};

/* Error payload (error.rs:7-63): `a`/`b` carry needed/available,
 * declared/actual, index/value, width/height, version, alignment, tile or
 * header size depending on the code. */
typedef struct {
    int32_t code;
    char what[28];
    uint64_t a;
    uint64_t b;
    uint16_t points[8];
    uint8_t tag[4];
} orc_error;

/* FrameHeader (header.rs:112-146). */
typedef struct {
    uint8_t version;
    uint8_t vendor[4];
    uint8_t bayer;        /* 0 RGGB, 1 GRBG, 2 GBRG, 3 BGGR */
    uint16_t width;
    uint16_t height;
    uint8_t crop[4];      /* left, right, top, bottom */
    uint16_t white_level;
    uint16_t wb_cct;
    uint16_t flags;
    uint8_t alignment;
    uint8_t quant_matrix[64];
    uint16_t curve[8];
    float wb_red;
    float wb_blue;
    float color_matrix[9];
    float gain;
} orc_header;

/* DemosaicParams (demosaic.rs:57-74). */
typedef struct {
    int32_t cfa_phase;
    int32_t black;
    float one_over_range;
    float noise[3];
    float clip[2];
    float black_over_range;
    float wb[2];
} orc_params;

/* --- primitives ------------------------------------------------------- */
void orc_codeword(uint32_t p, uint32_t cb, uint32_t *value, uint32_t *len);
/* Read up to `n` symbols with codebook list `cbs`, stopping at a zero view;
 * returns the count read (bits.rs tests drive BitReader this way). */
size_t orc_read_symbols(const uint8_t *data, size_t len, unsigned align,
                        const uint16_t *cbs, size_t n, uint32_t *out);
int orc_count_symbols(const uint8_t *data, size_t len, unsigned align, unsigned log2_blocks,
                      unsigned log2_coefs, uint64_t cap, uint64_t *ac_symbols,
                      uint64_t *consumed_bits);
size_t orc_frame_stream_census(const uint8_t *packet, size_t len, uint32_t *bytes,
                               uint32_t *symbols, size_t cap);
void orc_decode_component_reduced(const uint8_t *data, size_t len, unsigned align,
                                  unsigned log2_blocks, unsigned log2_coefs,
                                  int16_t *coefs);
void orc_apple_block(const int16_t *coefs, const int16_t *q, uint16_t *out);
void orc_ffmpeg_block(const int16_t *coefs, const int16_t *q, uint16_t *out);
void orc_apple_block_4x4(const int16_t *coefs, const int16_t *q, uint16_t *out);
void orc_apple_block_2x2(const int16_t *coefs, const int16_t *q, uint16_t *out);
uint16_t orc_output_code(float v);
uint16_t orc_code_value(int32_t centered);
int16_t orc_dequantize(int16_t c, int16_t q);
/* FNV-1a of the Apple transform of `n` testgen blocks (idct.rs:356-460). */
uint64_t orc_testgen_hash(uint64_t seed, uint64_t n, int ffmpeg);
void orc_testgen_block(uint64_t seed, uint64_t i, int16_t *c, int16_t *q);
uint16_t orc_linearize(const uint16_t *curve, uint16_t code);

/* --- frame structure ------------------------------------------------- */
int orc_header_len(const uint8_t *packet, size_t len, size_t *out, orc_error *err);
int orc_parse_header(const uint8_t *packet, size_t len, orc_header *out, orc_error *err);
size_t orc_tiles_per_row(unsigned width, unsigned alignment);
size_t orc_tile_count(unsigned width, unsigned height, unsigned alignment);
/* Geometry of every tile in frame order: x, y in sensels and log2_blocks. */
void orc_tile_geometry(unsigned width, unsigned height, unsigned alignment,
                       uint32_t *x, uint32_t *y, uint8_t *log2_blocks);
/* Tile::parse (tile.rs:126-154): 0 when valid, with qscale and the four
 * streams' offsets (relative to the tile start) and lengths. */
int orc_tile_parse(const uint8_t *data, size_t len, uint8_t *qscale,
                   uint32_t *offsets, uint32_t *lengths);
/* parse_frame (decoder.rs:76-123): tile start offsets (nb_tiles entries,
 * relative to the packet) and lengths, max_qscale. */
int orc_parse_frame(const uint8_t *packet, size_t len, orc_header *hdr,
                    uint64_t *tile_offsets, uint32_t *tile_sizes, size_t cap,
                    size_t *nb_tiles, uint8_t *max_qscale, orc_error *err);

/* --- decoders --------------------------------------------------------- */
/* decode_frame_with (decoder.rs:226-267): `samples` holds coded_w*coded_h;
 * `skipped` receives the skipped tile indices (capacity nb_tiles). */
int orc_decode_frame(const uint8_t *packet, size_t len, int ffmpeg, int threads,
                     uint16_t *samples, size_t samples_cap, uint32_t *skipped,
                     size_t skipped_cap, size_t *n_skipped, uint8_t *max_qscale,
                     orc_error *err);
/* Frame::linearize_into (decoder.rs:197-216). */
void orc_linearize_frame(const uint16_t *curve, const uint16_t *samples, size_t stride,
                         size_t width, size_t height, uint16_t *out, int threads);
/* decode_frame_reduced_with (reduced.rs:147-201): log2_factor 1..3,
 * average_greens 0/1; `rgb` holds 3*(coded_w>>s)*(coded_h>>s). */
int orc_decode_frame_reduced(const uint8_t *packet, size_t len, unsigned log2_factor,
                             int average_greens, int threads, uint16_t *rgb,
                             size_t rgb_cap, uint32_t *skipped, size_t skipped_cap,
                             size_t *n_skipped, uint8_t *max_qscale, orc_error *err);

/* --- develop ---------------------------------------------------------- */
void orc_params_with_white_balance(const orc_header *h, uint32_t quantizer, float red,
                                   float blue, orc_params *out);
void orc_output_matrix(const orc_header *h, int normalized, float *m);
/* develop_planar_into / develop_into (demosaic.rs:223-261), stage 0
 * Demosaiced, 1 WhiteBalanced, 2 Rec2020. Interleaved when g == b == NULL
 * (then r holds 3*w*h values). */
void orc_develop(const orc_header *h, const uint16_t *samples, size_t stride,
                 const orc_params *p, int stage, int threads, float *r, float *g,
                 float *b);
/* green_pass / chroma_pass / apply_matrix (demosaic.rs:453-476, 608-617, 692-696)
 * on a tightly packed linear mosaic, for the module's own unit tests. */
void orc_green_pass(const uint16_t *linear, size_t w, size_t h, const orc_params *p,
                    float *out2);
void orc_chroma_pass(const float *green2, size_t w, size_t h, const orc_params *p,
                     float *rgb);
void orc_apply_matrix(float *rgb, size_t n, const float *m);

#ifdef __cplusplus
}
#endif
#endif
