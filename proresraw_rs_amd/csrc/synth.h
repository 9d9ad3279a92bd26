/* Synthetic ProRes RAW frames: the random-payload frames of proresraw_rs's
 * synth module (src/synth.rs) and a seeded scene encoder that produces
 * decodable frames with real picture content. Input generation only; the
 * decoder, not this encoder, defines every expected output. */
#ifndef PRR_SYNTH_H
#define PRR_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* synth::Spec (synth.rs:34-89) plus the bitstream version, which the
 * crate's builder fixes at 1 (synth.rs:102). */
typedef struct {
    uint16_t width;
    uint16_t height;
    uint8_t alignment;
    uint8_t pattern;       /* 0 RGGB, 1 GRBG, 2 GBRG, 3 BGGR */
    uint8_t version;
    uint8_t valid_tiles;
    uint8_t has_qmat;
    uint8_t has_curve;
    uint8_t qmat[64];
    uint16_t curve[8];
    uint8_t vendor[4];
    uint8_t crop[4];
    uint16_t white_level;
    uint16_t wb_cct;
    float wb[2];
    float color_matrix[9];
    float gain;
} prs_spec;

/* Scene encoder settings. */
typedef struct {
    prs_spec spec;
    uint64_t seed;
    uint8_t qscale_min;     /* per-tile qscale uniform in [min, max] */
    uint8_t qscale_max;
    uint8_t n_gradients;    /* linear gradients summed into the background */
    uint8_t reserved;
    uint32_t n_discs;       /* hard-edged discs painted on top */
    float corrupt_fraction; /* tiles whose size entry and 4th stream are dropped */
    float dn_per_electron;  /* Poisson gain, DN per electron */
    float read_sigma;       /* Gaussian read noise, DN */
    int32_t threads;
} prs_scene;

void prs_spec_default(prs_spec *s, uint16_t width, uint16_t height, uint8_t pattern);
uint64_t prs_rng_new(uint64_t seed);
uint32_t prs_rng_next(uint64_t *state);
/* synth::specs (synth.rs:188-221): 20 specs. */
void prs_specs(uint64_t *rng, prs_spec *out20);
/* Byte length of the frame packet prs_random_frame / prs_frame_streams write. */
size_t prs_random_frame(const prs_spec *spec, uint64_t *rng, uint8_t *out, size_t cap);
/* synth::frame with every tile built by synth::tile(qscale, streams). */
size_t prs_frame_streams(const prs_spec *spec, uint8_t qscale, const uint8_t *const streams[4],
                         const size_t lens[4], uint8_t *out, size_t cap);
/* Seeded scene, noise, mosaic, inverse curve, forward DCT, quantization and
 * entropy coding into a frame packet. Returns its length (0 on failure;
 * pass out == NULL to get an upper bound of the length). */
size_t prs_encode_scene(const prs_scene *cfg, uint8_t *out, size_t cap);
/* Entropy code one component: `levels` holds 64 << log2_blocks quantized
 * coefficients, block after block in raster order (the layout
 * decode_component produces). Returns the stream length in bytes. */
size_t prs_encode_component(const int16_t *levels, unsigned log2_blocks, uint8_t *out, size_t cap);
/* The mosaic of linear values the encoder coded (before compression), for
 * comparing the decoded picture with its source: width*height values. */
void prs_scene_mosaic(const prs_scene *cfg, uint16_t *linear);

#ifdef __cplusplus
}
#endif
#endif
