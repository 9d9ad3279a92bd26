/* Synthetic ProRes RAW frames.
 *
 * Part 1 restates proresraw_rs's synth module (src/synth.rs): a valid frame
 * structure around random tile payloads, used by the crate's tests to drive
 * decoders through malformed streams. Part 2 is new: a seeded scene of
 * linear gradients and hard-edged discs with Poisson-Gaussian noise,
 * mosaicked, mapped through the inverse of the frame's linearization curve,
 * forward-transformed, quantized and entropy coded into a frame the decoder
 * reads back. Its codeword writer follows the crate's test encoder
 * (bits.rs:157-174) and its packet layout synth::frame / synth::tile
 * (synth.rs:99-184).
 */
#include "synth.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- */
/* xorshift64* (synth.rs:9-29). */

uint64_t prs_rng_new(uint64_t seed) { return (seed * 0x9E3779B97F4A7C15ull) | 1; }

uint32_t prs_rng_next(uint64_t *s)
{
    *s ^= *s >> 12;
    *s ^= *s << 25;
    *s ^= *s >> 27;
    return (uint32_t)((*s * 0x2545F4914F6CDD1Dull) >> 32);
}

static uint32_t rng_below(uint64_t *s, uint32_t bound) { return prs_rng_next(s) % bound; }

/* Spec::new (synth.rs:71-88), with version 1 as the crate writes. */
void prs_spec_default(prs_spec *s, uint16_t width, uint16_t height, uint8_t pattern)
{
    memset(s, 0, sizeof *s);
    s->width = width;
    s->height = height;
    s->alignment = 3;
    s->pattern = pattern;
    s->version = 1;
    memcpy(s->vendor, "test", 4);
    s->white_level = 0xFDFF + 256;
    s->wb[0] = 1.0f;
    s->wb[1] = 1.0f;
    s->color_matrix[0] = s->color_matrix[4] = s->color_matrix[8] = 1.0f;
    s->gain = 1.0f;
    s->wb_cct = 5600;
}

/* ---------------------------------------------------------------- */
/* Tile grid (tile.rs:36-107). */

static size_t tiles_per_row(unsigned width, unsigned a)
{
    size_t columns = (width + 15) / 16;
    return (columns >> a) + (size_t)__builtin_popcountll(columns & ((1u << a) - 1));
}

typedef struct {
    uint32_t x, y;
    uint8_t log2;
} tile_geo;

static size_t grid(const prs_spec *s, tile_geo **out)
{
    size_t columns = (s->width + 15) / 16, rows = (s->height + 15) / 16;
    size_t n = tiles_per_row(s->width, s->alignment) * rows, k = 0;
    tile_geo *g = (tile_geo *)malloc(sizeof(tile_geo) * (n + 1));
    for (size_t r = 0; r < rows; r++) {
        size_t col = 0, rem = columns;
        int l = s->alignment;
        while (rem > 0) {
            size_t unit = (size_t)1 << l;
            if (unit <= rem) {
                g[k].x = (uint32_t)(col * 16);
                g[k].y = (uint32_t)(r * 16);
                g[k].log2 = (uint8_t)l;
                k++;
                col += unit;
                rem -= unit;
            } else {
                l--;
            }
        }
    }
    *out = g;
    return n;
}

/* ---------------------------------------------------------------- */
/* Frame assembly (synth.rs:99-152). */

static void put_be16(uint8_t *p, uint16_t v)
{
    p[0] = (uint8_t)(v >> 8);
    p[1] = (uint8_t)v;
}
static void put_be32(uint8_t *p, uint32_t v)
{
    p[0] = (uint8_t)(v >> 24);
    p[1] = (uint8_t)(v >> 16);
    p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
}
static void put_bef32(uint8_t *p, float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    put_be32(p, u);
}

/* Header body: 70 fixed bytes plus the optional matrix and curve. */
static size_t header_body(const prs_spec *s, uint8_t *body)
{
    memset(body, 0, 70);
    uint16_t flags = (uint16_t)(s->alignment << 1);
    body[1] = s->version;
    memcpy(body + 2, s->vendor, 4);
    put_be16(body + 6, s->width);
    put_be16(body + 8, s->height);
    memcpy(body + 10, s->crop, 4);
    put_be16(body + 14, s->pattern);
    put_be16(body + 16, (uint16_t)(s->white_level - 256));
    put_bef32(body + 18, s->wb[0]);
    put_bef32(body + 22, s->wb[1]);
    for (int i = 0; i < 9; i++)
        put_bef32(body + 26 + 4 * i, s->color_matrix[i]);
    put_bef32(body + 62, s->gain);
    put_be16(body + 66, s->wb_cct);
    size_t n = 70;
    if (s->has_qmat) {
        flags |= 1;
        memcpy(body + n, s->qmat, 64);
        n += 64;
    }
    if (s->has_curve) {
        flags |= 1 << 4;
        for (int i = 0; i < 8; i++)
            put_be16(body + n + 2 * i, s->curve[i]);
        n += 16;
    }
    put_be16(body + 68, flags);
    return n;
}

typedef struct {
    uint8_t *data;
    size_t len, cap;
} bytes;

static void bytes_push(bytes *b, const uint8_t *p, size_t n)
{
    if (b->len + n > b->cap) {
        size_t c = b->cap ? b->cap : 64;
        while (c < b->len + n)
            c *= 2;
        b->data = (uint8_t *)realloc(b->data, c);
        b->cap = c;
    }
    memcpy(b->data + b->len, p, n);
    b->len += n;
}

/* Packet from per-tile payloads (synth.rs:138-151). */
static size_t assemble(const prs_spec *s, bytes *tiles, size_t n, uint8_t *out, size_t cap)
{
    uint8_t body[160];
    size_t bl = header_body(s, body);
    size_t total = 10 + bl + 2 * n;
    for (size_t i = 0; i < n; i++)
        total += tiles[i].len;
    if (!out)
        return total;
    if (total > cap)
        return 0;
    put_be32(out, (uint32_t)total);
    memcpy(out + 4, "prrf", 4);
    put_be16(out + 8, (uint16_t)(bl + 2));
    memcpy(out + 10, body, bl);
    size_t pos = 10 + bl;
    for (size_t i = 0; i < n; i++, pos += 2)
        put_be16(out + pos, (uint16_t)tiles[i].len);
    for (size_t i = 0; i < n; i++) {
        if (tiles[i].len)
            memcpy(out + pos, tiles[i].data, tiles[i].len);
        pos += tiles[i].len;
    }
    return total;
}

/* synth::tile (synth.rs:175-184). */
static void make_tile(bytes *t, uint8_t qscale, const uint8_t *const streams[4],
                      const size_t lens[4])
{
    uint8_t hdr[8] = {8 << 3, qscale};
    for (int i = 0; i < 3; i++)
        put_be16(hdr + 2 + 2 * i, (uint16_t)lens[i]);
    bytes_push(t, hdr, 8);
    for (int i = 0; i < 4; i++)
        if (lens[i])
            bytes_push(t, streams[i], lens[i]);
}

/* random_tile (synth.rs:158-170). */
static void random_tile(bytes *t, size_t blocks, int valid, uint64_t *rng)
{
    uint32_t floor = valid ? 1 : 0;
    uint8_t qscale = (uint8_t)(floor + rng_below(rng, 256 - floor));
    size_t sizes[4];
    for (int i = 0; i < 4; i++)
        sizes[i] = floor + rng_below(rng, (uint32_t)(48 * blocks + 1 - floor));
    uint8_t *buf[4];
    for (int i = 0; i < 4; i++) {
        buf[i] = (uint8_t *)malloc(sizes[i] + 1);
        for (size_t k = 0; k < sizes[i]; k++) {
            uint32_t a = prs_rng_next(rng);
            uint32_t b = prs_rng_next(rng);
            buf[i][k] = (uint8_t)(a & b);
        }
    }
    make_tile(t, qscale, (const uint8_t *const *)buf, sizes);
    for (int i = 0; i < 4; i++)
        free(buf[i]);
}

/* random_frame (synth.rs:92-94). */
size_t prs_random_frame(const prs_spec *spec, uint64_t *rng, uint8_t *out, size_t cap)
{
    tile_geo *g;
    size_t n = grid(spec, &g);
    bytes *tiles = (bytes *)calloc(n + 1, sizeof(bytes));
    for (size_t i = 0; i < n; i++)
        random_tile(&tiles[i], (size_t)1 << g[i].log2, spec->valid_tiles, rng);
    size_t total = assemble(spec, tiles, n, out, cap);
    for (size_t i = 0; i < n; i++)
        free(tiles[i].data);
    free(tiles);
    free(g);
    return total;
}

size_t prs_frame_streams(const prs_spec *spec, uint8_t qscale, const uint8_t *const streams[4],
                         const size_t lens[4], uint8_t *out, size_t cap)
{
    tile_geo *g;
    size_t n = grid(spec, &g);
    bytes *tiles = (bytes *)calloc(n + 1, sizeof(bytes));
    for (size_t i = 0; i < n; i++)
        make_tile(&tiles[i], qscale, streams, lens);
    size_t total = assemble(spec, tiles, n, out, cap);
    for (size_t i = 0; i < n; i++)
        free(tiles[i].data);
    free(tiles);
    free(g);
    return total;
}

/* synth::specs (synth.rs:188-221). */
void prs_specs(uint64_t *rng, prs_spec *out)
{
    int k = 0;
    for (uint8_t alignment = 0; alignment <= 4; alignment++) {
        for (uint8_t i = 0; i < 4; i++) {
            uint16_t width = (uint16_t)(2 * (1 + rng_below(rng, 300)));
            uint16_t height = (uint16_t)(2 * (1 + rng_below(rng, 40)));
            prs_spec *s = &out[k++];
            prs_spec_default(s, width, height, i);
            s->alignment = alignment;
            if (i == 1) {
                s->has_qmat = 1;
                memset(s->qmat, 63, 64);
            } else if (i >= 2) {
                s->has_qmat = 1;
                for (int q = 0; q < 64; q++)
                    s->qmat[q] = (uint8_t)(1 + rng_below(rng, 63));
            }
            if (i % 2 == 1) {
                uint16_t c[8];
                for (int q = 0; q < 8; q++)
                    c[q] = (uint16_t)rng_below(rng, 65536 - 8);
                for (int a = 1; a < 8; a++) /* sort_unstable: any sort gives the same order */
                    for (int b = a; b > 0 && c[b - 1] > c[b]; b--) {
                        uint16_t t = c[b];
                        c[b] = c[b - 1];
                        c[b - 1] = t;
                    }
                s->has_curve = 1;
                for (int q = 0; q < 8; q++)
                    s->curve[q] = (uint16_t)(c[q] + q);
            }
        }
    }
}

/* ---------------------------------------------------------------- */
/* Scene encoder (new). */

static const uint16_t DEFAULT_CURVE[8] = {0, 512, 1024, 2048, 4096, 8192, 16384, 32768};
static const uint8_t SCAN[64] = {
    0,  8,  1,  9,  16, 24, 17, 25, 2,  10, 3,  11, 18, 26, 19, 27, 32, 40, 33, 34, 41, 48,
    56, 49, 42, 35, 43, 50, 57, 58, 51, 59, 4,  12, 5,  6,  13, 20, 28, 21, 14, 7,  15, 22,
    29, 36, 44, 37, 30, 23, 31, 38, 45, 52, 60, 53, 46, 39, 47, 54, 61, 62, 55, 63};
static const uint16_t DC_CB[13] = {0x010, 0x021, 0x032, 0x033, 0x033, 0x033, 0x044,
                                   0x044, 0x044, 0x044, 0x044, 0x044, 0x076};
static const uint16_t RN_CB[28] = {0x200, 0x100, 0x000, 0x000, 0x211, 0x211, 0x111,
                                   0x111, 0x011, 0x011, 0x021, 0x021, 0x222, 0x022,
                                   0x022, 0x022, 0x022, 0x022, 0x022, 0x022, 0x022,
                                   0x022, 0x022, 0x022, 0x022, 0x032, 0x032, 0x044};
static const uint16_t LN_CB[15] = {0x100, 0x111, 0x222, 0x222, 0x122, 0x122, 0x433, 0x433,
                                   0x233, 0x233, 0x233, 0x233, 0x233, 0x233, 0x033};

static uint16_t ac_cb(uint32_t level)
{
    static const struct { uint8_t n; uint16_t v; } runs[10] = {
        {1, 0x000}, {1, 0x211}, {2, 0x111}, {3, 0x222}, {3, 0x122},
        {8, 0x233}, {2, 0x133}, {22, 0x244}, {52, 0x355}, {1, 0x166}};
    uint32_t i = level < 94 ? level : 94;
    for (int r = 0; r < 10; r++) {
        if (i < runs[r].n)
            return runs[r].v;
        i -= runs[r].n;
    }
    return 0x166;
}

static uint16_t linearize(const uint16_t *curve, uint32_t code)
{
    uint32_t seg = code >> 13, frac = code & 0x1FFF;
    uint32_t c0 = curve[seg], c1 = seg < 7 ? curve[seg + 1] : 65536u;
    uint32_t v = (c0 * 8192 + ((c1 - c0) & 0xFFFF) * frac + 4096) >> 13;
    return (uint16_t)(v < 65535 ? v : 65535);
}

/* Bit writer: MSB first, zero padded to a byte. */
typedef struct {
    bytes b;
    uint64_t acc;
    int nacc;
} bitw;

static void bw_put(bitw *w, uint32_t value, uint32_t n)
{
    while (n > 0) {
        uint32_t take = n > 24 ? 24 : n;
        uint32_t bits = (value >> (n - take)) & ((1u << take) - 1);
        w->acc = (w->acc << take) | bits;
        w->nacc += (int)take;
        n -= take;
        while (w->nacc >= 8) {
            uint8_t byte = (uint8_t)(w->acc >> (w->nacc - 8));
            bytes_push(&w->b, &byte, 1);
            w->nacc -= 8;
        }
    }
}

static void bw_finish(bitw *w)
{
    if (w->nacc > 0) {
        uint8_t byte = (uint8_t)(w->acc << (8 - w->nacc));
        bytes_push(&w->b, &byte, 1);
        w->nacc = 0;
    }
}

/* Length of the hybrid codeword of v under cb (bits.rs:157-174). */
static uint32_t sym_len(uint16_t cb, uint32_t v)
{
    uint32_t s = cb >> 8, e = (cb >> 4) & 15, r = cb & 15;
    uint32_t q = v >> r;
    if (q <= s)
        return q + 1 + r;
    uint64_t m = (uint64_t)v - ((uint64_t)(s + 1) << r) + ((uint64_t)1 << e);
    uint32_t bits = 64 - (uint32_t)__builtin_clzll(m);
    return s + 1 + (bits - 1 - e) + bits;
}

static void sym_put(bitw *w, uint16_t cb, uint32_t v)
{
    uint32_t s = cb >> 8, e = (cb >> 4) & 15, r = cb & 15;
    uint32_t q = v >> r;
    if (q <= s) {
        bw_put(w, 0, q);
        bw_put(w, 1, 1);
        bw_put(w, v & ((1u << r) - 1), r);
        return;
    }
    uint32_t m = v - ((s + 1) << r) + (1u << e);
    uint32_t bits = 32 - (uint32_t)__builtin_clz(m);
    bw_put(w, 0, s + 1 + (bits - 1 - e));
    bw_put(w, m, bits);
}

/* A level and its sign bit; the magnitude is reduced until the codeword is
 * at most 31 bits so that the decoder reads the sign after it. Returns the
 * level symbol written. */
static uint32_t level_put(bitw *w, uint16_t cb, int32_t v)
{
    uint32_t level = (uint32_t)(v < 0 ? -v : v) - 1;
    while (level > 0 && sym_len(cb, level) > 31)
        level = level * 3 / 4;
    sym_put(w, cb, level);
    bw_put(w, v < 0 ? 1 : 0, 1);
    return level;
}

/* Entropy code one component of 1 << log2 blocks of raster-order levels
 * (the inverse of coefficients.rs:58-145). */
static void encode_component(bitw *w, const int16_t *levels, unsigned log2)
{
    uint32_t blocks = 1u << log2, codes = 64u << log2, mask = blocks - 1;
#define AT(n) levels[((n)&mask) * 64 + SCAN[(n) >> log2]]
    /* DC: sign-folded first value, then deltas with a running sign. */
    int32_t dc = levels[0];
    sym_put(w, 0x2BC, dc >= 0 ? (uint32_t)(2 * dc) : (uint32_t)(-2 * dc - 1));
    uint16_t dc_cb = 0x064;
    uint32_t sign = 0;
    for (uint32_t b = 1; b < blocks; b++) {
        int32_t delta = levels[b * 64] - dc;
        uint32_t v;
        for (;;) {
            uint32_t mag = (uint32_t)(delta < 0 ? -delta : delta);
            uint32_t parity = mag ? (sign ^ (delta < 0 ? 1u : 0u)) : 0;
            v = parity ? 2 * mag - 1 : 2 * mag;
            if (sym_len(dc_cb, v) <= 32)
                break;
            delta = delta * 3 / 4;
        }
        sym_put(w, dc_cb, v);
        sign ^= v & 1;
        uint32_t mag = ((v & 1) + v) >> 1;
        sign = (mag && sign) ? 1 : 0;
        dc += sign ? -(int32_t)mag : (int32_t)mag;
        dc_cb = DC_CB[mag < 12 ? mag : 12];
    }
    /* AC: runs of levels and runs of zeros. */
    uint16_t ac = 0x031, rn = 0x000, ln = 0x042;
    uint32_t n = blocks;
    uint32_t last = 0; /* one past the last non-zero position */
    for (uint32_t k = blocks; k < codes; k++)
        if (AT(k))
            last = k + 1;
    while (n < codes && n < last) {
        uint32_t run = 0;
        while (n + run < codes && AT(n + run))
            run++;
        sym_put(w, ln, run);
        for (uint32_t k = 0; k < run; k++) {
            uint32_t lv = level_put(w, ac, AT(n));
            ac = ac_cb(lv);
            n++;
        }
        if (n >= codes || n >= last)
            break;
        uint32_t z = 0;
        while (!AT(n + z))
            z++;
        sym_put(w, rn, z - 1);
        rn = RN_CB[z - 1 < 27 ? z - 1 : 27];
        n += z;
        uint32_t lv = level_put(w, ac, AT(n));
        ac = ac_cb(lv);
        ln = LN_CB[lv < 14 ? lv : 14];
        n++;
    }
#undef AT
}

/* Orthonormal 8-point DCT-II basis. */
static double DCT[8][8];
static pthread_once_t dct_once = PTHREAD_ONCE_INIT;
static void dct_init(void)
{
    for (int k = 0; k < 8; k++)
        for (int n = 0; n < 8; n++)
            DCT[k][n] = (k ? 0.5 : sqrt(0.125)) * cos(M_PI * (2 * n + 1) * k / 16.0);
}

/* Coefficients the decoder's transform maps back to t: X / 8 of the
 * orthonormal 2-D DCT (a DC-only block of value d decodes to d everywhere). */
static void fdct(const double *t, double *coef)
{
    double tmp[64];
    for (int r = 0; r < 8; r++)
        for (int k = 0; k < 8; k++) {
            double s = 0;
            for (int n = 0; n < 8; n++)
                s += DCT[k][n] * t[8 * r + n];
            tmp[8 * r + k] = s;
        }
    for (int c = 0; c < 8; c++)
        for (int k = 0; k < 8; k++) {
            double s = 0;
            for (int n = 0; n < 8; n++)
                s += DCT[k][n] * tmp[8 * n + c];
            coef[8 * k + c] = s / 8.0;
        }
}

typedef struct {
    double nx, ny, off, len;
    float amp[3];
} gradient;

typedef struct {
    float cx, cy, r2;
    int x0, x1, y0, y1;
    float rgb[3];
} disc;

typedef struct {
    const prs_scene *cfg;
    size_t cw, ch, w, h;
    const uint16_t *curve;
    gradient *grads;
    disc *discs;
    float *mosaic;      /* w*h scene values in [0, 1] */
    uint16_t *linear;   /* w*h noisy linear values */
    uint16_t *codes;    /* cw*ch code values */
    const uint16_t *inverse;
    tile_geo *tiles;
    size_t n_tiles, per_row;
    bytes *payloads;
    uint8_t *qscales;
    const uint8_t *qmat;
    size_t next;
} scene_ctx;

static int color_at(uint8_t pattern, size_t y, size_t x)
{
    static const int red[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
    int rr = red[pattern][0], rc = red[pattern][1];
    int yy = (int)(y & 1), xx = (int)(x & 1);
    if (yy == rr && xx == rc)
        return 0;
    if (yy != rr && xx != rc)
        return 2;
    return 1;
}

static double gauss(uint64_t *s)
{
    double u1 = (prs_rng_next(s) + 1.0) / 4294967297.0;
    double u2 = prs_rng_next(s) / 4294967296.0;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static uint32_t poisson(uint64_t *s, double lambda)
{
    if (lambda <= 0)
        return 0;
    if (lambda < 30.0) {
        double l = exp(-lambda), p = 1.0;
        uint32_t k = 0;
        do {
            k++;
            p *= (prs_rng_next(s) + 0.5) / 4294967296.0;
        } while (p > l);
        return k - 1;
    }
    double v = floor(lambda + sqrt(lambda) * gauss(s) + 0.5);
    return v < 0 ? 0 : (uint32_t)v;
}

/* Scene value in [0, 1] of every visible sensel's colour, one row. */
static void render_row(scene_ctx *c, size_t y)
{
    const prs_scene *cfg = c->cfg;
    float *row = c->mosaic + y * c->w;
    for (size_t x = 0; x < c->w; x++) {
        int ch = color_at(cfg->spec.pattern, y, x);
        double v = 0;
        for (int k = 0; k < cfg->n_gradients; k++) {
            const gradient *g = &c->grads[k];
            double t = 0.5 + ((double)x * g->nx + (double)y * g->ny - g->off) / g->len;
            t = t < 0 ? 0 : t > 1 ? 1 : t;
            v += g->amp[ch] * t;
        }
        row[x] = (float)(cfg->n_gradients ? v / cfg->n_gradients : 0.0);
    }
    for (uint32_t k = 0; k < cfg->n_discs; k++) {
        const disc *d = &c->discs[k];
        if ((int)y < d->y0 || (int)y > d->y1)
            continue;
        float dy = (float)y - d->cy;
        for (int x = d->x0 < 0 ? 0 : d->x0; x <= d->x1 && x < (int)c->w; x++) {
            float dx = (float)x - d->cx;
            if (dx * dx + dy * dy <= d->r2)
                row[x] = d->rgb[color_at(cfg->spec.pattern, y, (size_t)x)];
        }
    }
}

/* Noise and the inverse curve for one visible row (its own generator). */
static void noise_row(scene_ctx *c, size_t y)
{
    const prs_scene *cfg = c->cfg;
    uint64_t rng = prs_rng_new(cfg->seed ^ (0xD1B54A32D192ED03ull * (y + 1)));
    double white = cfg->spec.white_level, gain = cfg->dn_per_electron;
    for (size_t x = 0; x < c->w; x++) {
        double clean = 256.0 + (double)c->mosaic[y * c->w + x] * (white - 256.0);
        double dn = 256.0;
        if (gain > 0)
            dn += poisson(&rng, (clean - 256.0) / gain) * gain;
        else
            dn = clean;
        if (cfg->read_sigma > 0)
            dn += cfg->read_sigma * gauss(&rng);
        double r = floor(dn + 0.5);
        c->linear[y * c->w + x] = (uint16_t)(r < 0 ? 0 : r > 65535 ? 65535 : r);
    }
}

/* Encode one tile into its payload; qscale doubles until it fits 64 KiB. */
static void encode_tile(scene_ctx *c, size_t t)
{
    const prs_scene *cfg = c->cfg;
    tile_geo g = c->tiles[t];
    unsigned blocks = 1u << g.log2;
    uint8_t qscale = c->qscales[t];
    static const int red[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
    int rr = red[cfg->spec.pattern][0], rc = red[cfg->spec.pattern][1];
    int phase[4][2] = {{1 - rr, rc}, {rr, 1 - rc}, {1 - rr, 1 - rc}, {rr, rc}};
    int16_t *levels = (int16_t *)malloc(sizeof(int16_t) * 64 * blocks);
    for (;;) {
        bitw streams[4];
        memset(streams, 0, sizeof streams);
        for (int s = 0; s < 4; s++) {
            for (unsigned b = 0; b < blocks; b++) {
                double tv[64], coef[64];
                for (int i = 0; i < 8; i++)
                    for (int j = 0; j < 8; j++) {
                        size_t y = g.y + 2 * i + (size_t)phase[s][0];
                        size_t x = g.x + 16 * b + 2 * j + (size_t)phase[s][1];
                        tv[8 * i + j] = (double)c->codes[y * c->cw + x] - 32768.0;
                    }
                fdct(tv, coef);
                for (int k = 0; k < 64; k++) {
                    double q = (double)c->qmat[k] * qscale;
                    double l = floor(coef[k] / q + 0.5);
                    l = l < -32767 ? -32767 : l > 32767 ? 32767 : l;
                    levels[64 * b + k] = (int16_t)l;
                }
            }
            encode_component(&streams[s], levels, g.log2);
            bw_finish(&streams[s]);
        }
        size_t lens[4], total = 8;
        const uint8_t *ptrs[4];
        for (int s = 0; s < 4; s++) {
            lens[s] = streams[s].b.len;
            ptrs[s] = streams[s].b.data;
            total += lens[s];
        }
        if (total <= 65535 || qscale == 255) {
            c->payloads[t].len = 0;
            make_tile(&c->payloads[t], qscale, ptrs, lens);
            c->qscales[t] = qscale;
            for (int s = 0; s < 4; s++)
                free(streams[s].b.data);
            break;
        }
        for (int s = 0; s < 4; s++)
            free(streams[s].b.data);
        qscale = qscale > 127 ? 255 : (uint8_t)(qscale * 2);
    }
    free(levels);
}

typedef void (*row_fn)(scene_ctx *, size_t);
typedef struct {
    scene_ctx *c;
    row_fn fn;
    size_t n;
} par_job;

static void *par_worker(void *arg)
{
    par_job *j = (par_job *)arg;
    for (;;) {
        size_t i = __atomic_fetch_add(&j->c->next, 1, __ATOMIC_RELAXED);
        if (i >= j->n)
            break;
        j->fn(j->c, i);
    }
    return NULL;
}

static void par_for(scene_ctx *c, size_t n, int threads, row_fn fn)
{
    par_job j = {c, fn, n};
    c->next = 0;
    if (threads < 1)
        threads = 1;
    if (threads > 64)
        threads = 64;
    pthread_t tids[64];
    int started = 0;
    for (int t = 1; t < threads; t++)
        if (pthread_create(&tids[started], NULL, par_worker, &j) == 0)
            started++;
    par_worker(&j);
    for (int t = 0; t < started; t++)
        pthread_join(tids[t], NULL);
}

static void encode_row(scene_ctx *c, size_t r)
{
    for (size_t i = 0; i < c->per_row; i++)
        encode_tile(c, r * c->per_row + i);
}

/* Scene layout from the seed: gradients and discs. */
static void scene_layout(scene_ctx *c)
{
    const prs_scene *cfg = c->cfg;
    uint64_t rng = prs_rng_new(cfg->seed);
    double w = (double)c->w, h = (double)c->h;
    c->grads = (gradient *)calloc(cfg->n_gradients + 1, sizeof(gradient));
    for (int k = 0; k < cfg->n_gradients; k++) {
        double th = 2.0 * M_PI * (prs_rng_next(&rng) / 4294967296.0);
        gradient *g = &c->grads[k];
        g->nx = cos(th);
        g->ny = sin(th);
        double cx = w * (prs_rng_next(&rng) / 4294967296.0);
        double cy = h * (prs_rng_next(&rng) / 4294967296.0);
        g->off = cx * g->nx + cy * g->ny;
        g->len = (0.3 + 1.2 * (prs_rng_next(&rng) / 4294967296.0)) * (w > h ? w : h);
        for (int ch = 0; ch < 3; ch++)
            g->amp[ch] = (float)(0.1 + 0.9 * (prs_rng_next(&rng) / 4294967296.0));
    }
    c->discs = (disc *)calloc(cfg->n_discs + 1, sizeof(disc));
    double m = w < h ? w : h;
    for (uint32_t k = 0; k < cfg->n_discs; k++) {
        disc *d = &c->discs[k];
        d->cx = (float)(w * (prs_rng_next(&rng) / 4294967296.0));
        d->cy = (float)(h * (prs_rng_next(&rng) / 4294967296.0));
        double r = m * (0.005 + 0.075 * (prs_rng_next(&rng) / 4294967296.0));
        if (r < 1.5)
            r = 1.5;
        d->r2 = (float)(r * r);
        d->x0 = (int)floor(d->cx - r);
        d->x1 = (int)ceil(d->cx + r);
        d->y0 = (int)floor(d->cy - r);
        d->y1 = (int)ceil(d->cy + r);
        for (int ch = 0; ch < 3; ch++)
            d->rgb[ch] = (float)(prs_rng_next(&rng) / 4294967296.0);
    }
}

static void render_and_noise(scene_ctx *c, size_t y)
{
    render_row(c, y);
    noise_row(c, y);
}

static int scene_prepare(scene_ctx *c, const prs_scene *cfg)
{
    memset(c, 0, sizeof *c);
    c->cfg = cfg;
    c->w = cfg->spec.width;
    c->h = cfg->spec.height;
    if (!c->w || !c->h)
        return 0;
    c->cw = (c->w + 15) & ~(size_t)15;
    c->ch = (c->h + 15) & ~(size_t)15;
    c->curve = cfg->spec.has_curve ? cfg->spec.curve : DEFAULT_CURVE;
    scene_layout(c);
    c->mosaic = (float *)malloc(sizeof(float) * c->w * c->h);
    c->linear = (uint16_t *)malloc(sizeof(uint16_t) * c->w * c->h);
    par_for(c, c->h, cfg->threads, render_and_noise);
    return 1;
}

static void scene_release(scene_ctx *c)
{
    free(c->grads);
    free(c->discs);
    free(c->mosaic);
    free(c->linear);
    free(c->codes);
    free((void *)c->inverse);
    free(c->tiles);
    free(c->qscales);
    if (c->payloads)
        for (size_t i = 0; i < c->n_tiles; i++)
            free(c->payloads[i].data);
    free(c->payloads);
}

void prs_scene_mosaic(const prs_scene *cfg, uint16_t *linear)
{
    scene_ctx c;
    if (!scene_prepare(&c, cfg))
        return;
    memcpy(linear, c.linear, sizeof(uint16_t) * c.w * c.h);
    scene_release(&c);
}

size_t prs_encode_scene(const prs_scene *cfg, uint8_t *out, size_t cap)
{
    pthread_once(&dct_once, dct_init);
    scene_ctx c;
    if (!scene_prepare(&c, cfg))
        return 0;
    /* Inverse curve: the code whose linear value is nearest. */
    uint16_t *inverse = (uint16_t *)malloc(sizeof(uint16_t) * 65536);
    {
        uint32_t u = 0;
        for (uint32_t l = 0; l < 65536; l++) {
            while (u < 65535 && linearize(c.curve, u) < l)
                u++;
            uint32_t best = u;
            if (u > 0 && (int)l - (int)linearize(c.curve, u - 1) <
                             (int)linearize(c.curve, u) - (int)l)
                best = u - 1;
            inverse[l] = (uint16_t)best;
        }
    }
    c.inverse = inverse;
    c.codes = (uint16_t *)malloc(sizeof(uint16_t) * c.cw * c.ch);
    /* Padding rows and columns repeat the last visible sensel of the same
     * Bayer phase (width and height are even). */
    for (size_t y = 0; y < c.ch; y++) {
        size_t sy = y < c.h ? y : c.h - 2 + (y & 1);
        for (size_t x = 0; x < c.cw; x++) {
            size_t sx = x < c.w ? x : c.w - 2 + (x & 1);
            c.codes[y * c.cw + x] = inverse[c.linear[sy * c.w + sx]];
        }
    }
    c.n_tiles = grid(&cfg->spec, &c.tiles);
    c.per_row = tiles_per_row(cfg->spec.width, cfg->spec.alignment);
    c.payloads = (bytes *)calloc(c.n_tiles + 1, sizeof(bytes));
    c.qscales = (uint8_t *)malloc(c.n_tiles + 1);
    static const uint8_t ones[64] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                     1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                     1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                     1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
    c.qmat = cfg->spec.has_qmat ? cfg->spec.qmat : ones;
    uint64_t qrng = prs_rng_new(cfg->seed ^ 0x5157u);
    uint32_t qlo = cfg->qscale_min ? cfg->qscale_min : 1;
    uint32_t qhi = cfg->qscale_max >= qlo ? cfg->qscale_max : qlo;
    for (size_t t = 0; t < c.n_tiles; t++)
        c.qscales[t] = (uint8_t)(qlo + rng_below(&qrng, qhi - qlo + 1));
    par_for(&c, (c.ch / 16), cfg->threads, encode_row);
    /* Corrupt a seeded choice of tiles: drop the fourth stream and shrink the
     * size entry to match, so exactly these tiles fail the tile header check. */
    size_t n_bad = (size_t)floor(cfg->corrupt_fraction * (double)c.n_tiles + 0.5);
    if (n_bad > c.n_tiles)
        n_bad = c.n_tiles;
    if (n_bad) {
        uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * c.n_tiles);
        for (size_t t = 0; t < c.n_tiles; t++)
            order[t] = (uint32_t)t;
        uint64_t crng = prs_rng_new(cfg->seed ^ 0xC0FFEEu);
        for (size_t k = 0; k < n_bad; k++) {
            size_t j = k + rng_below(&crng, (uint32_t)(c.n_tiles - k));
            uint32_t tmp = order[k];
            order[k] = order[j];
            order[j] = tmp;
            bytes *p = &c.payloads[order[k]];
            size_t used = 8 + (((size_t)p->data[2] << 8) | p->data[3]) +
                          (((size_t)p->data[4] << 8) | p->data[5]) +
                          (((size_t)p->data[6] << 8) | p->data[7]);
            p->len = used;
        }
        free(order);
    }
    size_t total = assemble(&cfg->spec, c.payloads, c.n_tiles, out, cap);
    scene_release(&c);
    return total;
}

size_t prs_encode_component(const int16_t *levels, unsigned log2_blocks, uint8_t *out, size_t cap)
{
    bitw w;
    memset(&w, 0, sizeof w);
    encode_component(&w, levels, log2_blocks);
    bw_finish(&w);
    size_t n = w.b.len;
    if (out && n <= cap)
        memcpy(out, w.b.data, n);
    free(w.b.data);
    return n;
}
