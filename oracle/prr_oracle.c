/* Oracle: CPU restatement of proresraw_rs's decode and develop path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py as the checker; never by the product.
 *
 * Each function names the reference lines it restates (src/<file>.rs:line of
 * the proresraw_rs crate). Floating point follows the crate exactly: every
 * operation rounded on its own (built with -ffp-contract=off), fmaf exactly
 * where the crate calls mul_add. Work is spread over threads the way the
 * crate's rayon code spreads it: tile rows for decoding, rows for
 * linearization, 32-row bands for develop.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Parallel for: dynamic distribution of [0, n) over `threads` threads. */

typedef void (*orc_task_fn)(void *ctx, size_t index, void *scratch);

typedef struct {
    orc_task_fn fn;
    void *ctx;
    size_t n;
    size_t next;
    size_t scratch_bytes;
} orc_pool;

static void *orc_worker(void *arg)
{
    orc_pool *pool = (orc_pool *)arg;
    void *scratch = pool->scratch_bytes ? malloc(pool->scratch_bytes) : NULL;
    for (;;) {
        size_t i = __atomic_fetch_add(&pool->next, 1, __ATOMIC_RELAXED);
        if (i >= pool->n)
            break;
        pool->fn(pool->ctx, i, scratch);
    }
    free(scratch);
    return NULL;
}

static void orc_parallel_for(size_t n, int threads, size_t scratch_bytes, orc_task_fn fn,
                             void *ctx)
{
    orc_pool pool = {fn, ctx, n, 0, scratch_bytes};
    if (threads <= 1 || n <= 1) {
        orc_worker(&pool);
        return;
    }
    if ((size_t)threads > n)
        threads = (int)n;
    pthread_t *tids = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    int started = 0;
    for (int t = 1; t < threads; t++)
        if (pthread_create(&tids[started], NULL, orc_worker, &pool) == 0)
            started++;
    orc_worker(&pool);
    for (int t = 0; t < started; t++)
        pthread_join(tids[t], NULL);
    free(tids);
}

/* ------------------------------------------------------------------ */
/* Codebooks and scans (codebook.rs:26-110). */

#define CB_SWITCH(cb) ((uint32_t)(cb) >> 8)
#define CB_EXP(cb) (((uint32_t)(cb) >> 4) & 0xF)
#define CB_RICE(cb) ((uint32_t)(cb)&0xF)

static const uint16_t FIRST_DC = 0x2BC, DC_RESET = 0x064, AC_INITIAL = 0x031,
                      RN_INITIAL = 0x000, LN_INITIAL = 0x042;
static const uint16_t DC_CB[13] = {0x010, 0x021, 0x032, 0x033, 0x033, 0x033, 0x044,
                                   0x044, 0x044, 0x044, 0x044, 0x044, 0x076};
static const uint16_t RN_CB[28] = {0x200, 0x100, 0x000, 0x000, 0x211, 0x211, 0x111,
                                   0x111, 0x011, 0x011, 0x021, 0x021, 0x222, 0x022,
                                   0x022, 0x022, 0x022, 0x022, 0x022, 0x022, 0x022,
                                   0x022, 0x022, 0x022, 0x022, 0x032, 0x032, 0x044};
static const uint16_t LN_CB[15] = {0x100, 0x111, 0x222, 0x222, 0x122, 0x122, 0x433, 0x433,
                                   0x233, 0x233, 0x233, 0x233, 0x233, 0x233, 0x033};

/* AC_CB as runs of (count, value), codebook.rs:57-85. */
static uint16_t ac_cb_entry(uint32_t i)
{
    static const struct { uint8_t count; uint16_t value; } runs[10] = {
        {1, 0x000}, {1, 0x211}, {2, 0x111}, {3, 0x222}, {3, 0x122},
        {8, 0x233}, {2, 0x133}, {22, 0x244}, {52, 0x355}, {1, 0x166}};
    for (int r = 0; r < 10; r++) {
        if (i < runs[r].count)
            return runs[r].value;
        i -= runs[r].count;
    }
    return 0x166;
}

static const uint8_t SCAN[64] = {
    0,  8,  1,  9,  16, 24, 17, 25, 2,  10, 3,  11, 18, 26, 19, 27, 32, 40, 33, 34, 41, 48,
    56, 49, 42, 35, 43, 50, 57, 58, 51, 59, 4,  12, 5,  6,  13, 20, 28, 21, 14, 7,  15, 22,
    29, 36, 44, 37, 30, 23, 31, 38, 45, 52, 60, 53, 46, 39, 47, 54, 61, 62, 55, 63};
static const uint8_t SCAN_4X4[16] = {0, 4, 1, 5, 8, 12, 9, 13, 2, 6, 3, 7, 10, 14, 11, 15};
static const uint8_t SCAN_2X2[4] = {0, 2, 1, 3};

/* ------------------------------------------------------------------ */
/* Bit reader (bits.rs:17-111). */

typedef struct {
    const uint8_t *data;
    size_t len;
    size_t pos;
    uint64_t window;
    int32_t extra;
} orc_reader;

/* Up to four bytes as the top bytes of a big-endian word (bits.rs:31-37). */
static uint32_t partial_word(const uint8_t *b, size_t n)
{
    uint32_t w = 0;
    for (size_t i = 0; i < n && i < 4; i++)
        w |= (uint32_t)b[i] << (24 - 8 * i);
    return w;
}

/* BitReader::new (bits.rs:44-76). */
static void reader_init(orc_reader *r, const uint8_t *data, size_t len, unsigned align)
{
    size_t remaining = len, pos, address = align;
    uint32_t first;
    if (remaining < 4) {
        first = partial_word(data, remaining);
        pos = remaining;
        address += remaining;
        remaining = 0;
    } else {
        first = partial_word(data, 4);
        pos = 4;
        address += 4;
        remaining -= 4;
    }
    uint32_t low = 0;
    int32_t extra = 0;
    size_t to_boundary = (4 - (address & 3)) & 3;
    if (to_boundary != 0 && to_boundary <= remaining) {
        low = partial_word(data + pos, to_boundary);
        pos += to_boundary;
        extra = 8 * (int32_t)to_boundary;
    }
    r->data = data;
    r->len = len;
    r->pos = pos;
    r->window = ((uint64_t)first << 32) | low;
    r->extra = extra;
}

/* BitReader::next_word (bits.rs:79-89): a whole word while four bytes
 * remain, else the zero-filled tail, then zeros. */
static uint32_t reader_next_word(orc_reader *r)
{
    size_t n = r->len - r->pos;
    if (n > 4)
        n = 4;
    uint32_t w = partial_word(r->data + r->pos, n);
    r->pos += n;
    return w;
}

static uint32_t reader_top(const orc_reader *r) { return (uint32_t)(r->window >> 32); }

/* BitReader::consume (bits.rs:97-105). */
static void reader_consume(orc_reader *r, uint32_t len)
{
    r->window <<= (len & 63);
    r->extra = (int32_t)((uint32_t)r->extra - len);
    if (r->extra < 0) {
        uint32_t shift = (uint32_t)(-(int32_t)(int8_t)(uint8_t)(uint32_t)r->extra) & 0x3f;
        r->window |= (uint64_t)reader_next_word(r) << shift;
        r->extra += 32;
    }
}

/* codeword (bits.rs:116-134). */
void orc_codeword(uint32_t p, uint32_t cb, uint32_t *value, uint32_t *len)
{
    uint32_t lz = p ? (uint32_t)__builtin_clz(p) : 32;
    uint32_t s = CB_SWITCH(cb), r = CB_RICE(cb);
    if ((int32_t)p < 0) {
        *value = (p & 0x7fffffffu) >> ((31u - r) & 31);
        *len = r + 1;
        return;
    }
    if (lz <= s) {
        uint32_t tail = ((p << ((lz + 1) & 31)) >> 1) >> ((31u - r) & 31);
        *value = (lz << r) + tail;
        *len = r + 1 + lz;
        return;
    }
    uint32_t e = CB_EXP(cb);
    uint32_t l = 2 * lz - s + e;
    *value = ((s + 1) << r) + (p >> ((32u - l) & 31)) - (1u << e);
    *len = l;
}

size_t orc_read_symbols(const uint8_t *data, size_t len, unsigned align,
                        const uint16_t *cbs, size_t n, uint32_t *out)
{
    orc_reader r;
    reader_init(&r, data, len, align);
    for (size_t i = 0; i < n; i++) {
        if (reader_top(&r) == 0)
            return i;
        uint32_t v, l;
        orc_codeword(reader_top(&r), cbs[i], &v, &l);
        reader_consume(&r, l);
        out[i] = v;
    }
    return n;
}

/* ------------------------------------------------------------------ */
/* Coefficient decoding (coefficients.rs:37-146). */

static uint16_t next_cb(const uint16_t *table, size_t n, uint32_t symbol)
{
    return table[symbol < n - 1 ? symbol : n - 1];
}

static uint16_t next_ac(uint32_t symbol) { return ac_cb_entry(symbol < 94 ? symbol : 94); }

/* The `level` closure (coefficients.rs:90-104): the symbol, and the signed
 * 16-bit value stored for it. */
static uint32_t read_level(orc_reader *r, uint16_t cb, int16_t *value)
{
    uint32_t top = reader_top(r), symbol, len;
    orc_codeword(top, cb, &symbol, &len);
    int negative = ((top << (len & 31)) >> 31) != 0;
    reader_consume(r, len + 1);
    uint16_t magnitude = (uint16_t)((uint16_t)symbol + 1);
    *value = (int16_t)(negative ? (uint16_t)(0u - magnitude) : magnitude);
    return symbol;
}

void orc_decode_component_reduced(const uint8_t *data, size_t len, unsigned align,
                                  unsigned log2_blocks, unsigned log2_coefs,
                                  int16_t *coefs)
{
    const uint8_t *scan = log2_coefs == 6 ? SCAN : log2_coefs == 4 ? SCAN_4X4 : SCAN_2X2;
    uint32_t per_block = 1u << log2_coefs;
    uint32_t blocks = 1u << log2_blocks;
    uint32_t codes = per_block * blocks;
    uint32_t block_mask = blocks - 1;
    memset(coefs, 0, sizeof(int16_t) * codes);
    orc_reader r;
    reader_init(&r, data, len, align);

    /* DC (coefficients.rs:58-82). */
    if (reader_top(&r) == 0)
        return;
    uint32_t first, l;
    orc_codeword(reader_top(&r), FIRST_DC, &first, &l);
    reader_consume(&r, l);
    uint32_t odd = first & 1;
    int32_t dc = (int32_t)((((odd + first) >> 1) ^ (0u - odd)) + odd);
    coefs[0] = (int16_t)dc;
    uint16_t dc_cb = DC_RESET;
    uint32_t sign = 0;
    for (uint32_t block = 1; block < blocks; block++) {
        if (reader_top(&r) == 0)
            return;
        uint32_t symbol;
        orc_codeword(reader_top(&r), dc_cb, &symbol, &l);
        sign ^= symbol & 1;
        uint32_t magnitude = ((symbol & 1) + symbol) >> 1;
        int32_t delta = (int32_t)(((0u - sign) ^ magnitude) + sign);
        dc_cb = next_cb(DC_CB, 13, magnitude);
        sign = delta < 0;
        reader_consume(&r, l);
        dc = (int32_t)((uint32_t)dc + (uint32_t)delta);
        coefs[per_block * block] = (int16_t)dc;
    }

    /* AC (coefficients.rs:84-145). */
#define PLACE(n) ((uint32_t)scan[(n) >> log2_blocks] + (((n)&block_mask) << log2_coefs))
    uint16_t ac_cb = AC_INITIAL, rn_cb = RN_INITIAL, ln_cb = LN_INITIAL;
    uint32_t n = blocks;
    if (n >= codes)
        return;
    while (reader_top(&r) != 0) {
        uint32_t run;
        orc_codeword(reader_top(&r), ln_cb, &run, &l);
        reader_consume(&r, l);
        for (uint32_t k = 0; k < run; k++) {
            if (reader_top(&r) == 0)
                return;
            int16_t value;
            uint32_t symbol = read_level(&r, ac_cb, &value);
            ac_cb = next_ac(symbol);
            coefs[PLACE(n)] = value;
            n += 1;
            if (n == codes)
                return;
        }
        if (reader_top(&r) == 0)
            return;
        uint32_t zeros;
        orc_codeword(reader_top(&r), rn_cb, &zeros, &l);
        rn_cb = next_cb(RN_CB, 28, zeros);
        reader_consume(&r, l);
        n = n + 1 + zeros;
        if (n >= codes || reader_top(&r) == 0)
            return;
        int16_t value;
        uint32_t symbol = read_level(&r, ac_cb, &value);
        ac_cb = next_ac(symbol);
        ln_cb = next_cb(LN_CB, 15, symbol);
        coefs[PLACE(n)] = value;
        n += 1;
        if (n >= codes)
            return;
    }
#undef PLACE
}

/* Instrumented copy of orc_decode_component_reduced (coefficients.rs:37-146):
 * the same loop, storing nothing, counting the symbols of its AC part (one
 * per codeword read after the DC values, the unit the device decoder's
 * symbol budget counts) and the bits its consumes ask for. It gives up after
 * `cap` AC symbols (returns 1), so a stream the loop would not leave cannot
 * hang a test. */
int orc_count_symbols(const uint8_t *data, size_t len, unsigned align, unsigned log2_blocks,
                      unsigned log2_coefs, uint64_t cap, uint64_t *ac_symbols,
                      uint64_t *consumed_bits)
{
    uint32_t per_block = 1u << log2_coefs;
    uint32_t blocks = 1u << log2_blocks;
    uint32_t codes = per_block * blocks;
    uint64_t count = 0, bits = 0;
    orc_reader r;
    reader_init(&r, data, len, align);
    *ac_symbols = 0;
    *consumed_bits = 0;
#define CONSUME(l) (bits += (l), reader_consume(&r, (l)))
    if (reader_top(&r) == 0)
        return 0;
    uint32_t first, l;
    orc_codeword(reader_top(&r), FIRST_DC, &first, &l);
    CONSUME(l);
    uint16_t dc_cb = DC_RESET;
    uint32_t sign = 0;
    for (uint32_t block = 1; block < blocks; block++) {
        if (reader_top(&r) == 0)
            goto out;
        uint32_t symbol;
        orc_codeword(reader_top(&r), dc_cb, &symbol, &l);
        sign ^= symbol & 1;
        uint32_t magnitude = ((symbol & 1) + symbol) >> 1;
        int32_t delta = (int32_t)(((0u - sign) ^ magnitude) + sign);
        dc_cb = next_cb(DC_CB, 13, magnitude);
        sign = delta < 0;
        CONSUME(l);
    }
    uint16_t ac_cb = AC_INITIAL, rn_cb = RN_INITIAL, ln_cb = LN_INITIAL;
    uint32_t n = blocks;
    if (n >= codes)
        goto out;
    while (reader_top(&r) != 0) {
        uint32_t run;
        if (++count > cap)
            goto capped;
        orc_codeword(reader_top(&r), ln_cb, &run, &l);
        CONSUME(l);
        for (uint32_t k = 0; k < run; k++) {
            if (reader_top(&r) == 0)
                goto out;
            if (++count > cap)
                goto capped;
            uint32_t symbol;
            orc_codeword(reader_top(&r), ac_cb, &symbol, &l);
            CONSUME(l + 1);
            ac_cb = next_ac(symbol);
            n += 1;
            if (n == codes)
                goto out;
        }
        if (reader_top(&r) == 0)
            goto out;
        uint32_t zeros;
        if (++count > cap)
            goto capped;
        orc_codeword(reader_top(&r), rn_cb, &zeros, &l);
        rn_cb = next_cb(RN_CB, 28, zeros);
        CONSUME(l);
        n = n + 1 + zeros;
        if (n >= codes || reader_top(&r) == 0)
            goto out;
        uint32_t symbol;
        if (++count > cap)
            goto capped;
        orc_codeword(reader_top(&r), ac_cb, &symbol, &l);
        CONSUME(l + 1);
        ac_cb = next_ac(symbol);
        ln_cb = next_cb(LN_CB, 15, symbol);
        n += 1;
        if (n >= codes)
            goto out;
    }
out:
    *ac_symbols = count;
    *consumed_bits = bits;
    return 0;
capped:
    *ac_symbols = cap;
    *consumed_bits = bits;
    return 1;
#undef CONSUME
}

/* ------------------------------------------------------------------ */
/* Inverse transforms (idct.rs). */

/* dequantize (idct.rs:41-43). */
int16_t orc_dequantize(int16_t c, int16_t q)
{
    int32_t v = (int32_t)c * (int32_t)q;
    return (int16_t)(v < -32768 ? -32768 : v > 32767 ? 32767 : v);
}

static const int32_t ROW_TABLES[4][4][8] = {
    {{4096, 5352, 4096, 2217, 4096, -2217, 4096, -5352},
     {4096, 2217, -4096, -5352, -4096, 5352, 4096, -2217},
     {5681, 4816, 4816, -1130, 3218, -5681, 1130, -3218},
     {3218, 1130, -5681, -3218, 1130, 4816, 4816, -5681}},
    {{5681, 7423, 5681, 3075, 5681, -3075, 5681, -7423},
     {5681, 3075, -5681, -7423, -5681, 7423, 5681, -3075},
     {7880, 6681, 6681, -1567, 4464, -7880, 1567, -4464},
     {4464, 1567, -7880, -4464, 1567, 6681, 6681, -7880}},
    {{5352, 6992, 5352, 2896, 5352, -2896, 5352, -6992},
     {5352, 2896, -5352, -6992, -5352, 6992, 5352, -2896},
     {7423, 6293, 6293, -1477, 4205, -7423, 1477, -4205},
     {4205, 1477, -7423, -4205, 1477, 6293, 6293, -7423}},
    {{4816, 6293, 4816, 2607, 4816, -2607, 4816, -6293},
     {4816, 2607, -4816, -6293, -4816, 6293, 4816, -2607},
     {6681, 5663, 5663, -1329, 3784, -6681, 1329, -3784},
     {3784, 1329, -6681, -3784, 1329, 5663, 5663, -6681}}};
static const int ROW_SET[8] = {0, 1, 2, 3, 0, 3, 2, 1};

static float f32_bits(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

#define SCALE f32_bits(0x39800000u)
#define BIAS f32_bits(0x47000080u)
#define TAN1 f32_bits(0x3e4bafafu)
#define TAN2 f32_bits(0x3ed413cdu)
#define TAN3 f32_bits(0x3f2b0dc1u)
#define COS4 f32_bits(0x3f3504f3u)
#define SCALE4 f32_bits(0x39000000u)

/* trunc_to_i32 + output_code (idct.rs:108-123). */
uint16_t orc_output_code(float v)
{
    int32_t t = (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)v : INT32_MIN;
    int32_t centred = (int32_t)((uint32_t)t - 32768u);
    if (centred < -32768)
        centred = -32768;
    if (centred > 32767)
        centred = 32767;
    return (uint16_t)((uint16_t)(int16_t)centred ^ 0x8000u);
}

/* apple_block (idct.rs:126-174) with row_pass (idct.rs:93-105). */
void orc_apple_block(const int16_t *coefs, const int16_t *q, uint16_t *out)
{
    float rows[8][8];
    for (int r = 0; r < 8; r++) {
        int32_t x[8];
        for (int c = 0; c < 8; c++)
            x[c] = orc_dequantize(coefs[8 * r + c], q[8 * r + c]);
        const int32_t(*t)[8] = ROW_TABLES[ROW_SET[r]];
        int32_t pairs[4][2] = {{x[0], x[2]}, {x[4], x[6]}, {x[1], x[3]}, {x[5], x[7]}};
        for (int k = 0; k < 4; k++) {
            int32_t d0 = pairs[0][0] * t[0][2 * k] + pairs[0][1] * t[0][2 * k + 1];
            int32_t d1 = pairs[1][0] * t[1][2 * k] + pairs[1][1] * t[1][2 * k + 1];
            int32_t d2 = pairs[2][0] * t[2][2 * k] + pairs[2][1] * t[2][2 * k + 1];
            int32_t d3 = pairs[3][0] * t[3][2 * k] + pairs[3][1] * t[3][2 * k + 1];
            int32_t even = d0 + d1, odd = d2 + d3;
            rows[r][k] = (float)(even + odd) * SCALE;
            rows[r][7 - k] = (float)(even - odd) * SCALE;
        }
    }
    for (int c = 0; c < 8; c++) {
        float r0 = rows[0][c], r1 = rows[1][c], r2 = rows[2][c], r3 = rows[3][c];
        float r4 = rows[4][c], r5 = rows[5][c], r6 = rows[6][c], r7 = rows[7][c];
        float p = r0 + BIAS;
        float ea = p + r4, eb = p - r4;
        float f = r6 * TAN2 + r2, g = r2 * TAN2 - r6;
        float e0 = ea + f, e3 = ea - f, e1 = eb + g, e2 = eb - g;
        float a = r7 * TAN1 + r1, b = r1 * TAN1 - r7;
        float cc = r5 * TAN3 + r3, d = r5 - r3 * TAN3;
        float o0 = a + cc, t = a - cc, o3 = b + d, u = b - d;
        float o1 = (t + u) * COS4, o2 = (t - u) * COS4;
        float y[8] = {e0 + o0, e1 + o1, e2 + o2, e3 + o3, e3 - o3, e2 - o2, e1 - o1, e0 - o0};
        for (int i = 0; i < 8; i++)
            out[8 * i + c] = orc_output_code(y[i]);
    }
}

/* saturate_code (idct.rs:189-191). */
static uint16_t saturate_code(int32_t v)
{
    if (v < -32768)
        v = -32768;
    if (v > 32767)
        v = 32767;
    return (uint16_t)((uint16_t)(int16_t)v ^ 0x8000u);
}

/* apple_block_4x4 (idct.rs:179-226). */
void orc_apple_block_4x4(const int16_t *coefs, const int16_t *q, uint16_t *out)
{
    static const int32_t K[2] = {8192, 10703};
    static const int32_t T[2][4] = {{10703, 4433, 4433, -10703}, {13985, 5793, 5793, -13985}};
    float rows[4][4];
    for (int r = 0; r < 4; r++) {
        int32_t x[4];
        for (int c = 0; c < 4; c++)
            x[c] = orc_dequantize(coefs[4 * r + c], q[8 * r + c]);
        int32_t k = K[r & 1];
        const int32_t *t = T[r & 1];
        int32_t e0 = k * x[0] + k * x[2], e1 = k * x[0] - k * x[2];
        int32_t o0 = t[0] * x[1] + t[1] * x[3], o1 = t[2] * x[1] + t[3] * x[3];
        rows[r][0] = (float)(e0 + o0) * SCALE4;
        rows[r][1] = (float)(e1 + o1) * SCALE4;
        rows[r][2] = (float)(e1 - o1) * SCALE4;
        rows[r][3] = (float)(e0 - o0) * SCALE4;
    }
    for (int c = 0; c < 4; c++) {
        float r0 = rows[0][c] + BIAS;
        float r1 = rows[1][c], r2 = rows[2][c], r3 = rows[3][c];
        float e0 = r2 + r0, e1 = r0 - r2;
        float f = r3 * TAN2 + r1, g = r1 * TAN2 - r3;
        out[0 + c] = orc_output_code(f + e0);
        out[4 + c] = orc_output_code(g + e1);
        out[8 + c] = orc_output_code(e1 - g);
        out[12 + c] = orc_output_code(e0 - f);
    }
}

/* apple_block_2x2 (idct.rs:232-247). */
void orc_apple_block_2x2(const int16_t *coefs, const int16_t *q, uint16_t *out)
{
    int32_t d0 = orc_dequantize(coefs[0], q[0]), d1 = orc_dequantize(coefs[1], q[1]);
    int32_t d2 = orc_dequantize(coefs[2], q[8]), d3 = orc_dequantize(coefs[3], q[9]);
    int32_t ts = d0 + d1, td = d0 - d1, bs = d2 + d3, bd = d2 - d3;
    out[0] = saturate_code(ts + bs);
    out[1] = saturate_code(td + bd);
    out[2] = saturate_code(ts - bs);
    out[3] = saturate_code(td - bd);
}

/* code_value (idct.rs:263-265). */
uint16_t orc_code_value(int32_t centered)
{
    int64_t v = (int64_t)centered + 32768;
    return (uint16_t)(v < 0 ? 0 : v > 65535 ? 65535 : v);
}

#define W1 22725
#define W2 21407
#define W3 19265
#define W4 16384
#define W5 12873
#define W6 8867
#define W7 4520

static uint32_t wm(int32_t k, int32_t v) { return (uint32_t)k * (uint32_t)v; }

/* butterfly + outputs (idct.rs:281-332), wrapping 32-bit arithmetic. */
static void ff_butterfly(const int32_t *x, int32_t round, int shift, int32_t *y)
{
    uint32_t base = wm(W4, x[0]) + (uint32_t)round;
    uint32_t a[4] = {base + wm(W2, x[2]) + wm(W4, x[4]) + wm(W6, x[6]),
                     base + wm(W6, x[2]) - wm(W4, x[4]) - wm(W2, x[6]),
                     base - wm(W6, x[2]) - wm(W4, x[4]) + wm(W2, x[6]),
                     base - wm(W2, x[2]) + wm(W4, x[4]) - wm(W6, x[6])};
    uint32_t b[4] = {wm(W1, x[1]) + wm(W3, x[3]) + wm(W5, x[5]) + wm(W7, x[7]),
                     wm(W3, x[1]) - wm(W7, x[3]) - wm(W1, x[5]) - wm(W5, x[7]),
                     wm(W5, x[1]) - wm(W1, x[3]) + wm(W7, x[5]) + wm(W3, x[7]),
                     wm(W7, x[1]) - wm(W5, x[3]) + wm(W3, x[5]) - wm(W1, x[7])};
    for (int k = 0; k < 4; k++) {
        y[k] = (int32_t)(a[k] + b[k]) >> shift;
        y[7 - k] = (int32_t)(a[k] - b[k]) >> shift;
    }
}

/* ffmpeg_block (idct.rs:268-351). */
void orc_ffmpeg_block(const int16_t *coefs, const int16_t *q, uint16_t *out)
{
    int32_t block[64];
    for (int i = 0; i < 64; i++)
        block[i] = orc_dequantize(coefs[i], q[i]);
    for (int r = 0; r < 8; r++) {
        int32_t y[8];
        ff_butterfly(block + 8 * r, 1 << 12, 13, y);
        memcpy(block + 8 * r, y, sizeof y);
    }
    for (int c = 0; c < 8; c++) {
        int32_t x[8], y[8];
        for (int i = 0; i < 8; i++)
            x[i] = block[8 * i + c];
        x[0] = (int32_t)((uint32_t)x[0] + 1u);
        ff_butterfly(x, 0, 15, y);
        for (int i = 0; i < 8; i++)
            block[8 * i + c] = y[i];
    }
    for (int i = 0; i < 64; i++)
        out[i] = orc_code_value(block[i]);
}

/* testgen::Gen (idct.rs:362-442): SplitMix64 and the six block kinds. */
typedef struct { uint64_t s; } orc_gen;

static uint64_t gen_next(orc_gen *g)
{
    g->s += 0x9e3779b97f4a7c15ull;
    uint64_t z = g->s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static uint64_t gen_below(orc_gen *g, uint64_t n) { return gen_next(g) % n; }
static int16_t gen_signed(orc_gen *g, int32_t m)
{
    return (int16_t)((int32_t)gen_below(g, 2 * (uint64_t)m + 1) - m);
}
static void gen_q(orc_gen *g, uint64_t qmat, uint64_t qscale, int16_t *q)
{
    uint64_t s = 1 + gen_below(g, qscale);
    for (int i = 0; i < 64; i++)
        q[i] = (int16_t)((1 + gen_below(g, qmat)) * s);
}
static void gen_block(orc_gen *g, uint64_t i, int16_t *c, int16_t *q)
{
    memset(c, 0, 128);
    switch (i % 6) {
    case 0:
        gen_q(g, 63, 255, q);
        c[0] = (int16_t)gen_next(g);
        break;
    case 1: {
        gen_q(g, 63, 255, q);
        c[0] = gen_signed(g, 2047);
        uint64_t n = 1 + gen_below(g, 9);
        for (uint64_t k = 0; k < n; k++) {
            size_t at = (size_t)gen_below(g, 64);
            c[at] = gen_signed(g, 2047);
        }
        break;
    }
    case 2:
        gen_q(g, 16, 8, q);
        for (int k = 0; k < 64; k++)
            c[k] = gen_signed(g, 255);
        break;
    case 3:
        gen_q(g, 8, 4, q);
        for (int k = 0; k < 64; k++) {
            int sh = (k / 8 + k % 8) / 2;
            c[k] = gen_signed(g, 4096 >> (sh < 12 ? sh : 12));
        }
        break;
    case 4:
        gen_q(g, 63, 255, q);
        for (int k = 0; k < 64; k++)
            c[k] = (int16_t)gen_next(g);
        break;
    default:
        for (int k = 0; k < 64; k++)
            q[k] = (int16_t)gen_next(g);
        for (int k = 0; k < 64; k++)
            c[k] = (int16_t)gen_next(g);
        break;
    }
}

void orc_testgen_block(uint64_t seed, uint64_t i, int16_t *c, int16_t *q)
{
    orc_gen g = {seed};
    int16_t tc[64], tq[64];
    for (uint64_t k = 0; k < i; k++)
        gen_block(&g, k, tc, tq);
    gen_block(&g, i, c, q);
}

/* FNV-1a over LE u16 outputs (idct.rs:444-459) of n blocks. */
uint64_t orc_testgen_hash(uint64_t seed, uint64_t n, int ffmpeg)
{
    orc_gen g = {seed};
    uint64_t h = 0xcbf29ce484222325ull;
    int16_t c[64], q[64];
    uint16_t out[64];
    for (uint64_t i = 0; i < n; i++) {
        gen_block(&g, i, c, q);
        if (ffmpeg)
            orc_ffmpeg_block(c, q, out);
        else
            orc_apple_block(c, q, out);
        for (int k = 0; k < 64; k++) {
            h ^= out[k] & 0xff;
            h *= 0x00000100000001b3ull;
            h ^= out[k] >> 8;
            h *= 0x00000100000001b3ull;
        }
    }
    return h;
}

/* ------------------------------------------------------------------ */
/* Curve (curve.rs:14-24). */

uint16_t orc_linearize(const uint16_t *curve, uint16_t code)
{
    uint32_t seg = code >> 13, frac = code & 0x1FFF;
    uint32_t c0 = curve[seg];
    uint32_t c1 = seg + 1 < 8 ? curve[seg + 1] : 65536u;
    uint32_t delta = (c1 - c0) & 0xFFFF;
    uint32_t v = (c0 * 8192 + delta * frac + 4096) >> 13;
    return (uint16_t)(v < 65535 ? v : 65535);
}

static void curve_lut(const uint16_t *curve, uint16_t *lut)
{
    for (uint32_t c = 0; c < 65536; c++)
        lut[c] = orc_linearize(curve, (uint16_t)c);
}

/* ------------------------------------------------------------------ */
/* Header and frame layout (decoder.rs:45-123, header.rs:156-246). */

static void set_err(orc_error *e, int code, const char *what, uint64_t a, uint64_t b)
{
    if (!e)
        return;
    memset(e, 0, sizeof *e);
    e->code = code;
    if (what)
        strncpy(e->what, what, sizeof e->what - 1);
    e->a = a;
    e->b = b;
}

static uint16_t be16(const uint8_t *p) { return (uint16_t)((p[0] << 8) | p[1]); }
static uint32_t be32(const uint8_t *p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static float bef32(const uint8_t *p)
{
    uint32_t u = be32(p);
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* header_len (decoder.rs:58-73). */
int orc_header_len(const uint8_t *packet, size_t len, size_t *out, orc_error *err)
{
    if (len < 10) {
        set_err(err, ORC_E_TRUNCATED, "frame prefix", 10, len);
        return ORC_E_TRUNCATED;
    }
    if (memcmp(packet + 4, "prrf", 4) != 0) {
        set_err(err, ORC_E_BAD_TAG, NULL, 0, 0);
        if (err)
            memcpy(err->tag, packet + 4, 4);
        return ORC_E_BAD_TAG;
    }
    uint16_t hs = be16(packet + 8);
    if (hs < 72) {
        set_err(err, ORC_E_HEADER_TOO_SHORT, NULL, hs, 0);
        return ORC_E_HEADER_TOO_SHORT;
    }
    *out = 10 + (size_t)hs - 2;
    return ORC_OK;
}

/* FrameHeader::parse (header.rs:156-246) on the body. */
static int parse_body(const uint8_t *b, size_t n, orc_header *h, orc_error *err)
{
    if (n < 70) {
        set_err(err, ORC_E_TRUNCATED, "frame header", 70, n);
        return ORC_E_TRUNCATED;
    }
    memset(h, 0, sizeof *h);
    h->version = b[1];
    if (h->version > 1) {
        set_err(err, ORC_E_UNSUPPORTED_VERSION, NULL, h->version, 0);
        return ORC_E_UNSUPPORTED_VERSION;
    }
    memcpy(h->vendor, b + 2, 4);
    h->width = be16(b + 6);
    h->height = be16(b + 8);
    if (h->width == 0 || h->height == 0 || (h->width & 1) || (h->height & 1)) {
        set_err(err, ORC_E_BAD_DIMENSIONS, NULL, h->width, h->height);
        return ORC_E_BAD_DIMENSIONS;
    }
    memcpy(h->crop, b + 10, 4);
    h->bayer = (uint8_t)(be16(b + 14) & 3);
    h->white_level = (uint16_t)(be16(b + 16) + 256u);
    h->wb_red = bef32(b + 18);
    h->wb_blue = bef32(b + 22);
    for (int i = 0; i < 9; i++)
        h->color_matrix[i] = bef32(b + 26 + 4 * i);
    h->gain = bef32(b + 62);
    h->wb_cct = be16(b + 66);
    h->flags = be16(b + 68);
    h->alignment = (uint8_t)((h->flags >> 1) & 7);
    if (h->alignment > 4) {
        set_err(err, ORC_E_BAD_ALIGNMENT, NULL, h->alignment, 0);
        return ORC_E_BAD_ALIGNMENT;
    }
    size_t pos = 70;
    memset(h->quant_matrix, 1, 64);
    if (h->flags & 1) {
        if (pos + 64 > n) {
            set_err(err, ORC_E_TRUNCATED, "quantization matrix", pos + 64, n);
            return ORC_E_TRUNCATED;
        }
        memcpy(h->quant_matrix, b + pos, 64);
        pos += 64;
        for (int i = 0; i < 64; i++)
            if (h->quant_matrix[i] < 1 || h->quant_matrix[i] > 63) {
                set_err(err, ORC_E_BAD_QUANT_MATRIX, NULL, (uint64_t)i, h->quant_matrix[i]);
                return ORC_E_BAD_QUANT_MATRIX;
            }
    }
    static const uint16_t DEFAULT_CURVE[8] = {0, 512, 1024, 2048, 4096, 8192, 16384, 32768};
    if ((h->flags >> 4) & 1) {
        if (pos + 16 > n) {
            set_err(err, ORC_E_TRUNCATED, "linearization curve", pos + 16, n);
            return ORC_E_TRUNCATED;
        }
        uint16_t pts[8];
        for (int i = 0; i < 8; i++)
            pts[i] = be16(b + pos + 2 * i);
        if (h->version == 0) {
            set_err(err, ORC_E_UNEXPECTED_CURVE, NULL, 0, 0);
            return ORC_E_UNEXPECTED_CURVE;
        }
        for (int i = 0; i < 7; i++)
            if (pts[i] >= pts[i + 1]) {
                set_err(err, ORC_E_BAD_CURVE, NULL, 0, 0);
                if (err)
                    memcpy(err->points, pts, sizeof pts);
                return ORC_E_BAD_CURVE;
            }
        memcpy(h->curve, pts, sizeof pts);
    } else {
        memcpy(h->curve, DEFAULT_CURVE, sizeof DEFAULT_CURVE);
    }
    return ORC_OK;
}

/* parse_header (decoder.rs:45-53). */
int orc_parse_header(const uint8_t *packet, size_t len, orc_header *out, orc_error *err)
{
    size_t body_end;
    int rc = orc_header_len(packet, len, &body_end, err);
    if (rc)
        return rc;
    if (body_end > len) {
        set_err(err, ORC_E_TRUNCATED, "frame header", body_end, len);
        return ORC_E_TRUNCATED;
    }
    return parse_body(packet + 10, body_end - 10, out, err);
}

/* TileGrid (tile.rs:36-107). */
size_t orc_tiles_per_row(unsigned width, unsigned alignment)
{
    size_t columns = (width + 15) / 16;
    size_t rem = columns & ((1u << alignment) - 1);
    return (columns >> alignment) + (size_t)__builtin_popcountll(rem);
}

size_t orc_tile_count(unsigned width, unsigned height, unsigned alignment)
{
    return orc_tiles_per_row(width, alignment) * ((height + 15) / 16);
}

void orc_tile_geometry(unsigned width, unsigned height, unsigned alignment, uint32_t *x,
                       uint32_t *y, uint8_t *log2_blocks)
{
    size_t columns = (width + 15) / 16, rows = (height + 15) / 16, k = 0;
    for (size_t row = 0; row < rows; row++) {
        size_t column = 0, remaining = columns;
        int log2 = (int)alignment;
        while (remaining > 0) {
            size_t unit = (size_t)1 << log2;
            if (unit <= remaining) {
                x[k] = (uint32_t)(column * 16);
                y[k] = (uint32_t)(row * 16);
                log2_blocks[k] = (uint8_t)log2;
                k++;
                column += unit;
                remaining -= unit;
            } else {
                log2--;
            }
        }
    }
}

/* Tile::parse (tile.rs:126-154). */
int orc_tile_parse(const uint8_t *data, size_t len, uint8_t *qscale, uint32_t *offsets,
                   uint32_t *lengths)
{
    if (len < 8)
        return ORC_E_BAD_TILE_HEADER;
    size_t header_len = data[0] >> 3;
    uint8_t qs = data[1];
    size_t s0 = be16(data + 2), s1 = be16(data + 4), s2 = be16(data + 6);
    if (header_len < 8 || qs == 0 || s0 == 0 || s1 == 0 || s2 == 0)
        return ORC_E_BAD_TILE_HEADER;
    size_t used = header_len + s0 + s1 + s2;
    if (used >= len)
        return ORC_E_BAD_TILE_HEADER;
    *qscale = qs;
    offsets[0] = (uint32_t)header_len;
    offsets[1] = (uint32_t)(header_len + s0);
    offsets[2] = (uint32_t)(header_len + s0 + s1);
    offsets[3] = (uint32_t)used;
    lengths[0] = (uint32_t)s0;
    lengths[1] = (uint32_t)s1;
    lengths[2] = (uint32_t)s2;
    lengths[3] = (uint32_t)(len - used);
    return ORC_OK;
}

/* parse_frame (decoder.rs:76-123). */
int orc_parse_frame(const uint8_t *packet, size_t len, orc_header *hdr,
                    uint64_t *tile_offsets, uint32_t *tile_sizes, size_t cap,
                    size_t *nb_tiles, uint8_t *max_qscale, orc_error *err)
{
    if (len < 10) {
        set_err(err, ORC_E_TRUNCATED, "frame prefix", 10, len);
        return ORC_E_TRUNCATED;
    }
    uint32_t declared = be32(packet);
    if (declared == 0 || (size_t)declared > len) {
        set_err(err, ORC_E_FRAME_SIZE_MISMATCH, NULL, declared, len);
        return ORC_E_FRAME_SIZE_MISMATCH;
    }
    size_t body_end;
    int rc = orc_header_len(packet, len, &body_end, err);
    if (rc)
        return rc;
    rc = orc_parse_header(packet, len, hdr, err);
    if (rc)
        return rc;
    size_t n = orc_tile_count(hdr->width, hdr->height, hdr->alignment);
    size_t table_end = body_end + 2 * n;
    if (table_end > len) {
        set_err(err, ORC_E_TRUNCATED, "tile size table", table_end, len);
        return ORC_E_TRUNCATED;
    }
    *nb_tiles = n;
    if (tile_offsets && cap < n) {
        set_err(err, ORC_E_TRUNCATED, "tile capacity", n, cap);
        return ORC_E_TRUNCATED;
    }
    uint64_t offset = table_end;
    uint8_t mq = 0;
    for (size_t i = 0; i < n; i++) {
        size_t size = be16(packet + body_end + 2 * i);
        uint8_t q = offset + 1 < len ? packet[offset + 1] : 0;
        if (q > mq)
            mq = q;
        if (offset + size > len) {
            set_err(err, ORC_E_TILE_OUT_OF_BOUNDS, NULL, i, 0);
            return ORC_E_TILE_OUT_OF_BOUNDS;
        }
        if (tile_offsets) {
            tile_offsets[i] = offset;
            tile_sizes[i] = (uint32_t)size;
        }
        offset += size;
    }
    *max_qscale = mq;
    return ORC_OK;
}

/* Stream census of a frame: for tile t and component s (decode order of
 * decoder.rs:286-306), entry 4t+s receives the stream's bytes and its
 * symbol count (DC values + AC symbols, counted by orc_count_symbols);
 * invalid tiles get 0 bytes and 0 symbols. Returns the tile count, or 0 on
 * a frame error. */
size_t orc_frame_stream_census(const uint8_t *packet, size_t len, uint32_t *bytes,
                               uint32_t *symbols, size_t cap)
{
    orc_header h;
    orc_error e;
    size_t n;
    uint8_t mq;
    if (orc_parse_frame(packet, len, &h, NULL, NULL, 0, &n, &mq, &e) || n > cap)
        return 0;
    uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    uint32_t *sz = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint32_t *tx = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint32_t *ty = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint8_t *tl = (uint8_t *)malloc(n + 1);
    orc_parse_frame(packet, len, &h, off, sz, n, &n, &mq, &e);
    orc_tile_geometry(h.width, h.height, h.alignment, tx, ty, tl);
    for (size_t t = 0; t < n; t++) {
        uint8_t qs;
        uint32_t so[4], sl[4];
        if (orc_tile_parse(packet + off[t], sz[t], &qs, so, sl)) {
            for (int k = 0; k < 4; k++)
                bytes[4 * t + k] = symbols[4 * t + k] = 0;
            continue;
        }
        for (int k = 0; k < 4; k++) {
            uint64_t start = off[t] + so[k], ac, bits;
            orc_count_symbols(packet + start, sl[k], (unsigned)(start & 3), tl[t], 6,
                              (uint64_t)1 << 40, &ac, &bits);
            bytes[4 * t + k] = sl[k];
            symbols[4 * t + k] = (uint32_t)(ac + ((size_t)1 << tl[t]));
        }
    }
    free(off);
    free(sz);
    free(tx);
    free(ty);
    free(tl);
    return n;
}

/* ------------------------------------------------------------------ */
/* Frame decoders (decoder.rs:221-307, reduced.rs:147-273). */

typedef struct {
    const uint8_t *packet;
    orc_header hdr;
    uint64_t *offsets;
    uint32_t *sizes;
    size_t per_row;
    size_t stride;      /* output row stride in values (full) or pixels (reduced) */
    int ffmpeg;
    unsigned shift;     /* reduced: log2 factor */
    int average;        /* reduced: average the greens */
    const uint16_t *lut;/* reduced: curve table */
    void *out;
    uint8_t *skipped;   /* per tile flag */
    uint32_t *tx, *ty;
    uint8_t *tl;
} decode_ctx;

/* BayerPattern::stream_phases (header.rs:79-82). */
static void stream_phases(uint8_t bayer, unsigned pr[4], unsigned pc[4])
{
    static const unsigned red[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
    unsigned r = red[bayer][0], c = red[bayer][1];
    pr[0] = 1 - r; pc[0] = c;
    pr[1] = r;     pc[1] = 1 - c;
    pr[2] = 1 - r; pc[2] = 1 - c;
    pr[3] = r;     pc[3] = c;
}

/* decode_tile (decoder.rs:272-307) for every tile of tile row `row`. */
static void decode_full_row(void *vctx, size_t row, void *scratch)
{
    (void)scratch;
    decode_ctx *d = (decode_ctx *)vctx;
    uint16_t *samples = (uint16_t *)d->out;
    unsigned pr[4], pc[4];
    stream_phases(d->hdr.bayer, pr, pc);
    int16_t coefs[64 * 16];
    for (size_t i = 0; i < d->per_row; i++) {
        size_t t = row * d->per_row + i;
        const uint8_t *tile = d->packet + d->offsets[t];
        uint8_t qscale;
        uint32_t so[4], sl[4];
        if (orc_tile_parse(tile, d->sizes[t], &qscale, so, sl)) {
            d->skipped[t] = 1;
            continue;
        }
        int16_t scaled[64];
        for (int k = 0; k < 64; k++)
            scaled[k] = (int16_t)(uint16_t)(d->hdr.quant_matrix[k] * qscale);
        unsigned log2 = d->tl[t];
        size_t blocks = (size_t)1 << log2;
        for (int s = 0; s < 4; s++) {
            uint64_t start = d->offsets[t] + so[s];
            orc_decode_component_reduced(d->packet + start, sl[s], (unsigned)(start & 3), log2, 6,
                                         coefs);
            for (size_t b = 0; b < blocks; b++) {
                uint16_t v[64];
                if (d->ffmpeg)
                    orc_ffmpeg_block(coefs + 64 * b, scaled, v);
                else
                    orc_apple_block(coefs + 64 * b, scaled, v);
                size_t x0 = d->tx[t] + b * 16 + pc[s];
                for (int r = 0; r < 8; r++) {
                    uint16_t *dst = samples + (d->ty[t] + 2 * r + pr[s]) * d->stride + x0;
                    for (int c = 0; c < 8; c++)
                        dst[2 * c] = v[8 * r + c];
                }
            }
        }
    }
}

/* TileContext::decode (reduced.rs:217-273) for every tile of a tile row. */
static void decode_reduced_row(void *vctx, size_t row, void *scratch)
{
    (void)scratch;
    decode_ctx *d = (decode_ctx *)vctx;
    uint16_t *rgb = (uint16_t *)d->out;
    unsigned log2_coefs = d->shift == 1 ? 6 : d->shift == 2 ? 4 : 2;
    size_t per_block = (size_t)1 << log2_coefs;
    size_t side = 16u >> d->shift;
    size_t rows_per_tile = 16u >> d->shift;
    int16_t coefs[64 * 16];
    for (size_t i = 0; i < d->per_row; i++) {
        size_t t = row * d->per_row + i;
        const uint8_t *tile = d->packet + d->offsets[t];
        uint8_t qscale;
        uint32_t so[4], sl[4];
        if (orc_tile_parse(tile, d->sizes[t], &qscale, so, sl)) {
            d->skipped[t] = 1;
            continue;
        }
        int16_t q[64];
        for (int k = 0; k < 64; k++)
            q[k] = (int16_t)(uint16_t)(d->hdr.quant_matrix[k] * qscale);
        unsigned log2 = d->tl[t];
        size_t blocks = (size_t)1 << log2;
        size_t x0 = d->tx[t] >> d->shift;
        uint16_t *rows = rgb + 3 * d->stride * rows_per_tile * row;
        static const int plan[4][3] = {{0, 3, 0}, {1, 0, 0}, {1, 1, 1}, {2, 2, 0}};
        for (int p = 0; p < 4; p++) {
            int channel = plan[p][0], stream = plan[p][1], average = plan[p][2];
            if (average && !d->average)
                continue;
            uint64_t start = d->offsets[t] + so[stream];
            orc_decode_component_reduced(d->packet + start, sl[stream], (unsigned)(start & 3),
                                         log2, log2_coefs, coefs);
            for (size_t b = 0; b < blocks; b++) {
                uint16_t v[64];
                const int16_t *blk = coefs + per_block * b;
                if (d->shift == 1)
                    orc_apple_block(blk, q, v);
                else if (d->shift == 2)
                    orc_apple_block_4x4(blk, q, v);
                else
                    orc_apple_block_2x2(blk, q, v);
                size_t bx = x0 + b * side;
                for (size_t li = 0; li < side; li++) {
                    uint16_t *dst = rows + 3 * (li * d->stride + bx) + channel;
                    for (size_t c = 0; c < side; c++) {
                        uint16_t lin = d->lut[v[li * side + c]];
                        if (average)
                            dst[3 * c] = (uint16_t)(((uint32_t)dst[3 * c] + lin + 1) / 2);
                        else
                            dst[3 * c] = lin;
                    }
                }
            }
        }
    }
}

static int run_decode(const uint8_t *packet, size_t len, int reduced, int ffmpeg,
                      unsigned shift, int average, int threads, void *out, size_t out_cap,
                      uint32_t *skipped, size_t skipped_cap, size_t *n_skipped,
                      uint8_t *max_qscale, orc_error *err)
{
    decode_ctx d;
    memset(&d, 0, sizeof d);
    size_t n;
    uint8_t mq;
    int rc = orc_parse_frame(packet, len, &d.hdr, NULL, NULL, 0, &n, &mq, err);
    if (rc)
        return rc;
    size_t cw = ((size_t)d.hdr.width + 15) & ~(size_t)15;
    size_t ch = ((size_t)d.hdr.height + 15) & ~(size_t)15;
    size_t need = reduced ? 3 * (cw >> shift) * (ch >> shift) : cw * ch;
    if (out_cap < need || skipped_cap < n) {
        set_err(err, ORC_E_TRUNCATED, "output capacity", need, out_cap);
        return ORC_E_TRUNCATED;
    }
    d.offsets = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    d.sizes = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    d.tx = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    d.ty = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    d.tl = (uint8_t *)malloc(n + 1);
    d.skipped = (uint8_t *)calloc(n + 1, 1);
    orc_parse_frame(packet, len, &d.hdr, d.offsets, d.sizes, n, &n, &mq, err);
    orc_tile_geometry(d.hdr.width, d.hdr.height, d.hdr.alignment, d.tx, d.ty, d.tl);
    d.packet = packet;
    d.per_row = orc_tiles_per_row(d.hdr.width, d.hdr.alignment);
    d.ffmpeg = ffmpeg;
    d.shift = shift;
    d.average = average;
    d.out = out;
    uint16_t *lut = NULL;
    memset(out, 0, need * sizeof(uint16_t));
    if (reduced) {
        d.stride = cw >> shift;
        lut = (uint16_t *)malloc(65536 * sizeof(uint16_t));
        curve_lut(d.hdr.curve, lut);
        d.lut = lut;
        orc_parallel_for(ch / 16, threads, 0, decode_reduced_row, &d);
    } else {
        d.stride = cw;
        orc_parallel_for(ch / 16, threads, 0, decode_full_row, &d);
    }
    size_t k = 0;
    for (size_t t = 0; t < n; t++)
        if (d.skipped[t])
            skipped[k++] = (uint32_t)t;
    *n_skipped = k;
    *max_qscale = mq;
    free(lut);
    free(d.offsets);
    free(d.sizes);
    free(d.tx);
    free(d.ty);
    free(d.tl);
    free(d.skipped);
    return ORC_OK;
}

int orc_decode_frame(const uint8_t *packet, size_t len, int ffmpeg, int threads,
                     uint16_t *samples, size_t samples_cap, uint32_t *skipped,
                     size_t skipped_cap, size_t *n_skipped, uint8_t *max_qscale,
                     orc_error *err)
{
    return run_decode(packet, len, 0, ffmpeg, 0, 0, threads, samples, samples_cap, skipped,
                      skipped_cap, n_skipped, max_qscale, err);
}

int orc_decode_frame_reduced(const uint8_t *packet, size_t len, unsigned log2_factor,
                             int average_greens, int threads, uint16_t *rgb,
                             size_t rgb_cap, uint32_t *skipped, size_t skipped_cap,
                             size_t *n_skipped, uint8_t *max_qscale, orc_error *err)
{
    return run_decode(packet, len, 1, 0, log2_factor, average_greens, threads, rgb, rgb_cap,
                      skipped, skipped_cap, n_skipped, max_qscale, err);
}

/* Frame::linearize_into (decoder.rs:197-216), rows spread over threads. */
typedef struct {
    const uint16_t *lut, *samples;
    size_t stride, width;
    uint16_t *out;
} lin_ctx;

static void lin_row(void *vctx, size_t y, void *scratch)
{
    (void)scratch;
    lin_ctx *c = (lin_ctx *)vctx;
    const uint16_t *src = c->samples + y * c->stride;
    uint16_t *dst = c->out + y * c->width;
    for (size_t x = 0; x < c->width; x++)
        dst[x] = c->lut[src[x]];
}

void orc_linearize_frame(const uint16_t *curve, const uint16_t *samples, size_t stride,
                         size_t width, size_t height, uint16_t *out, int threads)
{
    if (width == 0)
        return;
    uint16_t *lut = (uint16_t *)malloc(65536 * sizeof(uint16_t));
    curve_lut(curve, lut);
    lin_ctx c = {lut, samples, stride, width, out};
    orc_parallel_for(height, threads, 0, lin_row, &c);
    free(lut);
}

/* ------------------------------------------------------------------ */
/* Develop (demosaic.rs). */

static const double XYZ_TO_REC2020[3][3] = {
    {1.716651187971268, -0.355670783776393, -0.25336628137366},
    {-0.666684351832489, 1.616481236634939, 0.015768545813911},
    {0.017639857445311, -0.042770613257809, 0.942103121235474}};

/* DemosaicParams::with_white_balance (demosaic.rs:86-106). */
void orc_params_with_white_balance(const orc_header *h, uint32_t quantizer, float red,
                                   float blue, orc_params *p)
{
    uint32_t range = (uint16_t)(h->white_level - 256u);
    double q = (double)quantizer;
    double base = q * q * 5.0 + 64.0 / 3.0;
    double range2 = (double)range * (double)range;
    float rangef = (float)range;
    p->cfa_phase = h->bayer;
    p->black = 256;
    p->one_over_range = 1.0f / rangef;
    p->noise[0] = (float)(base / 150994944.0);
    p->noise[1] = (float)(base / (range2 * 256.0));
    p->noise[2] = (float)(64.0 / 3.0 / range2);
    p->clip[0] = red == 0.0f ? 1.0f : 1.0f / red;
    p->clip[1] = blue == 0.0f ? 1.0f : 1.0f / blue;
    p->black_over_range = 256.0f / rangef;
    p->wb[0] = red;
    p->wb[1] = blue;
}

/* output_matrix (demosaic.rs:138-153). */
void orc_output_matrix(const orc_header *h, int normalized, float *m)
{
    double gain = (double)h->gain;
    if (normalized)
        gain = gain * 10.0 / 9.0;
    for (int i = 0; i < 9; i++) {
        int row = i / 3, col = i % 3;
        double v = XYZ_TO_REC2020[row][0] * (double)h->color_matrix[0 * 3 + col] +
                   XYZ_TO_REC2020[row][1] * (double)h->color_matrix[1 * 3 + col] +
                   XYZ_TO_REC2020[row][2] * (double)h->color_matrix[2 * 3 + col];
        m[i] = (float)(v * gain);
    }
}

/* variance (demosaic.rs:423-426). */
static inline float variance(float v, float off, const float *n)
{
    float t = off + v;
    return n[2] + fmaxf(n[1], n[0] * (t * t));
}

/* dot4 (demosaic.rs:431-433). */
static inline float dot4(const float *a, const float *b)
{
    return fmaf(a[3], b[3], fmaf(a[2], b[2], fmaf(a[0], b[0], a[1] * b[1])));
}

/* mirror (demosaic.rs:547-557). */
static inline void mirror4(float *d, size_t pos, size_t len, float c)
{
    if (pos == 0) {
        float a = d[1], b = d[3];
        d[0] = a; d[1] = a; d[2] = b; d[3] = b;
    } else if (pos == 1) {
        d[2] = c;
    } else if (pos == len - 1) {
        float a = d[0], b = d[2];
        d[0] = a; d[1] = a; d[2] = b; d[3] = b;
    } else if (pos == len - 2) {
        d[3] = c;
    }
}

/* green_row_impl (demosaic.rs:481-541). rows[0..5] are rows y-2..y+2
 * clamped; out holds [rb, g] pairs. */
static void green_row(const float *rows[5], size_t y, size_t h, size_t w,
                      const orc_params *p, float *out)
{
    float bs = (float)p->black * p->one_over_range;
    size_t red_row = (size_t)(p->cfa_phase >> 1), red_col = (size_t)(p->cfa_phase & 1);
    size_t last_x = w - 1;
    const float *centre = rows[2];
    for (size_t x = 0; x < w; x++) {
        out[2 * x] = 0.0f;
        out[2 * x + 1] = centre[x];
    }
    int phase = (y % 2) == red_row;
    size_t first = phase ? red_col : 1 - red_col;
    float clip = phase ? p->clip[0] : p->clip[1];
    float limit[4] = {1.0f, 1.0f, clip, clip};
    for (size_t x = first; x < w; x += 2) {
        float c = fminf(centre[x], clip);
#define AT(dx) centre[((ptrdiff_t)x + (dx)) < 0 ? 0 : ((size_t)((ptrdiff_t)x + (dx)) > last_x ? last_x : (size_t)((ptrdiff_t)x + (dx)))]
        float hh[4] = {AT(-1), AT(1), AT(-2), AT(2)};
#undef AT
        float vv[4] = {rows[1][x], rows[3][x], rows[0][x], rows[4][x]};
        for (int i = 0; i < 4; i++) {
            hh[i] = fminf(hh[i], limit[i]);
            vv[i] = fminf(vv[i], limit[i]);
        }
        mirror4(hh, x, w, c);
        mirror4(vv, y, h, c);
        float var_h[4], var_v[4];
        for (int i = 0; i < 4; i++) {
            var_h[i] = variance(hh[i], bs, p->noise);
            var_v[i] = variance(vv[i], bs, p->noise);
        }
        float sums[4] = {hh[0] + hh[1], vv[0] + vv[1], hh[2] + hh[3], vv[2] + vv[3]};
        float ratio[2] = {0.0f, 0.0f};
        if (c > sqrtf(variance(c, bs, p->noise))) {
            float t[4] = {c * 2.0f + sums[2], c * 2.0f + sums[3], c * -2.0f + sums[2],
                          c * -2.0f + sums[3]};
            for (int dir = 0; dir < 2; dir++) {
                const float *d = dir ? vv : hh;
                const float *vd = dir ? var_v : var_h;
                float sd[4];
                for (int i = 0; i < 4; i++)
                    sd[i] = sqrtf(vd[i]);
                float a = d[0] - d[1], b = d[2] - d[3];
                int any = a * b > 0.0f || (d[2] - c) * (d[3] - c) > 0.0f ||
                          sd[0] + sd[1] > fabsf(a) || sd[2] + sd[3] > fabsf(b);
                int all = d[0] > sd[0] && d[1] > sd[1] && d[2] > sd[2] && d[3] > sd[3];
                if (any && all)
                    ratio[dir] = t[2 + dir] / (t[dir] * 2.0f);
            }
        }
        float gh = sums[0] * 0.5f * (1.0f - ratio[0]);
        float gv = sums[1] * 0.5f * (1.0f - ratio[1]);
        float dh[4] = {hh[0] - gh, hh[1] - gh, hh[2] - c, hh[3] - c};
        float dv[4] = {vv[0] - gv, vv[1] - gv, vv[2] - c, vv[3] - c};
        float eh = dot4(dh, dh) + (var_h[0] + var_h[1]);
        float ev = dot4(dv, dv) + (var_v[0] + var_v[1]);
        out[2 * x] = c;
        out[2 * x + 1] = gv + (gh - gv) * ev / (eh + ev);
    }
}

/* estimate (demosaic.rs:572-604). */
static float estimate(const float *xx, const float *gg, float cg, float cs, float off,
                      const float *noise)
{
    static const float quarter[4] = {0.25f, 0.25f, 0.25f, 0.25f};
    static const float ones[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    float mean_x = dot4(xx, quarter);
    if (!(cs < cg))
        return mean_x;
    float var_x[4], sx[4], sg[4];
    int above = 1;
    for (int i = 0; i < 4; i++) {
        var_x[i] = variance(xx[i], off, noise);
        sx[i] = sqrtf(var_x[i]);
        sg[i] = sqrtf(variance(gg[i], off, noise));
        above &= (sx[i] < xx[i]) & (sg[i] < gg[i]);
    }
    if (!above)
        return mean_x;
    float lo[4], mid[4], hi[4], ml[4], hm[4];
    for (int i = 0; i < 4; i++) {
        lo[i] = ((cg - cs) * (xx[i] - sx[i])) / (sg[i] + gg[i]);
        mid[i] = (cg * xx[i]) / gg[i];
        hi[i] = ((cs + cg) * (sx[i] + xx[i])) / (gg[i] - sg[i]);
    }
    for (int i = 0; i < 4; i++) {
        ml[i] = mid[i] - lo[i];
        hm[i] = hi[i] - mid[i];
    }
    float spread = dot4(ml, hm);
    float mean_mid = dot4(mid, quarter);
    float dm[4], dx[4];
    for (int i = 0; i < 4; i++) {
        dm[i] = mid[i] - mean_mid;
        dx[i] = xx[i] - mean_x;
    }
    float b = fmaf(spread, 0.75f, dot4(dm, dm));
    float a = fmaf(dot4(var_x, ones), 0.75f, dot4(dx, dx));
    float weight = a / (b + a);
    return fmaf(mean_mid - mean_x, weight, mean_x);
}

static inline void mirrored(size_t y, size_t len, size_t *ym, size_t *yp)
{
    *ym = y > 0 ? y - 1 : y + 1;
    *yp = y < len - 1 ? y + 1 : y - 1;
}

/* chroma_row_with (demosaic.rs:648-688): rows are the green-pass rows
 * above, at and below y; writes rgb via put (planar or interleaved). */
static void chroma_row(const float *up, const float *mid, const float *down, size_t y,
                       size_t w, const orc_params *p, float *r, float *g, float *b,
                       size_t step)
{
    float off = p->black_over_range;
    int on_red_row = (y % 2) == (size_t)(p->cfa_phase >> 1);
    size_t red_col = (size_t)(p->cfa_phase & 1);
    size_t rb_col = on_red_row ? red_col : 1 - red_col;
    for (size_t x = 0; x < w; x++) {
        size_t xm, xp;
        mirrored(x, w, &xm, &xp);
        float c = mid[2 * x], gv = mid[2 * x + 1];
        float sigma = sqrtf(variance(gv, off, p->noise));
        float rr, bb;
        if (x % 2 == rb_col) {
            float xx[4] = {up[2 * xm], down[2 * xm], up[2 * xp], down[2 * xp]};
            float gg[4] = {up[2 * xm + 1], down[2 * xm + 1], up[2 * xp + 1], down[2 * xp + 1]};
            float e = estimate(xx, gg, gv, sigma, off, p->noise);
            if (on_red_row) { rr = c; bb = e; } else { rr = e; bb = c; }
        } else {
            float hx[4] = {mid[2 * xm], mid[2 * xp], mid[2 * xm], mid[2 * xp]};
            float hg[4] = {mid[2 * xm + 1], mid[2 * xp + 1], mid[2 * xm + 1], mid[2 * xp + 1]};
            float vx[4] = {up[2 * x], down[2 * x], up[2 * x], down[2 * x]};
            float vg[4] = {up[2 * x + 1], down[2 * x + 1], up[2 * x + 1], down[2 * x + 1]};
            float eh = estimate(hx, hg, gv, sigma, off, p->noise);
            float ev = estimate(vx, vg, gv, sigma, off, p->noise);
            if (on_red_row) { rr = eh; bb = ev; } else { rr = ev; bb = eh; }
        }
        r[x * step] = fminf(rr * p->wb[0], 1.0f);
        g[x * step] = fminf(gv, 1.0f);
        b[x * step] = fminf(bb * p->wb[1], 1.0f);
    }
}

/* matrix_planes_impl / matrix_rows_impl (demosaic.rs:700-720). */
static inline void matrix_px(float *r, float *g, float *b, const float *m)
{
    float rv = *r, gv = *g, bv = *b;
    *r = fmaf(bv, m[2], fmaf(rv, m[0], gv * m[1]));
    *g = fmaf(bv, m[5], fmaf(rv, m[3], gv * m[4]));
    *b = fmaf(bv, m[8], fmaf(rv, m[6], gv * m[7]));
}

#define BAND 32

typedef struct {
    const uint16_t *samples;
    size_t stride, w, h;
    orc_params p;
    const float *lut;
    int has_matrix;
    float m[9];
    float *r, *g, *b; /* g == NULL: interleaved into r */
} develop_ctx;

/* Conversion::band (demosaic.rs:356-398). Scratch holds the normalized
 * rows and the green rows of the band. */
static void develop_band(void *vctx, size_t index, void *scratch)
{
    develop_ctx *d = (develop_ctx *)vctx;
    size_t w = d->w, h = d->h;
    size_t y0 = index * BAND, y1 = y0 + BAND < h ? y0 + BAND : h;
    size_t g0 = y0 > 0 ? y0 - 1 : 0, g1 = y1 + 1 < h ? y1 + 1 : h;
    size_t n0 = g0 > 2 ? g0 - 2 : 0, n1 = g1 + 2 < h ? g1 + 2 : h;
    float *norm = (float *)scratch;
    float *green = norm + (size_t)(BAND + 6) * w;
    for (size_t y = n0; y < n1; y++) {
        const uint16_t *src = d->samples + y * d->stride;
        float *dst = norm + (y - n0) * w;
        for (size_t x = 0; x < w; x++)
            dst[x] = d->lut[src[x]];
    }
    for (size_t y = g0; y < g1; y++) {
        const float *rows[5];
        for (int k = 0; k < 5; k++) {
            ptrdiff_t yy = (ptrdiff_t)y + k - 2;
            if (yy < 0)
                yy = 0;
            if (yy > (ptrdiff_t)h - 1)
                yy = (ptrdiff_t)h - 1;
            rows[k] = norm + ((size_t)yy - n0) * w;
        }
        green_row(rows, y, h, w, &d->p, green + (y - g0) * 2 * w);
    }
    for (size_t y = y0; y < y1; y++) {
        size_t ym, yp;
        mirrored(y, h, &ym, &yp);
        const float *up = green + (ym - g0) * 2 * w, *mid = green + (y - g0) * 2 * w,
                    *down = green + (yp - g0) * 2 * w;
        if (d->g) {
            float *r = d->r + y * w, *g = d->g + y * w, *b = d->b + y * w;
            chroma_row(up, mid, down, y, w, &d->p, r, g, b, 1);
            if (d->has_matrix)
                for (size_t x = 0; x < w; x++)
                    matrix_px(r + x, g + x, b + x, d->m);
        } else {
            float *o = d->r + 3 * y * w;
            chroma_row(up, mid, down, y, w, &d->p, o, o + 1, o + 2, 3);
            if (d->has_matrix)
                for (size_t x = 0; x < w; x++)
                    matrix_px(o + 3 * x, o + 3 * x + 1, o + 3 * x + 2, d->m);
        }
    }
}

/* develop_into / develop_planar_into via Conversion (demosaic.rs:223-304). */
void orc_develop(const orc_header *h, const uint16_t *samples, size_t stride,
                 const orc_params *params, int stage, int threads, float *r, float *g,
                 float *b)
{
    develop_ctx d;
    memset(&d, 0, sizeof d);
    d.samples = samples;
    d.stride = stride;
    d.w = h->width;
    d.h = h->height;
    d.p = *params;
    if (stage == 0) {
        d.p.wb[0] = d.p.wb[1] = 1.0f;
        d.p.clip[0] = d.p.clip[1] = 1.0f;
    }
    if (stage == 2) {
        d.has_matrix = 1;
        orc_output_matrix(h, 1, d.m);
    }
    uint16_t *clut = (uint16_t *)malloc(65536 * sizeof(uint16_t));
    float *lut = (float *)malloc(65536 * sizeof(float));
    curve_lut(h->curve, clut);
    for (size_t c = 0; c < 65536; c++)
        lut[c] = d.p.one_over_range * (float)((int32_t)clut[c] - d.p.black);
    d.lut = lut;
    d.r = r;
    d.g = g;
    d.b = b;
    size_t bands = (d.h + BAND - 1) / BAND;
    size_t scratch = sizeof(float) * ((size_t)(BAND + 6) * d.w + (size_t)(BAND + 2) * 2 * d.w);
    orc_parallel_for(bands, threads, scratch, develop_band, &d);
    free(lut);
    free(clut);
}

/* green_pass (demosaic.rs:453-476). */
void orc_green_pass(const uint16_t *linear, size_t w, size_t h, const orc_params *p,
                    float *out2)
{
    float *norm = (float *)malloc(sizeof(float) * w * h);
    for (size_t i = 0; i < w * h; i++)
        norm[i] = p->one_over_range * (float)((int32_t)linear[i] - p->black);
    for (size_t y = 0; y < h; y++) {
        const float *rows[5];
        for (int k = 0; k < 5; k++) {
            ptrdiff_t yy = (ptrdiff_t)y + k - 2;
            if (yy < 0)
                yy = 0;
            if (yy > (ptrdiff_t)h - 1)
                yy = (ptrdiff_t)h - 1;
            rows[k] = norm + (size_t)yy * w;
        }
        green_row(rows, y, h, w, p, out2 + 2 * y * w);
    }
    free(norm);
}

/* chroma_pass (demosaic.rs:608-617). */
void orc_chroma_pass(const float *green2, size_t w, size_t h, const orc_params *p,
                     float *rgb)
{
    for (size_t y = 0; y < h; y++) {
        size_t ym, yp;
        mirrored(y, h, &ym, &yp);
        float *o = rgb + 3 * y * w;
        chroma_row(green2 + 2 * ym * w, green2 + 2 * y * w, green2 + 2 * yp * w, y, w, p, o,
                   o + 1, o + 2, 3);
    }
}

/* apply_matrix (demosaic.rs:692-696) in place on n interleaved pixels. */
void orc_apply_matrix(float *rgb, size_t n, const float *m)
{
    for (size_t i = 0; i < n; i++)
        matrix_px(rgb + 3 * i, rgb + 3 * i + 1, rgb + 3 * i + 2, m);
}
