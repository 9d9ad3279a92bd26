/* HIP kernels for gfx950 (MI355X): ProRes RAW decode and develop.
 *
 *  k_prepare      status reset (no hipMemsetAsync: graph safe), size table
 *                 -> tile offsets, max_qscale, first OOB tile (parse_frame,
 *                 decoder.rs:104-116), tile groups ordered longest first
 *  k_decode<F>    one wave per group of tiles: one lane per component
 *                 stream runs the entropy decoder (bits.rs,
 *                 coefficients.rs) over the packet, staging coefficients
 *                 in the tile's own region of the code-value buffer; a
 *                 chunk of blocks at a time goes to LDS and the wave runs
 *                 the 8x8 inverse transform
 *                 (idct.rs Apple or FFmpeg), places each stream on its
 *                 Bayer phase (decoder.rs:272-307) through an LDS tile and
 *                 writes code values and, fused, linear values (curve.rs)
 *                 with coalesced stores
 *  k_reduced      the same front end with 8x8 / 4x4 / 2x2 transforms to
 *                 linear interleaved RGB (reduced.rs:217-273)
 *  k_linearize    Frame::linearize_into (decoder.rs:197-216)
 *  k_develop      green pass + chroma pass + white balance + output matrix
 *                 in one kernel over 64x16 pixel tiles with LDS halos
 *                 (demosaic.rs:356-720)
 *
 * Floating point follows the crate operation for operation: this file is
 * compiled with -ffp-contract=off, fmaf appears exactly where the crate
 * calls mul_add, and division and square root are correctly rounded (the
 * defaults, or div_in_range where its bounds are shown to hold).
 */
#include "prr_internal.h"
#include "prr_net.h"

#include <mutex>

namespace prr {

/* ------------------------------------------------------------------ */
/* Tables (codebook.rs:26-110, idct.rs:51-87). */

/* The codebooks as restated (codebook.rs:26-110): AC 0..94, RN 96..123,
 * LN 128..142, DC 144..156. */
enum { TBL_AC = 0, TBL_RN = 96, TBL_LN = 128, TBL_DC = 144, TBL_N = 160 };
/* The table the decode loop reads from LDS, one dword per entry: entry i
 * below 95 holds AC_CB[i] in its low half and LN_CB[min(i, 14)] in its high
 * half (the two codebooks a level after a run of zeros updates), then
 * RN_CB at K_RN and DC_CB at K_DC. */
enum { K_RN = 95, K_DC = 123, K_N = 136 };
/* Bytes before the per-tile records: the LDS table, curve, quant matrix. */
enum { LDS_CURVE = K_N * 4, LDS_QMAT = 592, LDS_HEAD = 656 };

__device__ const uint16_t g_codebooks[TBL_N] = {
    /* AC_CB: runs 1x000, 1x211, 2x111, 3x222, 3x122, 8x233, 2x133, 22x244, 52x355, 1x166 */
    0x000, 0x211, 0x111, 0x111, 0x222, 0x222, 0x222, 0x122, 0x122, 0x122,
    0x233, 0x233, 0x233, 0x233, 0x233, 0x233, 0x233, 0x233, 0x133, 0x133,
    0x244, 0x244, 0x244, 0x244, 0x244, 0x244, 0x244, 0x244, 0x244, 0x244,
    0x244, 0x244, 0x244, 0x244, 0x244, 0x244, 0x244, 0x244, 0x244, 0x244,
    0x244, 0x244, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355,
    0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355,
    0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355,
    0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355,
    0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355, 0x355,
    0x355, 0x355, 0x355, 0x355, 0x166, 0, /* 96: RN_CB */
    0x200, 0x100, 0x000, 0x000, 0x211, 0x211, 0x111, 0x111, 0x011, 0x011,
    0x021, 0x021, 0x222, 0x022, 0x022, 0x022, 0x022, 0x022, 0x022, 0x022,
    0x022, 0x022, 0x022, 0x022, 0x022, 0x032, 0x032, 0x044, 0, 0, 0, 0, /* 128: LN_CB */
    0x100, 0x111, 0x222, 0x222, 0x122, 0x122, 0x433, 0x433, 0x233, 0x233,
    0x233, 0x233, 0x233, 0x233, 0x033, 0, /* 144: DC_CB */
    0x010, 0x021, 0x032, 0x033, 0x033, 0x033, 0x044, 0x044, 0x044, 0x044,
    0x044, 0x044, 0x076, 0, 0, 0};

/* Scan orders (codebook.rs:98-110): position -> raster index of the block, or of its
 * 4x4 or 2x2 corner. The decode loop stores coefficients by scan position;
 * the transforms read raster index k from position kInv*.v[k]. */
constexpr uint8_t kScan8[64] = {
    0,  8,  1,  9,  16, 24, 17, 25, 2,  10, 3,  11, 18, 26, 19, 27, 32, 40, 33, 34, 41, 48,
    56, 49, 42, 35, 43, 50, 57, 58, 51, 59, 4,  12, 5,  6,  13, 20, 28, 21, 14, 7,  15, 22,
    29, 36, 44, 37, 30, 23, 31, 38, 45, 52, 60, 53, 46, 39, 47, 54, 61, 62, 55, 63};
constexpr uint8_t kScan4[16] = {0, 4, 1, 5, 8, 12, 9, 13, 2, 6, 3, 7, 10, 14, 11, 15};
constexpr uint8_t kScan2[4] = {0, 2, 1, 3};

struct ScanInv {
    uint8_t v[64];
};
template <int N>
constexpr ScanInv invert_scan(const uint8_t (&scan)[N])
{
    ScanInv r{};
    for (int i = 0; i < N; i++)
        r.v[scan[i]] = (uint8_t)i;
    return r;
}
__device__ constexpr ScanInv kInv8 = invert_scan(kScan8);
__device__ constexpr ScanInv kInv4 = invert_scan(kScan4);
__device__ constexpr ScanInv kInv2 = invert_scan(kScan2);

__device__ constexpr int32_t kRowTables[4][4][8] = {
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
__device__ constexpr int kRowSet[8] = {0, 1, 2, 3, 0, 3, 2, 1};

__device__ inline float bits_f32(uint32_t u) { return __uint_as_float(u); }

/* ------------------------------------------------------------------ */
/* Small helpers. */

__device__ inline uint32_t be16_at(const uint8_t *p)
{
    return ((uint32_t)p[0] << 8) | p[1];
}

/* Curve::linearize (curve.rs:14-24) from the 9-point curve in LDS
 * (c[8] = 65536). */
__device__ inline uint32_t linearize(const uint32_t *c, uint32_t code)
{
    uint32_t seg = code >> 13, frac = code & 0x1FFFu;
    uint32_t c0 = c[seg], c1 = c[seg + 1];
    uint32_t v = (c0 * 8192u + ((c1 - c0) & 0xFFFFu) * frac + 4096u) >> 13;
    return v < 65535u ? v : 65535u;
}

/* Tile geometry from its frame-order index (TileGrid::row, tile.rs:78-102). */
__device__ inline void tile_geometry(const FrameDesc &f, uint32_t t, uint32_t &x, uint32_t &y,
                                     uint32_t &log2)
{
    uint32_t row = t / f.tiles_per_row;
    uint32_t k = t - row * f.tiles_per_row;
    uint32_t a = f.alignment;
    uint32_t full = f.columns >> a;
    y = row * 16;
    if (k < full) {
        x = (k << a) * 16;
        log2 = a;
        return;
    }
    uint32_t j = k - full, rem = f.columns & ((1u << a) - 1), col = full << a;
    log2 = 0;
    x = col * 16;
    for (int e = (int)a - 1; e >= 0; e--) {
        if (rem & (1u << e)) {
            if (j == 0) {
                log2 = (uint32_t)e;
                x = col * 16;
                return;
            }
            j--;
            col += 1u << e;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Bit reader (bits.rs:17-111) over the packet in global memory, addressed
 * by dword from a 4-byte aligned base; positions are byte offsets from it.
 * Reads go through the vector memory counter, apart from the LDS traffic
 * of the decode loop (codebook reads, coefficient stores), so waiting for
 * a codebook entry never waits for a refill word. The pointer carries its
 * address space so the loads are global_load, never flat. */

#define PRR_LDS __attribute__((address_space(3)))
#define PRR_GLOBAL __attribute__((address_space(1)))
typedef const PRR_GLOBAL uint32_t *global_words;

template <typename P>
__device__ inline uint32_t load_be(P buf, uint32_t pos)
{
    uint32_t i = pos >> 2;
    uint32_t lo = buf[i], hi = buf[i + 1];
    uint32_t w = __builtin_amdgcn_alignbyte(hi, lo, pos & 3u);
    return __builtin_bswap32(w);
}

/* The top `n` bytes of w, zero below (n clamped to 0..4). */
__device__ inline uint32_t keep_bytes(uint32_t w, int32_t n)
{
    if (n >= 4)
        return w;
    if (n <= 0)
        return 0;
    return w & (0xFFFFFFFFu << (32 - 8 * n));
}

/* Reader state. After BitReader::new every word a refill reads starts on a
 * 4-byte boundary of the packet (new() reads up to the boundary first), so
 * the word after the one in `nxt` is one aligned dword: `raw`, loaded one
 * refill ahead so its latency overlaps the symbols in between. Words at
 * or past the end of the stream read as zero (next_word, bits.rs:79-89);
 * the one partial word, at `ql`, is kept whole in `lw`. */
template <typename P>
struct Bits {
    P buf;
    uint32_t hi, lo; /* the 64-bit window */
    int32_t extra;
    uint32_t nxt;    /* the next word a refill ORs in */
    uint32_t raw;    /* source dword at min(q, ql) */
    uint32_t q;      /* byte position of the word after nxt */
    uint32_t ql;     /* stop & ~3 */
    uint32_t lw;     /* the partial word at ql, zero filled (0 if none) */
};

/* BitReader::new (bits.rs:44-76) over bytes [start, stop) of buf; align is
 * the packet address of `start` mod 4. */
template <typename P>
__device__ inline void bits_init(Bits<P> &b, P buf, uint32_t start, uint32_t stop, uint32_t align)
{
    b.buf = buf;
    uint32_t remaining = stop - start, address = align, pos, first;
    if (remaining < 4) {
        first = keep_bytes(load_be(buf, start), (int32_t)remaining);
        pos = start + remaining;
        address += remaining;
        remaining = 0;
    } else {
        first = load_be(buf, start);
        pos = start + 4;
        address += 4;
        remaining -= 4;
    }
    uint32_t low = 0;
    b.extra = 0;
    const uint32_t to_boundary = (4u - (address & 3u)) & 3u;
    if (to_boundary != 0 && to_boundary <= remaining) {
        low = keep_bytes(load_be(buf, pos), (int32_t)to_boundary);
        pos += to_boundary;
        b.extra = 8 * (int32_t)to_boundary;
    }
    b.hi = first;
    b.lo = low;
    /* The first next_word(). When pos is not on a boundary here (a stream
     * too short to reach one), fewer than 4 bytes are left, so q below is
     * past the end and every later word is zero. */
    const int32_t n = (int32_t)(stop - pos);
    b.nxt = n > 0 ? keep_bytes(load_be(buf, pos), n) : 0u;
    b.q = pos + 4;
    b.ql = stop & ~3u;
    b.lw = (stop & 3u) ? keep_bytes(__builtin_bswap32(buf[b.ql >> 2]), (int32_t)(stop & 3u)) : 0u;
    b.raw = buf[min(b.q, b.ql) >> 2];
}

/* consume (bits.rs:97-105). */
template <typename P>
__device__ inline void bits_consume(Bits<P> &b, uint32_t len)
{
    uint64_t w = (((uint64_t)b.hi << 32) | b.lo) << (len & 63u);
    b.extra = (int32_t)((uint32_t)b.extra - len);
    if (b.extra < 0) {
        const uint32_t shift = (uint32_t)(-(int32_t)(int8_t)(uint8_t)(uint32_t)b.extra) & 0x3fu;
        w |= (uint64_t)b.nxt << shift;
        b.extra += 32;
        b.nxt = b.q < b.ql ? __builtin_bswap32(b.raw) : b.q == b.ql ? b.lw : 0u;
        b.q += 4;
        b.raw = b.buf[min(b.q, b.ql) >> 2];
    }
    b.hi = (uint32_t)(w >> 32);
    b.lo = (uint32_t)w;
}

/* codeword (bits.rs:116-134) of a non-zero view. The first branch of the
 * reference (top bit set) is the Rice branch with zero leading zeros. */
__device__ inline void codeword(uint32_t p, uint32_t cb, uint32_t &value, uint32_t &len)
{
    uint32_t lz = (uint32_t)__builtin_clz(p);
    uint32_t s = cb >> 8, e = (cb >> 4) & 15u, r = cb & 15u;
    uint32_t tail = ((p << ((lz + 1) & 31u)) >> 1) >> ((31u - r) & 31u);
    uint32_t rice_value = (lz << r) + tail, rice_len = r + 1 + lz;
    uint32_t l = 2 * lz - s + e;
    uint32_t golomb_value = ((s + 1) << r) + (p >> ((32u - l) & 31u)) - (1u << e);
    const bool rice = lz <= s;
    value = rice ? rice_value : golomb_value;
    len = rice ? rice_len : l;
}

enum { S_RUN = 0, S_ZEROS = 1, S_LEVEL = 2, S_AFTER_ZEROS = 3 };

/* Where a lane's decoded coefficients go. Coefficient n of a stream of
 * 2^lb blocks is scan position n >> lb of block n & (2^lb - 1); both sinks
 * keep that frequency-major order, which the transforms read without bank
 * conflicts once the blocks are in LDS. */

/* In LDS, at coefs[((n >> lb) << kshift) + (n & mask)] (kshift = the
 * frame's alignment, so every stream of a group has the same stride);
 * symbols that store nothing write coefs[dummy], a padding slot. */
struct LdsSink {
    PRR_LDS int16_t *coefs;
    uint32_t lb, kshift, mask, dummy;
    __device__ void dc(uint32_t blk, uint32_t v) const { coefs[blk] = (int16_t)v; }
    __device__ void ac(bool level, uint32_t n, uint32_t v) const
    {
        coefs[level ? ((n >> lb) << kshift) + (n & mask) : dummy] = (int16_t)v;
    }
};

/* In the frame's code-value buffer, over the stream's own share of its
 * tile: a tile of 2^lb blocks a stream is 16 rows of 16 << lb sensels, the
 * same number of values as its four streams' coefficients, so stream s
 * takes rows 4s .. 4s + 3 of the tile and coefficient n sits at row
 * n >> (4 + lb), column n & (16 << lb) - 1 of them. The region is zeroed
 * before the decode; only levels are stored. `row0` is the u16 offset of
 * the stream's first sensel, gap = coded width - tile width. */
struct GlobalSink {
    uint16_t *samples;
    uint32_t row0, sh, gap;
    __device__ void dc(uint32_t blk, uint32_t v) const { samples[row0 + blk] = (uint16_t)v; }
    __device__ void ac(bool level, uint32_t n, uint32_t v) const
    {
        /* n >> sh is a row (0..3) and gap below 2^16: a 24-bit product,
         * full rate where a 32-bit multiply issues at quarter rate. */
        const uint32_t at = row0 + n + __mul24(n >> sh, gap);
        if (level)
            samples[at] = (uint16_t)v;
    }
};

/* decode_component_reduced (coefficients.rs:37-146) of one stream by one
 * lane. The DC values are a short loop of their own; the AC part runs one
 * symbol per iteration through four states (run length, zero run, level,
 * level after zeros) with selects rather than branches, so lanes decoding
 * different streams stay in step. Coefficients go to `sink`. Returns true
 * if the symbol budget ran out. */
template <typename P, class Sink>
__device__ inline bool decode_stream(Bits<P> &b, uint32_t lb, uint32_t log2_coefs,
                                     const PRR_LDS uint32_t *tbl, const Sink &sink,
                                     uint32_t budget)
{
    const uint32_t blocks = 1u << lb;
    const uint32_t codes = 1u << (lb + log2_coefs);
    uint32_t sym, len;
    if (b.hi == 0)
        return false;
    codeword(b.hi, 0x2BCu, sym, len);
    bits_consume(b, len);
    const uint32_t odd = sym & 1u;
    int32_t dc = (int32_t)((((odd + sym) >> 1) ^ (0u - odd)) + odd);
    sink.dc(0, (uint32_t)dc);
    uint32_t dc_cb = 0x064u, sign = 0;
    for (uint32_t blk = 1; blk < blocks; blk++) {
        if (b.hi == 0)
            return false;
        codeword(b.hi, dc_cb, sym, len);
        sign ^= sym & 1u;
        const uint32_t magnitude = ((sym & 1u) + sym) >> 1;
        const int32_t delta = (int32_t)(((0u - sign) ^ magnitude) + sign);
        dc_cb = tbl[K_DC + min(magnitude, 12u)];
        sign = delta < 0 ? 1u : 0u;
        bits_consume(b, len);
        dc = (int32_t)((uint32_t)dc + (uint32_t)delta);
        sink.dc(blk, (uint32_t)dc);
    }

    uint32_t n = blocks, state = S_RUN, left = 0;
    uint32_t ac_cb = 0x031u, rn_cb = 0x000u, ln_cb = 0x042u, cb = 0x042u;
    bool done = false;
    /* One exit test per symbol: an empty view, the last coefficient, or
     * the budget. */
    for (uint32_t it = blocks; !((b.hi == 0) | done | (it >= budget)); ++it) {
        const uint32_t top = b.hi;
        codeword(top, cb, sym, len);
        const bool zeros = state == S_ZEROS, level = state >= S_LEVEL;
        const uint32_t tv = tbl[zeros ? K_RN + min(sym, 27u) : min(sym, 94u)];
        bits_consume(b, level ? len + 1 : len);
        const bool negative = (int32_t)(top << (len & 31u)) < 0;
        const uint32_t magnitude = (sym + 1u) & 0xFFFFu;
        const uint32_t value = negative ? 0u - magnitude : magnitude;
        sink.ac(level, n, value);
        const uint32_t n2 = zeros ? n + 1u + sym : level ? n + 1u : n;
        const uint32_t left2 = state == S_RUN ? sym : left - 1u;
        const uint32_t next = zeros                    ? (uint32_t)S_AFTER_ZEROS
                              : state == S_AFTER_ZEROS ? (uint32_t)S_RUN
                              : left2                  ? (uint32_t)S_LEVEL
                                                       : (uint32_t)S_ZEROS;
        ac_cb = level ? (tv & 0xFFFFu) : ac_cb;
        rn_cb = zeros ? tv : rn_cb;
        ln_cb = state == S_AFTER_ZEROS ? tv >> 16 : ln_cb;
        cb = next == S_RUN ? ln_cb : next == S_ZEROS ? rn_cb : ac_cb;
        state = next;
        n = n2;
        left = left2;
        done = n2 >= codes;
    }
    return b.hi != 0 && !done;
}

/* ------------------------------------------------------------------ */
/* Inverse transforms (idct.rs). d holds dequantized coefficients. */

/* c is a 16-bit coefficient and q a product of two bytes (quantization
 * matrix entry and tile qscale), so both fit the 24-bit multiplier and the
 * product fits 32 bits: __mul24 is exact and runs at full rate, where a
 * 32-bit multiply issues at quarter rate. */
__device__ inline int32_t dequantize(int32_t c, int32_t q)
{
    int32_t v = __mul24(c, q);
    return v < -32768 ? -32768 : v > 32767 ? 32767 : v;
}

/* The quantizer step of a coefficient: matrix entry qm times the tile's
 * qscale, both bytes. The empty asm makes q one opaque value; otherwise the
 * compiler reassociates c * (qm * qs) into (c * qs) * qm and issues the
 * second product as a quarter-rate 32-bit multiply. */
__device__ inline int32_t quant_step(uint32_t qm, uint32_t qs)
{
    int32_t q = (int32_t)((qm & 0xFFu) * (qs & 0xFFu));
    asm("" : "+v"(q));
    return q;
}

/* trunc_to_i32 + output_code (idct.rs:108-123). */
__device__ inline uint32_t output_code(float v)
{
    int32_t t = (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)v : INT32_MIN;
    int32_t centred = (int32_t)((uint32_t)t - 32768u);
    centred = centred < -32768 ? -32768 : centred > 32767 ? 32767 : centred;
    return ((uint32_t)(uint16_t)(int16_t)centred) ^ 0x8000u;
}

/* apple_block (idct.rs:126-174): integer rows, unfused f32 columns. */
__device__ inline void apple_8x8(const int32_t *d, uint32_t *out)
{
    float rows[64];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int set = kRowSet[r];
        const int32_t *x = d + 8 * r;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int32_t even = (x[0] * kRowTables[set][0][2 * k] + x[2] * kRowTables[set][0][2 * k + 1]) +
                           (x[4] * kRowTables[set][1][2 * k] + x[6] * kRowTables[set][1][2 * k + 1]);
            int32_t odd = (x[1] * kRowTables[set][2][2 * k] + x[3] * kRowTables[set][2][2 * k + 1]) +
                          (x[5] * kRowTables[set][3][2 * k] + x[7] * kRowTables[set][3][2 * k + 1]);
            rows[8 * r + k] = (float)(even + odd) * bits_f32(0x39800000u);
            rows[8 * r + 7 - k] = (float)(even - odd) * bits_f32(0x39800000u);
        }
    }
    const float BIAS = bits_f32(0x47000080u), TAN1 = bits_f32(0x3e4bafafu),
                TAN2 = bits_f32(0x3ed413cdu), TAN3 = bits_f32(0x3f2b0dc1u),
                COS4 = bits_f32(0x3f3504f3u);
#pragma unroll
    for (int c = 0; c < 8; c++) {
        float r0 = rows[c], r1 = rows[8 + c], r2 = rows[16 + c], r3 = rows[24 + c];
        float r4 = rows[32 + c], r5 = rows[40 + c], r6 = rows[48 + c], r7 = rows[56 + c];
        float p = r0 + BIAS;
        float ea = p + r4, eb = p - r4;
        float f = r6 * TAN2 + r2, g = r2 * TAN2 - r6;
        float e0 = ea + f, e3 = ea - f, e1 = eb + g, e2 = eb - g;
        float a = r7 * TAN1 + r1, b = r1 * TAN1 - r7;
        float cc = r5 * TAN3 + r3, dd = r5 - r3 * TAN3;
        float o0 = a + cc, t = a - cc, o3 = b + dd, u = b - dd;
        float o1 = (t + u) * COS4, o2 = (t - u) * COS4;
        out[c] = output_code(e0 + o0);
        out[8 + c] = output_code(e1 + o1);
        out[16 + c] = output_code(e2 + o2);
        out[24 + c] = output_code(e3 + o3);
        out[32 + c] = output_code(e3 - o3);
        out[40 + c] = output_code(e2 - o2);
        out[48 + c] = output_code(e1 - o1);
        out[56 + c] = output_code(e0 - o0);
    }
}

/* butterfly + outputs (idct.rs:281-332), wrapping 32-bit arithmetic. */
__device__ inline void ff_pass(const int32_t *x, int32_t round, int shift, int32_t *y)
{
    auto w = [](int32_t k, int32_t v) { return (uint32_t)k * (uint32_t)v; };
    uint32_t base = w(16384, x[0]) + (uint32_t)round;
    uint32_t a0 = base + w(21407, x[2]) + w(16384, x[4]) + w(8867, x[6]);
    uint32_t a1 = base + w(8867, x[2]) - w(16384, x[4]) - w(21407, x[6]);
    uint32_t a2 = base - w(8867, x[2]) - w(16384, x[4]) + w(21407, x[6]);
    uint32_t a3 = base - w(21407, x[2]) + w(16384, x[4]) - w(8867, x[6]);
    uint32_t b0 = w(22725, x[1]) + w(19265, x[3]) + w(12873, x[5]) + w(4520, x[7]);
    uint32_t b1 = w(19265, x[1]) - w(4520, x[3]) - w(22725, x[5]) - w(12873, x[7]);
    uint32_t b2 = w(12873, x[1]) - w(22725, x[3]) + w(4520, x[5]) + w(19265, x[7]);
    uint32_t b3 = w(4520, x[1]) - w(12873, x[3]) + w(19265, x[5]) - w(22725, x[7]);
    y[0] = (int32_t)(a0 + b0) >> shift;
    y[1] = (int32_t)(a1 + b1) >> shift;
    y[2] = (int32_t)(a2 + b2) >> shift;
    y[3] = (int32_t)(a3 + b3) >> shift;
    y[4] = (int32_t)(a3 - b3) >> shift;
    y[5] = (int32_t)(a2 - b2) >> shift;
    y[6] = (int32_t)(a1 - b1) >> shift;
    y[7] = (int32_t)(a0 - b0) >> shift;
}

/* ffmpeg_block (idct.rs:268-351) with code_value (idct.rs:263-265). */
__device__ inline void ffmpeg_8x8(const int32_t *d, uint32_t *out)
{
    int32_t blk[64];
#pragma unroll
    for (int r = 0; r < 8; r++)
        ff_pass(d + 8 * r, 1 << 12, 13, blk + 8 * r);
#pragma unroll
    for (int c = 0; c < 8; c++) {
        int32_t x[8], y[8];
#pragma unroll
        for (int i = 0; i < 8; i++)
            x[i] = blk[8 * i + c];
        x[0] = (int32_t)((uint32_t)x[0] + 1u);
        ff_pass(x, 0, 15, y);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            int64_t v = (int64_t)y[i] + 32768;
            out[8 * i + c] = (uint32_t)(v < 0 ? 0 : v > 65535 ? 65535 : v);
        }
    }
}

/* apple_block_4x4 (idct.rs:179-226); d holds the 4x4 corner. */
__device__ inline void apple_4x4(const int32_t *d, uint32_t *out)
{
    const float SCALE4 = bits_f32(0x39000000u), BIAS = bits_f32(0x47000080u),
                TAN2 = bits_f32(0x3ed413cdu);
    float rows[16];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int32_t k = (r & 1) ? 10703 : 8192;
        const int32_t t0 = (r & 1) ? 13985 : 10703, t1 = (r & 1) ? 5793 : 4433;
        const int32_t t2 = (r & 1) ? 5793 : 4433, t3 = (r & 1) ? -13985 : -10703;
        const int32_t *x = d + 4 * r;
        int32_t e0 = k * x[0] + k * x[2], e1 = k * x[0] - k * x[2];
        int32_t o0 = t0 * x[1] + t1 * x[3], o1 = t2 * x[1] + t3 * x[3];
        rows[4 * r + 0] = (float)(e0 + o0) * SCALE4;
        rows[4 * r + 1] = (float)(e1 + o1) * SCALE4;
        rows[4 * r + 2] = (float)(e1 - o1) * SCALE4;
        rows[4 * r + 3] = (float)(e0 - o0) * SCALE4;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float r0 = rows[c] + BIAS;
        float r1 = rows[4 + c], r2 = rows[8 + c], r3 = rows[12 + c];
        float e0 = r2 + r0, e1 = r0 - r2;
        float f = r3 * TAN2 + r1, g = r1 * TAN2 - r3;
        out[c] = output_code(f + e0);
        out[4 + c] = output_code(g + e1);
        out[8 + c] = output_code(e1 - g);
        out[12 + c] = output_code(e0 - f);
    }
}

__device__ inline uint32_t saturate_code(int32_t v)
{
    v = v < -32768 ? -32768 : v > 32767 ? 32767 : v;
    return ((uint32_t)(uint16_t)(int16_t)v) ^ 0x8000u;
}

/* apple_block_2x2 (idct.rs:232-247). */
__device__ inline void apple_2x2(const int32_t *d, uint32_t *out)
{
    int32_t ts = d[0] + d[1], td = d[0] - d[1], bs = d[2] + d[3], bd = d[2] - d[3];
    out[0] = saturate_code(ts + bs);
    out[1] = saturate_code(td + bd);
    out[2] = saturate_code(ts - bs);
    out[3] = saturate_code(td - bd);
}

/* ------------------------------------------------------------------ */
/* Shared front end of the decode kernels: LDS layout, tile headers and
 * the entropy decode of every stream of the group. */

struct TileInfo {
    uint32_t x, y, log2, qscale, valid;
};

/* Padding (in int16) after each stream's coefficients so that the lanes
 * of one transform chunk read distinct LDS banks. */
__device__ __host__ inline uint32_t stream_stride(uint32_t per_block, uint32_t amax)
{
    uint32_t blocks = 1u << amax;
    return (per_block << amax) + (blocks < 2 ? 2 : blocks);
}

struct Group {
    uint32_t *tbl;       /* K_N entries */
    uint32_t *curve;     /* 9 entries */
    uint8_t *qmat;       /* 64 entries */
    TileInfo *tiles;     /* up to 16 */
    int16_t *coefs;
    uint32_t n, T;       /* tiles in the group, capacity */
};

/* The transform stage of the chunk whose first slot is c0: an LDS copy of
 * the chunk's tiles, placed over the chunk's own coefficient blocks. Each
 * lane has read its block's coefficients into registers before the wave
 * writes any stage value (a wave's LDS operations run in order), and the
 * stage a chunk needs (16 rows of its tiles) never exceeds the blocks of
 * those tiles: 4 streams x stride >= 4 x (per_block << alignment) values a
 * tile, which is the tile's stage size. */
__device__ inline uint16_t *chunk_stage(const Group &g, uint32_t c0, uint32_t amax,
                                        uint32_t stride)
{
    return (uint16_t *)(g.coefs + (c0 >> amax) * stride);
}

/* The tiles of one group: list[0 .. count) when list is set (tiles in the
 * order a queue chose), else t0 .. t0 + count - 1. */
struct TileSet {
    const uint32_t *list;
    uint32_t t0, count;
};

__device__ inline uint32_t tile_at(const TileSet &ts, uint32_t i)
{
    return ts.list ? ts.list[i] : ts.t0 + i;
}

/* Carve the dynamic LDS. Layout: [table 544 B][curve 36 B][qmat 64 B at
 * 592][tiles 16 x 20 B at 656][coefs]. */
__device__ inline Group carve(uint32_t *smem, uint32_t spw, uint32_t stride)
{
    Group g;
    uint8_t *base = (uint8_t *)smem;
    g.tbl = (uint32_t *)base;
    g.curve = (uint32_t *)(base + LDS_CURVE);
    g.qmat = base + LDS_QMAT;
    g.tiles = (TileInfo *)(base + LDS_HEAD);
    g.coefs = (int16_t *)(base + LDS_HEAD + 16 * sizeof(TileInfo));
    return g;
}

__host__ inline size_t group_lds_bytes(uint32_t spw, uint32_t stride)
{
    size_t coef_bytes = ((size_t)spw * stride * 2 + 15) & ~(size_t)15;
    return LDS_HEAD + 16 * sizeof(TileInfo) + coef_bytes;
}

/* LDS of a full decode (coefficients staged in the code-value buffer):
 * the chunk copy holds, for each of the chunk's tiles and streams, four
 * rows of 8 << alignment dwords, padded to `staged_stream_dwords` so that
 * the lanes reading one scan position of every block hit distinct banks
 * (pad = half the blocks, at least 4 so each copy is a 16-byte LDS store).
 * The transform stage (16 rows of the chunk's tiles, 8 KiB) is placed over
 * the same bytes. */
__device__ __host__ inline uint32_t staged_stream_dwords(uint32_t amax)
{
    const uint32_t half = (1u << amax) >> 1;
    return (32u << amax) + (half > 4 ? half : 4);
}

__device__ __host__ inline uint32_t chunk_tiles(uint32_t amax)
{
    return 64u / (4u << amax) ? 64u / (4u << amax) : 1u;
}

__host__ inline size_t staged_lds_bytes(uint32_t amax)
{
    return LDS_HEAD + 16 * sizeof(TileInfo) +
           (size_t)16 * chunk_tiles(amax) * staged_stream_dwords(amax);
}

/* Everything up to and including the entropy decode. Each lane `l` below
 * 4T decodes stream l & 3 of the group's tile l >> 2 whose streams are
 * listed in `want` (bit s set = decode stream s): into LDS (reduced
 * decodes) or, STAGED (full decodes), into the tile's region of `samples`
 * (GlobalSink; zeroed here first).
 * Returns the lane's overrun flag. */
template <bool STAGED>
__device__ inline bool front_end(const FrameDesc &f, Group &g, const TileSet &ts, uint32_t spw,
                                 uint32_t stride, uint32_t log2_coefs,
                                 uint32_t want, uint8_t *flags, prr_device_status *status,
                                 uint16_t *samples)
{
    const uint32_t lane = threadIdx.x;
    g.T = spw >> 2;
    g.n = ts.count;

    for (uint32_t i = lane; i < K_N; i += 64) {
        uint32_t v;
        if (i < K_RN)
            v = g_codebooks[TBL_AC + i] | ((uint32_t)g_codebooks[TBL_LN + min(i, 14u)] << 16);
        else if (i < K_DC)
            v = g_codebooks[TBL_RN + (i - K_RN)];
        else
            v = g_codebooks[TBL_DC + (i - K_DC)];
        g.tbl[i] = v;
    }
    /* Selects with constant indices keep `f` in scalar registers when it
     * is a register copy (the batch kernel's frame record). */
    {
        uint32_t cv = 65536u, qv = 0;
#pragma unroll
        for (uint32_t k = 0; k < 8; k++)
            cv = lane == k ? (uint32_t)f.curve[k] : cv;
#pragma unroll
        for (uint32_t k = 0; k < 16; k++)
            qv = lane == k ? ((uint32_t)f.qmat[4 * k] | ((uint32_t)f.qmat[4 * k + 1] << 8) |
                              ((uint32_t)f.qmat[4 * k + 2] << 16) |
                              ((uint32_t)f.qmat[4 * k + 3] << 24))
                           : qv;
        if (lane < 9)
            g.curve[lane] = cv;
        if (lane < 16)
            ((uint32_t *)g.qmat)[lane] = qv;
    }
    if constexpr (!STAGED) {
        uint32_t zero_words = (spw * stride) / 2;
        for (uint32_t i = lane; i < zero_words; i += 64)
            ((uint32_t *)g.coefs)[i] = 0;
    }

    /* Tile::parse (tile.rs:126-154), by the four lanes of each tile. */
    const uint32_t tl = lane >> 2, s = lane & 3;
    bool have = lane < 4 * g.T && tl < g.n;
    const uint32_t t = have ? tile_at(ts, tl) : 0u;
    uint64_t toff = 0, tend = 0;
    bool valid = false;
    uint32_t hl = 0, sz[3] = {0, 0, 0}, qs = 0;
    if (have) {
        toff = f.offsets[t];
        tend = f.offsets[t + 1];
        if (tend <= f.len && tend - toff >= 8) {
            const uint8_t *h = f.packet + toff;
            hl = h[0] >> 3;
            qs = h[1];
            sz[0] = be16_at(h + 2);
            sz[1] = be16_at(h + 4);
            sz[2] = be16_at(h + 6);
            valid = hl >= 8 && qs != 0 && sz[0] && sz[1] && sz[2] &&
                    (uint64_t)(hl + sz[0] + sz[1] + sz[2]) < tend - toff;
        }
    }
    uint32_t tx = 0, ty = 0, tlog2 = 0;
    if (have)
        tile_geometry(f, t, tx, ty, tlog2);
    if (have && s == 0) {
        TileInfo info = {tx, ty, tlog2, qs, (uint32_t)(valid ? 1 : 0)};
        if (tend > f.len)
            info.valid = 2; /* out of bounds: the frame is an error, write nothing */
        g.tiles[tl] = info;
        if (flags)
            flags[t] = valid ? 0 : 1;
        if (!valid && tend <= f.len)
            atomicAdd(&status->n_skipped, 1u);
    }
    __syncthreads();
    if constexpr (STAGED) {
        /* Zero the regions of the group's decodable tiles, 8 sensels a
         * store; the barrier's release orders these stores before the
         * coefficient stores of every lane. */
        const uint32_t sh_row = 1 + f.alignment; /* log2 of 16-byte pieces a tile row */
        const uint32_t total = g.n << (4 + sh_row);
        for (uint32_t i = lane; i < total; i += 64) {
            const uint32_t tli = i >> (4 + sh_row), row = (i >> sh_row) & 15u;
            const uint32_t q = i & ((1u << sh_row) - 1);
            const TileInfo ti = g.tiles[tli];
            if (ti.valid == 1 && q < (2u << ti.log2))
                *(uint4 *)(samples + (size_t)(ti.y + row) * f.coded_width + ti.x + 8 * q) =
                    make_uint4(0, 0, 0, 0);
        }
        __syncthreads();
    }

    bool overrun = false;
    if (valid && ((want >> s) & 1u)) {
        uint64_t start = toff + hl;
        for (uint32_t k = 0; k < s; k++)
            start += sz[k];
        uint64_t stop = s < 3 ? start + sz[s] : tend;
        uint32_t align = (uint32_t)(start & 3);
        uint32_t codes = 1u << (tlog2 + log2_coefs);
        /* The symbol budget: above the most AC symbols the crate's loop can
         * read from a stream of this many bytes and coefficients (8.25
         * bytes + codes + 1250, from the reader's refill and shift rules),
         * so it ends only streams the crate would not leave. */
        uint32_t budget = 16u * (8u * (uint32_t)(stop - start) + codes) + 2048u;
        const PRR_LDS uint32_t *tbl = (const PRR_LDS uint32_t *)g.tbl;
        const uint64_t base = start & ~(uint64_t)3;
        Bits<global_words> rd;
        bits_init(rd, (global_words)(f.packet + base), (uint32_t)(start - base),
                  (uint32_t)(stop - base), align);
        if constexpr (STAGED) {
            const GlobalSink sink{samples, (ty + 4 * s) * f.coded_width + tx, 4 + tlog2,
                                  f.coded_width - (16u << tlog2)};
            overrun = decode_stream(rd, tlog2, log2_coefs, tbl, sink, budget);
        } else {
            const LdsSink sink{(PRR_LDS int16_t *)(g.coefs + lane * stride), tlog2, f.alignment,
                               (1u << tlog2) - 1,
                               (1u << log2_coefs) << f.alignment /* first padding slot */};
            overrun = decode_stream(rd, tlog2, log2_coefs, tbl, sink, budget);
        }
    }
    if (overrun)
        atomicOr(&status->overrun, 1u);
    __syncthreads();
    return overrun;
}

/* ------------------------------------------------------------------ */

/* Exclusive prefix sum of one value per lane over a workgroup of up to
 * 1024 lanes: within each wave with shuffles, then across the waves
 * through `sums` (32 entries of LDS). Returns the lane's exclusive prefix;
 * `total` gets the sum of all. */
__device__ inline uint64_t block_exclusive(uint64_t v, uint64_t *sums, uint64_t &total)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    uint64_t incl = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint64_t u = __shfl_up(incl, d);
        if (lane >= d)
            incl += u;
    }
    if (lane == 63)
        sums[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        uint64_t w = lane < waves ? sums[lane] : 0;
#pragma unroll
        for (uint32_t d = 1; d < 16; d <<= 1) {
            const uint64_t u = __shfl_up(w, d);
            if (lane >= d)
                w += u;
        }
        if (lane < 16)
            sums[16 + lane] = w;
    }
    __syncthreads();
    total = sums[31];
    return (incl - v) + (wave ? sums[16 + wave - 1] : 0);
}

/* Size-table walk of parse_frame (decoder.rs:104-116) by one 1024-lane
 * workgroup: the tile offsets (a prefix sum of the size table), the first
 * tile that runs past the packet, and the largest qscale byte. The frame is
 * walked in passes of 1024 * PER tiles, each lane over PER consecutive tiles
 * of a pass with their sizes in registers, so the size table is read once
 * and every qscale byte of a pass is requested before any is waited for.
 * When gbytes is not NULL (frames of one pass), each tile's size is also
 * added to gbytes[tile / T], its group's coded bytes. `f` is a copy in
 * registers: the offset stores cannot alias it. */
enum { PREP_PER = 8, PREP_SMALL = 1024 * PREP_PER };
template <uint32_t PER>
__device__ inline void scan_frame(const FrameDesc f, prr_device_status *status, uint64_t *sums,
                                  uint32_t *gbytes, uint32_t T)
{
    constexpr uint32_t PASS = 1024 * PER;
    const uint32_t n = f.nb_tiles, t = threadIdx.x;
    const uint8_t *tab = f.packet + f.table_off;
    uint64_t base = (uint64_t)f.table_off + 2ull * n;
    uint32_t oob = 0xFFFFFFFFu, mq = 0;
    for (uint32_t p0 = 0; p0 < n; p0 += PASS) {
        const uint32_t m = min(n - p0, PASS);
        const uint32_t per = (m + 1023) / 1024;
        const uint32_t lo = p0 + min(t * per, m), hi = p0 + min(t * per + per, m);
        uint32_t sz[PER];
        uint64_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            sz[k] = lo + k < hi ? be16_at(tab + 2 * (size_t)(lo + k)) : 0u;
            sum += sz[k];
        }
        /* block_exclusive's reads of sums end before any lane reaches its
         * next call's first barrier, so passes need no barrier of their own. */
        uint64_t total;
        const uint64_t first = base + block_exclusive(sum, sums, total);
        base += total;
        /* A lane's tiles start at first + rel, rel < PER * 65536. */
        uint32_t q[PER];
        {
            uint32_t rel = 0;
#pragma unroll
            for (uint32_t k = 0; k < PER; k++) {
                q[k] = f.packet[min(first + rel + 1, f.len - 1)];
                rel += sz[k];
            }
        }
        uint32_t rel = 0;
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            const uint32_t i = lo + k;
            const uint64_t off = first + rel;
            rel += sz[k];
            if (i < hi) {
                const uint32_t qk = off + 1 < f.len ? q[k] : 0u;
                mq = qk > mq ? qk : mq;
                if (off + sz[k] > f.len && oob == 0xFFFFFFFFu)
                    oob = i;
                f.offsets[i] = off;
                if (gbytes)
                    atomicAdd(&gbytes[i / T], sz[k]);
            }
        }
    }
    if (t == 0)
        f.offsets[n] = base;
    /* One atomic per wave. */
#pragma unroll
    for (uint32_t d = 32; d > 0; d >>= 1) {
        oob = min(oob, (uint32_t)__shfl_xor((int)oob, (int)d));
        mq = max(mq, (uint32_t)__shfl_xor((int)mq, (int)d));
    }
    if ((t & 63) == 0) {
        if (oob != 0xFFFFFFFFu)
            atomicMin(&status->oob_tile, oob);
        if (mq)
            atomicMax(&status->max_qscale, mq);
    }
}

/* The longest-first order's bucket of a tile group of `bytes` coded bytes:
 * 64 buckets on a half-octave log scale, the longest in bucket 0. */
enum { QUEUE_BUCKETS = 64 };
__device__ inline uint32_t bytes_bucket(uint64_t bytes)
{
    const float l2 = bytes ? __log2f((float)bytes) : 0.0f;
    int b = (int)(2.0f * l2);
    b = b < 0 ? 0 : b > QUEUE_BUCKETS - 1 ? QUEUE_BUCKETS - 1 : b;
    return (uint32_t)(QUEUE_BUCKETS - 1 - b);
}

/* Exclusive prefix sum over the 64 lanes of a wave. */
__device__ inline uint32_t wave_exclusive(uint32_t v)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t incl = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(incl, d);
        if (lane >= d)
            incl += u;
    }
    return incl - v;
}

/* Tile groups ordered by coded length, longest first, into queue[0 ..
 * groups): a bucket sort on a half-octave log scale of each group's bytes
 * (`bytes_of(g)`), by one 1024-lane workgroup. Waves that start first get
 * the longest chains, and the waves that share a SIMD at the end of the
 * launch are the short ones. */
template <class Bytes>
__device__ inline void order_groups(uint32_t groups, uint32_t *queue, uint32_t *hist,
                                    uint32_t *base, Bytes bytes_of)
{
    constexpr uint32_t NB = QUEUE_BUCKETS;
    const uint32_t tid = threadIdx.x;
    auto bucket_of = [&](uint32_t g) { return bytes_bucket(bytes_of(g)); };
    for (uint32_t g = tid; g < groups; g += 1024)
        atomicAdd(&hist[bucket_of(g)], 1u);
    __syncthreads();
    if (tid < NB)
        base[tid] = wave_exclusive(hist[tid]);
    __syncthreads();
    for (uint32_t g = tid; g < groups; g += 1024)
        queue[atomicAdd(&base[bucket_of(g)], 1u)] = g;
}

/* Status reset, size-table scan and the order of the frame's groups of T
 * tiles (T = 1: the tiles themselves, longest first). */
__global__ __launch_bounds__(1024) void k_prepare(FrameDesc f, prr_device_status *status,
                                                  uint32_t *queue, uint32_t T)
{
    __shared__ uint64_t sums[32];
    __shared__ uint32_t hist[64], base[64];
    __shared__ uint32_t gbytes[PREP_SMALL];
    const uint32_t tid = threadIdx.x, n = f.nb_tiles, groups = (n + T - 1) / T;
    if (tid == 0) {
        status->oob_tile = 0xFFFFFFFFu;
        status->max_qscale = 0;
        status->n_skipped = 0;
        status->overrun = 0;
    }
    if (tid < 64)
        hist[tid] = 0;
    if (n <= PREP_SMALL) {
        for (uint32_t g = tid; g < groups; g += 1024)
            gbytes[g] = 0;
        __syncthreads();
        scan_frame<PREP_PER>(f, status, sums, gbytes, T);
        __syncthreads();
        order_groups(groups, queue, hist, base, [&](uint32_t g) { return (uint64_t)gbytes[g]; });
    } else {
        __syncthreads();
        scan_frame<PREP_PER>(f, status, sums, nullptr, T);
        __syncthreads();
        order_groups(groups, queue, hist, base, [&](uint32_t g) {
            const uint32_t t0 = g * T, t1 = min(t0 + T, n);
            return f.offsets[t1] - f.offsets[t0];
        });
    }
}

/* Full decode (decoder.rs:226-307) of tile group `group` (spw/4
 * consecutive tiles) by one 64-lane wave. The coefficients are decoded
 * into the tiles' own regions of `samples` and copied into LDS a chunk at
 * a time before the transforms, so a wave holds a few KiB of LDS whatever
 * spw is: 64-stream waves at four a SIMD. (With every stream's
 * coefficients in LDS, a 64-stream wave took 68 KiB.) */
template <int FFMPEG>
__device__ inline void decode_group_full(const FrameDesc &f, const TileSet &ts, uint16_t *samples,
                                         uint16_t *linear, uint8_t *flags,
                                         prr_device_status *status, uint32_t spw,
                                         uint32_t *smem)
{
    const uint32_t amax = f.alignment;
    Group g = carve(smem, spw, 0);
    front_end<true>(f, g, ts, spw, 0, 6, 0xFu, flags, status, samples);

    const uint32_t lane = threadIdx.x;
    const uint32_t blocks_max = 1u << amax;
    const uint32_t W = 16u << amax;                 /* stage tile width */
    const uint32_t tiles_per_chunk = chunk_tiles(amax);
    const uint32_t SS = staged_stream_dwords(amax);
    /* Bayer phases of the streams (header.rs:79-82). */
    const uint32_t rr = f.bayer >> 1, rc = f.bayer & 1;
    const uint32_t slots = (4u * g.T) << amax;
    for (uint32_t c0 = 0; c0 < slots; c0 += 64) {
        const uint32_t slot = c0 + lane;
        const uint32_t ls = slot >> amax, b = slot & (blocks_max - 1);
        const uint32_t tl = ls >> 2, s = ls & 3;
        const uint32_t tc = tl - (c0 >> amax) / 4;
        /* The transform stage (16 rows of the chunk's tiles) over the
         * chunk's coefficients: each lane has read its block's coefficients
         * into registers before the wave writes any stage value (a wave's
         * LDS operations run in order). */
        uint16_t *const stage = (uint16_t *)g.coefs;
        {
            /* The chunk's coefficients into LDS: each decodable tile's
             * region, 16 bytes a lane and load, tile t's stream s rows at
             * (t * 4 + s) * SS dwords. */
            uint32_t *area = (uint32_t *)g.coefs;
            const uint32_t first = (c0 >> amax) / 4;
            const uint32_t sh_row = 1 + amax; /* log2 of 16-byte pieces a tile row */
            const uint32_t total = tiles_per_chunk << (4 + sh_row);
            for (uint32_t i = lane; i < total; i += 64) {
                const uint32_t tcc = i >> (4 + sh_row), row = (i >> sh_row) & 15u;
                const uint32_t q = i & ((1u << sh_row) - 1);
                if (first + tcc >= g.n)
                    continue;
                const TileInfo ti = g.tiles[first + tcc];
                if (ti.valid != 1 || q >= (2u << ti.log2))
                    continue;
                *(uint4 *)(area + (tcc * 4 + (row >> 2)) * SS + (row & 3) * (8u << amax) + 4 * q) =
                    *(const uint4 *)(samples + (size_t)(ti.y + row) * f.coded_width + ti.x + 8 * q);
            }
            __syncthreads();
        }
        if (tl < g.n) {
            TileInfo ti = g.tiles[tl];
            if (ti.valid == 1 && b < (1u << ti.log2)) {
                /* Scan position p of block b: row p >> 4 of the stream's
                 * four, column ((p & 15) << log2) + b. */
                const uint16_t *src =
                    (const uint16_t *)((uint32_t *)g.coefs + (tc * 4 + s) * SS) + b;
                int32_t d[64];
#pragma unroll
                for (int k = 0; k < 64; k++) {
                    const uint32_t p = kInv8.v[k];
                    d[k] = dequantize((int16_t)src[(p >> 4) * W + ((p & 15u) << ti.log2)],
                                      quant_step(g.qmat[k], ti.qscale));
                }
                uint32_t v[64];
                if (FFMPEG)
                    ffmpeg_8x8(d, v);
                else
                    apple_8x8(d, v);
                const uint32_t pr = (s == 0 || s == 2) ? 1 - rr : rr;
                const uint32_t pc = (s == 1 || s == 2) ? 1 - rc : rc;
                uint16_t *dst = stage + tc * 16 * W + pr * W + 16 * b + pc;
#pragma unroll
                for (int i = 0; i < 8; i++)
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        dst[2 * i * W + 2 * j] = (uint16_t)v[8 * i + j];
            }
        }
        __syncthreads();
        /* Coalesced copy of the chunk's tiles: two sensels per lane and
         * store, code values and (fused) linear values. */
        const uint32_t first_tile = (c0 >> amax) / 4;
        const uint32_t sh_row = 3 + amax;          /* log2 of dwords per stage row */
        const uint32_t sh_tile = 4 + sh_row;       /* log2 of dwords per stage tile */
        const uint32_t total = tiles_per_chunk << sh_tile;
#pragma unroll 8
        for (uint32_t i = lane; i < total; i += 64) {
            const uint32_t tcc = i >> sh_tile;
            const uint32_t row = (i >> sh_row) & 15u, dcol = i & ((1u << sh_row) - 1);
            const uint32_t tli = first_tile + tcc;
            if (tli >= g.n)
                continue;
            TileInfo ti = g.tiles[tli];
            if (ti.valid == 2 || dcol >= (8u << ti.log2))
                continue;
            uint32_t pair = 0;
            if (ti.valid == 1)
                pair = *(const uint32_t *)(stage + tcc * 16 * W + row * W + 2 * dcol);
            const uint32_t y = ti.y + row, x = ti.x + 2 * dcol;
            *(uint32_t *)(samples + (size_t)y * f.coded_width + x) = pair;
            if (linear && y < f.height && x < f.width) {
                uint32_t l0 = linearize(g.curve, pair & 0xFFFFu);
                uint32_t l1 = linearize(g.curve, pair >> 16);
                *(uint32_t *)(linear + (size_t)y * f.width + x) = l0 | (l1 << 16);
            }
        }
        __syncthreads();
    }
}

/* Reduced decode (reduced.rs:147-273) of one tile group: streams 3, 0,
 * (1,) 2 as R, G, (G,) B; shift 1/2/3 for Half/Quarter/Eighth. */
template <uint32_t shift>
__device__ inline void decode_group_reduced(const FrameDesc &f, const TileSet &ts, uint16_t *rgb,
                                            uint8_t *flags, prr_device_status *status,
                                            uint32_t spw, uint32_t average,
                                            uint32_t *smem)
{
    const uint32_t amax = f.alignment;
    constexpr uint32_t log2_coefs = shift == 1 ? 6u : shift == 2 ? 4u : 2u;
    constexpr uint32_t per_block = 1u << log2_coefs;
    const uint32_t stride = stream_stride(per_block, amax);
    Group g = carve(smem, spw, stride);
    const uint32_t want = average ? 0xFu : 0xDu;
    front_end<false>(f, g, ts, spw, stride, log2_coefs, want, flags, status, nullptr);

    const uint32_t lane = threadIdx.x;
    const uint32_t blocks_max = 1u << amax;
    constexpr uint32_t side = 16u >> shift;         /* pixels per block side */
    const uint32_t W = side << amax;                /* stage row width in pixels */
    const uint32_t tiles_per_chunk = 64u / (4u << amax) ? 64u / (4u << amax) : 1u;
    const uint32_t out_stride = f.coded_width >> shift; /* pixels */
    const uint32_t slots = (4u * g.T) << amax;
    for (uint32_t c0 = 0; c0 < slots; c0 += 64) {
        const uint32_t slot = c0 + lane;
        const uint32_t ls = slot >> amax, b = slot & (blocks_max - 1);
        const uint32_t tl = ls >> 2, s = ls & 3;
        const uint32_t tc = tl - (c0 >> amax) / 4;
        uint16_t *const stage = chunk_stage(g, c0, amax, stride);
        if (tl < g.n && ((want >> s) & 1u)) {
            TileInfo ti = g.tiles[tl];
            if (ti.valid == 1 && b < (1u << ti.log2)) {
                const int16_t *src = g.coefs + ls * stride + b;
                uint32_t v[64];
                if constexpr (shift == 1) {
                    int32_t d[64];
#pragma unroll
                    for (int k = 0; k < 64; k++)
                        d[k] = dequantize(src[kInv8.v[k] << amax],
                                          quant_step(g.qmat[k], ti.qscale));
                    apple_8x8(d, v);
                } else if constexpr (shift == 2) {
                    int32_t d[16];
#pragma unroll
                    for (int k = 0; k < 16; k++)
                        d[k] = dequantize(src[kInv4.v[k] << amax],
                                          quant_step(g.qmat[8 * (k >> 2) + (k & 3)], ti.qscale));
                    apple_4x4(d, v);
                } else {
                    int32_t d[4];
                    const int qi[4] = {0, 1, 8, 9};
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        d[k] = dequantize(src[kInv2.v[k] << amax],
                                          quant_step(g.qmat[qi[k]], ti.qscale));
                    apple_2x2(d, v);
                }
                uint16_t *dst = stage + ((tc * 4 + s) * side) * W + b * side;
#pragma unroll
                for (uint32_t i = 0; i < side; i++)
#pragma unroll
                    for (uint32_t j = 0; j < side; j++)
                        dst[i * W + j] = (uint16_t)linearize(g.curve, v[i * side + j]);
            }
        }
        __syncthreads();
        /* Interleaved RGB rows: two values per dword store. */
        const uint32_t first_tile = (c0 >> amax) / 4;
        const uint32_t dw_per_row = (3 * W) / 2;
        for (uint32_t tr = 0; tr < tiles_per_chunk * side; tr++)
        for (uint32_t dcol = lane; dcol < dw_per_row; dcol += 64) {
            const uint32_t tcc = tr / side, row = tr % side; /* side is a constant */
            const uint32_t tli = first_tile + tcc;
            if (tli >= g.n)
                continue;
            TileInfo ti = g.tiles[tli];
            const uint32_t tw = side << ti.log2;
            if (ti.valid == 2 || dcol >= (3 * tw) / 2)
                continue;
            uint32_t pair = 0;
            if (ti.valid == 1) {
                uint32_t vals[2];
                for (int h = 0; h < 2; h++) {
                    uint32_t e = 2 * dcol + h, px = e / 3, ch = e - px * 3;
                    const uint16_t *base = stage + (tcc * 4 * side + row) * W + px;
                    uint32_t v;
                    if (ch == 0) {
                        v = base[3 * side * W];
                    } else if (ch == 2) {
                        v = base[2 * side * W];
                    } else {
                        v = base[0];
                        if (average)
                            v = (v + base[side * W] + 1) >> 1;
                    }
                    vals[h] = v;
                }
                pair = vals[0] | (vals[1] << 16);
            }
            const uint32_t y = (ti.y >> shift) + row;
            const size_t at = 3 * ((size_t)y * out_stride + (ti.x >> shift)) + 2 * dcol;
            *(uint32_t *)(rgb + at) = pair;
        }
        __syncthreads();
    }
}

/* The tiles of workgroup `item` of a single-frame launch over `order`
 * (the frame's tiles, longest first): the first `solo` workgroups take one
 * tile each, so the longest chains end with the shortest tail (one tile's
 * transforms and stores), and the rest take spw / 4 tiles each. */
__device__ inline TileSet item_tiles(const uint32_t *order, uint32_t n, uint32_t solo,
                                     uint32_t spw, uint32_t item)
{
    if (item < solo)
        return TileSet{order + item, 0, 1};
    const uint32_t T = spw >> 2, start = solo + (item - solo) * T;
    return TileSet{order + start, 0, min(T, n - start)};
}

template <int FFMPEG>
__global__ __launch_bounds__(64) void k_decode(FrameDesc f, uint16_t *samples, uint16_t *linear,
                                                uint8_t *flags, prr_device_status *status,
                                                uint32_t spw, const uint32_t *order,
                                                uint32_t solo)
{
    extern __shared__ uint32_t smem[];
    const TileSet ts = item_tiles(order, f.nb_tiles, solo, spw, blockIdx.x);
    decode_group_full<FFMPEG>(f, ts, samples, linear, flags, status, spw, smem);
}

template <uint32_t shift>
__global__ __launch_bounds__(64) void k_reduced(FrameDesc f, uint16_t *rgb, uint8_t *flags,
                                                 prr_device_status *status, uint32_t spw,
                                                 uint32_t average, const uint32_t *order,
                                                 uint32_t solo)
{
    extern __shared__ uint32_t smem[];
    const TileSet ts = item_tiles(order, f.nb_tiles, solo, spw, blockIdx.x);
    decode_group_reduced<shift>(f, ts, rgb, flags, status, spw, average, smem);
}

/* Frame::linearize_into (decoder.rs:197-216), two values per lane. */
__global__ __launch_bounds__(256) void k_linearize(const uint16_t *samples, uint32_t stride,
                                                   uint32_t width, uint32_t height,
                                                   uint4 curve_lo, uint4 curve_hi, uint16_t *out)
{
    __shared__ uint32_t c[9];
    if (threadIdx.x < 9) {
        uint32_t w[8] = {curve_lo.x, curve_lo.y, curve_lo.z, curve_lo.w,
                         curve_hi.x, curve_hi.y, curve_hi.z, curve_hi.w};
        c[threadIdx.x] = threadIdx.x < 8 ? w[threadIdx.x] : 65536u;
    }
    __syncthreads();
    const uint32_t pairs_per_row = width / 2;
    const size_t total = (size_t)pairs_per_row * height;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        uint32_t y = (uint32_t)(i / pairs_per_row), xp = (uint32_t)(i - (size_t)y * pairs_per_row);
        uint32_t pair = *(const uint32_t *)(samples + (size_t)y * stride + 2 * xp);
        uint32_t l0 = linearize(c, pair & 0xFFFFu), l1 = linearize(c, pair >> 16);
        *(uint32_t *)(out + (size_t)y * width + 2 * xp) = l0 | (l1 << 16);
    }
}

/* ------------------------------------------------------------------ */
/* Develop (demosaic.rs). */

/* variance (demosaic.rs:423-426). */
__device__ inline float variance(float v, float off, float n0, float n1, float n2)
{
    float t = off + v;
    return n2 + fmaxf(n1, n0 * (t * t));
}

/* dot4 (demosaic.rs:431-433). */
__device__ inline float dot4(float a0, float a1, float a2, float a3, float b0, float b1, float b2,
                             float b3)
{
    return fmaf(a3, b3, fmaf(a2, b2, fmaf(a0, b0, a1 * b1)));
}

/* mirror (demosaic.rs:547-557), as selects (the branches of the crate's
 * if / else-if chain, in order) so the four values stay in registers. */
__device__ inline void mirror4(float *d, uint32_t pos, uint32_t len, float c)
{
    const bool p0 = pos == 0;
    const bool p1 = !p0 && pos == 1;
    const bool pl = !p0 && !p1 && pos == len - 1;
    const bool pl2 = !p0 && !p1 && !pl && pos == len - 2;
    const float d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3];
    d[0] = p0 ? d1 : d0;
    d[1] = p0 ? d1 : pl ? d0 : d1;
    d[2] = p0 ? d3 : p1 ? c : d2;
    d[3] = p0 ? d3 : pl ? d2 : pl2 ? c : d3;
}

/* n / d correctly rounded, as the general division sequence computes it
 * (v_div_scale of both operands, v_rcp, five FMAs, v_div_fmas,
 * v_div_fixup) without the three steps that only act outside these bounds:
 * for 2^-103 <= n, 2^-43 <= d <= 2^12, n / d >= 2^-126 and n < 2^96 d,
 * v_div_scale returns its operand unchanged with no scale flag, v_div_fmas
 * is a plain FMA and v_div_fixup passes the quotient through, so the
 * remaining operations, in the same order, give the same bits. */
__device__ __forceinline__ float div_in_range(float n, float d)
{
    const float r = __builtin_amdgcn_rcpf(d);
    const float r1 = fmaf(fmaf(-d, r, 1.0f), r, r);
    const float q = n * r1;
    const float q1 = fmaf(fmaf(-d, q, n), r1, q);
    return fmaf(fmaf(-d, q1, n), r1, q1);
}

/* The reciprocal refinement of div_in_range: it depends on the divisor
 * alone, so a divisor shared by several quotients is refined once. */
__device__ __forceinline__ float refined_rcp(float d)
{
    const float r = __builtin_amdgcn_rcpf(d);
    return fmaf(fmaf(-d, r, 1.0f), r, r);
}

/* div_in_range(n, d) given r1 = refined_rcp(d): the same operations on the
 * same values, so the same bits. */
__device__ __forceinline__ float div_in_range_r(float n, float d, float r1)
{
    const float q = n * r1;
    const float q1 = fmaf(fmaf(-d, q, n), r1, q);
    return fmaf(fmaf(-d, q1, n), r1, q1);
}

/* sqrt(x) correctly rounded, as the general square root sequence computes
 * it (scale below 2^-96, v_sqrt, the two one-ulp corrections by FMA
 * residuals, unscale, pass zeros and +inf through) without the steps that
 * only act on x < 2^-96, zeros and +inf: for x >= 2^-96 the remaining
 * operations, in the same order, give the same bits. */
__device__ __forceinline__ float sqrt_in_range(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float down = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    const float r = fmaf(-down, s, x) <= 0.0f ? down : s;
    return fmaf(-up, s, x) > 0.0f ? up : r;
}

/* estimate (demosaic.rs:572-604) with sx[i] = sqrt(variance(xx[i])) and
 * sg[i] = sqrt(variance(gg[i])) already known (the same operations on the
 * same values, so the same bits). IN_RANGE: the frame meets
 * chroma_in_range, and the lo / mid / hi quotients use div_in_range. */
template <bool IN_RANGE>
__device__ inline float estimate_pre(const float *xx, const float *gg, const float *sx,
                                     const float *sg, const float4 *rc, float cg, float cs,
                                     float off, float n0, float n1, float n2)
{
    auto div = [](float n, float d) { return IN_RANGE ? div_in_range(n, d) : n / d; };
    float mean_x = dot4(xx[0], xx[1], xx[2], xx[3], 0.25f, 0.25f, 0.25f, 0.25f);
    if (!(cs < cg))
        return mean_x;
    float dx[4], dg[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        dx[i] = xx[i] - sx[i];
        dg[i] = gg[i] - sg[i];
    }
    if (IN_RANGE) {
        /* Every operand is finite here, and for finite a, b the IEEE
         * difference a - b (with gradual underflow) is positive exactly when
         * b < a: the eight tests are one minimum, over the differences the
         * quotients below need anyway. */
        const float m = fminf(fminf(fminf(dx[0], dx[1]), fminf(dx[2], dx[3])),
                              fminf(fminf(dg[0], dg[1]), fminf(dg[2], dg[3])));
        if (!(m > 0.0f))
            return mean_x;
    } else {
        bool above = true;
#pragma unroll
        for (int i = 0; i < 4; i++)
            above = above & (sx[i] < xx[i]) & (sg[i] < gg[i]);
        if (!above)
            return mean_x;
    }
    float var_x[4], lo[4], mid[4], hi[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (IN_RANGE) {
            /* rc[i] = refined reciprocals of gg, sg + gg, gg - sg (x, y, z)
             * and variance(xx, off) (w). */
            var_x[i] = rc[i].w;
            lo[i] = div_in_range_r((cg - cs) * dx[i], sg[i] + gg[i], rc[i].y);
            mid[i] = div_in_range_r(cg * xx[i], gg[i], rc[i].x);
            hi[i] = div_in_range_r((cs + cg) * (sx[i] + xx[i]), dg[i], rc[i].z);
        } else {
            var_x[i] = variance(xx[i], off, n0, n1, n2);
            lo[i] = div((cg - cs) * dx[i], sg[i] + gg[i]);
            mid[i] = div(cg * xx[i], gg[i]);
            hi[i] = div((cs + cg) * (sx[i] + xx[i]), dg[i]);
        }
    }
    float spread = dot4(mid[0] - lo[0], mid[1] - lo[1], mid[2] - lo[2], mid[3] - lo[3],
                        hi[0] - mid[0], hi[1] - mid[1], hi[2] - mid[2], hi[3] - mid[3]);
    float mean_mid = dot4(mid[0], mid[1], mid[2], mid[3], 0.25f, 0.25f, 0.25f, 0.25f);
    float dm0 = mid[0] - mean_mid, dm1 = mid[1] - mean_mid, dm2 = mid[2] - mean_mid,
          dm3 = mid[3] - mean_mid;
    float b = fmaf(spread, 0.75f, dot4(dm0, dm1, dm2, dm3, dm0, dm1, dm2, dm3));
    float dx0 = xx[0] - mean_x, dx1 = xx[1] - mean_x, dx2 = xx[2] - mean_x, dx3 = xx[3] - mean_x;
    float a = fmaf(dot4(var_x[0], var_x[1], var_x[2], var_x[3], 1.0f, 1.0f, 1.0f, 1.0f), 0.75f,
                   dot4(dx0, dx1, dx2, dx3, dx0, dx1, dx2, dx3));
    float weight = a / (b + a);
    return fmaf(mean_mid - mean_x, weight, mean_x);
}

/* Develop tiles of 64 x 16 pixels: a tile's LDS (27 KB) lets six workgroups
 * share a CU. 64 x 32 tiles (49 KB, three a CU) recompute less halo but
 * ran 10 % slower; 64 x 8 tiles recompute more halo than occupancy repays. */
enum { DTW = 64, DTH = 16, NORM_W = DTW + 6, NORM_H = DTH + 6, GRN_W = DTW + 4, GRN_H = DTH + 2 };

static_assert(DTW == DEVELOP_TILE_W && DTH == DEVELOP_TILE_H, "the host's tile size");

/* Lanes of a develop workgroup: four waves, six workgroups a CU. 512 lanes
 * (eight waves, four workgroups a CU: more waves a SIMD, half the work a
 * lane) ran 8 % slower. */
constexpr uint32_t DEV_LANES = 256;
static_assert(DEV_LANES % 64 == 0 && (DEV_LANES / 64) % 2 == 0 && DTH % (DEV_LANES / 32) == 0,
              "rows a pass are even (one Bayer phase a lane) and divide the tile");

/* LDS of one develop tile: normalized values, and beside each the noise
 * deviation it has as a green-pass neighbour (sqrt(variance) at the green
 * pass offset of the value clipped at 1 for a green sensel, at its row's
 * clip for a red or blue one), computed once per sensel instead of once per
 * use. The green-pass cells of a row alternate red/blue and green, so each
 * kind has its own array indexed by column / 2: a red/blue cell holds its
 * value, its green and the two deviations the chroma estimates reading it
 * need (sqrt(variance(rb)) and sqrt(variance(g)) at the chroma offset); a
 * green cell holds its value and sqrt(variance(g)). */
struct DevelopLds {
    union {
        struct {
            float norm[NORM_H][NORM_W];
            float dev[NORM_H][NORM_W];
        } in;
        /* After the green pass (the normalized values are dead by then),
         * on in-range frames: each red/blue cell's refined reciprocals of the
         * three chroma denominators it is a neighbour with (green, green +
         * its deviation, green - its deviation) and the variance of its
         * value at the chroma offset. */
        float4 rcp[GRN_H][GRN_W / 2];
    } u;
    float4 rb[GRN_H][GRN_W / 2];
    float2 gr[GRN_H][GRN_W / 2];
    uint32_t curve[9];
};

/* The noise model of DemosaicParams::with_white_balance (demosaic.rs:86-101)
 * from the frame's quantizer, in f64 as the crate computes it. */
__device__ inline void noise_from_quantizer(uint32_t quantizer, uint32_t range, float *noise)
{
    const double q = (double)quantizer;
    const double base = q * q * 5.0 + 64.0 / 3.0;
    const double range2 = (double)range * (double)range;
    noise[0] = (float)(base / 150994944.0);
    noise[1] = (float)(base / (range2 * 256.0));
    noise[2] = (float)(64.0 / 3.0 / range2);
}

/* Per-frame constants of the develop passes. */
struct DevK {
    float n0, n1, n2; /* noise model */
    float bs;         /* black in normalized units: the green pass offset */
    float coff;       /* black_over_range: the chroma pass offset */
    uint32_t red_row, red_col;
    uint32_t w, h;
    bool in_range; /* chroma_in_range */
};

__device__ __forceinline__ DevK dev_consts(const DevelopDesc &d)
{
    float noise[3] = {d.noise[0], d.noise[1], d.noise[2]};
    if (d.qstatus)
        noise_from_quantizer(d.qmat0 * d.qstatus->max_qscale, d.range, noise);
    DevK k;
    k.n0 = noise[0];
    k.n1 = noise[1];
    k.n2 = noise[2];
    k.bs = (float)d.black * d.one_over_range;
    k.coff = d.black_over_range;
    k.red_row = (uint32_t)(d.cfa_phase >> 1);
    k.red_col = (uint32_t)(d.cfa_phase & 1);
    k.w = d.width;
    k.h = d.height;
    /* Whether every lo / mid / hi quotient of the chroma estimates meets
     * div_in_range's bounds (and every variance sqrt_in_range's). They are taken only when cs < cg, sx < xx and
     * sg < gg, with cs, sx, sg = sqrt(variance) >= sqrt(n2 + n1) >= 2^-20;
     * so every difference cg - cs, xx - sx, gg - sg is >= 2^-43 (an exact
     * difference of floats >= 2^-20, or more than half the larger one). A
     * normalized value is one_over_range * (linear - black), below 2^9 in
     * magnitude when one_over_range <= 2^-8 and |black| <= 2^16; a red/blue
     * value xx <= 2^9 and a green estimate cg, gg (between two halves of
     * neighbour sums of values clipped at 1) < 2^11. Then numerators lie in
     * [2^-86, 2^22], denominators in [2^-43, 2^12], quotients in [2^-98,
     * 2^65]. */
    k.in_range = k.n1 + k.n2 >= 0x1p-40f && d.one_over_range > 0.0f &&
                 d.one_over_range <= 0x1p-8f && d.black >= -65536 && d.black <= 65536 &&
                 d.clip[0] == d.clip[0] && d.clip[1] == d.clip[1];
    return k;
}

/* sqrt(variance(v)) at offset `off`. Every variance is at least n2 + n1
 * (variance's max is at least n1), so on an in-range frame at least 2^-40
 * and sqrt_in_range applies. */
template <bool IN_RANGE>
__device__ __forceinline__ float deviation(float v, float off, const DevK &k)
{
    const float var = variance(v, off, k.n0, k.n1, k.n2);
    return IN_RANGE ? sqrt_in_range(var) : sqrtf(var);
}

/* Normalized value of a code value (Conversion::new folded table,
 * demosaic.rs:293-298), or of a linearized value when the input is the
 * linearized picture. */
__device__ __forceinline__ float norm_of(const DevelopDesc &d, const uint32_t *curve, uint32_t v)
{
    const int32_t lin = d.linear_input ? (int32_t)v : (int32_t)linearize(curve, v);
    return d.one_over_range * (float)(lin - d.black);
}

/* min(v, c) for a normalized value v. On an in-range frame v is finite and
 * the clips are not NaN, and v_med3(v, -inf, c) returns min(v, c) in one
 * instruction, where fminf of a value read from LDS first quiets it with a
 * second one. The -inf is opaque to the compiler, which would otherwise
 * fold the median back into that fminf. */
template <bool IN_RANGE>
__device__ __forceinline__ float clip_norm(float v, float c)
{
    if constexpr (IN_RANGE) {
        float lo = -__builtin_inff();
        asm("" : "+s"(lo));
        return __builtin_amdgcn_fmed3f(v, lo, c);
    } else {
        return fminf(v, c);
    }
}

/* Green pass (green_row_impl, demosaic.rs:481-541) of one red/blue sensel
 * (gx, gy): `norm(y, x)` reads normalized values and `dev(y, x)` their
 * deviations as neighbours at picture coordinates (rows and columns outside
 * the picture hold clamped copies, which the mirroring replaces; it can
 * only act when edge_x / edge_y say the tile reaches within two sensels of
 * the picture's side). Returns the cell (rb, green, sqrt(variance(rb)),
 * sqrt(variance(green))) at the chroma offset. */
template <bool IN_RANGE, class Norm, class Dev>
__device__ __forceinline__ float4 green_rb(const DevK &k, int32_t gx, int32_t gy, float clip,
                                           Norm norm, Dev dev, bool edge_x, bool edge_y)
{
    const float n0 = k.n0, n1 = k.n1, n2 = k.n2, bs = k.bs;
    const float centre = norm(gy, gx);
    const float c = clip_norm<IN_RANGE>(centre, clip);
    const float sc = dev(gy, gx);
    float hh[4] = {norm(gy, gx - 1), norm(gy, gx + 1), norm(gy, gx - 2), norm(gy, gx + 2)};
    float vv[4] = {norm(gy - 1, gx), norm(gy + 1, gx), norm(gy - 2, gx), norm(gy + 2, gx)};
    float sh[4] = {dev(gy, gx - 1), dev(gy, gx + 1), dev(gy, gx - 2), dev(gy, gx + 2)};
    float sv[4] = {dev(gy - 1, gx), dev(gy + 1, gx), dev(gy - 2, gx), dev(gy + 2, gx)};
    hh[0] = clip_norm<IN_RANGE>(hh[0], 1.0f);
    hh[1] = clip_norm<IN_RANGE>(hh[1], 1.0f);
    hh[2] = clip_norm<IN_RANGE>(hh[2], clip);
    hh[3] = clip_norm<IN_RANGE>(hh[3], clip);
    vv[0] = clip_norm<IN_RANGE>(vv[0], 1.0f);
    vv[1] = clip_norm<IN_RANGE>(vv[1], 1.0f);
    vv[2] = clip_norm<IN_RANGE>(vv[2], clip);
    vv[3] = clip_norm<IN_RANGE>(vv[3], clip);
    if (edge_x) {
        mirror4(hh, (uint32_t)gx, k.w, c);
        mirror4(sh, (uint32_t)gx, k.w, sc);
    }
    if (edge_y) {
        mirror4(vv, (uint32_t)gy, k.h, c);
        mirror4(sv, (uint32_t)gy, k.h, sc);
    }
    float var_h[2], var_v[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        var_h[q] = variance(hh[q], bs, n0, n1, n2);
        var_v[q] = variance(vv[q], bs, n0, n1, n2);
    }
    const float sums0 = hh[0] + hh[1], sums1 = vv[0] + vv[1];
    const float sums2 = hh[2] + hh[3], sums3 = vv[2] + vv[3];
    float ratio0 = 0.0f, ratio1 = 0.0f;
    if (c > sc) {
        const float t0 = c * 2.0f + sums2, t1 = c * 2.0f + sums3;
        const float t2 = c * -2.0f + sums2, t3 = c * -2.0f + sums3;
#pragma unroll
        for (int dir = 0; dir < 2; dir++) {
            const float d0 = dir ? vv[0] : hh[0], d1 = dir ? vv[1] : hh[1];
            const float d2 = dir ? vv[2] : hh[2], d3 = dir ? vv[3] : hh[3];
            const float sd0 = dir ? sv[0] : sh[0], sd1 = dir ? sv[1] : sh[1];
            const float sd2 = dir ? sv[2] : sh[2], sd3 = dir ? sv[3] : sh[3];
            float a = d0 - d1, b = d2 - d3;
            bool any = a * b > 0.0f || (d2 - c) * (d3 - c) > 0.0f || sd0 + sd1 > fabsf(a) ||
                       sd2 + sd3 > fabsf(b);
            bool all = d0 > sd0 && d1 > sd1 && d2 > sd2 && d3 > sd3;
            if (any && all) {
                if (dir == 0)
                    ratio0 = t2 / (t0 * 2.0f);
                else
                    ratio1 = t3 / (t1 * 2.0f);
            }
        }
    }
    const float gh = sums0 * 0.5f * (1.0f - ratio0);
    const float gv = sums1 * 0.5f * (1.0f - ratio1);
    const float dh0 = hh[0] - gh, dh1 = hh[1] - gh, dh2 = hh[2] - c, dh3 = hh[3] - c;
    const float dv0 = vv[0] - gv, dv1 = vv[1] - gv, dv2 = vv[2] - c, dv3 = vv[3] - c;
    const float eh = dot4(dh0, dh1, dh2, dh3, dh0, dh1, dh2, dh3) + (var_h[0] + var_h[1]);
    const float ev = dot4(dv0, dv1, dv2, dv3, dv0, dv1, dv2, dv3) + (var_v[0] + var_v[1]);
    const float gout = gv + (gh - gv) * ev / (eh + ev);
    return make_float4(c, gout, deviation<IN_RANGE>(c, k.coff, k),
                       deviation<IN_RANGE>(gout, k.coff, k));
}

/* The green-pass cell of a green sensel: its own value as green, and
 * sqrt(variance(green)) at the chroma offset. */
template <bool IN_RANGE>
__device__ __forceinline__ float2 green_g(const DevK &k, float centre)
{
    return make_float2(centre, deviation<IN_RANGE>(centre, k.coff, k));
}

/* LDS cell indices (row * GRN_W / 2 + column) of one chroma pair: the rows
 * above, at and below it (mirrored at the picture's top and bottom, as
 * chroma_row_with does) and the cells of its red/blue pixel, its green
 * pixel and their left and right neighbours (mirrored at the sides). */
struct PairCells {
    uint32_t rm, r0, rp;          /* row offsets, in cells */
    uint32_t rb_x, rb_xm, rb_xp;  /* columns of the red/blue pixel's cells */
    uint32_t g_x, g_xm, g_xp;     /* columns of the green pixel's cells */
};

/* Chroma pass (chroma_row_with, demosaic.rs:648-688), white balance and
 * the output matrix (matrix_planes_impl, demosaic.rs:711-720) of the pair
 * of pixels (xb, y), (xb + 1, y), xb even, whose red/blue pixel is at xb +
 * rb_col: `pc` locates its green-pass cells in the red/blue (`rbcell`) and
 * green (`gcell`) cell arrays and the reciprocal cells (`rcpcell`).
 * res[i] is pixel xb + i. */
template <bool IN_RANGE>
__device__ __forceinline__ void chroma_pair(const DevelopDesc &d, const DevK &k, bool on_red_row,
                                            uint32_t rb_col, const PairCells &pc,
                                            const float4 *rbcell, const float2 *gcell,
                                            const float4 *rcpcell, float (&res)[2][3])
{
    const float n0 = k.n0, n1 = k.n1, n2 = k.n2, off = k.coff;
    /* The neighbour cells' refined reciprocals, read only where they were
     * computed (in-range frames). */
    auto rcp_of = [&](uint32_t i) {
        if constexpr (IN_RANGE)
            return rcpcell[i];
        else
            return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    };
    auto finish = [&](float rr, float gval, float bb, float *out) {
        float r = fminf(rr * d.wb[0], 1.0f);
        float g = fminf(gval, 1.0f);
        float b = fminf(bb * d.wb[1], 1.0f);
        if (d.has_matrix) {
            const float r2 = fmaf(b, d.m[2], fmaf(r, d.m[0], g * d.m[1]));
            const float g2 = fmaf(b, d.m[5], fmaf(r, d.m[3], g * d.m[4]));
            const float b2 = fmaf(b, d.m[8], fmaf(r, d.m[6], g * d.m[7]));
            r = r2;
            g = g2;
            b = b2;
        }
        out[0] = r;
        out[1] = g;
        out[2] = b;
    };
    float res_rb[3], res_g[3];
    {
        /* The red or blue pixel: diagonal neighbours. */
        const float4 mid = rbcell[pc.r0 + pc.rb_x];
        const float4 a = rbcell[pc.rm + pc.rb_xm], b = rbcell[pc.rp + pc.rb_xm];
        const float4 c2 = rbcell[pc.rm + pc.rb_xp], e = rbcell[pc.rp + pc.rb_xp];
        const float xx[4] = {a.x, b.x, c2.x, e.x};
        const float gg[4] = {a.y, b.y, c2.y, e.y};
        const float sx[4] = {a.z, b.z, c2.z, e.z};
        const float sg[4] = {a.w, b.w, c2.w, e.w};
        const float4 rc[4] = {rcp_of(pc.rm + pc.rb_xm), rcp_of(pc.rp + pc.rb_xm),
                              rcp_of(pc.rm + pc.rb_xp), rcp_of(pc.rp + pc.rb_xp)};
        const float est =
            estimate_pre<IN_RANGE>(xx, gg, sx, sg, rc, mid.y, mid.w, off, n0, n1, n2);
        finish(on_red_row ? mid.x : est, mid.y, on_red_row ? est : mid.x, res_rb);
    }
    {
        /* The green pixel: horizontal and vertical neighbour pairs (red or
         * blue cells). */
        const float2 gm = gcell[pc.r0 + pc.g_x];
        const float4 mid = make_float4(0.0f, gm.x, 0.0f, gm.y);
        const float4 l = rbcell[pc.r0 + pc.g_xm], r = rbcell[pc.r0 + pc.g_xp];
        const float4 u = rbcell[pc.rm + pc.g_x], dn = rbcell[pc.rp + pc.g_x];
        const float hx[4] = {l.x, r.x, l.x, r.x}, hg[4] = {l.y, r.y, l.y, r.y};
        const float hsx[4] = {l.z, r.z, l.z, r.z}, hsg[4] = {l.w, r.w, l.w, r.w};
        const float vx[4] = {u.x, dn.x, u.x, dn.x}, vg[4] = {u.y, dn.y, u.y, dn.y};
        const float vsx[4] = {u.z, dn.z, u.z, dn.z}, vsg[4] = {u.w, dn.w, u.w, dn.w};
        const float4 rl = rcp_of(pc.r0 + pc.g_xm), rr = rcp_of(pc.r0 + pc.g_xp);
        const float4 ru = rcp_of(pc.rm + pc.g_x), rd = rcp_of(pc.rp + pc.g_x);
        const float4 hrc[4] = {rl, rr, rl, rr}, vrc[4] = {ru, rd, ru, rd};
        const float eh =
            estimate_pre<IN_RANGE>(hx, hg, hsx, hsg, hrc, mid.y, mid.w, off, n0, n1, n2);
        const float ev =
            estimate_pre<IN_RANGE>(vx, vg, vsx, vsg, vrc, mid.y, mid.w, off, n0, n1, n2);
        finish(on_red_row ? eh : ev, mid.y, on_red_row ? ev : eh, res_g);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        res[0][c] = rb_col == 0 ? res_rb[c] : res_g[c];
        res[1][c] = rb_col == 0 ? res_g[c] : res_rb[c];
    }
}

/* Store the pair (xb, xb + 1) at `at` (the index of xb): planar with
 * float2 stores, or interleaved RGB when out_g is NULL. */
typedef float f32x2 __attribute__((ext_vector_type(2)));

/* The planes are written once and not read again by the kernel: streaming
 * (non-temporal) stores keep them from filling the L2 with dirty lines the
 * end of the launch would have to write back. */
__device__ __forceinline__ void store_pair(float *out_r, float *out_g, float *out_b, size_t at,
                                           uint32_t xb, uint32_t w, const float (&res)[2][3])
{
    if (out_g) {
        if (xb + 1 < w) {
            __builtin_nontemporal_store((f32x2){res[0][0], res[1][0]}, (f32x2 *)(out_r + at));
            __builtin_nontemporal_store((f32x2){res[0][1], res[1][1]}, (f32x2 *)(out_g + at));
            __builtin_nontemporal_store((f32x2){res[0][2], res[1][2]}, (f32x2 *)(out_b + at));
        } else if (xb < w) {
            __builtin_nontemporal_store(res[0][0], out_r + at);
            __builtin_nontemporal_store(res[0][1], out_g + at);
            __builtin_nontemporal_store(res[0][2], out_b + at);
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < 2; k++)
            if (xb + k < w) {
                __builtin_nontemporal_store(res[k][0], out_r + 3 * (at + k) + 0);
                __builtin_nontemporal_store(res[k][1], out_r + 3 * (at + k) + 1);
                __builtin_nontemporal_store(res[k][2], out_r + 3 * (at + k) + 2);
            }
    }
}

/* One 64x16 tile of output per DEV_LANES-lane workgroup: normalized code values
 * with a 3-sensel halo and green-pass cells with a 1-sensel halo live in
 * LDS; each lane handles a red/blue + green pair of sensels so every lane
 * does the same work. Output sensel (x, y) goes to (y - oy) * ostride +
 * (x - ox) of the planes (ox = oy = 0, ostride = width for a whole frame). */
template <bool IN_RANGE>
__device__ inline void develop_tile_as(const DevelopDesc &d, const DevK &k, float *out_r,
                                       float *out_g, float *out_b, uint32_t bx, uint32_t by,
                                       DevelopLds &lds, uint32_t ox, uint32_t oy, uint32_t ostride)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t w = k.w, h = k.h;
    const int32_t x0 = (int32_t)(bx * DTW), y0 = (int32_t)(by * DTH);
    if (tid < 9)
        lds.curve[tid] = tid < 8 ? d.curve[tid] : 65536u;

    /* Normalized values, rows and columns clamped to the picture, and each
     * one's deviation as a green-pass neighbour (outside the picture, by the
     * Bayer phase of the position: those entries are never used unmirrored). */
    const float clip_red = d.clip[0], clip_blue = d.clip[1];
    /* Code value at tile position (ly, lx), clamped to the picture. Rows are
     * addressed from the tile's first clamped row, so a lane's offsets are
     * 32-bit: (row - ylo) < NORM_H + 2 and the stride < 2^17 (header widths
     * are 16-bit). */
    const int32_t ylo = y0 - 3 < 0 ? 0 : y0 - 3;
    const uint8_t *rows = reinterpret_cast<const uint8_t *>(d.samples + (size_t)ylo * d.stride);
    auto code_at = [&](int32_t ly, int32_t lx) -> uint32_t {
        const int32_t py = y0 - 3 + ly, px = x0 - 3 + lx;
        const int32_t y = py < 0 ? 0 : py > (int32_t)h - 1 ? (int32_t)h - 1 : py;
        const int32_t x = px < 0 ? 0 : px > (int32_t)w - 1 ? (int32_t)w - 1 : px;
        const uint32_t off = __umul24((uint32_t)(y - ylo), d.stride) + (uint32_t)x;
        return *reinterpret_cast<const uint16_t *>(rows + 2u * off);
    };
    auto store = [&](int32_t ly, int32_t lx, uint32_t code, float cap) {
        const float v = norm_of(d, lds.curve, code);
        lds.u.in.norm[ly][lx] = v;
        lds.u.in.dev[ly][lx] = deviation<IN_RANGE>(clip_norm<IN_RANGE>(v, cap), k.bs, k);
    };
    /* The clip of a neighbour by its Bayer phase (a red/blue sensel at its
     * row's clip, a green one at 1). */
    auto cap_of = [&](int32_t ly, int32_t lx) {
        const int32_t py = y0 - 3 + ly, px = x0 - 3 + lx;
        const bool red_row = ((uint32_t)py & 1u) == k.red_row;
        const bool rb = ((uint32_t)px & 1u) == (red_row ? k.red_col : 1u - k.red_col);
        return rb ? (red_row ? clip_red : clip_blue) : 1.0f;
    };
    {
        /* The first 64 columns: lane tid keeps column tid % 64 and takes
         * rows tid / 64 + STEP j (STEP = lanes / 64, even), so its column
         * clamp and Bayer phase are fixed; the remaining columns, one
         * sensel a lane.
         * Every code value a lane needs is requested before the first is
         * used (rows past the tile read a clamped row and are not kept), so
         * a lane waits for memory once, and before the barrier that makes
         * the curve points visible, so the loads overlap it. */
        constexpr int32_t STEP = DEV_LANES / 64, ROWS = (NORM_H + STEP - 1) / STEP;
        constexpr uint32_t EXTRA = NORM_H * (NORM_W - 64);
        static_assert(NORM_W >= 64 && EXTRA <= DEV_LANES, "the halo fits one pass of the lanes");
        const int32_t lx = (int32_t)(tid & 63), ly0 = (int32_t)(tid >> 6);
        uint32_t code[ROWS];
#pragma unroll
        for (int32_t j = 0; j < ROWS; j++)
            code[j] = code_at(ly0 + STEP * j, lx);
        const uint32_t ei = tid < EXTRA ? tid : 0u;
        const int32_t ely = (int32_t)(ei / (NORM_W - 64));
        const int32_t elx = 64 + (int32_t)(ei - (uint32_t)ely * (NORM_W - 64));
        const uint32_t ecode = code_at(ely, elx);
        const float cap = cap_of(ly0, lx);
        __syncthreads();
#pragma unroll
        for (int32_t j = 0; j < ROWS; j++)
            if (ly0 + STEP * j < NORM_H)
                store(ly0 + STEP * j, lx, code[j], cap);
        if (tid < EXTRA)
            store(ely, elx, ecode, cap_of(ely, elx));
    }
    __syncthreads();

    /* Green pass. Each lane takes one (green, red/blue) pair of a row and
     * does the green sensel, then the red/blue one, so every lane of a wave
     * runs the same code whatever the Bayer phase of its row. */
    auto norm = [&](int32_t y, int32_t x) { return lds.u.in.norm[y - (y0 - 3)][x - (x0 - 3)]; };
    auto dev = [&](int32_t y, int32_t x) { return lds.u.in.dev[y - (y0 - 3)][x - (x0 - 3)]; };
    /* Green-pass cells lie in [x0 - 1, x0 + DTW] x [y0 - 1, y0 + DTH]; the
     * mirroring acts only within two sensels of the picture's sides. */
    const bool edge_x = x0 - 1 <= 1 || x0 + DTW >= (int32_t)w - 2;
    const bool edge_y = y0 - 1 <= 1 || y0 + DTH >= (int32_t)h - 2;
    auto cell_in_tile = [&](int32_t gx) {
        return gx >= x0 - 1 && gx <= x0 + DTW && gx >= 0 && gx < (int32_t)w;
    };
    /* Green-pass cells of pair column lp of row ly: the green sensel's own
     * cell and the red/blue one's (green_row_impl). */
    auto green_cells = [&](int32_t ly, int32_t lp) {
        const int32_t gy = y0 - 1 + ly;
        if (gy < 0 || gy >= (int32_t)h)
            return;
        const bool phase = ((uint32_t)gy & 1u) == k.red_row;
        const int32_t first = (int32_t)(phase ? k.red_col : 1 - k.red_col);
        const float clip = phase ? d.clip[0] : d.clip[1];
        const int32_t gxg = x0 - 2 + 2 * lp + (1 - first);
        if (cell_in_tile(gxg))
            lds.gr[ly][lp] = green_g<IN_RANGE>(k, norm(gy, gxg));
        const int32_t gx = x0 - 2 + 2 * lp + first;
        if (cell_in_tile(gx))
            lds.rb[ly][lp] = green_rb<IN_RANGE>(k, gx, gy, clip, norm, dev, edge_x, edge_y);
    };
    {
        /* Pair columns 0..31: lane tid keeps column tid % 32 over rows
         * tid / 32 + (lanes / 32) j (one Bayer phase); columns 32 and 33 go
         * to lanes 64..99, whose wave has one row fewer than the first. */
        static_assert(GRN_W / 2 == 34 && GRN_H == 18, "green-pass cells of a 64x16 tile");
        static_assert(GRN_H % (DEV_LANES / 32) == 2, "the first wave's rows are the ones with a row more");
        const int32_t lp = (int32_t)(tid & 31), ly0 = (int32_t)(tid >> 5);
#pragma unroll 1
        for (int32_t ly = ly0; ly < GRN_H; ly += DEV_LANES / 32)
            green_cells(ly, lp);
        if (tid >= 64 && tid < 64 + 2 * GRN_H) {
            const int32_t i = (int32_t)tid - 64;
            green_cells(i >> 1, 32 + (i & 1));
        }
    }
    __syncthreads();

    if constexpr (IN_RANGE) {
        /* Each red/blue cell's refined reciprocals of its three chroma
         * denominators and the variance of its value at the chroma offset,
         * once per cell instead of once per use (a cell is a neighbour of up
         * to eight pixels). Cells outside the picture hold whatever the LDS
         * held; no estimate reads them. */
        for (uint32_t i = tid; i < GRN_H * (GRN_W / 2); i += DEV_LANES) {
            const uint32_t ly = i / (GRN_W / 2), lp = i - ly * (GRN_W / 2);
            const float4 c = lds.rb[ly][lp];
            lds.u.rcp[ly][lp] = make_float4(refined_rcp(c.y), refined_rcp(c.w + c.y),
                                            refined_rcp(c.y - c.w),
                                            variance(c.x, k.coff, k.n0, k.n1, k.n2));
        }
        __syncthreads();
    }

    /* Chroma pass, white balance and output matrix: each lane takes the
     * red/blue pixel and the green pixel of one pair. Lane tid takes pair
     * column tid % 32 of rows tid / 32 + (lanes / 32) j: the column, its
     * cells and (an even number of rows apart) the Bayer phase stay the same
     * for the lane. */
    constexpr uint32_t CROWS = DEV_LANES / 32;
    static_assert(DTW / 2 == 32 && DTH % CROWS == 0, "chroma lanes cover whole rows of 32 pairs");
    constexpr uint32_t RW = GRN_W / 2; /* cells a row */
    const uint32_t lp = tid % (DTW / 2), ly0 = tid / (DTW / 2);
    const uint32_t xb = (uint32_t)x0 + 2 * lp;
    const bool on_red_row = ((uint32_t)(y0 + (int32_t)ly0) & 1u) == k.red_row;
    const uint32_t rb_col = on_red_row ? k.red_col : 1 - k.red_col;
    /* Cell column of picture column x (mirrored neighbours included). */
    auto col = [&](uint32_t x) { return (uint32_t)(((int32_t)x - (x0 - 2)) >> 1); };
    PairCells pc;
    {
        const uint32_t x = xb + rb_col;
        pc.rb_x = col(x);
        pc.rb_xm = col(x > 0 ? x - 1 : x + 1);
        pc.rb_xp = col(x < w - 1 ? x + 1 : x - 1);
        const uint32_t gxp = xb + (1 - rb_col);
        pc.g_x = col(gxp);
        pc.g_xm = col(gxp > 0 ? gxp - 1 : gxp + 1);
        pc.g_xp = col(gxp < w - 1 ? gxp + 1 : gxp - 1);
    }
    const float4 *rbcell = &lds.rb[0][0];
    const float2 *gcell = &lds.gr[0][0];
    const float4 *rcpcell = &lds.u.rcp[0][0];
#pragma unroll
    for (uint32_t j = 0; j < DTH / CROWS; j++) {
        const uint32_t y = (uint32_t)y0 + ly0 + CROWS * j;
        if (y >= h || xb >= w)
            continue;
        /* Row y is cell row ly0 + 8j + 1. */
        pc.r0 = (ly0 + CROWS * j + 1) * RW;
        pc.rm = y > 0 ? pc.r0 - RW : pc.r0 + RW;
        pc.rp = y < h - 1 ? pc.r0 + RW : pc.r0 - RW;
        float res[2][3];
        chroma_pair<IN_RANGE>(d, k, on_red_row, rb_col, pc, rbcell, gcell, rcpcell, res);
        store_pair(out_r, out_g, out_b, (size_t)(y - oy) * ostride + (xb - ox), xb, w, res);
    }
}

/* One 64x16 tile of develop: the in-range form when the frame's constants
 * allow it (the same bits either way). */
__device__ inline void develop_tile(const DevelopDesc &d, float *out_r, float *out_g,
                                    float *out_b, uint32_t bx, uint32_t by, DevelopLds &lds,
                                    uint32_t ox, uint32_t oy, uint32_t ostride)
{
    const DevK k = dev_consts(d);
    if (k.in_range)
        develop_tile_as<true>(d, k, out_r, out_g, out_b, bx, by, lds, ox, oy, ostride);
    else
        develop_tile_as<false>(d, k, out_r, out_g, out_b, bx, by, lds, ox, oy, ostride);
}

/* The develop tile of this workgroup in its grid's x-y plane. Workgroups go
 * to the 8 XCDs in turn (workgroup b to XCD b % 8); this order gives each
 * XCD one contiguous run of tiles in raster order, so the halo rows and edge
 * lines that neighbouring tiles share are read into one XCD's L2 (develop's
 * HBM reads fall from 2.9x to 1.0x the picture). A bijection whatever the
 * dispatch does. */
__device__ __forceinline__ uint32_t xcd_tile()
{
    const uint32_t nb = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t per = nb / 8, extra = nb % 8, xcd = b % 8;
    return (xcd < extra ? xcd * (per + 1) : extra * (per + 1) + (xcd - extra) * per) + b / 8;
}

__global__ __launch_bounds__(DEV_LANES) void k_develop(DevelopDesc d, float *out_r, float *out_g,
                                                 float *out_b)
{
    __shared__ DevelopLds lds;
    if (d.gate && *d.gate == 0)
        return;
    if (d.status_out && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 4)
        reinterpret_cast<uint32_t *>(d.status_out)[threadIdx.x] =
            reinterpret_cast<const uint32_t *>(d.qstatus)[threadIdx.x];
    const uint32_t t = xcd_tile();
    develop_tile(d, out_r, out_g, out_b, t % gridDim.x, t / gridDim.x, lds, 0, 0, d.width);
}

/* Plain develop of sampled 64x64 tiles (SAMPLE_PARTS develop tiles each)
 * into scratch[k] = r, g, b planes of 64x64, for the learned path's check. */
constexpr uint32_t SAMPLE_PARTS = 64 / DTH;
static_assert(DTW == 64 && 64 % DTH == 0, "a 64x64 sample is whole develop tiles");
__global__ __launch_bounds__(DEV_LANES) void k_develop_sampled(DevelopDesc d, const uint32_t *tiles,
                                                         float *scratch)
{
    __shared__ DevelopLds lds;
    const uint32_t k = blockIdx.x / SAMPLE_PARTS, part = blockIdx.x % SAMPLE_PARTS;
    const uint32_t bx = tiles[2 * k], by64 = tiles[2 * k + 1];
    const uint32_t by = SAMPLE_PARTS * by64 + part;
    if (by >= (d.height + DTH - 1) / DTH)
        return;
    float *base = scratch + (size_t)k * 3 * 4096;
    develop_tile(d, base, base + 4096, base + 8192, bx, by, lds, DTW * bx, 64 * by64, 64);
}

/* Load a struct the whole wave agrees on into scalar registers. */
template <class T>
__device__ inline T load_uniform(const T *p)
{
    static_assert(sizeof(T) % 4 == 0, "dword-sized struct");
    T out;
    const uint32_t *src = (const uint32_t *)p;
    uint32_t *dst = (uint32_t *)&out;
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; i++)
        dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return out;
}

/* ------------------------------------------------------------------ */
/* Batched decode: one persistent launch over tile groups of many frames. */

/* Bucket of tile group g (tiles g*T .. g*T+T-1) of a frame, from its
 * scanned offsets. */
__device__ inline uint32_t group_bucket(const FrameDesc &f, uint32_t T, uint32_t g)
{
    const uint32_t t0 = g * T, t1 = min(t0 + T, f.nb_tiles);
    return bytes_bucket(f.offsets[t1] - f.offsets[t0]);
}

/* The batch's size-table scans run in chunks of SCAN_CHUNK tiles, so that a
 * large frame's scan is spread over many workgroups (one CU alone cannot
 * keep enough of a frame's scattered qscale reads in flight: a 73,000-tile
 * frame took 0.27 ms in one workgroup). chunks[c] = (frame << 16) | chunk
 * index in the frame; a frame's chunks are consecutive entries. */
enum { SCAN_LANES = 256, SCAN_PER = SCAN_CHUNK / SCAN_LANES };

/* Pass 1: the coded bytes of each chunk's tiles. Chunk 0 of a frame also
 * resets the frame's status and group histogram, entry 0 the queue head. */
__global__ __launch_bounds__(SCAN_LANES) void k_batch_chunk_bytes(const BatchFrame *frames,
                                                                  const uint32_t *chunks,
                                                                  uint32_t *chunk_bytes,
                                                                  uint32_t *hist, uint32_t *head)
{
    __shared__ uint32_t part[SCAN_LANES / 64];
    const uint32_t e = chunks[blockIdx.x], fr = e >> 16, c = e & 0xFFFFu, tid = threadIdx.x;
    const BatchFrame bf = load_uniform(&frames[fr]);
    if (c == 0) {
        if (tid == 0) {
            bf.status->oob_tile = 0xFFFFFFFFu;
            bf.status->max_qscale = 0;
            bf.status->n_skipped = 0;
            bf.status->overrun = 0;
            if (blockIdx.x == 0)
                *head = 0;
        }
        if (tid < QUEUE_BUCKETS)
            hist[fr * QUEUE_BUCKETS + tid] = 0;
    }
    const uint32_t n = bf.f.nb_tiles;
    const uint8_t *tab = bf.f.packet + bf.f.table_off;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER; k++) {
        const uint32_t i = c * SCAN_CHUNK + k * SCAN_LANES + tid;
        sum += i < n ? be16_at(tab + 2 * (size_t)i) : 0u;
    }
#pragma unroll
    for (uint32_t d = 32; d > 0; d >>= 1)
        sum += (uint32_t)__shfl_xor((int)sum, (int)d);
    if ((tid & 63) == 0)
        part[tid >> 6] = sum;
    __syncthreads();
    if (tid == 0) {
        uint32_t all = 0;
#pragma unroll
        for (uint32_t w = 0; w < SCAN_LANES / 64; w++)
            all += part[w];
        chunk_bytes[blockIdx.x] = all;
    }
}

/* Pass 2: the walk of parse_frame (decoder.rs:104-116) over one chunk,
 * starting after the bytes of the frame's earlier chunks: tile offsets,
 * first tile past the packet, largest qscale byte; then the chunk's tile
 * groups (T tiles each; T divides SCAN_CHUNK, so no group spans two
 * chunks) counted per length bucket into the frame's histogram. */
__global__ __launch_bounds__(SCAN_LANES) void k_batch_scan(const BatchFrame *frames,
                                                           const uint32_t *chunks,
                                                           const uint32_t *chunk_bytes,
                                                           uint32_t *hist)
{
    __shared__ uint64_t sums[32];
    __shared__ uint32_t gbytes[SCAN_CHUNK];
    __shared__ uint32_t count[QUEUE_BUCKETS];
    const uint32_t e = chunks[blockIdx.x], fr = e >> 16, c = e & 0xFFFFu, tid = threadIdx.x;
    const BatchFrame bf = load_uniform(&frames[fr]);
    const FrameDesc &f = bf.f;
    const uint32_t n = f.nb_tiles, T = bf.spw >> 2;
    const uint32_t c0 = c * SCAN_CHUNK, m = c0 < n ? min(n - c0, (uint32_t)SCAN_CHUNK) : 0u;
    for (uint32_t g = tid; g < SCAN_CHUNK; g += SCAN_LANES)
        gbytes[g] = 0;
    if (tid < QUEUE_BUCKETS)
        count[tid] = 0;
    uint64_t before = 0;
    for (uint32_t j = tid; j < c; j += SCAN_LANES)
        before += chunk_bytes[blockIdx.x - c + j];
    uint64_t earlier;
    (void)block_exclusive(before, sums, earlier);
    const uint64_t base = (uint64_t)f.table_off + 2ull * n + earlier;

    const uint32_t per = (m + SCAN_LANES - 1) / SCAN_LANES;
    const uint32_t lo = c0 + min(tid * per, m), hi = c0 + min(tid * per + per, m);
    const uint8_t *tab = f.packet + f.table_off;
    uint32_t sz[SCAN_PER];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER; k++) {
        sz[k] = lo + k < hi ? be16_at(tab + 2 * (size_t)(lo + k)) : 0u;
        sum += sz[k];
    }
    /* The first call's reads of sums end before any lane reaches this
     * call's first barrier. */
    uint64_t total;
    const uint64_t first = base + block_exclusive(sum, sums, total);
    uint32_t q[SCAN_PER];
    {
        uint32_t rel = 0;
#pragma unroll
        for (uint32_t k = 0; k < SCAN_PER; k++) {
            q[k] = f.packet[min(first + rel + 1, f.len - 1)];
            rel += sz[k];
        }
    }
    uint32_t oob = 0xFFFFFFFFu, mq = 0, rel = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER; k++) {
        const uint32_t i = lo + k;
        const uint64_t off = first + rel;
        rel += sz[k];
        if (i < hi) {
            const uint32_t qk = off + 1 < f.len ? q[k] : 0u;
            mq = qk > mq ? qk : mq;
            if (off + sz[k] > f.len && oob == 0xFFFFFFFFu)
                oob = i;
            f.offsets[i] = off;
            atomicAdd(&gbytes[(i - c0) / T], sz[k]);
        }
    }
    if (tid == 0 && c0 + SCAN_CHUNK >= n)
        f.offsets[n] = base + total;
#pragma unroll
    for (uint32_t d = 32; d > 0; d >>= 1) {
        oob = min(oob, (uint32_t)__shfl_xor((int)oob, (int)d));
        mq = max(mq, (uint32_t)__shfl_xor((int)mq, (int)d));
    }
    if ((tid & 63) == 0) {
        if (oob != 0xFFFFFFFFu)
            atomicMin(&bf.status->oob_tile, oob);
        if (mq)
            atomicMax(&bf.status->max_qscale, mq);
    }
    __syncthreads();
    const uint32_t groups = (m + T - 1) / T;
    for (uint32_t g = tid; g < groups; g += SCAN_LANES)
        atomicAdd(&count[bytes_bucket(gbytes[g])], 1u);
    __syncthreads();
    if (tid < QUEUE_BUCKETS && count[tid])
        atomicAdd(&hist[fr * QUEUE_BUCKETS + tid], count[tid]);
}

/* Place the batch's tile groups in the queue, longest first: the groups of
 * bucket b of frame f start after every group of the buckets before b and
 * the bucket-b groups of the frames before f. One workgroup per frame;
 * entries are (frame << 20) | group. */
__global__ __launch_bounds__(1024) void k_build_queue(const BatchFrame *frames, uint32_t n_frames,
                                                      const uint32_t *hist, uint32_t *queue)
{
    __shared__ uint32_t part_all[16][QUEUE_BUCKETS], part_before[16][QUEUE_BUCKETS];
    __shared__ uint32_t base[QUEUE_BUCKETS];
    const uint32_t tid = threadIdx.x, bucket = tid & (QUEUE_BUCKETS - 1), part = tid >> 6;
    const uint32_t fr = blockIdx.x;
    uint32_t all = 0, before = 0;
    for (uint32_t i = part; i < n_frames; i += 16) {
        const uint32_t v = hist[i * QUEUE_BUCKETS + bucket];
        all += v;
        before += i < fr ? v : 0u;
    }
    part_all[part][bucket] = all;
    part_before[part][bucket] = before;
    __syncthreads();
    if (tid < QUEUE_BUCKETS) {
        uint32_t in_bucket = 0, ahead = 0;
#pragma unroll
        for (uint32_t k = 0; k < 16; k++) {
            in_bucket += part_all[k][tid];
            ahead += part_before[k][tid];
        }
        base[tid] = wave_exclusive(in_bucket) + ahead;
    }
    __syncthreads();
    const BatchFrame bf = load_uniform(&frames[fr]);
    const uint32_t T = bf.spw >> 2, groups = (bf.f.nb_tiles + T - 1) / T;
    for (uint32_t g = tid; g < groups; g += 1024)
        queue[atomicAdd(&base[group_bucket(bf.f, T, g)], 1u)] = (fr << 20) | g;
}

/* The persistent decode: every wave pulls `chunk` queue entries at a time
 * until the queue is empty, and decodes each entry's tile group in the
 * frame's mode. */
__global__ __launch_bounds__(64) void k_batch_decode(const BatchFrame *__restrict__ frames,
                                                     const uint32_t *__restrict__ queue, uint32_t n_items,
                                                     uint32_t *head, uint32_t chunk)
{
    extern __shared__ uint32_t smem[];
    __shared__ uint32_t s_base;
    for (;;) {
        if (threadIdx.x == 0)
            s_base = atomicAdd(head, chunk);
        __syncthreads();
        /* Uniform across the wave: keep it in scalar registers so the
         * frame table is read with scalar loads. */
        const uint32_t start = __builtin_amdgcn_readfirstlane(s_base);
        __syncthreads();
        if (start >= n_items)
            break;
        const uint32_t stop = min(start + chunk, n_items);
        for (uint32_t i = start; i < stop; i++) {
            const uint32_t item = __builtin_amdgcn_readfirstlane(queue[i]);
            const BatchFrame bf = load_uniform(frames + (item >> 20));
            const uint32_t T = bf.spw >> 2, t0 = (item & 0xFFFFFu) * T;
            const TileSet g{nullptr, t0, min(T, bf.f.nb_tiles - t0)};
            switch (bf.mode) {
            case BATCH_FULL_APPLE:
                decode_group_full<0>(bf.f, g, bf.samples, bf.linear, bf.flags, bf.status, bf.spw,
                                     smem);
                break;
            case BATCH_FULL_FFMPEG:
                decode_group_full<1>(bf.f, g, bf.samples, bf.linear, bf.flags, bf.status, bf.spw,
                                     smem);
                break;
            case BATCH_HALF:
                decode_group_reduced<1>(bf.f, g, bf.rgb, bf.flags, bf.status, bf.spw,
                                        bf.average, smem);
                break;
            case BATCH_QUARTER:
                decode_group_reduced<2>(bf.f, g, bf.rgb, bf.flags, bf.status, bf.spw,
                                        bf.average, smem);
                break;
            default:
                decode_group_reduced<3>(bf.f, g, bf.rgb, bf.flags, bf.status, bf.spw,
                                        bf.average, smem);
                break;
            }
        }
    }
}

/* Develop every full-decode frame of a batch: one workgroup per develop
 * tile of the batch's frames (their tiles numbered frame after frame), in
 * the XCD-wise order of k_develop, so that no workgroup is launched for
 * nothing and each XCD's run of tiles shares its halo rows in its L2. The
 * frame of tile t is the last job whose first tile is at most t (jobs are
 * in tile order). */
__global__ __launch_bounds__(DEV_LANES) void k_batch_develop(const BatchDevelop *jobs,
                                                             uint32_t n_jobs)
{
    __shared__ DevelopLds lds;
    const uint32_t t = xcd_tile();
    uint32_t lo = 0, hi = n_jobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (jobs[mid].first_tile <= t)
            lo = mid;
        else
            hi = mid;
    }
    const BatchDevelop &j = jobs[lo];
    const uint32_t k = t - j.first_tile;
    develop_tile(j.d, j.r, j.g, j.b, k % j.tiles_x, k / j.tiles_x, lds, 0, 0, j.d.width);
}

/* ------------------------------------------------------------------ */
/* Launchers. */

static uint32_t pick_spw(const DecodeLaunch &cfg, uint32_t per_block, uint32_t amax, size_t &lds)
{
    uint32_t spw = cfg.streams_per_wg;
    for (;;) {
        lds = group_lds_bytes(spw, stream_stride(per_block, amax));
        if (lds <= cfg.lds_limit || spw <= 4)
            return spw;
        spw >>= 1;
    }
}

/* Each decode launch is preceded by k_prepare: status reset, size-table
 * scan, and the frame's tiles ordered longest first into `order`
 * (nb_tiles entries). Workgroups: `solo` single tiles, then groups of
 * spw / 4. */
static uint32_t single_frame_items(uint32_t n, uint32_t spw, uint32_t solo)
{
    const uint32_t T = spw / 4;
    return solo + (n - solo + T - 1) / T;
}

hipError_t launch_decode(const FrameDesc &f, int ffmpeg, uint16_t *samples, uint16_t *linear,
                         uint8_t *flags, prr_device_status *status, uint32_t *order,
                         const DecodeLaunch &cfg, hipStream_t s)
{
    const uint32_t spw = cfg.streams_per_wg;
    const size_t lds = staged_lds_bytes(f.alignment);
    const uint32_t n = f.nb_tiles, solo = min(n, cfg.solo_tiles);
    const uint32_t items = single_frame_items(n, spw, solo);
    hipLaunchKernelGGL(k_prepare, dim3(1), dim3(1024), 0, s, f, status, order, 1u);
    if (items == 0)
        return hipGetLastError();
    const uint32_t *o = order;
    if (ffmpeg)
        hipLaunchKernelGGL(k_decode<1>, dim3(items), dim3(64), lds, s, f, samples, linear, flags,
                           status, spw, o, solo);
    else
        hipLaunchKernelGGL(k_decode<0>, dim3(items), dim3(64), lds, s, f, samples, linear, flags,
                           status, spw, o, solo);
    return hipGetLastError();
}

hipError_t launch_decode_reduced(const FrameDesc &f, int shift, int average, uint16_t *rgb,
                                 uint8_t *flags, prr_device_status *status, uint32_t *order,
                                 const DecodeLaunch &cfg, hipStream_t s)
{
    const uint32_t per_block = shift == 1 ? 64u : shift == 2 ? 16u : 4u;
    size_t lds;
    const uint32_t spw = pick_spw(cfg, per_block, f.alignment, lds);
    const uint32_t n = f.nb_tiles, solo = min(n, cfg.solo_tiles);
    const uint32_t items = single_frame_items(n, spw, solo);
    hipLaunchKernelGGL(k_prepare, dim3(1), dim3(1024), 0, s, f, status, order, 1u);
    if (items == 0)
        return hipGetLastError();
    const uint32_t *o = order;
    if (shift == 1)
        hipLaunchKernelGGL(k_reduced<1>, dim3(items), dim3(64), lds, s, f, rgb, flags, status,
                           spw, (uint32_t)average, o, solo);
    else if (shift == 2)
        hipLaunchKernelGGL(k_reduced<2>, dim3(items), dim3(64), lds, s, f, rgb, flags, status,
                           spw, (uint32_t)average, o, solo);
    else
        hipLaunchKernelGGL(k_reduced<3>, dim3(items), dim3(64), lds, s, f, rgb, flags, status,
                           spw, (uint32_t)average, o, solo);
    return hipGetLastError();
}

hipError_t launch_linearize(const uint16_t *samples, uint32_t stride, uint32_t width,
                            uint32_t height, const uint16_t *curve, uint16_t *out, hipStream_t s)
{
    size_t total = (size_t)(width / 2) * height;
    if (total == 0)
        return hipSuccess;
    uint4 lo = make_uint4(curve[0], curve[1], curve[2], curve[3]);
    uint4 hi = make_uint4(curve[4], curve[5], curve[6], curve[7]);
    size_t blocks = (total + 255) / 256;
    if (blocks > 65535)
        blocks = 65535;
    hipLaunchKernelGGL(k_linearize, dim3((uint32_t)blocks), dim3(256), 0, s, samples, stride,
                       width, height, lo, hi, out);
    return hipGetLastError();
}

hipError_t launch_develop(const DevelopDesc &d, float *r, float *g, float *b, hipStream_t s)
{
    if (d.width == 0 || d.height == 0)
        return hipSuccess;
    dim3 grid((d.width + DTW - 1) / DTW, (d.height + DTH - 1) / DTH);
    hipLaunchKernelGGL(k_develop, grid, dim3(DEV_LANES), 0, s, d, r, g, b);
    return hipGetLastError();
}

hipError_t launch_batch(const BatchFrame *d_frames, uint32_t n_frames, uint32_t n_items,
                        uint32_t *d_queue, uint32_t *d_head, uint32_t *d_hist,
                        const uint32_t *d_chunks, uint32_t n_chunks, uint32_t *d_chunk_bytes,
                        const BatchDevelop *d_jobs,
                        uint32_t n_jobs, uint32_t dev_tiles, uint32_t grid,
                        uint32_t chunk, size_t lds, hipStream_t s)
{
    if (n_frames == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_batch_chunk_bytes, dim3(n_chunks), dim3(SCAN_LANES), 0, s, d_frames,
                       d_chunks, d_chunk_bytes, d_hist, d_head);
    hipLaunchKernelGGL(k_batch_scan, dim3(n_chunks), dim3(SCAN_LANES), 0, s, d_frames, d_chunks,
                       (const uint32_t *)d_chunk_bytes, d_hist);
    if (n_items) {
        hipLaunchKernelGGL(k_build_queue, dim3(n_frames), dim3(1024), 0, s, d_frames, n_frames,
                           (const uint32_t *)d_hist, d_queue);
        hipLaunchKernelGGL(k_batch_decode, dim3(grid), dim3(64), lds, s, d_frames, d_queue,
                           n_items, d_head, chunk);
    }
    if (n_jobs && dev_tiles)
        hipLaunchKernelGGL(k_batch_develop, dim3(dev_tiles), dim3(DEV_LANES), 0, s, d_jobs, n_jobs);
    return hipGetLastError();
}

const void *batch_decode_kernel() { return (const void *)k_batch_decode; }


namespace net {
hipError_t launch_develop_sampled(const DevelopDesc &d, const uint32_t *tiles, uint32_t n_tiles,
                                  float *scratch, hipStream_t s)
{
    if (n_tiles == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_develop_sampled, dim3(SAMPLE_PARTS * n_tiles), dim3(DEV_LANES), 0, s, d, tiles,
                       scratch);
    return hipGetLastError();
}
} // namespace net

size_t decode_lds_bytes(uint32_t spw, uint32_t per_block, uint32_t amax)
{
    return group_lds_bytes(spw, stream_stride(per_block, amax));
}

size_t staged_decode_lds_bytes(uint32_t amax) { return staged_lds_bytes(amax); }

/* ------------------------------------------------------------------ */
/* Device scratch of the allocation-free decode
 * (prr_decode_frame_device_noalloc): one frame's tile offsets (nb_tiles + 1
 * uint64, written whole by k_prepare) followed by its tile order (nb_tiles
 * uint32, likewise), for frames of at most PRR_NOALLOC_MAX_TILES tiles:
 * 3 MiB. Each GPU holds its own copy from the first context created on it
 * until the process ends; every call on one GPU shares that copy, so calls
 * that overlap on one GPU race. Nothing in it has to start at zero. */
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mutex;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
} // namespace

/* The device address of `device`'s copy of `scratch`, looked up once per
 * device with that device current (the caller has made it current). On the
 * host the array's name is only a stand-in for the symbol; kernels and
 * copies take this address. */
hipError_t noalloc_scratch(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

} // namespace prr
