/* Learned demosaic for gfx950: a small residual convolutional network,
 * trained and run with hand-written HIP kernels.
 *
 * Network (half resolution, NHWC, 32 channels):
 *   x0      the RGGB mosaic packed 4 channels per 2x2 cell (R, G on the
 *           red row, G on the blue row, B) in the normalized units develop
 *           works on, red and blue clipped at the stage's clip values; the
 *           noise feature 128*sqrt(noise[0]) in channel 4; the noise-level
 *           map 128*sigma(v) of the cell's four sensels in channels 5-8,
 *           sigma(v) the noise deviation develop's tests compare v with
 *           (demosaic.rs:420-425); in channels 9-12 the green pass's
 *           curvature ratios (horizontal, vertical) at the red sensel, then
 *           at the blue one (demosaic.rs:495-532: 0 where its noise tests
 *           leave them out); in channels 13-16 whether each of the four
 *           sensels lies above its noise deviation at the chroma offset (1
 *           or 0, the test of demosaic.rs:576-587); zeros after
 *   layer 0 h0 = relu(conv(x0) + b)
 *   1..6    h = h + relu(conv(h) + b)
 *   7       out = conv(h) + b; channels 3*(2*dy+dx)+c are the colours of
 *           the cell's 2x2 sensels, added to a bilinear demosaic
 * Weights: fp32 master copy (Adam state) used as bf16 on the matrix cores
 * (v_mfma_f32_16x16x32_bf16, fp32 accumulation); activations and their
 * gradients are bf16. Weight gradients are an fp32 reduction with
 * per-workgroup partials summed in a fixed order, so training is
 * deterministic.
 *
 * The network has no counterpart in the proresraw_rs crate; it is trained
 * against the crate's develop(Stage::Demosaiced) (demosaic.rs:190-402) on
 * the normalized values of Conversion::new (demosaic.rs:275-304).
 */
#include "prr_net.h"

#include <algorithm>

namespace prr {
namespace net {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f32x2v __attribute__((ext_vector_type(2)));

/* fp32 -> bf16, round to nearest even: the compiler's conversion, one
 * v_cvt_pk_bf16_f32 for two values. On every non-NaN float it equals adding
 * 0x7FFF plus the lowest kept bit and truncating (checked over all 2^32 bit
 * patterns on gfx950, denormals and infinities included); NaN stays NaN. */
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

/* Two values, the first in the low half. */
__device__ inline uint32_t f2bf2(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2v){a, b}, bf16x2v));
}

/* The bf16 in the low / high half of a packed pair, as fp32. */
__device__ inline float bf_lo(uint32_t p) { return __uint_as_float(p << 16); }
__device__ inline float bf_hi(uint32_t p) { return __uint_as_float(p & 0xFFFF0000u); }

/* Four channels of a residual layer's output, bf16(resid + bf16(relu(v))),
 * packed (resid: the input's four bf16 channels); *r gets bf16(relu(v)). */
__device__ inline uint2 resid_relu4(uint2 resid, const float *v, uint2 *r = nullptr)
{
    const uint32_t q0 = f2bf2(fmaxf(v[0], 0.0f), fmaxf(v[1], 0.0f));
    const uint32_t q1 = f2bf2(fmaxf(v[2], 0.0f), fmaxf(v[3], 0.0f));
    if (r)
        *r = make_uint2(q0, q1);
    return make_uint2(f2bf2(bf_lo(resid.x) + bf_lo(q0), bf_hi(resid.x) + bf_hi(q0)),
                      f2bf2(bf_lo(resid.y) + bf_lo(q1), bf_hi(resid.y) + bf_hi(q1)));
}

__device__ inline uint32_t pack2(uint16_t a, uint16_t b) { return (uint32_t)a | ((uint32_t)b << 16); }

/* ------------------------------------------------------------------ */
/* Parameters -> bf16 forward weights W[l][t][co][ci] and the flipped,
 * transposed weights of the input gradient Wt[l][t][ci][co] =
 * W[l][8-t][co][ci]. */

__global__ void k_weights(const float *__restrict__ params, uint16_t *__restrict__ wf,
                          uint16_t *__restrict__ wt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= LAYERS * LAYER_W)
        return;
    const uint32_t l = i / LAYER_W, r = i - l * LAYER_W;
    const uint32_t t = r / (C * C), co = (r / C) % C, ci = r % C;
    const uint16_t v = f2bf(params[l * LAYER_PARAMS + r]);
    wf[i] = v;
    wt[l * LAYER_W + (TAPS - 1 - t) * C * C + ci * C + co] = v;
}

/* ------------------------------------------------------------------ */
/* 3x3 convolution, 32 -> 32 channels, on the matrix cores.
 *
 * One 256-lane workgroup computes 4 rows x 64 pixels x 32 output channels,
 * one row per wave. The input rows with a one-pixel halo (zero outside the
 * image) and the 9x32x32 weights are staged in LDS with a 96-byte pixel
 * stride. For each tap a wave issues 8 MFMAs: 2 output-channel tiles x 4
 * pixel tiles of 16, K = the 32 input channels. Operand layout of
 * v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)..+7] (A = weights,
 * rows = output channels), B[8(l/16)..+7][l%16] (B = input, columns =
 * pixels) and gets D[4(l/16)+i][l%16], i = 0..3. */

/* Pixels in LDS are CP bf16 channels apart (96 bytes): measured on gfx950
 * with the B-operand read pattern below (16 consecutive pixels, four 16-byte
 * quarters), a 64- or 80-byte stride costs 4 bank-conflict cycles a read,
 * 72 bytes none but unaligned 16-byte reads, 96 bytes none and the fastest. */
enum { CT_ROWS = 4, CT_COLS = 64, CP = 48 };

template <int EPI>
__global__ __launch_bounds__(256) void k_conv(const uint16_t *__restrict__ in, int H, int W,
                                              const uint16_t *__restrict__ wgt,
                                              const float *__restrict__ bias,
                                              const uint16_t *__restrict__ resid,
                                              uint16_t *__restrict__ out,
                                              uint16_t *__restrict__ aux,
                                              float *__restrict__ out_f32)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_in[CT_ROWS + 2][CT_COLS + 2][CP];
    __shared__ __attribute__((aligned(16))) uint16_t s_w[TAPS][C][CP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x0 = blockIdx.x * CT_COLS, y0 = blockIdx.y * CT_ROWS;
    const size_t img = (size_t)blockIdx.z * H * W;

    for (int i = tid; i < TAPS * C * 4; i += 256) {
        const int row = i >> 2, q = i & 3;
        const uint4 v = ((const uint4 *)(wgt + (size_t)row * C))[q];
        *(uint4 *)&s_w[row / C][row % C][8 * q] = v;
    }
    for (int i = tid; i < (CT_ROWS + 2) * (CT_COLS + 2) * 4; i += 256) {
        const int q = i & 3, p = i >> 2;
        const int ly = p / (CT_COLS + 2), lx = p - ly * (CT_COLS + 2);
        const int y = y0 + ly - 1, x = x0 + lx - 1;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (y >= 0 && y < H && x >= 0 && x < W)
            v = ((const uint4 *)(in + (img + (size_t)y * W + x) * C))[q];
        *(uint4 *)&s_in[ly][lx][8 * q] = v;
    }
    __syncthreads();

    f32x4 acc[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
            acc[mt][nt] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
    for (int t = 0; t < TAPS; t++) {
        const int ky = t / 3, kx = t % 3;
        bf16x8 a[2], b[4];
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
            a[mt] = *(const bf16x8 *)&s_w[t][16 * mt + r16][kq];
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
            b[nt] = *(const bf16x8 *)&s_in[wave + ky][16 * nt + r16 + kx][kq];
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b[nt], acc[mt][nt], 0,
                                                                      0, 0);
    }

    const int y = y0 + wave;
    if (y >= H)
        return;
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const int x = x0 + 16 * nt + r16;
        if (x >= W)
            continue;
        const size_t px = img + (size_t)y * W + x;
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            const int co = 16 * mt + (lane >> 4) * 4;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; i++)
                v[i] = acc[mt][nt][i] + (bias ? bias[co + i] : 0.0f);
            if (EPI == EPI_LINEAR_F32) {
                *(float4 *)(out_f32 + px * C + co) = make_float4(v[0], v[1], v[2], v[3]);
                continue;
            }
            uint2 o;
            if (EPI == EPI_RELU) {
                o = make_uint2(f2bf2(fmaxf(v[0], 0.0f), fmaxf(v[1], 0.0f)),
                               f2bf2(fmaxf(v[2], 0.0f), fmaxf(v[3], 0.0f)));
            } else if (EPI == EPI_RESID_RELU) {
                const uint2 rv = *(const uint2 *)(resid + px * C + co);
                uint2 rr;
                o = resid_relu4(rv, v, &rr);
                if (aux)
                    *(uint2 *)(aux + px * C + co) = rr;
            } else if (EPI == EPI_ADD) {
                const uint2 rv = *(const uint2 *)(resid + px * C + co);
                o = make_uint2(f2bf2(bf_lo(rv.x) + v[0], bf_hi(rv.x) + v[1]),
                               f2bf2(bf_lo(rv.y) + v[2], bf_hi(rv.y) + v[3]));
            } else { /* EPI_LINEAR */
                o = make_uint2(f2bf2(v[0], v[1]), f2bf2(v[2], v[3]));
            }
            *(uint2 *)(out + px * C + co) = o;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Inference convolution: the arithmetic of k_conv (the same MFMAs in the
 * same order, so the same bits), in a persistent workgroup that keeps the
 * layer's weights in registers (A operands of all 9 taps, loaded once),
 * walks a contiguous run of the frame's 4 x 64 tiles down its columns, and
 * loads the next tile's input into registers while the current one is on
 * the matrix cores.
 *   IM_HEAD  h = relu(conv(x0) + b)
 *   IM_RES   h' = h + relu(conv(h) + b) with h read from the staged tile
 *   IM_TAIL  out[px][0..11] = conv(h) + b in fp32, the 12 channels the
 *            composition reads (output-channel tile 0 only)
 * A non-zero *skip (the fallback flag) makes the launch do nothing. */

enum { IT_SLOTS = (CT_ROWS + 2) * (CT_COLS + 2) * 4, IT_PER = (IT_SLOTS + 255) / 256 };

template <int MODE>
__global__ __launch_bounds__(256) void k_conv_infer(const uint16_t *__restrict__ in, int H, int W,
                                                    const uint16_t *__restrict__ wgt,
                                                    const float *__restrict__ bias,
                                                    uint16_t *__restrict__ out,
                                                    float *__restrict__ out12,
                                                    const uint32_t *__restrict__ skip)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_in[CT_ROWS + 2][CT_COLS + 2][CP];
    if (skip && *skip)
        return;
    constexpr int MT = MODE == IM_TAIL ? 1 : 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    const int tiles_x = (W + CT_COLS - 1) / CT_COLS, tiles_y = (H + CT_ROWS - 1) / CT_ROWS;
    const int n_tiles = tiles_x * tiles_y;
    /* Tiles in column-major order, a contiguous run per workgroup: the next
     * tile is usually the one below, whose top halo rows this tile just
     * loaded (cache hits rather than HBM reads). */
    const int per = (n_tiles + (int)gridDim.x - 1) / (int)gridDim.x;
    const int first = (int)blockIdx.x * per, last = min(first + per, n_tiles);

    bf16x8 wa[TAPS][MT];
#pragma unroll
    for (int t = 0; t < TAPS; t++)
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
            wa[t][mt] = *(const bf16x8 *)(wgt + ((size_t)(t * C + 16 * mt + r16) * C + kq));
    float bv[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            bv[mt][i] = bias[16 * mt + (lane >> 4) * 4 + i];

    uint4 pre[IT_PER];
    auto fetch = [&](int tile) {
        const int tx = tile / tiles_y, ty = tile - tx * tiles_y;
        const int x0 = tx * CT_COLS, y0 = ty * CT_ROWS;
#pragma unroll
        for (int k = 0; k < IT_PER; k++) {
            const int i = tid + 256 * k;
            const int q = i & 3, p = i >> 2;
            const int ly = p / (CT_COLS + 2), lx = p - ly * (CT_COLS + 2);
            const int y = y0 + ly - 1, x = x0 + lx - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (i < IT_SLOTS && y >= 0 && y < H && x >= 0 && x < W)
                v = ((const uint4 *)(in + ((size_t)y * W + x) * C))[q];
            pre[k] = v;
        }
    };
    if (first < last)
        fetch(first);
    for (int tile = first; tile < last; tile++) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < IT_PER; k++) {
            const int i = tid + 256 * k;
            if (i < IT_SLOTS) {
                const int q = i & 3, p = i >> 2;
                const int ly = p / (CT_COLS + 2), lx = p - ly * (CT_COLS + 2);
                *(uint4 *)&s_in[ly][lx][8 * q] = pre[k];
            }
        }
        __syncthreads();
        const int tx = tile / tiles_y, ty = tile - tx * tiles_y;
        const int x0 = tx * CT_COLS, y0 = ty * CT_ROWS;
        if (tile + 1 < last)
            fetch(tile + 1);

        f32x4 acc[MT][4];
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
                acc[mt][nt] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < TAPS; t++) {
            const int ky = t / 3, kx = t % 3;
            bf16x8 b[4];
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
                b[nt] = *(const bf16x8 *)&s_in[wave + ky][16 * nt + r16 + kx][kq];
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
#pragma unroll
                for (int nt = 0; nt < 4; nt++)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[t][mt], b[nt],
                                                                          acc[mt][nt], 0, 0, 0);
        }

        const int y = y0 + wave;
        if (y >= H)
            continue;
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const int x = x0 + 16 * nt + r16;
            if (x >= W)
                continue;
            const size_t px = (size_t)y * W + x;
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                const int co = 16 * mt + (lane >> 4) * 4;
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; i++)
                    v[i] = acc[mt][nt][i] + bv[mt][i];
                if (MODE == IM_TAIL) {
                    if (co < OUT_CH)
                        *(float4 *)(out12 + px * OUT_CH + co) = make_float4(v[0], v[1], v[2], v[3]);
                    continue;
                }
                uint2 o;
                if (MODE == IM_HEAD)
                    o = make_uint2(f2bf2(fmaxf(v[0], 0.0f), fmaxf(v[1], 0.0f)),
                                   f2bf2(fmaxf(v[2], 0.0f), fmaxf(v[3], 0.0f)));
                else
                    o = resid_relu4(*(const uint2 *)&s_in[wave + 1][16 * nt + r16 + 1][co], v);
                *(uint2 *)(out + px * C + co) = o;
            }
        }
    }
}

/* Two residual layers in one pass (inference): h2 = h1 + relu(conv_B(h1) +
 * b_B) with h1 = h0 + relu(conv_A(h0) + b_A), h1 kept in LDS. A workgroup
 * takes a strip of PS_W output columns and PS_SEG rows and walks it row by
 * row: rings of RING rows hold h0 (columns x0 - 2 .. x0 + 65 of the strip)
 * and h1 (the same columns); at step t waves 0 and 1 compute h1 row t (64
 * columns, x0 - 1 .. x0 + 62, two 16-pixel tiles each) from h0 rows t - 1 ..
 * t + 1, while waves 2 and 3 compute h2 row t - 2 (x0 .. x0 + 63, of which
 * the first PS_W are kept) from h1 rows t - 3 .. t - 1, so each wave holds
 * one layer's weights in registers, and h0 row t + 2 is loaded into
 * registers meanwhile. h1 outside the picture is zero (the next layer's
 * padding). Every value is computed with k_conv_infer's operations in its
 * order (the same MFMAs, taps in order, the same epilogue), so h2 equals
 * two k_conv_infer<IM_RES> launches bit for bit, with half their HBM
 * traffic and 3 % more matrix work (64 columns computed for 62). */
enum { PS_W = 62, PS_COLS = 68, PS_SEG = 64, RING = 4, PS_SLOTS = 66 * 4 };

template <int AM, int BM>
__global__ __launch_bounds__(256) void k_conv_pair(const uint16_t *__restrict__ in, int H, int W,
                                                   const uint16_t *__restrict__ wA,
                                                   const float *__restrict__ bA,
                                                   const uint16_t *__restrict__ wB,
                                                   const float *__restrict__ bB,
                                                   uint16_t *__restrict__ out,
                                                   float *__restrict__ out12,
                                                   const uint32_t *__restrict__ skip)
{
    static_assert(AM == IM_HEAD || AM == IM_RES, "layer A: the first or a residual layer");
    static_assert(BM == IM_RES || BM == IM_TAIL, "layer B: a residual or the last layer");
    /* The last layer computes one 16-channel tile (its 12 outputs). */
    constexpr int MTB = BM == IM_TAIL ? 1 : 2;
    /* Rows in LDS quarter-major: [row][channel quarter][pixel][8 channels],
     * so the B-operand reads (16 consecutive pixels of one quarter), the
     * epilogue's 8-byte reads and writes and the row stores are all
     * consecutive 16- or 8-byte accesses (a pixel-major row with a 96-byte
     * stride gave those many bank conflicts on gfx950). */
    __shared__ __attribute__((aligned(16))) uint16_t s_x[RING][4][PS_COLS][8];
    __shared__ __attribute__((aligned(16))) uint16_t s_h[RING][4][PS_COLS][8];
    if (skip && *skip)
        return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    const int strips = (W + PS_W - 1) / PS_W;
    const int strip = (int)blockIdx.x % strips, seg = (int)blockIdx.x / strips;
    const int x0 = strip * PS_W, y0 = seg * PS_SEG, y1 = min(y0 + PS_SEG, H);
    const bool second = wave >= 2;  /* layer B */
    const int nt0 = 2 * (wave & 1); /* this wave's first 16-pixel tile */

    const uint16_t *wgt = second ? wB : wA;
    const float *bias = second ? bB : bA;
    bf16x8 wa[TAPS][2];
#pragma unroll
    for (int t = 0; t < TAPS; t++)
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
            wa[t][mt] = *(const bf16x8 *)(wgt + ((size_t)(t * C + 16 * mt + r16) * C + kq));
    float bv[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            bv[mt][i] = bias[16 * mt + (lane >> 4) * 4 + i];

    for (int i = tid; i < RING * 4 * PS_COLS; i += 256) {
        ((uint4 *)&s_x[0][0][0])[i] = make_uint4(0, 0, 0, 0);
        ((uint4 *)&s_h[0][0][0])[i] = make_uint4(0, 0, 0, 0);
    }
    /* h0 row y, columns x0 - 2 .. x0 + 63 (zero outside the picture), into
     * registers: 66 pixels x 4 16-byte quarters, two a lane; loaded a step
     * before it is stored to the ring. */
    uint4 next[2];
    auto fetch = [&](int y, uint4 (&pre)[2]) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int i = tid + 256 * k, q = i & 3, c = i >> 2, x = x0 - 2 + c;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (i < PS_SLOTS && y >= 0 && y < H && x >= 0 && x < W)
                v = ((const uint4 *)(in + ((size_t)y * W + x) * C))[q];
            pre[k] = v;
        }
    };
    auto stash = [&](int y, const uint4 (&pre)[2]) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int i = tid + 256 * k;
            if (i < PS_SLOTS)
                *(uint4 *)&s_x[y & (RING - 1)][i & 3][i >> 2][0] = pre[k];
        }
    };
    __syncthreads();
    fetch(y0 - 2, next);
    stash(y0 - 2, next);
    fetch(y0 - 1, next);
    stash(y0 - 1, next);
    fetch(y0, next);
    stash(y0, next);
    fetch(y0 + 1, next);

    /* One barrier a step: at step t, layer A reads h0 rows t - 1 .. t + 1
     * (stored by step t - 1 or before) while row t + 2 goes to the slot of
     * row t - 2, which step t - 1 was the last to read; layer B reads h1
     * rows t - 3 .. t - 1 while A writes row t. The barrier that opens step
     * t + 1 orders all of these. */
    for (int t = y0 - 1; t <= y1 + 1; t++) {
        __syncthreads();
        stash(t + 2, next);
        if (t + 3 <= y1 + 1)
            fetch(t + 3, next);
        f32x4 acc[2][2];
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int n = 0; n < 2; n++)
                acc[mt][n] = (f32x4){0.f, 0.f, 0.f, 0.f};
        if (!second) {
            if (t > y1)
                continue;
            /* h1 row t: pixel x0 - 1 + 16 n + r16 sits at ring column
             * 16 n + r16 + 1; tap (ky, kx) reads h0 column 16 n + r16 + kx. */
            /* Each tap's B fragments are read from LDS while the previous
             * tap's MFMAs run. */
            auto frag = [&](int tap, bf16x8 (&b)[2]) {
                const int ky = tap / 3, kx = tap % 3;
#pragma unroll
                for (int n = 0; n < 2; n++)
                    b[n] = *(const bf16x8 *)&s_x[(t - 1 + ky) & (RING - 1)][lane >> 4]
                                                [16 * (nt0 + n) + r16 + kx][0];
            };
            bf16x8 bc[2], bn[2];
            frag(0, bc);
#pragma unroll
            for (int tap = 0; tap < TAPS; tap++) {
                if (tap + 1 < TAPS)
                    frag(tap + 1, bn);
#pragma unroll
                for (int mt = 0; mt < 2; mt++)
#pragma unroll
                    for (int n = 0; n < 2; n++)
                        acc[mt][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[tap][mt], bc[n],
                                                                             acc[mt][n], 0, 0, 0);
#pragma unroll
                for (int n = 0; n < 2; n++)
                    bc[n] = bn[n];
            }
            const bool row_in = t >= 0 && t < H;
#pragma unroll
            for (int n = 0; n < 2; n++) {
                const int col = 16 * (nt0 + n) + r16 + 1, x = x0 - 2 + col;
                const bool in_pic = row_in && x >= 0 && x < W;
#pragma unroll
                for (int mt = 0; mt < 2; mt++) {
                    const int co = 16 * mt + (lane >> 4) * 4;
                    float v[4];
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        v[i] = acc[mt][n][i] + bv[mt][i];
                    uint2 o;
                    if (AM == IM_HEAD)
                        o = make_uint2(f2bf2(fmaxf(v[0], 0.0f), fmaxf(v[1], 0.0f)),
                                       f2bf2(fmaxf(v[2], 0.0f), fmaxf(v[3], 0.0f)));
                    else
                        o = resid_relu4(*(const uint2 *)&s_x[t & (RING - 1)][co >> 3][col][co & 7], v);
                    *(uint2 *)&s_h[t & (RING - 1)][co >> 3][col][co & 7] = in_pic ? o : make_uint2(0, 0);
                }
            }
        } else {
            const int r = t - 2;
            if (r < y0 || r >= y1)
                continue;
            /* h2 row r: pixel x0 + 16 n + r16 sits at ring column
             * 16 n + r16 + 2; tap (ky, kx) reads h1 column 16 n + r16 + kx + 1. */
            auto frag = [&](int tap, bf16x8 (&b)[2]) {
                const int ky = tap / 3, kx = tap % 3;
#pragma unroll
                for (int n = 0; n < 2; n++)
                    b[n] = *(const bf16x8 *)&s_h[(r - 1 + ky) & (RING - 1)][lane >> 4]
                                                [16 * (nt0 + n) + r16 + kx + 1][0];
            };
            bf16x8 bc[2], bn[2];
            frag(0, bc);
#pragma unroll
            for (int tap = 0; tap < TAPS; tap++) {
                if (tap + 1 < TAPS)
                    frag(tap + 1, bn);
#pragma unroll
                for (int mt = 0; mt < MTB; mt++)
#pragma unroll
                    for (int n = 0; n < 2; n++)
                        acc[mt][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[tap][mt], bc[n],
                                                                             acc[mt][n], 0, 0, 0);
#pragma unroll
                for (int n = 0; n < 2; n++)
                    bc[n] = bn[n];
            }
#pragma unroll
            for (int n = 0; n < 2; n++) {
                const int px = 16 * (nt0 + n) + r16, x = x0 + px;
                if (px >= PS_W || x >= W)
                    continue;
#pragma unroll
                for (int mt = 0; mt < MTB; mt++) {
                    const int co = 16 * mt + (lane >> 4) * 4;
                    float v[4];
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        v[i] = acc[mt][n][i] + bv[mt][i];
                    if (BM == IM_TAIL) {
                        if (co < OUT_CH)
                            *(float4 *)(out12 + ((size_t)r * W + x) * OUT_CH + co) =
                                make_float4(v[0], v[1], v[2], v[3]);
                    } else {
                        *(uint2 *)(out + ((size_t)r * W + x) * C + co) =
                            resid_relu4(*(const uint2 *)&s_h[r & (RING - 1)][co >> 3][px + 2][co & 7], v);
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Weight and bias gradients: dW[t][co][ci] = sum_p g[p][co] x[p+off(t)][ci]
 * and db[co] = sum_p g[p][co], over a tile of 8 x 32 pixels per workgroup.
 * Lanes 0..287 own one (tap, ci) each and all 32 co; lanes 288..319 one
 * bias each. Partials go to part[tile][LAYER_PARAMS]; k_reduce sums them
 * in tile order. */

enum { GT_ROWS = 8, GT_COLS = 32, GT_THREADS = 320 };

__global__ __launch_bounds__(GT_THREADS) void k_wgrad(const uint16_t *__restrict__ x,
                                                      const uint16_t *__restrict__ g, int H, int W,
                                                      float *__restrict__ part)
{
    __shared__ __attribute__((aligned(16))) float s_g[GT_ROWS * GT_COLS][C];
    __shared__ uint16_t s_x[GT_ROWS + 2][GT_COLS + 2][C];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * GT_COLS, y0 = blockIdx.y * GT_ROWS;
    const size_t img = (size_t)blockIdx.z * H * W;
    const size_t tile = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;

    for (int i = tid; i < GT_ROWS * GT_COLS * (C / 2); i += GT_THREADS) {
        const int p = i / (C / 2), c2 = i - p * (C / 2);
        const int y = y0 + p / GT_COLS, xx = x0 + p % GT_COLS;
        uint32_t v = 0;
        if (y < H && xx < W)
            v = ((const uint32_t *)(g + (img + (size_t)y * W + xx) * C))[c2];
        s_g[p][2 * c2] = bf2f((uint16_t)v);
        s_g[p][2 * c2 + 1] = bf2f((uint16_t)(v >> 16));
    }
    for (int i = tid; i < (GT_ROWS + 2) * (GT_COLS + 2) * (C / 2); i += GT_THREADS) {
        const int c2 = i % (C / 2), p = i / (C / 2);
        const int ly = p / (GT_COLS + 2), lx = p - ly * (GT_COLS + 2);
        const int y = y0 + ly - 1, xx = x0 + lx - 1;
        uint32_t v = 0;
        if (y >= 0 && y < H && xx >= 0 && xx < W)
            v = ((const uint32_t *)(x + (img + (size_t)y * W + xx) * C))[c2];
        *(uint32_t *)&s_x[ly][lx][2 * c2] = v;
    }
    __syncthreads();

    float *dst = part + tile * LAYER_PARAMS;
    if (tid < TAPS * C) {
        const int t = tid / C, ci = tid % C, ky = t / 3, kx = t % 3;
        float acc[C];
#pragma unroll
        for (int co = 0; co < C; co++)
            acc[co] = 0.0f;
        for (int p = 0; p < GT_ROWS * GT_COLS; p++) {
            const float xv = bf2f(s_x[p / GT_COLS + ky][p % GT_COLS + kx][ci]);
#pragma unroll
            for (int c4 = 0; c4 < C / 4; c4++) {
                const float4 gv = *(const float4 *)&s_g[p][4 * c4];
                acc[4 * c4 + 0] = fmaf(gv.x, xv, acc[4 * c4 + 0]);
                acc[4 * c4 + 1] = fmaf(gv.y, xv, acc[4 * c4 + 1]);
                acc[4 * c4 + 2] = fmaf(gv.z, xv, acc[4 * c4 + 2]);
                acc[4 * c4 + 3] = fmaf(gv.w, xv, acc[4 * c4 + 3]);
            }
        }
#pragma unroll
        for (int co = 0; co < C; co++)
            dst[t * C * C + co * C + ci] = acc[co];
    } else {
        const int co = tid - TAPS * C;
        float acc = 0.0f;
        for (int p = 0; p < GT_ROWS * GT_COLS; p++)
            acc += s_g[p][co];
        dst[LAYER_W + co] = acc;
    }
}

__global__ void k_reduce(const float *__restrict__ part, uint32_t tiles, float *__restrict__ grad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= LAYER_PARAMS)
        return;
    float s = 0.0f;
    for (uint32_t t = 0; t < tiles; t++)
        s += part[(size_t)t * LAYER_PARAMS + i];
    grad[i] = s;
}

/* dz = dh where the layer's ReLU output is positive, else 0. */
__global__ void k_relu_grad(const uint16_t *__restrict__ dh, const uint16_t *__restrict__ mask,
                            uint16_t *__restrict__ dz, size_t count)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const uint16_t m = mask[i];
    dz[i] = (m != 0 && !(m & 0x8000u)) ? dh[i] : (uint16_t)0;
}

/* Adam (betas 0.9 / 0.999, eps 1e-8); the step count is read on the
 * device so a captured graph replays correctly. With decay > 0 the rate
 * follows lr_min + (lr - lr_min) * (1 + cos(pi * min(step, decay) / decay)) / 2. */
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, const uint32_t *__restrict__ step, float lr0,
                       float lr_min, uint32_t decay)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= PARAMS)
        return;
    float lr = lr0;
    if (decay) {
        const double u = (double)min(*step, decay) / (double)decay;
        lr = (float)((double)lr_min +
                     ((double)lr0 - (double)lr_min) * 0.5 * (1.0 + cos(3.141592653589793 * u)));
    }
    const double t = (double)(*step + 1);
    const float c1 = (float)(1.0 - pow(0.9, t));
    const float c2 = (float)(1.0 - pow(0.999, t));
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * (gi * gi);
    m[i] = mi;
    v[i] = vi;
    const float num = lr * (mi / c1);
    const float den = sqrtf(vi / c2) + 1e-8f;
    p[i] = p[i] - num / den;
}

__global__ void k_step_inc(uint32_t *step)
{
    if (threadIdx.x == 0)
        *step += 1;
}

__global__ void k_set(uint32_t *flag, uint32_t value)
{
    if (threadIdx.x == 0)
        *flag = value;
}

__device__ inline float feature_of(const NetFrame &f);

__global__ void k_noise_gate(NetFrame f, float lo, float hi, uint32_t *use_plain)
{
    if (threadIdx.x != 0)
        return;
    const float v = feature_of(f);
    if (!(v >= lo && v <= hi))
        *use_plain = 1u;
}

/* ------------------------------------------------------------------ */
/* Mosaic access in RGGB order. */

/* Curve::linearize (curve.rs:14-24) through the 9 curve points `curve`
 * (the frame's, or a copy in LDS). */
__device__ inline uint32_t lin_at(const uint32_t *curve, uint32_t code)
{
    const uint32_t seg = code >> 13, frac = code & 0x1FFFu;
    const uint32_t c0 = curve[seg], c1 = curve[seg + 1];
    const uint32_t v = (c0 * 8192u + ((c1 - c0) & 0xFFFFu) * frac + 4096u) >> 13;
    return v < 65535u ? v : 65535u;
}

/* The code value of the sensel at RGGB-order (y, x). */
__device__ inline uint32_t code_at(const NetFrame &f, int y, int x)
{
    const uint32_t sy = f.flip_y ? f.height - 1 - (uint32_t)y : (uint32_t)y;
    const uint32_t sx = f.flip_x ? f.width - 1 - (uint32_t)x : (uint32_t)x;
    return f.samples[(size_t)sy * f.stride + sx];
}

/* Normalized value of code value `code` at RGGB-order (y, x), red and blue
 * clipped. */
__device__ inline float norm_of_code(const NetFrame &f, const uint32_t *curve, uint32_t code, int y,
                                     int x)
{
    float v = f.one_over_range * (float)((int32_t)lin_at(curve, code) - f.black);
    if (((y | x) & 1) == 0)
        v = fminf(v, f.clip_r);
    else if ((y & x & 1) == 1)
        v = fminf(v, f.clip_b);
    return v;
}

/* Normalized value of the sensel at RGGB-order (y, x), red and blue
 * clipped. `curve`: the frame's 9 curve points, in LDS (stage_curve) where
 * the frame is a kernel argument (indexing the argument at a run-time index
 * would copy it to scratch memory), or the frame's own in global memory. */
__device__ inline float norm_at(const NetFrame &f, int y, int x, const uint32_t *curve)
{
    return norm_of_code(f, curve, code_at(f, y, x), y, x);
}

/* The frame's curve points into LDS (constant indices; the caller
 * synchronizes before use). */
__device__ inline void stage_curve(const NetFrame &f, uint32_t *s_curve)
{
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < 9; j++)
            s_curve[j] = f.curve[j];
    }
}

__device__ inline int mirror(int i, int n) { return i < 0 ? -i : i >= n ? 2 * n - 2 - i : i; }

/* Bilinear demosaic at (y, x) of the window of rows/columns [oy, oy+hh) x
 * [ox, ox+ww), mirrored at the window's edges. */
__device__ inline void bilinear(const NetFrame &f, const uint32_t *curve, int oy, int ox, int hh,
                                int ww, int y, int x, float *rgb)
{
    auto v = [&](int dy, int dx) {
        return norm_at(f, oy + mirror(y + dy, hh), ox + mirror(x + dx, ww), curve);
    };
    const int py = y & 1, px = x & 1;
    const float c = v(0, 0);
    if (py == px) {
        const float cross = ((v(-1, 0) + v(1, 0)) + (v(0, -1) + v(0, 1))) * 0.25f;
        const float diag = ((v(-1, -1) + v(-1, 1)) + (v(1, -1) + v(1, 1))) * 0.25f;
        rgb[0] = py ? diag : c;
        rgb[1] = cross;
        rgb[2] = py ? c : diag;
    } else {
        const float horiz = (v(0, -1) + v(0, 1)) * 0.5f;
        const float vert = (v(-1, 0) + v(1, 0)) * 0.5f;
        rgb[0] = py ? vert : horiz;
        rgb[1] = c;
        rgb[2] = py ? horiz : vert;
    }
}

/* noise[0] of DemosaicParams::with_white_balance (demosaic.rs:86-101): from
 * the quantizer read on the device, or as given. */
__device__ inline float noise_a_of(const NetFrame &f)
{
    if (!f.qstatus)
        return f.noise_a;
    const double q = (double)(f.qmat0 * f.qstatus->max_qscale);
    return (float)((q * q * 5.0 + 64.0 / 3.0) / 150994944.0);
}

/* The noise feature 128*sqrt(noise[0]). */
__device__ inline float feature_of(const NetFrame &f)
{
    if (!f.qstatus)
        return f.feature;
    return 128.0f * sqrtf(noise_a_of(f));
}

/* The noise deviation sqrt(variance(v)) at offset `off` (demosaic.rs:420-425,
 * in its operation order). */
__device__ inline float deviation(float a, const NetFrame &f, float off, float v)
{
    const float t = off + v;
    return sqrtf(f.noise_c + fmaxf(f.noise_b, a * (t * t)));
}

/* The green pass's edge rule (mirror, demosaic.rs:547-557) on the four
 * neighbours d of a sensel at `pos` of a line of `len`, centre c. */
__device__ inline void mirror_line(float (&d)[4], int pos, int len, float c)
{
    /* The crate's if / else-if chain as selects, so the values stay in
     * registers. */
    const bool p0 = pos == 0, p1 = !p0 && pos == 1;
    const bool pl = !p0 && !p1 && pos == len - 1, pl2 = !p0 && !p1 && !pl && pos == len - 2;
    const float d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3];
    d[0] = p0 ? d1 : d0;
    d[1] = pl ? d0 : d1;
    d[2] = p0 ? d3 : p1 ? c : d2;
    d[3] = p0 ? d3 : pl ? d2 : pl2 ? c : d3;
}

/* One direction's ratio of the green pass (demosaic.rs:520-530): its four
 * neighbours d with their deviations sd, the centre c, and t[2 + dir] /
 * (t[dir] * 2) as tn / (td * 2) when the noise tests let it apply. */
__device__ inline float direction_ratio(const float (&d)[4], const float (&sd)[4], float c,
                                        float tn, float td)
{
    const float p = d[0] - d[1], q = d[2] - d[3];
    const bool any = p * q > 0.0f || (d[2] - c) * (d[3] - c) > 0.0f ||
                     sd[0] + sd[1] > fabsf(p) || sd[2] + sd[3] > fabsf(q);
    const bool all = d[0] > sd[0] && d[1] > sd[1] && d[2] > sd[2] && d[3] > sd[3];
    return any && all ? tn / (td * 2.0f) : 0.0f;
}

/* The green pass's curvature ratios at the red or blue sensel (y, x) of the
 * RGGB-order mosaic (green_row_impl, demosaic.rs:495-532): horizontal and
 * vertical, t[2 + dir] / (t[dir] * 2) where the crate's noise tests apply
 * it, else 0. Neighbours are read at clamped rows and columns, greens
 * clipped at 1 and same-colour sensels at `clip` (the row's), and mirrored
 * at the picture's edges, as the crate reads them; flips to RGGB order
 * only swap the two neighbours of each pair, which leaves every test and
 * ratio unchanged. */
__device__ inline float2 curvature_ratios(const NetFrame &f, const uint32_t *curve, float a, int y,
                                          int x, float clip)
{
    const int w = (int)f.width, h = (int)f.height;
    auto at = [&](int yy, int xx) {
        return norm_at(f, min(max(yy, 0), h - 1), min(max(xx, 0), w - 1), curve);
    };
    const float off = f.gp_off;
    const float c = fminf(at(y, x), clip);
    float hh[4] = {fminf(at(y, x - 1), 1.0f), fminf(at(y, x + 1), 1.0f),
                   fminf(at(y, x - 2), clip), fminf(at(y, x + 2), clip)};
    float vv[4] = {fminf(at(y - 1, x), 1.0f), fminf(at(y + 1, x), 1.0f),
                   fminf(at(y - 2, x), clip), fminf(at(y + 2, x), clip)};
    mirror_line(hh, x, w, c);
    mirror_line(vv, y, h, c);
    float2 ratio = make_float2(0.0f, 0.0f);
    if (c > deviation(a, f, off, c)) {
        const float sums2 = hh[2] + hh[3], sums3 = vv[2] + vv[3];
        float sh[4], sv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            sh[i] = deviation(a, f, off, hh[i]);
            sv[i] = deviation(a, f, off, vv[i]);
        }
        ratio.x = direction_ratio(hh, sh, c, c * -2.0f + sums2, c * 2.0f + sums2);
        ratio.y = direction_ratio(vv, sv, c, c * -2.0f + sums3, c * 2.0f + sums3);
    }
    return ratio;
}

/* One packed cell, 32 bf16 channels (the layout in this file's header),
 * from the normalized values of its four sensels and the curvature ratios
 * at its red and blue sensels. */
__device__ inline void write_cell(const NetFrame &f, float a, float vr, float vg1, float vg2,
                                  float vb, float2 rr, float2 rb, uint16_t *dst)
{
    const float dr = deviation(a, f, f.noise_off, vr), dg1 = deviation(a, f, f.noise_off, vg1);
    const float dg2 = deviation(a, f, f.noise_off, vg2), db = deviation(a, f, f.noise_off, vb);
    /* 1 (bf16 0x3F80) where a sensel lies above its noise deviation at the
     * chroma offset, the test the chroma estimates make (demosaic.rs:576-587). */
    auto above = [](float v, float dev) { return dev < v ? (uint16_t)0x3F80u : (uint16_t)0u; };
    uint4 *d = (uint4 *)dst;
    d[0] = make_uint4(f2bf2(vr, vg1), f2bf2(vg2, vb), f2bf2(feature_of(f), 128.0f * dr),
                      f2bf2(128.0f * dg1, 128.0f * dg2));
    d[1] = make_uint4(f2bf2(128.0f * db, rr.x), f2bf2(rr.y, rb.x),
                      pack2(f2bf(rb.y), above(vr, dr)), pack2(above(vg1, dg1), above(vg2, dg2)));
    d[2] = make_uint4(above(vb, db), 0, 0, 0);
    d[3] = make_uint4(0, 0, 0, 0);
}

/* One packed cell of the RGGB-order mosaic at (y, x), every value read from
 * the frame (the crops and check windows). */
__device__ inline void pack_cell(const NetFrame &f, const uint32_t *curve, int y, int x,
                                 uint16_t *dst)
{
    const float vr = norm_at(f, y, x, curve), vg1 = norm_at(f, y, x + 1, curve);
    const float vg2 = norm_at(f, y + 1, x, curve), vb = norm_at(f, y + 1, x + 1, curve);
    const float a = noise_a_of(f);
    const float2 rr = curvature_ratios(f, curve, a, y, x, f.clip_r);
    const float2 rb = curvature_ratios(f, curve, a, y + 1, x + 1, f.clip_b);
    write_cell(f, a, vr, vg1, vg2, vb, rr, rb, dst);
}

__global__ void k_pack_crops(const NetFrame *__restrict__ frames, const Crop *__restrict__ crops,
                             const uint32_t *__restrict__ step, uint32_t steps, int n, int h2,
                             int w2, uint16_t *__restrict__ x0)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t cells = (size_t)h2 * w2;
    if (i >= cells * n)
        return;
    const int k = (int)(i / cells), c = (int)(i - (size_t)k * cells);
    const Crop cr = crops[(size_t)(*step % steps) * n + k];
    const NetFrame &f = frames[cr.frame];
    const int ci = c / w2, cj = c - ci * w2;
    pack_cell(f, f.curve, (int)cr.y0 + 2 * ci, (int)cr.x0 + 2 * cj, x0 + i * C);
}

/* The whole frame, one 256-lane workgroup per tile of PT_W x PT_H cells:
 * each sensel of the tile and of a 2-sensel halo is normalized once into
 * LDS with its value as a green-pass neighbour (greens capped at 1) and that
 * value's deviation, and each lane assembles one cell from them. The
 * operations on each value are pack_cell's, so the channels are the same
 * bits (pack_cell packs the crops and check windows). Halo positions
 * outside the picture hold the clamped sensel; the edge rule replaces every
 * such neighbour before a test reads it. */
enum { PT_W = 32, PT_H = 8, PT_SW = 2 * PT_W + 4, PT_SH = 2 * PT_H + 4 };

__global__ __launch_bounds__(256) void k_pack_frame(NetFrame f, uint16_t *__restrict__ x0,
                                                    const uint32_t *__restrict__ skip)
{
    __shared__ float s_raw[PT_SH][PT_SW], s_lim[PT_SH][PT_SW], s_dev[PT_SH][PT_SW];
    __shared__ uint32_t s_curve[9];
    if (skip && *skip)
        return;
    const int w = (int)f.width, h = (int)f.height, w2 = w / 2, h2 = h / 2;
    const int cx0 = (int)blockIdx.x * PT_W, cy0 = (int)blockIdx.y * PT_H;
    const int sx0 = 2 * cx0 - 2, sy0 = 2 * cy0 - 2;
    stage_curve(f, s_curve);
    /* Every code value a lane needs is loaded before any is used. */
    constexpr int N = PT_SH * PT_SW, PER = (N + 255) / 256;
    uint32_t code[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = (int)threadIdx.x + 256 * k;
        const int ly = i / PT_SW, lx = i - ly * PT_SW;
        const int y = min(max(sy0 + ly, 0), h - 1), x = min(max(sx0 + lx, 0), w - 1);
        code[k] = i < N ? code_at(f, y, x) : 0u;
    }
    const float a = noise_a_of(f);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = (int)threadIdx.x + 256 * k;
        if (i >= N)
            break;
        const int ly = i / PT_SW, lx = i - ly * PT_SW;
        const int y = min(max(sy0 + ly, 0), h - 1), x = min(max(sx0 + lx, 0), w - 1);
        const float v = norm_of_code(f, s_curve, code[k], y, x);
        /* Red and blue values are clipped (at their row's clip, the green
         * pass's limit for same-colour neighbours); greens are capped at 1
         * as neighbours. */
        const float lim = ((y ^ x) & 1) ? fminf(v, 1.0f) : v;
        s_raw[ly][lx] = v;
        s_lim[ly][lx] = lim;
        s_dev[ly][lx] = deviation(a, f, f.gp_off, lim);
    }
    __syncthreads();
    const int lj = (int)threadIdx.x % PT_W, li = (int)threadIdx.x / PT_W;
    const int ci = cy0 + li, cj = cx0 + lj;
    if (ci >= h2 || cj >= w2)
        return;
    const int ry = 2 + 2 * li, rx = 2 + 2 * lj; /* the red sensel in LDS */
    /* curvature_ratios from the LDS values: neighbours and their
     * deviations mirrored alike (a replacement's deviation is that of the
     * value it copies; the centre's is sc). */
    auto ratios = [&](int ly, int lx, int y, int x) {
        const float c = s_lim[ly][lx], sc = s_dev[ly][lx];
        float hh[4] = {s_lim[ly][lx - 1], s_lim[ly][lx + 1], s_lim[ly][lx - 2], s_lim[ly][lx + 2]};
        float vv[4] = {s_lim[ly - 1][lx], s_lim[ly + 1][lx], s_lim[ly - 2][lx], s_lim[ly + 2][lx]};
        float sh[4] = {s_dev[ly][lx - 1], s_dev[ly][lx + 1], s_dev[ly][lx - 2], s_dev[ly][lx + 2]};
        float sv[4] = {s_dev[ly - 1][lx], s_dev[ly + 1][lx], s_dev[ly - 2][lx], s_dev[ly + 2][lx]};
        mirror_line(hh, x, w, c);
        mirror_line(sh, x, w, sc);
        mirror_line(vv, y, h, c);
        mirror_line(sv, y, h, sc);
        float2 ratio = make_float2(0.0f, 0.0f);
        if (c > sc) {
            const float sums2 = hh[2] + hh[3], sums3 = vv[2] + vv[3];
            ratio.x = direction_ratio(hh, sh, c, c * -2.0f + sums2, c * 2.0f + sums2);
            ratio.y = direction_ratio(vv, sv, c, c * -2.0f + sums3, c * 2.0f + sums3);
        }
        return ratio;
    };
    const float vr = s_raw[ry][rx], vg1 = s_raw[ry][rx + 1];
    const float vg2 = s_raw[ry + 1][rx], vb = s_raw[ry + 1][rx + 1];
    const float2 rr = ratios(ry, rx, 2 * ci, 2 * cj);
    const float2 rb = ratios(ry + 1, rx + 1, 2 * ci + 1, 2 * cj + 1);
    write_cell(f, a, vr, vg1, vg2, vb, rr, rb, x0 + ((size_t)ci * w2 + cj) * C);
}

/* The window of S_y x S_x cells (at most WIN x WIN) around sampled 64x64
 * tile (bx, by): 8 cells of context on each side (the network's receptive
 * field), shifted to stay inside the frame. Where the window meets the
 * frame's edge, the convolution's zero padding at the window's edge is the
 * frame's own; elsewhere the tile's cells are 8 or more cells from the
 * window's edge, so the network's output on the tile equals the
 * whole-frame output. */
__device__ inline void window_origin(const NetFrame &f, uint32_t bx, uint32_t by, int sy, int sx,
                                     int &wy, int &wx)
{
    /* The tile's first cell in the network's (RGGB-order) coordinates. */
    const int h = (int)f.height, w = (int)f.width;
    const int ly = f.flip_y ? (h - min(64 * (int)by + 64, h)) / 2 : 32 * (int)by;
    const int lx = f.flip_x ? (w - min(64 * (int)bx + 64, w)) / 2 : 32 * (int)bx;
    wy = min(max(ly - 8, 0), h / 2 - sy);
    wx = min(max(lx - 8, 0), w / 2 - sx);
}

__global__ void k_pack_windows(NetFrame f, const uint32_t *__restrict__ tiles, int n, int sy,
                               int sx, uint16_t *__restrict__ x0)
{
    __shared__ uint32_t s_curve[9];
    stage_curve(f, s_curve);
    __syncthreads();
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t cells = (size_t)sy * sx;
    if (i >= cells * n)
        return;
    const int k = (int)(i / cells), c = (int)(i - (size_t)k * cells);
    int wy, wx;
    window_origin(f, tiles[2 * k], tiles[2 * k + 1], sy, sx, wy, wx);
    const int ci = c / sx, cj = c - ci * sx;
    pack_cell(f, s_curve, 2 * (wy + ci), 2 * (wx + cj), x0 + i * C);
}

/* ------------------------------------------------------------------ */
/* Loss: mean squared error between out + bilinear and the target over
 * the cells at least `margin` from the crop's edge, and its gradient with
 * respect to out (bf16). One lane per cell. A crop side on the picture's
 * border keeps no margin: there the zero padding of the convolution and the
 * mirrored bilinear base are the whole-frame ones, so the cells along the
 * picture's sides train as inference sees them. The gradient keeps the
 * scale of the interior count. */

__global__ __launch_bounds__(256) void k_loss(const NetFrame *__restrict__ frames,
                                              const Crop *__restrict__ crops,
                                              const uint32_t *__restrict__ step, uint32_t steps,
                                              int n, int h2, int w2, int margin, float inv_tau2,
                                              const float *__restrict__ out,
                                              uint16_t *__restrict__ dout,
                                              float *__restrict__ loss_part)
{
    __shared__ float s_sum[256];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t cells = (size_t)h2 * w2;
    float sq = 0.0f;
    if (i < cells * n) {
        const int k = (int)(i / cells), c = (int)(i - (size_t)k * cells);
        const int ci = c / w2, cj = c - ci * w2;
        float d[C];
#pragma unroll
        for (int ch = 0; ch < C; ch++)
            d[ch] = 0.0f;
        const Crop cr = crops[(size_t)(*step % steps) * n + k];
        const NetFrame &f = frames[cr.frame];
        const int m_top = cr.y0 == 0 ? 0 : margin, m_left = cr.x0 == 0 ? 0 : margin;
        const int m_bottom = cr.y0 + 2u * (uint32_t)h2 >= f.height ? 0 : margin;
        const int m_right = cr.x0 + 2u * (uint32_t)w2 >= f.width ? 0 : margin;
        if (ci >= m_top && ci < h2 - m_bottom && cj >= m_left && cj < w2 - m_right) {
            const size_t count = (size_t)n * 3 * (size_t)(2 * (h2 - 2 * margin)) *
                                 (size_t)(2 * (w2 - 2 * margin));
            const float scale = 2.0f / (float)count;
            const size_t plane = (size_t)f.width * f.height;
#pragma unroll
            for (int dy = 0; dy < 2; dy++)
#pragma unroll
                for (int dx = 0; dx < 2; dx++) {
                    float base[3];
                    const int y = 2 * ci + dy, x = 2 * cj + dx;
                    bilinear(f, f.curve, (int)cr.y0, (int)cr.x0, 2 * h2, 2 * w2, y, x, base);
                    const size_t at = (size_t)((int)cr.y0 + y) * f.width + (size_t)((int)cr.x0 + x);
#pragma unroll
                    for (int col = 0; col < 3; col++) {
                        const int ch = 3 * (2 * dy + dx) + col;
                        const float pred = out[i * C + ch] + base[col];
                        const float diff = pred - f.target[col * plane + at];
                        /* d^2 + d^4 / tau^2 (inv_tau2 = 0: the squared
                         * error) and its derivative 2d (1 + 2 d^2 / tau^2). */
                        const float d2 = diff * diff;
                        sq += d2 * (1.0f + d2 * inv_tau2);
                        d[ch] = diff * scale * (1.0f + 2.0f * d2 * inv_tau2);
                    }
                }
        }
        uint4 *dst = (uint4 *)(dout + i * C);
#pragma unroll
        for (int q = 0; q < 4; q++)
            dst[q] = make_uint4(f2bf2(d[8 * q + 0], d[8 * q + 1]), f2bf2(d[8 * q + 2], d[8 * q + 3]),
                                f2bf2(d[8 * q + 4], d[8 * q + 5]), f2bf2(d[8 * q + 6], d[8 * q + 7]));
    }
    s_sum[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        loss_part[blockIdx.x] = s_sum[0];
}

__global__ void k_loss_total(const float *__restrict__ part, uint32_t n, float inv_count,
                             float *__restrict__ loss)
{
    if (threadIdx.x != 0)
        return;
    double s = 0.0;
    for (uint32_t i = 0; i < n; i++)
        s += part[i];
    *loss = (float)(s * (double)inv_count);
}

/* ------------------------------------------------------------------ */
/* Inference output: out + bilinear, then white balance with the clip at 1
 * and the output matrix as develop's last steps do them (chroma_row_with
 * and matrix_planes_impl, demosaic.rs:648-720), written in sensor order.
 * One lane per cell; nothing when *skip is set (the frame falls back). */

__global__ __launch_bounds__(256) void k_compose(NetFrame f, Finish fin,
                                                 const float *__restrict__ out,
                                                 float *__restrict__ pr, float *__restrict__ pg,
                                                 float *__restrict__ pb,
                                                 const uint32_t *__restrict__ skip)
{
    __shared__ uint32_t s_curve[9];
    if (skip && *skip)
        return;
    stage_curve(f, s_curve);
    __syncthreads();
    const int w2 = (int)f.width / 2, h2 = (int)f.height / 2;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)h2 * w2)
        return;
    const int ci = (int)(i / w2), cj = (int)(i - (size_t)ci * w2);
    const int hh = (int)f.height, ww = (int)f.width;
    /* The cell's 4x4 neighbourhood of normalized values (rows 2ci - 1 ..
     * 2ci + 2, mirrored at the picture's edges): every value the four
     * sensels' bilinear estimates read, each read once. */
    float nb[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int y = mirror(2 * ci - 1 + r, hh);
#pragma unroll
        for (int c = 0; c < 4; c++)
            nb[r][c] = norm_at(f, y, mirror(2 * cj - 1 + c, ww), s_curve);
    }
    float o[OUT_CH];
#pragma unroll
    for (int q = 0; q < OUT_CH / 4; q++) {
        const float4 v = ((const float4 *)(out + i * OUT_CH))[q];
        o[4 * q] = v.x;
        o[4 * q + 1] = v.y;
        o[4 * q + 2] = v.z;
        o[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
        float res[2][3];
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            /* bilinear() at (2ci + dy, 2cj + dx), from the neighbourhood. */
            auto v = [&](int ky, int kx) { return nb[dy + 1 + ky][dx + 1 + kx]; };
            float base[3];
            const float c = v(0, 0);
            if (dy == dx) {
                const float cross = ((v(-1, 0) + v(1, 0)) + (v(0, -1) + v(0, 1))) * 0.25f;
                const float diag = ((v(-1, -1) + v(-1, 1)) + (v(1, -1) + v(1, 1))) * 0.25f;
                base[0] = dy ? diag : c;
                base[1] = cross;
                base[2] = dy ? c : diag;
            } else {
                const float horiz = (v(0, -1) + v(0, 1)) * 0.5f;
                const float vert = (v(-1, 0) + v(1, 0)) * 0.5f;
                base[0] = dy ? vert : horiz;
                base[1] = c;
                base[2] = dy ? horiz : vert;
            }
            const int ch = 3 * (2 * dy + dx);
            float r = fminf((o[ch + 0] + base[0]) * fin.wb[0], 1.0f);
            float g = fminf(o[ch + 1] + base[1], 1.0f);
            float b = fminf((o[ch + 2] + base[2]) * fin.wb[1], 1.0f);
            if (fin.has_matrix) {
                const float r2 = fmaf(b, fin.m[2], fmaf(r, fin.m[0], g * fin.m[1]));
                const float g2 = fmaf(b, fin.m[5], fmaf(r, fin.m[3], g * fin.m[4]));
                const float b2 = fmaf(b, fin.m[8], fmaf(r, fin.m[6], g * fin.m[7]));
                r = r2;
                g = g2;
                b = b2;
            }
            res[dx][0] = r;
            res[dx][1] = g;
            res[dx][2] = b;
        }
        /* The pair's two sensels are neighbours in sensor order too (the
         * flip only swaps them): one 8-byte store a plane. */
        const int y = 2 * ci + dy;
        const uint32_t sy = f.flip_y ? f.height - 1 - (uint32_t)y : (uint32_t)y;
        const uint32_t sx0 = f.flip_x ? f.width - 2 - (uint32_t)(2 * cj) : (uint32_t)(2 * cj);
        const int lo = f.flip_x ? 1 : 0;
        const size_t at = (size_t)sy * f.width + sx0;
        *(float2 *)(pr + at) = make_float2(res[lo][0], res[1 - lo][0]);
        *(float2 *)(pg + at) = make_float2(res[lo][1], res[1 - lo][1]);
        *(float2 *)(pb + at) = make_float2(res[lo][2], res[1 - lo][2]);
    }
}

/* Sampled-tile check: per tile, the network's output at each sensel
 * (out + bilinear, white balance and clip at 1, no matrix) against the
 * plain develop in scratch; PSNR with peak 1. One 1,024-lane workgroup per
 * tile, four sensels a lane, summed in a fixed order. */
enum { CK_LANES = 1024 };

__global__ __launch_bounds__(CK_LANES) void k_check_psnr(NetFrame f, Finish fin,
                                                    const float *__restrict__ wout, int sy, int sx,
                                                    const uint32_t *__restrict__ tiles,
                                                    const float *__restrict__ scratch,
                                                    float min_db, float *__restrict__ psnr,
                                                    uint32_t *__restrict__ use_plain)
{
    __shared__ double s_sum[CK_LANES];
    __shared__ uint32_t s_n[CK_LANES];
    __shared__ uint32_t s_curve[9];
    stage_curve(f, s_curve);
    __syncthreads();
    const uint32_t k = blockIdx.x;
    const uint32_t x0 = 64 * tiles[2 * k], y0 = 64 * tiles[2 * k + 1];
    const float *plain = scratch + (size_t)k * 3 * 4096;
    int wy, wx;
    window_origin(f, tiles[2 * k], tiles[2 * k + 1], sy, sx, wy, wx);
    const float *out = wout + (size_t)k * sy * sx * C;
    double sum = 0.0;
    uint32_t n = 0;
    for (uint32_t i = threadIdx.x; i < 4096; i += CK_LANES) {
        const uint32_t py = y0 + i / 64, px = x0 + i % 64; /* sensor order */
        if (py >= f.height || px >= f.width)
            continue;
        const int y = (int)(f.flip_y ? f.height - 1 - py : py);
        const int x = (int)(f.flip_x ? f.width - 1 - px : px);
        float base[3];
        bilinear(f, s_curve, 0, 0, (int)f.height, (int)f.width, y, x, base);
        const size_t cell = (size_t)((y >> 1) - wy) * sx + ((x >> 1) - wx);
        const int ch = 3 * (2 * (y & 1) + (x & 1));
        const float r = fminf((out[cell * C + ch + 0] + base[0]) * fin.wb[0], 1.0f);
        const float g = fminf(out[cell * C + ch + 1] + base[1], 1.0f);
        const float b = fminf((out[cell * C + ch + 2] + base[2]) * fin.wb[1], 1.0f);
        const double e0 = (double)r - plain[i], e1 = (double)g - plain[4096 + i],
                     e2 = (double)b - plain[8192 + i];
        sum += e0 * e0 + e1 * e1 + e2 * e2;
        n += 3;
    }
    s_sum[threadIdx.x] = sum;
    s_n[threadIdx.x] = n;
    __syncthreads();
    for (uint32_t s = CK_LANES / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
            s_n[threadIdx.x] += s_n[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double mse = s_n[0] ? s_sum[0] / s_n[0] : 0.0;
        const float db = mse > 0.0 ? (float)(-10.0 * log10(mse)) : 200.0f;
        psnr[k] = db;
        if (!(db >= min_db))
            atomicOr(use_plain, 1u);
    }
}

/* ------------------------------------------------------------------ */

hipError_t launch_pack_windows(const NetFrame &f, const uint32_t *tiles, int n, int sy, int sx,
                               uint16_t *x0, hipStream_t s)
{
    const size_t cells = (size_t)n * sy * sx;
    hipLaunchKernelGGL(k_pack_windows, dim3((uint32_t)((cells + 255) / 256)), dim3(256), 0, s, f,
                       tiles, n, sy, sx, x0);
    return hipGetLastError();
}

hipError_t launch_check(const DevelopDesc &d, const NetFrame &f, const Finish &fin,
                        const float *wout, int sy, int sx, const uint32_t *tiles,
                        uint32_t n_tiles, float *scratch, float min_db, float *psnr,
                        uint32_t *use_plain, hipStream_t s)
{
    if (n_tiles == 0)
        return hipSuccess;
    hipError_t e = launch_develop_sampled(d, tiles, n_tiles, scratch, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_check_psnr, dim3(n_tiles), dim3(CK_LANES), 0, s, f, fin, wout, sy, sx,
                       tiles, scratch, min_db, psnr, use_plain);
    return hipGetLastError();
}

hipError_t launch_conv_infer(int mode, const uint16_t *in, int h, int w, const uint16_t *wgt,
                             const float *bias, uint16_t *out, float *out12, const uint32_t *skip,
                             uint32_t grid, hipStream_t s)
{
    if (h <= 0 || w <= 0)
        return hipSuccess;
    const uint32_t tiles = (uint32_t)(((w + CT_COLS - 1) / CT_COLS) * ((h + CT_ROWS - 1) / CT_ROWS));
    const dim3 g(std::min(grid, tiles));
    if (mode == IM_HEAD)
        hipLaunchKernelGGL(k_conv_infer<IM_HEAD>, g, dim3(256), 0, s, in, h, w, wgt, bias, out,
                           out12, skip);
    else if (mode == IM_RES)
        hipLaunchKernelGGL(k_conv_infer<IM_RES>, g, dim3(256), 0, s, in, h, w, wgt, bias, out,
                           out12, skip);
    else
        hipLaunchKernelGGL(k_conv_infer<IM_TAIL>, g, dim3(256), 0, s, in, h, w, wgt, bias, out,
                           out12, skip);
    return hipGetLastError();
}

const void *conv_infer_kernel() { return (const void *)k_conv_infer<IM_RES>; }

hipError_t launch_conv_pair(int first, int last, const uint16_t *in, int h, int w,
                            const uint16_t *wa, const float *ba, const uint16_t *wb,
                            const float *bb, uint16_t *out, float *out12, const uint32_t *skip,
                            hipStream_t s)
{
    if (h <= 0 || w <= 0)
        return hipSuccess;
    const dim3 grid((uint32_t)((w + PS_W - 1) / PS_W) * (uint32_t)((h + PS_SEG - 1) / PS_SEG));
    if (first && last)
        hipLaunchKernelGGL((k_conv_pair<IM_HEAD, IM_TAIL>), grid, dim3(256), 0, s, in, h, w, wa,
                           ba, wb, bb, out, out12, skip);
    else if (first)
        hipLaunchKernelGGL((k_conv_pair<IM_HEAD, IM_RES>), grid, dim3(256), 0, s, in, h, w, wa, ba,
                           wb, bb, out, out12, skip);
    else if (last)
        hipLaunchKernelGGL((k_conv_pair<IM_RES, IM_TAIL>), grid, dim3(256), 0, s, in, h, w, wa, ba,
                           wb, bb, out, out12, skip);
    else
        hipLaunchKernelGGL((k_conv_pair<IM_RES, IM_RES>), grid, dim3(256), 0, s, in, h, w, wa, ba,
                           wb, bb, out, out12, skip);
    return hipGetLastError();
}


hipError_t launch_weights(const float *params, uint16_t *wf, uint16_t *wt, hipStream_t s)
{
    const uint32_t n = LAYERS * LAYER_W;
    hipLaunchKernelGGL(k_weights, dim3((n + 255) / 256), dim3(256), 0, s, params, wf, wt);
    return hipGetLastError();
}

hipError_t launch_conv(int epi, const uint16_t *in, int n, int h, int w, const uint16_t *wgt,
                       const float *bias, const uint16_t *resid, uint16_t *out, uint16_t *aux,
                       float *out_f32, hipStream_t s)
{
    if (n <= 0 || h <= 0 || w <= 0)
        return hipSuccess;
    const dim3 grid((w + CT_COLS - 1) / CT_COLS, (h + CT_ROWS - 1) / CT_ROWS, n);
    switch (epi) {
    case EPI_RELU:
        hipLaunchKernelGGL(k_conv<EPI_RELU>, grid, dim3(256), 0, s, in, h, w, wgt, bias, resid,
                           out, aux, out_f32);
        break;
    case EPI_RESID_RELU:
        hipLaunchKernelGGL(k_conv<EPI_RESID_RELU>, grid, dim3(256), 0, s, in, h, w, wgt, bias,
                           resid, out, aux, out_f32);
        break;
    case EPI_LINEAR_F32:
        hipLaunchKernelGGL(k_conv<EPI_LINEAR_F32>, grid, dim3(256), 0, s, in, h, w, wgt, bias,
                           resid, out, aux, out_f32);
        break;
    case EPI_LINEAR:
        hipLaunchKernelGGL(k_conv<EPI_LINEAR>, grid, dim3(256), 0, s, in, h, w, wgt, bias, resid,
                           out, aux, out_f32);
        break;
    case EPI_ADD:
        hipLaunchKernelGGL(k_conv<EPI_ADD>, grid, dim3(256), 0, s, in, h, w, wgt, bias, resid,
                           out, aux, out_f32);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t wgrad_tiles(int n, int h, int w)
{
    return (size_t)n * ((h + GT_ROWS - 1) / GT_ROWS) * ((w + GT_COLS - 1) / GT_COLS);
}

hipError_t launch_wgrad(const uint16_t *x, const uint16_t *g, int n, int h, int w, float *part,
                        float *grad, hipStream_t s)
{
    const dim3 grid((w + GT_COLS - 1) / GT_COLS, (h + GT_ROWS - 1) / GT_ROWS, n);
    hipLaunchKernelGGL(k_wgrad, grid, dim3(GT_THREADS), 0, s, x, g, h, w, part);
    const uint32_t tiles = (uint32_t)wgrad_tiles(n, h, w);
    hipLaunchKernelGGL(k_reduce, dim3((LAYER_PARAMS + 255) / 256), dim3(256), 0, s, part, tiles,
                       grad);
    return hipGetLastError();
}

hipError_t launch_relu_grad(const uint16_t *dh, const uint16_t *mask, uint16_t *dz, size_t count,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_relu_grad, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, s, dh,
                       mask, dz, count);
    return hipGetLastError();
}

hipError_t launch_adam(float *params, const float *grads, float *m, float *v, uint32_t *step,
                       float lr, float lr_min, uint32_t decay, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3((PARAMS + 255) / 256), dim3(256), 0, s, params, grads, m, v,
                       step, lr, lr_min, decay);
    return hipGetLastError();
}

hipError_t launch_step_inc(uint32_t *step, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, step);
    return hipGetLastError();
}

hipError_t launch_gate_clear(uint32_t *flag, uint32_t value, hipStream_t s)
{
    hipLaunchKernelGGL(k_set, dim3(1), dim3(64), 0, s, flag, value);
    return hipGetLastError();
}

hipError_t launch_noise_gate(const NetFrame &f, float lo, float hi, uint32_t *use_plain,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_noise_gate, dim3(1), dim3(64), 0, s, f, lo, hi, use_plain);
    return hipGetLastError();
}

hipError_t launch_pack_crops(const NetFrame *frames, const Crop *crops, const uint32_t *step,
                             uint32_t steps, int n, int h2, int w2, uint16_t *x0, hipStream_t s)
{
    const size_t cells = (size_t)n * h2 * w2;
    hipLaunchKernelGGL(k_pack_crops, dim3((uint32_t)((cells + 255) / 256)), dim3(256), 0, s,
                       frames, crops, step, steps, n, h2, w2, x0);
    return hipGetLastError();
}

hipError_t launch_loss(const NetFrame *frames, const Crop *crops, const uint32_t *step,
                       uint32_t steps, int n, int h2, int w2, int margin, float inv_tau2,
                       const float *out, uint16_t *dout, float *loss_part, float *loss,
                       hipStream_t s)
{
    const size_t cells = (size_t)n * h2 * w2;
    const uint32_t blocks = (uint32_t)((cells + 255) / 256);
    hipLaunchKernelGGL(k_loss, dim3(blocks), dim3(256), 0, s, frames, crops, step, steps, n, h2,
                       w2, margin, inv_tau2, out, dout, loss_part);
    const double count = (double)n * 3.0 * (2.0 * (h2 - 2 * margin)) * (2.0 * (w2 - 2 * margin));
    hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(64), 0, s, loss_part, blocks,
                       (float)(1.0 / count), loss);
    return hipGetLastError();
}

hipError_t launch_pack_frame(const NetFrame &f, uint16_t *x0, const uint32_t *skip, hipStream_t s)
{
    const dim3 grid((f.width / 2 + PT_W - 1) / PT_W, (f.height / 2 + PT_H - 1) / PT_H);
    hipLaunchKernelGGL(k_pack_frame, grid, dim3(256), 0, s, f, x0, skip);
    return hipGetLastError();
}

hipError_t launch_compose(const NetFrame &f, const Finish &fin, const float *out, float *r,
                          float *g, float *b, const uint32_t *skip, hipStream_t s)
{
    const size_t cells = (size_t)(f.width / 2) * (f.height / 2);
    hipLaunchKernelGGL(k_compose, dim3((uint32_t)((cells + 255) / 256)), dim3(256), 0, s, f, fin,
                       out, r, g, b, skip);
    return hipGetLastError();
}

} // namespace net
} // namespace prr
