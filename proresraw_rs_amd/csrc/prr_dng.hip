/* Lossless JPEG tiles of a DNG raw image on gfx950: the LosslessJpeg branch
 * of the DNG writer (examples/prr2dng/dng.rs:92-113) with its `tile`
 * padding (dng.rs:115-128) and lj92::encode (examples/prr2dng/lj92.rs:14-93)
 * for every 256x256 tile of a picture, two components, predictor 1.
 *
 * Two kernels, one 1024-lane workgroup per tile:
 *
 * k_lj92_encode
 *   1. the tile (65,536 u16, edge sensels repeating the last row or column
 *      of the same CFA phase) is loaded with coalesced reads into LDS
 *      (128 KiB; one workgroup per CU);
 *   2. lane t owns the 64 samples of row t/4, columns 64*(t%4)..+63: it
 *      computes each prediction difference and its category (bit length)
 *      and counts categories in five packed byte counters, which are summed
 *      over the wave by shuffles and over the workgroup by LDS atomics;
 *   3. wave 0 builds the optimal table (HuffmanTable::optimal,
 *      lj92.rs:105-206) with one lane per symbol: every merge takes the two
 *      smallest frequencies by a wave-wide minimum of (frequency, 31 - symbol)
 *      keys, which picks exactly the symbols the crate's sequential scan
 *      picks (ties to the larger symbol), and lengthens every symbol of both
 *      merged groups; the JPEG length limit (Annex K.3) runs on one lane when
 *      a length exceeds 16; canonical codes follow from each symbol's rank in
 *      (length, symbol) order. Lane 0 writes the SOI/SOF3/DHT/SOS header;
 *   4. each lane sums its samples' code + extra bits, a workgroup scan gives
 *      its first bit, and it emits its bits MSB first through a 64-bit
 *      accumulator as big-endian words into the tile's raw stream; a word a
 *      lane shares with its neighbour (at most one at each end, since a lane
 *      emits at least 64 bits) is merged through LDS. The 0xFF bytes of the
 *      stream are counted for the stuffing.
 * k_lj92_pack
 *   each tile's place in the output is the sum of the earlier tiles'
 *   lengths; the header is copied, the raw bytes are copied in 16
 *   wave-owned segments with a zero byte after every 0xFF (a wave-wide scan
 *   of the 0xFF counts gives each byte's shift), and EOI ends the stream.
 *
 * Bounds: the image's bytes are read once (2 B/sensel) and each tile's
 * stream is written twice (raw, then stuffed), about 3.5 B/sensel of HBM
 * traffic at 12-bit content; the work per sample is a few dozen VALU
 * operations, so the kernels are latency/issue bound, not HBM bound.
 */
#include <hip/hip_runtime.h>

#include "prr_internal.h"

namespace prr {

namespace {

constexpr int TILE = LJ92_TILE;
constexpr int LANES = 1024;
constexpr int RUN = 64; /* samples a lane codes */

__device__ __forceinline__ uint32_t clamp_phase(uint32_t i, uint32_t n)
{
    return i < n ? i : n - 2 + ((i - n) & 1);
}

__device__ __forceinline__ uint32_t category(int32_t d)
{
    uint32_t a = (uint32_t)(d < 0 ? -d : d);
    return a ? 32u - (uint32_t)__clz(a) : 0u;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t x)
{
    for (int o = 32; o; o >>= 1)
        x = min(x, (uint32_t)__shfl_xor((int)x, o));
    return x;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
    for (int o = 32; o; o >>= 1)
        x += (uint32_t)__shfl_xor((int)x, o);
    return x;
}

/* Inclusive scan over the 64 lanes of a wave. */
__device__ __forceinline__ uint32_t wave_scan(uint32_t x, uint32_t lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)x, o);
        if (lane >= (uint32_t)o)
            x += y;
    }
    return x;
}

__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

__device__ __forceinline__ uint32_t ff_bytes(uint32_t w, uint32_t nbytes)
{
    uint32_t n = 0;
    for (uint32_t b = 0; b < nbytes; b++)
        n += ((w >> (24 - 8 * b)) & 0xFF) == 0xFF;
    return n;
}

struct EncodeLds {
    alignas(16) uint16_t tile[TILE * TILE];
    uint32_t hist[17];
    uint32_t table[17];      /* (code << 16) | length, per category */
    uint32_t wave_total[16];
    uint32_t wave_ff[16];
    uint32_t tail[LANES];    /* a lane's last, partial word (MSB aligned) */
    uint32_t size[18];       /* each symbol's Huffman length before limiting */
    uint32_t bits[33];       /* code-length counts while the table is built */
    uint32_t first[18];      /* canonical first code of each length */
    uint32_t cum[18];        /* codes shorter than each length */
    uint8_t huffval[17];
    uint32_t nv;
};

/* Lane t's two predictions before its run: the samples two columns to the
 * left of its first two (same row), or for a run at column 0 the samples
 * above, or 32768 for the tile's first row (lj92.rs:72-87). */
__device__ __forceinline__ uint32_t pred_pair(const EncodeLds &s, uint32_t row, uint32_t c0)
{
    if (c0 >= 2)
        return (uint32_t)s.tile[row * TILE + c0 - 2] | ((uint32_t)s.tile[row * TILE + c0 - 1] << 16);
    if (row > 0)
        return (uint32_t)s.tile[(row - 1) * TILE] | ((uint32_t)s.tile[(row - 1) * TILE + 1] << 16);
    return 0x80008000u;
}

/* Calls fn(d) for each prediction difference of the lane's run, in order,
 * reading the run from LDS eight samples at a time; only the two previous
 * samples are carried (the prediction is two columns back). */
template <class Fn>
__device__ __forceinline__ void for_each_difference(const EncodeLds &s, uint32_t row, uint32_t c0,
                                                    Fn fn)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(&s.tile[row * TILE + c0]);
    const uint32_t prev = pred_pair(s, row, c0);
    uint32_t p0 = prev & 0xFFFF, p1 = prev >> 16; /* samples k-2 and k-1 */
#pragma unroll 1
    for (int q = 0; q < RUN / 8; q++) {
        const uint4 v = src[q];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t x = k & 1 ? w[k / 2] >> 16 : w[k / 2] & 0xFFFF;
            fn((int32_t)(int16_t)(uint16_t)(x - p0));
            p0 = p1;
            p1 = x;
        }
    }
}

} // namespace

__global__ void __launch_bounds__(LANES) k_lj92_encode(const uint16_t *__restrict__ values,
                                                       uint32_t stride, uint32_t width,
                                                       uint32_t height, uint32_t across,
                                                       uint32_t *__restrict__ raw,
                                                       Lj92Record *__restrict__ records)
{
    __shared__ EncodeLds s;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t tile = blockIdx.x;
    const uint32_t left = (tile % across) * TILE, top = (tile / across) * TILE;

    if (t < 17)
        s.hist[t] = 0;
    /* 1. The padded tile: inside the picture at a 16-byte aligned stride,
     * 16-byte loads (a wave reads two rows of 512 bytes an instruction);
     * at the picture's edges, one sensel a lane with the CFA-phase padding. */
    if (left + TILE <= width && top + TILE <= height && (stride & 7) == 0 &&
        (reinterpret_cast<uintptr_t>(values) & 15) == 0) {
        uint4 *dst = reinterpret_cast<uint4 *>(s.tile);
        for (uint32_t i = t; i < TILE * TILE / 8; i += LANES) {
            const uint32_t y = i / (TILE / 8), c = i % (TILE / 8);
            dst[i] = *reinterpret_cast<const uint4 *>(values + (size_t)(top + y) * stride + left +
                                                      8 * c);
        }
    } else {
        for (uint32_t i = t; i < TILE * TILE; i += LANES) {
            uint32_t y = i / TILE, x = i % TILE;
            uint32_t sy = clamp_phase(top + y, height), sx = clamp_phase(left + x, width);
            s.tile[i] = values[(size_t)sy * stride + sx];
        }
    }
    __syncthreads();

    /* 2. Differences, categories, counts. */
    const uint32_t row = t >> 2, c0 = (t & 3) * RUN;
    uint32_t h0 = 0, h1 = 0, h2 = 0, h3 = 0, h4 = 0;
    for_each_difference(s, row, c0, [&](int32_t d) {
        uint32_t c = category(d);
        uint32_t inc = 1u << ((c & 3) * 8), g = c >> 2;
        h0 += g == 0 ? inc : 0;
        h1 += g == 1 ? inc : 0;
        h2 += g == 2 ? inc : 0;
        h3 += g == 3 ? inc : 0;
        h4 += g == 4 ? inc : 0;
    });
    {
        /* Byte fields (at most 64 each) widened to 16-bit fields, then summed
         * over the wave (at most 4096 each). */
        uint32_t w16[9];
        const uint32_t hh[5] = {h0, h1, h2, h3, h4};
#pragma unroll
        for (int q = 0; q < 5; q++) {
            if (2 * q < 9)
                w16[2 * q] = (hh[q] & 0xFF) | ((hh[q] & 0xFF00) << 8);
            if (2 * q + 1 < 9)
                w16[2 * q + 1] = ((hh[q] >> 16) & 0xFF) | ((hh[q] >> 8) & 0xFF0000);
        }
#pragma unroll
        for (int q = 0; q < 9; q++)
            w16[q] = wave_sum(w16[q]);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < 9; q++) {
                atomicAdd(&s.hist[2 * q], w16[q] & 0xFFFF);
                if (2 * q + 1 < 17)
                    atomicAdd(&s.hist[2 * q + 1], w16[q] >> 16);
            }
        }
    }
    __syncthreads();

    /* 3. The table: wave 0, lane v holds symbol v (17 = the reserved one);
     * each hand-over between lanes goes through LDS and a barrier. */
    if (wave == 0) {
        const uint32_t v = lane;
        uint32_t f = v < 17 ? s.hist[v] : v == 17 ? 1u : 0u;
        uint32_t size = 0, grp = v;
        for (;;) {
            uint32_t key = (v < 18 && f) ? (f << 5) | (31 - v) : 0xFFFFFFFFu;
            uint32_t k1 = wave_min(key);
            uint32_t k2 = wave_min(key == k1 ? 0xFFFFFFFFu : key);
            if (k2 == 0xFFFFFFFFu)
                break;
            uint32_t v1 = 31 - (k1 & 31), v2 = 31 - (k2 & 31);
            if (grp == v1 || grp == v2) {
                size++;
                grp = v1;
            }
            if (v == v1)
                f += k2 >> 5;
            if (v == v2)
                f = 0;
        }
        if (v < 18)
            s.size[v] = size;
    }
    __syncthreads();
    if (t == 0) {
        for (int i = 0; i < 33; i++)
            s.bits[i] = 0;
        for (int v = 0; v < 18; v++)
            if (s.size[v])
                s.bits[s.size[v]]++;
        /* Lengths past 16 (JPEG Annex K.3, Figure K.3), then the reserved
         * code (the longest) dropped. */
        for (int i = 32; i > 16; i--) {
            while (s.bits[i] > 0) {
                int j = i - 2;
                while (j > 0 && s.bits[j] == 0)
                    j--;
                s.bits[i] -= 2;
                s.bits[i - 1] += 1;
                s.bits[j + 1] += 2;
                s.bits[j] -= 1;
            }
        }
        int i = 16;
        while (i > 0 && s.bits[i] == 0)
            i--;
        if (s.bits[i])
            s.bits[i]--;
        uint32_t code = 0, cum = 0;
        for (int len = 1; len <= 16; len++) {
            s.first[len] = code;
            s.cum[len] = cum;
            code = (code + s.bits[len]) << 1;
            cum += s.bits[len];
        }
        s.cum[17] = cum;
        s.nv = cum;
    }
    __syncthreads();
    if (t < 17) {
        /* Rank in (unlimited length, symbol) order among the used symbols,
         * then the canonical code of that rank. */
        const uint32_t v = t, size = s.size[v];
        if (size) {
            uint32_t rank = 0;
            for (uint32_t u = 0; u < 17; u++) {
                uint32_t su = s.size[u];
                rank += (su && (su < size || (su == size && u < v))) ? 1u : 0u;
            }
            int len = 1;
            while (len < 16 && s.cum[len + 1] <= rank)
                len++;
            uint32_t code = s.first[len] + (rank - s.cum[len]);
            s.table[v] = (code << 16) | (uint32_t)len;
            s.huffval[rank] = (uint8_t)v;
        } else {
            s.table[v] = 0;
        }
    }
    __syncthreads();
    if (t == 0) {
        uint8_t *h = records[tile].header;
        uint32_t n = 0;
        auto put = [&](uint32_t b) { h[n++] = (uint8_t)b; };
        auto put16 = [&](uint32_t b) {
            put(b >> 8);
            put(b & 0xFF);
        };
        put16(0xFFD8);                     /* SOI */
        put16(0xFFC3);                     /* SOF3, lossless Huffman */
        put16(8 + 3 * 2);
        put(16);                           /* precision */
        put16(TILE);                       /* lines */
        put16(TILE / 2);                   /* samples a line, per component */
        put(2);
        for (uint32_t id = 0; id < 2; id++) {
            put(id);
            put(0x11);
            put(0);
        }
        put16(0xFFC4);                     /* DHT: DC table 0 */
        put16(2 + 1 + 16 + s.nv);
        put(0);
        for (int len = 1; len <= 16; len++)
            put(s.bits[len]);
        for (uint32_t k = 0; k < s.nv; k++)
            put(s.huffval[k]);
        put16(0xFFDA);                     /* SOS: predictor 1, no point transform */
        put16(6 + 2 * 2);
        put(2);
        for (uint32_t id = 0; id < 2; id++) {
            put(id);
            put(0);
        }
        put(1);
        put(0);
        put(0);
        records[tile].header_len = n;
    }

    /* 4. Bit counts, their scan, the words. */
    uint32_t nbits = 0;
    for_each_difference(s, row, c0, [&](int32_t d) {
        uint32_t c = category(d);
        nbits += (s.table[c] & 0xFF) + ((c - 1u) < 15u ? c : 0u);
    });
    uint32_t incl = wave_scan(nbits, lane);
    if (lane == 63)
        s.wave_total[wave] = incl;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (uint32_t w = 0; w < 16; w++) {
        uint32_t x = s.wave_total[w];
        base += w < wave ? x : 0;
        total += x;
    }
    const uint32_t first_bit = base + incl - nbits;

    uint32_t *out = raw + (size_t)tile * LJ92_RAW_WORDS;
    uint64_t acc = 0;
    uint32_t nacc = first_bit & 31; /* leading bits belong to the lane before */
    uint32_t word = first_bit >> 5;
    uint32_t head = 0, ff = 0;
    bool have_head = false;
    const bool shared_head = (first_bit & 31) != 0;
    for_each_difference(s, row, c0, [&](int32_t d) {
        uint32_t c = category(d);
        uint32_t e = s.table[c];
        uint32_t len = e & 0xFF, code = e >> 16;
        uint32_t xbits = (c - 1u) < 15u ? c : 0u;
        uint32_t extra = (uint32_t)(d < 0 ? d - 1 : d) & ((1u << xbits) - 1u);
        acc = (acc << (len + xbits)) | ((uint64_t)code << xbits) | extra;
        nacc += len + xbits;
        if (nacc >= 32) {
            nacc -= 32;
            uint32_t wv = (uint32_t)(acc >> nacc);
            acc &= (1ull << nacc) - 1;
            if (shared_head && !have_head) {
                head = wv;
                have_head = true;
            } else {
                out[word] = bswap32(wv);
                ff += ff_bytes(wv, 4);
            }
            word++;
        }
    });
    /* The last partial word: MSB aligned, handed to the next lane. */
    const uint32_t tail = nacc ? (uint32_t)(acc << (32 - nacc)) : 0u;
    s.tail[t] = tail;
    __syncthreads();
    if (shared_head) {
        uint32_t wv = head | s.tail[t - 1];
        out[first_bit >> 5] = bswap32(wv);
        ff += ff_bytes(wv, 4);
    }
    const uint32_t raw_bytes = (total + 7) >> 3;
    if (t == LANES - 1 && nacc) {
        /* The stream's end: the last byte padded with one bits. */
        uint32_t pad_bits = (8 - (total & 7)) & 7;
        uint32_t wv = tail | (pad_bits ? (((1u << pad_bits) - 1u) << (32 - nacc - pad_bits)) : 0u);
        out[word] = bswap32(wv);
        ff += ff_bytes(wv, (nacc + pad_bits) >> 3);
    }
    ff = wave_sum(ff);
    if (lane == 0)
        s.wave_ff[wave] = ff;
    __syncthreads();
    if (t == 0) {
        uint32_t all = 0;
        for (int w = 0; w < 16; w++)
            all += s.wave_ff[w];
        Lj92Record &r = records[tile];
        r.raw_bits = total;
        r.ff = all;
        r.total = r.header_len + raw_bytes + all + 2;
    }
}

__global__ void __launch_bounds__(LANES) k_lj92_pack(const uint32_t *__restrict__ raw,
                                                     const Lj92Record *__restrict__ records,
                                                     uint8_t *__restrict__ out,
                                                     uint32_t *__restrict__ offsets,
                                                     uint32_t *__restrict__ lengths)
{
    __shared__ uint32_t part[16];
    __shared__ uint32_t seg_ff[16];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t tile = blockIdx.x;

    uint32_t before = 0;
    for (uint32_t u = t; u < tile; u += LANES)
        before += records[u].total;
    before = wave_sum(before);
    if (lane == 0)
        part[wave] = before;
    __syncthreads();
    uint32_t base = 0;
    for (int w = 0; w < 16; w++)
        base += part[w];

    const Lj92Record &r = records[tile];
    const uint32_t hlen = r.header_len, raw_bytes = (r.raw_bits + 7) >> 3;
    uint8_t *dst = out + base;
    if (t == 0) {
        offsets[tile] = base;
        lengths[tile] = r.total;
        dst[r.total - 2] = 0xFF; /* EOI */
        dst[r.total - 1] = 0xD9;
    }
    if (t < hlen)
        dst[t] = r.header[t];

    /* Wave w copies bytes [w*seg, (w+1)*seg) of the raw stream, 256 a step. */
    const uint32_t *src = raw + (size_t)tile * LJ92_RAW_WORDS;
    const uint32_t seg = ((raw_bytes + 16 * 256 - 1) / (16 * 256)) * 256;
    const uint32_t lo = wave * seg, hi = min(raw_bytes, lo + seg);
    uint32_t mine = 0;
    for (uint32_t at = lo + 4 * lane; at < hi; at += 256)
        mine += ff_bytes(bswap32(src[at >> 2]), min(4u, hi - at));
    mine = wave_sum(mine);
    if (lane == 0)
        seg_ff[wave] = mine;
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t w = 0; w < wave; w++)
        carry += seg_ff[w];
    uint8_t *data = dst + hlen;
    for (uint32_t at0 = lo; at0 < hi; at0 += 256) {
        const uint32_t at = at0 + 4 * lane;
        const uint32_t n = at < hi ? min(4u, hi - at) : 0u;
        const uint32_t wv = n ? bswap32(src[at >> 2]) : 0u;
        const uint32_t cnt = ff_bytes(wv, n);
        const uint32_t incl = wave_scan(cnt, lane);
        uint32_t pos = at + carry + incl - cnt;
        for (uint32_t b = 0; b < n; b++) {
            uint8_t byte = (uint8_t)(wv >> (24 - 8 * b));
            data[pos++] = byte;
            if (byte == 0xFF)
                data[pos++] = 0;
        }
        carry += (uint32_t)__shfl((int)incl, 63);
    }
}

hipError_t launch_lj92(const uint16_t *values, uint32_t stride, uint32_t width, uint32_t height,
                       uint32_t *raw, Lj92Record *records, uint8_t *out, uint32_t *offsets,
                       uint32_t *lengths, hipStream_t s)
{
    const uint32_t across = (width + TILE - 1) / TILE, down = (height + TILE - 1) / TILE;
    const uint32_t n = across * down;
    hipLaunchKernelGGL(k_lj92_encode, dim3(n), dim3(LANES), 0, s, values, stride, width, height,
                       across, raw, records);
    hipLaunchKernelGGL(k_lj92_pack, dim3(n), dim3(LANES), 0, s, raw, records, out, offsets, lengths);
    return hipGetLastError();
}

} // namespace prr
