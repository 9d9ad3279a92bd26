/* Crop and orientation of a developed or CFA picture on gfx950.
 *
 * The crate records how a track's stored frames are turned for display
 * (mov::Orientation, the eight EXIF symmetries, src/mov.rs:121-157) and the
 * crop margins the camera recommends (header::Crop, src/header.rs:98-109,
 * applied by readers as DNG DefaultCrop, examples/prr2dng/dng.rs:190-196).
 * This kernel applies both: the picture is cropped, then turned upright.
 *
 * Output pixel (x, y) of an ow x oh output comes from cropped pixel
 *   1 Identity           (x, y)
 *   2 MirrorHorizontal   (cw-1-x, y)
 *   3 Rotate180          (cw-1-x, ch-1-y)
 *   4 MirrorVertical     (x, ch-1-y)
 *   5 Transpose          (y, x)
 *   6 Rotate90 (cw)      (y, ch-1-x)
 *   7 Transverse         (cw-1-y, ch-1-x)
 *   8 Rotate270 (ccw)    (cw-1-y, x)
 * where cw x ch is the cropped size (swapped for 5-8).
 *
 * One 256-lane workgroup per 64x64 output tile and plane: the tile's
 * source is itself a 64x64 block of the cropped picture, read row by row
 * (coalesced) into a padded LDS tile, then written row by row in output
 * order (coalesced), each lane reading LDS at the turned index. The kernel
 * moves each byte once in and once out: HBM bound.
 */
#include <hip/hip_runtime.h>

#include "prr_internal.h"

namespace prr {

namespace {

constexpr int OT = 64;        /* output tile side */
constexpr int OLANES = 256;

__device__ __forceinline__ void source_of(uint32_t orient, uint32_t cw, uint32_t ch, uint32_t x,
                                          uint32_t y, uint32_t &sx, uint32_t &sy)
{
    switch (orient) {
    case 2: sx = cw - 1 - x; sy = y; break;
    case 3: sx = cw - 1 - x; sy = ch - 1 - y; break;
    case 4: sx = x; sy = ch - 1 - y; break;
    case 5: sx = y; sy = x; break;
    case 6: sx = y; sy = ch - 1 - x; break;
    case 7: sx = cw - 1 - y; sy = ch - 1 - x; break;
    case 8: sx = cw - 1 - y; sy = x; break;
    default: sx = x; sy = y;
    }
}

} // namespace

template <class T>
__device__ __forceinline__ void orient_tile(const OrientDesc &d, const T *__restrict__ in,
                                            T *__restrict__ out, uint32_t (&tile)[OT][OT + 1],
                                            uint32_t ox, uint32_t oy, uint32_t tw, uint32_t th,
                                            uint32_t sx0, uint32_t sy0, uint32_t sw, uint32_t sh)
{
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (uint32_t j = ty; j < sh; j += OLANES / OT) {
        if (tx < sw) {
            const T v = in[(size_t)(d.top + sy0 + j) * d.in_stride + d.left + sx0 + tx];
            if constexpr (sizeof(T) == 4)
                tile[j][tx] = __builtin_bit_cast(uint32_t, v);
            else
                tile[j][tx] = (uint32_t)v;
        }
    }
    __syncthreads();
    for (uint32_t v = ty; v < th; v += OLANES / OT) {
        if (tx < tw) {
            uint32_t sx, sy;
            source_of(d.orient, d.cw, d.ch, ox + tx, oy + v, sx, sy);
            uint32_t bits = tile[sy - sy0][sx - sx0];
            T value;
            if constexpr (sizeof(T) == 4)
                value = __builtin_bit_cast(T, bits);
            else
                value = (T)bits;
            out[(size_t)(oy + v) * d.out_stride + ox + tx] = value;
        }
    }
}

/* VEC: every row of the cropped input and of the output starts 16-byte
 * aligned and the cropped width is a multiple of 4 (checked on the host),
 * so a full tile's source block and output rows move as float4. */
template <class T, bool VEC>
__global__ void __launch_bounds__(OLANES) k_orient(OrientDesc d)
{
    __shared__ uint32_t tile[OT][OT + 1];
    const T *__restrict__ in = static_cast<const T *>(d.in[blockIdx.z]);
    T *__restrict__ out = static_cast<T *>(d.out[blockIdx.z]);
    const uint32_t ox = blockIdx.x * OT, oy = blockIdx.y * OT;
    const uint32_t tw = min((uint32_t)OT, d.ow - ox), th = min((uint32_t)OT, d.oh - oy);
    /* The source block: each source coordinate follows one output
     * coordinate with slope +-1, so its range comes from two corners. */
    uint32_t ax, ay, bx, by;
    source_of(d.orient, d.cw, d.ch, ox, oy, ax, ay);
    source_of(d.orient, d.cw, d.ch, ox + tw - 1, oy + th - 1, bx, by);
    const uint32_t sx0 = min(ax, bx), sy0 = min(ay, by);
    const uint32_t sw = max(ax, bx) - sx0 + 1, sh = max(ay, by) - sy0 + 1;
    if constexpr (VEC && sizeof(T) == 4) {
        if (tw == OT && th == OT) {
            /* sw == sh == 64 too, and sx0 is a multiple of 4. */
            const uint32_t t = threadIdx.x;
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t i = t + OLANES * j, r = i / (OT / 4), c = 4 * (i % (OT / 4));
                const float4 v = *reinterpret_cast<const float4 *>(
                    in + (size_t)(d.top + sy0 + r) * d.in_stride + d.left + sx0 + c);
                tile[r][c] = __float_as_uint(v.x);
                tile[r][c + 1] = __float_as_uint(v.y);
                tile[r][c + 2] = __float_as_uint(v.z);
                tile[r][c + 3] = __float_as_uint(v.w);
            }
            __syncthreads();
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t i = t + OLANES * j, r = i / (OT / 4), c = 4 * (i % (OT / 4));
                uint32_t o[4];
#pragma unroll
                for (uint32_t k = 0; k < 4; k++) {
                    uint32_t sx, sy;
                    source_of(d.orient, d.cw, d.ch, ox + c + k, oy + r, sx, sy);
                    o[k] = tile[sy - sy0][sx - sx0];
                }
                *reinterpret_cast<float4 *>(out + (size_t)(oy + r) * d.out_stride + ox + c) =
                    make_float4(__uint_as_float(o[0]), __uint_as_float(o[1]),
                                __uint_as_float(o[2]), __uint_as_float(o[3]));
            }
            return;
        }
    }
    orient_tile<T>(d, in, out, tile, ox, oy, tw, th, sx0, sy0, sw, sh);
}

hipError_t launch_orient(const OrientDesc &d, int element_bytes, hipStream_t s)
{
    dim3 grid((d.ow + OT - 1) / OT, (d.oh + OT - 1) / OT, d.planes);
    bool vec = element_bytes == 4 && d.in_stride % 4 == 0 && d.out_stride % 4 == 0 &&
               d.left % 4 == 0 && d.cw % 4 == 0;
    for (uint32_t p = 0; p < d.planes; p++)
        vec = vec && ((uintptr_t)d.in[p] & 15) == 0 && ((uintptr_t)d.out[p] & 15) == 0;
    if (element_bytes == 4 && vec)
        hipLaunchKernelGGL((k_orient<float, true>), grid, dim3(OLANES), 0, s, d);
    else if (element_bytes == 4)
        hipLaunchKernelGGL((k_orient<float, false>), grid, dim3(OLANES), 0, s, d);
    else
        hipLaunchKernelGGL((k_orient<uint16_t, false>), grid, dim3(OLANES), 0, s, d);
    return hipGetLastError();
}

} // namespace prr
