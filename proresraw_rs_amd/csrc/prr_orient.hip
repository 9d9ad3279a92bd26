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
__global__ void __launch_bounds__(OLANES) k_orient(OrientDesc d)
{
    __shared__ uint32_t tile[OT][OT + 1];
    const T *__restrict__ in = static_cast<const T *>(d.in[blockIdx.z]);
    T *__restrict__ out = static_cast<T *>(d.out[blockIdx.z]);
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const uint32_t ox = blockIdx.x * OT, oy = blockIdx.y * OT;
    const uint32_t tw = min((uint32_t)OT, d.ow - ox), th = min((uint32_t)OT, d.oh - oy);
    /* The source block: each source coordinate follows one output
     * coordinate with slope +-1, so its range comes from two corners. */
    uint32_t ax, ay, bx, by;
    source_of(d.orient, d.cw, d.ch, ox, oy, ax, ay);
    source_of(d.orient, d.cw, d.ch, ox + tw - 1, oy + th - 1, bx, by);
    const uint32_t sx0 = min(ax, bx), sy0 = min(ay, by);
    const uint32_t sw = max(ax, bx) - sx0 + 1, sh = max(ay, by) - sy0 + 1;
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

hipError_t launch_orient(const OrientDesc &d, int element_bytes, hipStream_t s)
{
    dim3 grid((d.ow + OT - 1) / OT, (d.oh + OT - 1) / OT, d.planes);
    if (element_bytes == 4)
        hipLaunchKernelGGL(k_orient<float>, grid, dim3(OLANES), 0, s, d);
    else
        hipLaunchKernelGGL(k_orient<uint16_t>, grid, dim3(OLANES), 0, s, d);
    return hipGetLastError();
}

} // namespace prr
