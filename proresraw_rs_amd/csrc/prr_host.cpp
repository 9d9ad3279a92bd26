/* Host layer of proresraw_amd: the device context and the C-ABI entry
 * points of include/proresraw_amd.h that run on the device (frame structure
 * parsing is prr_parse.cpp). Everything per tile or per sensel runs in the
 * HIP kernels of prr_kernels.hip. No path here computes a decode or develop
 * on the CPU: without a usable HIP device the entry points fail with
 * PRR_E_DEVICE.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "prr_fail.h"
#include "prr_internal.h"

using namespace prr;

/* ------------------------------------------------------------------ */
/* Errors. */

static int device_fail(prr_error *err, hipError_t e, const char *what)
{
    return fail(err, PRR_E_DEVICE, what, (uint64_t)e, 0);
}

#define HIP_TRY(call, what)                                                                        \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return device_fail(err, e_, what);                                                     \
    } while (0)


/* ------------------------------------------------------------------ */
/* Context: a stream and grow-only device scratch. */

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct prr_context {
    int device = 0;
    hipStream_t stream = nullptr;
    DecodeLaunch launch{8, 65536, 256};
    /* The decode width follows the frame (frame_launch) unless the caller
     * fixed it; `in_flight` frames' decodes share the chip (2 in a
     * pipeline's decode contexts). */
    bool spw_auto = true;
    uint32_t in_flight = 1;
    uint32_t simds = 1024;
    DevBuf packet, offsets, queue, status, flags, samples, out0, out1, out2;
    DevBuf lj_raw, lj_records, lj_values, lj_out, lj_meta; /* DNG tile streams */
    std::vector<uint8_t> host_flags;
    void *noalloc = nullptr; /* this device's static scratch (prr_decode_frame_device_noalloc) */
};

static int ensure(prr_context *ctx, DevBuf &b, size_t bytes, prr_error *err)
{
    if (b.bytes >= bytes && b.ptr)
        return PRR_OK;
    if (b.ptr) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    HIP_TRY(hipMalloc(&b.ptr, want), "hipMalloc");
    b.bytes = want;
    return PRR_OK;
}

extern "C" int prr_context_create(int device, prr_context **out, prr_error *err)
{
    if (!out)
        return fail(err, PRR_E_INVALID_ARGUMENT, "out");
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count)
        return fail(err, PRR_E_DEVICE, "no such device", (uint64_t)device, (uint64_t)count);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    prr_context *ctx = new prr_context();
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return device_fail(err, e, "hipStreamCreate");
    }
    /* The allocation-free decode's static scratch: its address on this
     * device, looked up once per device with the device current. */
    e = noalloc_scratch(device, &ctx->noalloc);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return device_fail(err, e, "hipGetSymbolAddress");
    }
    ctx->launch.lds_limit = (uint32_t)std::min<size_t>(prop.sharedMemPerBlock, 65536);
    ctx->simds = 4u * (uint32_t)prop.multiProcessorCount;
    if (const char *v = std::getenv("PRR_STREAMS_PER_WG")) {
        const uint32_t spw = (uint32_t)std::atoi(v);
        if (spw == 8 || spw == 16 || spw == 32 || spw == 64) {
            ctx->launch.streams_per_wg = spw;
            ctx->spw_auto = false;
        }
    }
    *out = ctx;
    return PRR_OK;
}

extern "C" void prr_context_destroy(prr_context *ctx)
{
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (DevBuf *b : {&ctx->packet, &ctx->offsets, &ctx->queue, &ctx->status, &ctx->flags,
                      &ctx->samples, &ctx->out0, &ctx->out1, &ctx->out2, &ctx->lj_raw,
                      &ctx->lj_records, &ctx->lj_values, &ctx->lj_out, &ctx->lj_meta})
        if (b->ptr)
            (void)hipFree(b->ptr);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" void *prr_context_stream(prr_context *ctx) { return ctx ? ctx->stream : nullptr; }

int prr::context_device(const prr_context *ctx) { return ctx->device; }

/* Decode launch parameter: streams per workgroup (8, 16, 32 or 64); 0
 * returns to the automatic width. */
extern "C" int prr_set_decode_config(prr_context *ctx, uint32_t streams_per_wg, prr_error *err)
{
    if (!ctx)
        return fail(err, PRR_E_INVALID_ARGUMENT, "ctx");
    if (streams_per_wg == 0) {
        ctx->spw_auto = true;
        return PRR_OK;
    }
    if (streams_per_wg != 8 && streams_per_wg != 16 && streams_per_wg != 32 &&
        streams_per_wg != 64)
        return fail(err, PRR_E_INVALID_ARGUMENT, "streams_per_wg", streams_per_wg);
    ctx->launch.streams_per_wg = streams_per_wg;
    ctx->spw_auto = false;
    return PRR_OK;
}

/* The decode launch of a frame of nb_tiles tiles. The automatic width is
 * the narrowest of 8, 16, 32 and 64 streams per workgroup that keeps the
 * decode's waves at 1.25 a SIMD or fewer, counting the streams of every
 * frame decoding at once: a wave's
 * loop runs as long as its longest stream, so wider groups issue fewer
 * wave instructions for the same symbols (the streams of a frame are of
 * similar length), while the waves left must still hide the loop's
 * latency from one another. On 5760x3240 frames (36,540 streams) this is
 * 32 alone and 64 with two frames decoding, the fastest widths measured
 * for each. */
static DecodeLaunch frame_launch(const prr_context *ctx, uint32_t nb_tiles)
{
    DecodeLaunch l = ctx->launch;
    if (ctx->spw_auto) {
        const uint64_t streams = 4ull * nb_tiles * ctx->in_flight;
        const uint64_t max_waves = (uint64_t)ctx->simds * 5 / 4;
        uint32_t spw = 8;
        while (spw < 64 && streams / spw > max_waves)
            spw <<= 1;
        l.streams_per_wg = spw;
    }
    return l;
}

/* ------------------------------------------------------------------ */
/* Device-resident entry points. */

static FrameDesc frame_desc(const prr_header *hdr, const prr_frame_info *info,
                            const uint8_t *d_packet, size_t len, uint64_t *d_offsets)
{
    FrameDesc f;
    std::memset(&f, 0, sizeof f);
    f.packet = d_packet;
    f.len = len;
    f.offsets = d_offsets;
    f.table_off = (uint32_t)info->header_end;
    f.nb_tiles = info->nb_tiles;
    f.tiles_per_row = info->tiles_per_row;
    f.columns = ((uint32_t)hdr->width + 15) / 16;
    f.alignment = hdr->alignment;
    f.coded_width = info->coded_width;
    f.coded_height = info->coded_height;
    f.width = hdr->width;
    f.height = hdr->height;
    f.bayer = hdr->bayer;
    std::memcpy(f.qmat, hdr->quant_matrix, 64);
    std::memcpy(f.curve, hdr->curve, 16);
    return f;
}

static int check_device_args(prr_context *ctx, const prr_header *hdr, const prr_frame_info *info,
                             const uint8_t *d_packet, prr_device_status *d_status,
                             prr_error *err)
{
    if (!ctx || !hdr || !info || !d_packet || !d_status)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    if ((uintptr_t)d_packet & 3)
        return fail(err, PRR_E_INVALID_ARGUMENT, "packet alignment");
    return PRR_OK;
}

extern "C" int prr_decode_frame_device(prr_context *ctx, const prr_header *hdr,
                                       const prr_frame_info *info, const uint8_t *d_packet,
                                       size_t len, int idct, uint16_t *d_samples,
                                       uint16_t *d_linear, uint8_t *d_tile_flags,
                                       prr_device_status *d_status, void *stream,
                                       prr_error *err)
{
    int rc = check_device_args(ctx, hdr, info, d_packet, d_status, err);
    if (rc)
        return rc;
    if (!d_samples || (idct != PRR_IDCT_APPLE && idct != PRR_IDCT_FFMPEG))
        return fail(err, PRR_E_INVALID_ARGUMENT, "samples or idct");
    if (((uintptr_t)d_samples & 15) || ((uintptr_t)d_linear & 3))
        return fail(err, PRR_E_INVALID_ARGUMENT, "samples or linear alignment");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    rc = ensure(ctx, ctx->offsets, sizeof(uint64_t) * ((size_t)info->nb_tiles + 1), err);
    if (!rc)
        rc = ensure(ctx, ctx->queue, sizeof(uint32_t) * std::max<size_t>(info->nb_tiles, 1), err);
    if (rc)
        return rc;
    FrameDesc f = frame_desc(hdr, info, d_packet, len, (uint64_t *)ctx->offsets.ptr);
    HIP_TRY(launch_decode(f, idct, d_samples, d_linear, d_tile_flags, d_status,
                          (uint32_t *)ctx->queue.ptr, frame_launch(ctx, info->nb_tiles), s),
            "decode kernels");
    return PRR_OK;
}

extern "C" int prr_decode_frame_device_noalloc(prr_context *ctx, const prr_header *hdr,
                                               const prr_frame_info *info,
                                               const uint8_t *d_packet, size_t len, int idct,
                                               uint16_t *d_samples, uint16_t *d_linear,
                                               uint8_t *d_tile_flags,
                                               prr_device_status *d_status, void *stream,
                                               prr_error *err)
{
    int rc = check_device_args(ctx, hdr, info, d_packet, d_status, err);
    if (rc)
        return rc;
    if (!d_samples || (idct != PRR_IDCT_APPLE && idct != PRR_IDCT_FFMPEG))
        return fail(err, PRR_E_INVALID_ARGUMENT, "samples or idct");
    if (((uintptr_t)d_samples & 15) || ((uintptr_t)d_linear & 3))
        return fail(err, PRR_E_INVALID_ARGUMENT, "samples or linear alignment");
    if (info->nb_tiles > PRR_NOALLOC_MAX_TILES)
        return fail(err, PRR_E_INVALID_ARGUMENT, "tiles above static scratch", info->nb_tiles,
                    PRR_NOALLOC_MAX_TILES);
    if (!ctx->noalloc)
        return fail(err, PRR_E_DEVICE, "no static scratch");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    /* The static scratch, by the address looked up for this device: the
     * offsets, then the tile order. */
    uint64_t *offsets = (uint64_t *)ctx->noalloc;
    uint32_t *order = (uint32_t *)((uint8_t *)ctx->noalloc + NOALLOC_OFFSETS_BYTES);
    FrameDesc f = frame_desc(hdr, info, d_packet, len, offsets);
    HIP_TRY(launch_decode(f, idct, d_samples, d_linear, d_tile_flags, d_status, order,
                          frame_launch(ctx, info->nb_tiles), (hipStream_t)stream),
            "decode kernels");
    return PRR_OK;
}

extern "C" int prr_decode_frame_reduced_device(prr_context *ctx, const prr_header *hdr,
                                               const prr_frame_info *info,
                                               const uint8_t *d_packet, size_t len, int scale,
                                               int greens, uint16_t *d_rgb,
                                               uint8_t *d_tile_flags,
                                               prr_device_status *d_status, void *stream,
                                               prr_error *err)
{
    int rc = check_device_args(ctx, hdr, info, d_packet, d_status, err);
    if (rc)
        return rc;
    if (!d_rgb || scale < PRR_SCALE_HALF || scale > PRR_SCALE_EIGHTH ||
        (greens != PRR_GREENS_FIRST && greens != PRR_GREENS_AVERAGE))
        return fail(err, PRR_E_INVALID_ARGUMENT, "rgb, scale or greens");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    rc = ensure(ctx, ctx->offsets, sizeof(uint64_t) * ((size_t)info->nb_tiles + 1), err);
    if (!rc)
        rc = ensure(ctx, ctx->queue, sizeof(uint32_t) * std::max<size_t>(info->nb_tiles, 1), err);
    if (rc)
        return rc;
    FrameDesc f = frame_desc(hdr, info, d_packet, len, (uint64_t *)ctx->offsets.ptr);
    HIP_TRY(launch_decode_reduced(f, scale, greens, d_rgb, d_tile_flags, d_status,
                                  (uint32_t *)ctx->queue.ptr, frame_launch(ctx, info->nb_tiles),
                                  s),
            "reduced kernels");
    return PRR_OK;
}

extern "C" int prr_linearize_device(prr_context *ctx, const prr_header *hdr,
                                    const uint16_t *d_samples, uint16_t *d_out, void *stream,
                                    prr_error *err)
{
    if (!ctx || !hdr || !d_samples || !d_out)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    uint32_t stride = ((uint32_t)hdr->width + 15) & ~15u;
    HIP_TRY(launch_linearize(d_samples, stride, hdr->width, hdr->height, hdr->curve, d_out, s),
            "linearize kernel");
    return PRR_OK;
}

/* Conversion::new (demosaic.rs:286-304): the stage folded into the params. */
DevelopDesc prr::develop_desc(const prr_header *hdr, const uint16_t *d_samples,
                              const prr_demosaic_params *p, int stage)
{
    DevelopDesc d;
    std::memset(&d, 0, sizeof d);
    d.samples = d_samples;
    d.stride = ((uint32_t)hdr->width + 15) & ~15u;
    d.width = hdr->width;
    d.height = hdr->height;
    d.cfa_phase = p->cfa_phase;
    d.black = p->black;
    d.one_over_range = p->one_over_range;
    for (int i = 0; i < 3; i++)
        d.noise[i] = p->noise[i];
    d.clip[0] = p->clip[0];
    d.clip[1] = p->clip[1];
    d.black_over_range = p->black_over_range;
    d.wb[0] = p->wb[0];
    d.wb[1] = p->wb[1];
    if (stage == PRR_STAGE_DEMOSAICED) {
        d.wb[0] = d.wb[1] = 1.0f;
        d.clip[0] = d.clip[1] = 1.0f;
    }
    if (stage == PRR_STAGE_REC2020) {
        d.has_matrix = 1;
        prr_output_matrix(hdr, PRR_EXPOSURE_NORMALIZED, d.m);
    }
    std::memcpy(d.curve, hdr->curve, 16);
    return d;
}

/* develop(&frame, stage) with the noise model taken on the device from the
 * frame's status (its max_qscale), so no readback is needed between decode
 * and develop. explicit_wb selects with_white_balance(red, blue) over the
 * header's gains. */
DevelopDesc prr::develop_desc_device(const prr_header *hdr, const uint16_t *d_samples,
                                     const prr_device_status *d_status, int explicit_wb,
                                     float red, float blue, int stage)
{
    prr_demosaic_params p;
    if (explicit_wb)
        prr_demosaic_params_with_white_balance(hdr, 0, red, blue, &p);
    else
        prr_demosaic_params_new(hdr, 0, &p);
    DevelopDesc d = develop_desc(hdr, d_samples, &p, stage);
    d.qstatus = d_status;
    d.qmat0 = hdr->quant_matrix[0];
    d.range = (uint16_t)(hdr->white_level - 256u);
    return d;
}

extern "C" int prr_develop_frame_device(prr_context *ctx, const prr_header *hdr,
                                        const uint16_t *d_samples,
                                        const prr_device_status *d_status, int explicit_wb,
                                        float wb_red, float wb_blue, int stage, float *d_r,
                                        float *d_g, float *d_b, void *stream, prr_error *err)
{
    if (!ctx || !hdr || !d_samples || !d_status || !d_r)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    if ((d_g == nullptr) != (d_b == nullptr))
        return fail(err, PRR_E_INVALID_ARGUMENT, "planes");
    if (stage < PRR_STAGE_DEMOSAICED || stage > PRR_STAGE_REC2020)
        return fail(err, PRR_E_INVALID_ARGUMENT, "stage", (uint64_t)stage);
    if (((uintptr_t)d_r | (uintptr_t)d_g | (uintptr_t)d_b) & 7)
        return fail(err, PRR_E_INVALID_ARGUMENT, "plane alignment");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    DevelopDesc d = develop_desc_device(hdr, d_samples, d_status, explicit_wb, wb_red, wb_blue,
                                        stage);
    HIP_TRY(launch_develop(d, d_r, d_g, d_b, s), "develop kernel");
    return PRR_OK;
}

/* develop(&frame, stage) from the linear picture; the develop kernel also
 * copies the frame's status to status_out when that is set. */
static int develop_linear(prr_context *ctx, const prr_header *hdr, const uint16_t *d_linear,
                          const prr_device_status *d_status, int explicit_wb, float wb_red,
                          float wb_blue, int stage, float *d_r, float *d_g, float *d_b,
                          prr_device_status *status_out, void *stream, prr_error *err);

extern "C" int prr_develop_linear_device(prr_context *ctx, const prr_header *hdr,
                                         const uint16_t *d_linear,
                                         const prr_device_status *d_status, int explicit_wb,
                                         float wb_red, float wb_blue, int stage, float *d_r,
                                         float *d_g, float *d_b, void *stream, prr_error *err)
{
    return develop_linear(ctx, hdr, d_linear, d_status, explicit_wb, wb_red, wb_blue, stage, d_r,
                          d_g, d_b, nullptr, stream, err);
}

static int develop_linear(prr_context *ctx, const prr_header *hdr, const uint16_t *d_linear,
                          const prr_device_status *d_status, int explicit_wb, float wb_red,
                          float wb_blue, int stage, float *d_r, float *d_g, float *d_b,
                          prr_device_status *status_out, void *stream, prr_error *err)
{
    if (!ctx || !hdr || !d_linear || !d_status || !d_r)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    if ((d_g == nullptr) != (d_b == nullptr))
        return fail(err, PRR_E_INVALID_ARGUMENT, "planes");
    if (stage < PRR_STAGE_DEMOSAICED || stage > PRR_STAGE_REC2020)
        return fail(err, PRR_E_INVALID_ARGUMENT, "stage", (uint64_t)stage);
    if (((uintptr_t)d_r | (uintptr_t)d_g | (uintptr_t)d_b) & 7)
        return fail(err, PRR_E_INVALID_ARGUMENT, "plane alignment");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    DevelopDesc d = develop_desc_device(hdr, d_linear, d_status, explicit_wb, wb_red, wb_blue,
                                        stage);
    d.stride = hdr->width;
    d.linear_input = 1;
    d.status_out = status_out;
    HIP_TRY(launch_develop(d, d_r, d_g, d_b, s), "develop kernel");
    return PRR_OK;
}

extern "C" int prr_develop_device(prr_context *ctx, const prr_header *hdr,
                                  const uint16_t *d_samples, const prr_demosaic_params *params,
                                  int stage, float *d_r, float *d_g, float *d_b, void *stream,
                                  prr_error *err)
{
    if (!ctx || !hdr || !d_samples || !params || !d_r)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    if ((d_g == nullptr) != (d_b == nullptr))
        return fail(err, PRR_E_INVALID_ARGUMENT, "planes");
    if (stage < PRR_STAGE_DEMOSAICED || stage > PRR_STAGE_REC2020)
        return fail(err, PRR_E_INVALID_ARGUMENT, "stage", (uint64_t)stage);
    if (((uintptr_t)d_r | (uintptr_t)d_g | (uintptr_t)d_b) & 7)
        return fail(err, PRR_E_INVALID_ARGUMENT, "plane alignment");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    DevelopDesc d = develop_desc(hdr, d_samples, params, stage);
    HIP_TRY(launch_develop(d, d_r, d_g, d_b, s), "develop kernel");
    return PRR_OK;
}

/* ------------------------------------------------------------------ */
/* Host-buffer entry points: copy in, run the device path, copy out. */

static int upload_packet(prr_context *ctx, const uint8_t *packet, size_t len, prr_error *err)
{
    size_t padded = ((len + 3) & ~(size_t)3) + 16;
    int rc = ensure(ctx, ctx->packet, padded, err);
    if (rc)
        return rc;
    HIP_TRY(hipMemcpyAsync(ctx->packet.ptr, packet, len, hipMemcpyHostToDevice, ctx->stream),
            "hipMemcpy packet");
    return PRR_OK;
}

/* Turn the device status into the crate's errors and the skipped list. */
static int finish_frame(prr_context *ctx, const prr_frame_info *info, uint32_t *skipped,
                        size_t skipped_cap, size_t *n_skipped, uint8_t *max_qscale,
                        prr_error *err)
{
    prr_device_status st;
    ctx->host_flags.resize(info->nb_tiles);
    HIP_TRY(hipMemcpyAsync(&st, ctx->status.ptr, sizeof st, hipMemcpyDeviceToHost, ctx->stream),
            "hipMemcpy status");
    if (info->nb_tiles)
        HIP_TRY(hipMemcpyAsync(ctx->host_flags.data(), ctx->flags.ptr, info->nb_tiles,
                               hipMemcpyDeviceToHost, ctx->stream),
                "hipMemcpy flags");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (st.oob_tile != 0xFFFFFFFFu)
        return fail(err, PRR_E_TILE_OUT_OF_BOUNDS, nullptr, st.oob_tile);
    if (st.overrun)
        return fail(err, PRR_E_DECODE_OVERRUN, "symbol bound");
    size_t k = 0;
    for (uint32_t t = 0; t < info->nb_tiles; t++)
        if (ctx->host_flags[t]) {
            if (k >= skipped_cap)
                return fail(err, PRR_E_OUTPUT_SIZE, "skipped capacity", k + 1, skipped_cap);
            skipped[k++] = t;
        }
    if (n_skipped)
        *n_skipped = k;
    if (max_qscale)
        *max_qscale = (uint8_t)st.max_qscale;
    return PRR_OK;
}

extern "C" int prr_decode_frame(prr_context *ctx, const uint8_t *packet, size_t len, int idct,
                                prr_header *hdr, uint16_t *samples, size_t samples_len,
                                uint32_t *skipped, size_t skipped_cap, size_t *n_skipped,
                                uint8_t *max_qscale, prr_error *err)
{
    if (!ctx || !hdr || (!samples && samples_len))
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    prr_frame_info info;
    int rc = prr_parse_frame_info(packet, len, hdr, &info, err);
    if (rc)
        return rc;
    size_t n = (size_t)info.coded_width * info.coded_height;
    if (samples_len != n)
        return fail(err, PRR_E_OUTPUT_SIZE, "output size", n, samples_len);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    if ((rc = upload_packet(ctx, packet, len, err)) ||
        (rc = ensure(ctx, ctx->status, sizeof(prr_device_status), err)) ||
        (rc = ensure(ctx, ctx->flags, info.nb_tiles + 1, err)) ||
        (rc = ensure(ctx, ctx->samples, n * 2 + 16, err)))
        return rc;
    rc = prr_decode_frame_device(ctx, hdr, &info, (const uint8_t *)ctx->packet.ptr, len, idct,
                                 (uint16_t *)ctx->samples.ptr, nullptr,
                                 (uint8_t *)ctx->flags.ptr,
                                 (prr_device_status *)ctx->status.ptr, ctx->stream, err);
    if (rc)
        return rc;
    rc = finish_frame(ctx, &info, skipped, skipped_cap, n_skipped, max_qscale, err);
    if (rc)
        return rc;
    HIP_TRY(hipMemcpy(samples, ctx->samples.ptr, n * 2, hipMemcpyDeviceToHost), "hipMemcpy samples");
    return PRR_OK;
}

extern "C" int prr_decode_frame_reduced(prr_context *ctx, const uint8_t *packet, size_t len,
                                        int scale, int greens, prr_header *hdr, uint16_t *rgb,
                                        size_t rgb_len, uint32_t *skipped, size_t skipped_cap,
                                        size_t *n_skipped, uint8_t *max_qscale, prr_error *err)
{
    if (!ctx || !hdr || (!rgb && rgb_len))
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    if (scale < PRR_SCALE_HALF || scale > PRR_SCALE_EIGHTH)
        return fail(err, PRR_E_INVALID_ARGUMENT, "scale", (uint64_t)scale);
    prr_frame_info info;
    int rc = prr_parse_frame_info(packet, len, hdr, &info, err);
    if (rc)
        return rc;
    size_t n = 3 * (size_t)(info.coded_width >> scale) * (info.coded_height >> scale);
    if (rgb_len != n)
        return fail(err, PRR_E_OUTPUT_SIZE, "output size", n, rgb_len);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    if ((rc = upload_packet(ctx, packet, len, err)) ||
        (rc = ensure(ctx, ctx->status, sizeof(prr_device_status), err)) ||
        (rc = ensure(ctx, ctx->flags, info.nb_tiles + 1, err)) ||
        (rc = ensure(ctx, ctx->out0, n * 2 + 16, err)))
        return rc;
    rc = prr_decode_frame_reduced_device(ctx, hdr, &info, (const uint8_t *)ctx->packet.ptr, len,
                                         scale, greens, (uint16_t *)ctx->out0.ptr,
                                         (uint8_t *)ctx->flags.ptr,
                                         (prr_device_status *)ctx->status.ptr, ctx->stream, err);
    if (rc)
        return rc;
    rc = finish_frame(ctx, &info, skipped, skipped_cap, n_skipped, max_qscale, err);
    if (rc)
        return rc;
    HIP_TRY(hipMemcpy(rgb, ctx->out0.ptr, n * 2, hipMemcpyDeviceToHost), "hipMemcpy rgb");
    return PRR_OK;
}

extern "C" int prr_linearize_into(prr_context *ctx, const prr_header *hdr,
                                  const uint16_t *samples, size_t samples_len, uint16_t *out,
                                  size_t out_len, prr_error *err)
{
    if (!ctx || !hdr || !samples || (!out && out_len))
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    size_t cw = ((size_t)hdr->width + 15) & ~(size_t)15, ch = ((size_t)hdr->height + 15) & ~(size_t)15;
    if (samples_len != cw * ch)
        return fail(err, PRR_E_INVALID_ARGUMENT, "samples size", cw * ch, samples_len);
    size_t n = (size_t)hdr->width * hdr->height;
    if (out_len != n)
        return fail(err, PRR_E_OUTPUT_SIZE, "output size", n, out_len);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    int rc;
    if ((rc = ensure(ctx, ctx->samples, cw * ch * 2 + 16, err)) ||
        (rc = ensure(ctx, ctx->out0, n * 2 + 16, err)))
        return rc;
    HIP_TRY(hipMemcpyAsync(ctx->samples.ptr, samples, cw * ch * 2, hipMemcpyHostToDevice,
                           ctx->stream),
            "hipMemcpy samples");
    rc = prr_linearize_device(ctx, hdr, (const uint16_t *)ctx->samples.ptr,
                              (uint16_t *)ctx->out0.ptr, ctx->stream, err);
    if (rc)
        return rc;
    HIP_TRY(hipMemcpyAsync(out, ctx->out0.ptr, n * 2, hipMemcpyDeviceToHost, ctx->stream),
            "hipMemcpy linear");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return PRR_OK;
}

static int develop_host(prr_context *ctx, const prr_header *hdr, const uint16_t *samples,
                        size_t samples_len, const prr_demosaic_params *params, int stage,
                        float *r, float *g, float *b, size_t out_len, int planar,
                        prr_error *err)
{
    if (!ctx || !hdr || !samples || !params || !r || (planar && (!g || !b)))
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    size_t cw = ((size_t)hdr->width + 15) & ~(size_t)15, ch = ((size_t)hdr->height + 15) & ~(size_t)15;
    if (samples_len != cw * ch)
        return fail(err, PRR_E_INVALID_ARGUMENT, "samples size", cw * ch, samples_len);
    size_t n = (size_t)hdr->width * hdr->height;
    if (out_len != (planar ? n : 3 * n))
        return fail(err, PRR_E_OUTPUT_SIZE, planar ? "plane size" : "output size",
                    planar ? n : 3 * n, out_len);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    int rc;
    if ((rc = ensure(ctx, ctx->samples, cw * ch * 2 + 16, err)) ||
        (rc = ensure(ctx, ctx->out0, (planar ? n : 3 * n) * 4 + 16, err)) ||
        (planar && (rc = ensure(ctx, ctx->out1, n * 4 + 16, err))) ||
        (planar && (rc = ensure(ctx, ctx->out2, n * 4 + 16, err))))
        return rc;
    HIP_TRY(hipMemcpyAsync(ctx->samples.ptr, samples, cw * ch * 2, hipMemcpyHostToDevice,
                           ctx->stream),
            "hipMemcpy samples");
    rc = prr_develop_device(ctx, hdr, (const uint16_t *)ctx->samples.ptr, params, stage,
                            (float *)ctx->out0.ptr, planar ? (float *)ctx->out1.ptr : nullptr,
                            planar ? (float *)ctx->out2.ptr : nullptr, ctx->stream, err);
    if (rc)
        return rc;
    if (planar) {
        HIP_TRY(hipMemcpyAsync(r, ctx->out0.ptr, n * 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy r");
        HIP_TRY(hipMemcpyAsync(g, ctx->out1.ptr, n * 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy g");
        HIP_TRY(hipMemcpyAsync(b, ctx->out2.ptr, n * 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy b");
    } else {
        HIP_TRY(hipMemcpyAsync(r, ctx->out0.ptr, 3 * n * 4, hipMemcpyDeviceToHost, ctx->stream),
                "hipMemcpy rgb");
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return PRR_OK;
}

extern "C" int prr_develop_planar_into(prr_context *ctx, const prr_header *hdr,
                                       const uint16_t *samples, size_t samples_len,
                                       const prr_demosaic_params *params, int stage, float *r,
                                       float *g, float *b, size_t plane_len, prr_error *err)
{
    return develop_host(ctx, hdr, samples, samples_len, params, stage, r, g, b, plane_len, 1, err);
}

extern "C" int prr_develop_into(prr_context *ctx, const prr_header *hdr, const uint16_t *samples,
                                size_t samples_len, const prr_demosaic_params *params, int stage,
                                float *rgb, size_t rgb_len, prr_error *err)
{
    return develop_host(ctx, hdr, samples, samples_len, params, stage, rgb, nullptr, nullptr,
                        rgb_len, 0, err);
}

/* ------------------------------------------------------------------ */
/* Batches: many frames, one persistent decode launch. */

/* What a frame's tile-group layout depends on, kept so the layout can be
 * redone when the streams per workgroup change. */
struct BatchShape {
    uint32_t per_block, alignment, nb_tiles, full;
};

struct prr_batch {
    prr_context *ctx = nullptr;
    std::vector<BatchFrame> frames;
    std::vector<BatchShape> shapes;
    std::vector<BatchDevelop> jobs;
    std::vector<size_t> offset_base; /* per frame, in u64 entries */
    size_t offsets_total = 0;
    std::vector<uint32_t> chunks; /* (frame << 16) | chunk of the size-table scan */
    DevBuf d_frames, d_jobs, d_queue, d_head, d_hist, d_offsets, d_chunks, d_chunk_bytes;
    uint32_t n_items = 0;
    uint32_t dev_tiles = 0; /* develop tiles of the jobs so far */
    uint32_t spw = 64, chunk = 1, waves_per_cu = 0;
    /* LDS a batch workgroup may use: 32 KiB keeps every alignment's group at
     * or below 18 KiB, so 8 persistent workgroups fit a CU (measured best on
     * the mixed batch: 60.8 Gsensels/s against 53.1 at 64 KiB and 56.5 at
     * 16 KiB). */
    uint32_t lds_budget = 32768;
    size_t lds = 0;
    uint32_t grid = 0;
    bool uploaded = false;
};

/* ------------------------------------------------------------------ */
/* Frame pipeline: decode of frame i + 1 beside develop of frame i. */

namespace {
struct PipeSet {
    DevBuf samples, linear, flags, status;
    hipEvent_t decoded = nullptr, developed = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr, t2 = nullptr, t3 = nullptr; /* timing */
    bool used = false, timed = false;
};
} // namespace

/* Decode contexts of a pipeline: consecutive frames decode on alternate
 * ones, so two frames' decodes overlap (a frame alone has too few streams
 * to fill the chip with wide decode waves). */
enum { PIPE_DECODERS = 2 };

struct prr_pipeline {
    prr_context *ctx = nullptr;   /* the caller's: device and launch shape */
    /* Decode streams and their scratch; frame k decodes on dctx[k % 2]. */
    prr_context *dctx[PIPE_DECODERS] = {nullptr, nullptr};
    hipStream_t develop = nullptr;
    hipEvent_t caller = nullptr;  /* recorded on a caller's stream by prr_pipeline_wait */
    std::vector<PipeSet> sets;
    uint64_t submitted = 0;
    bool timing = false;
    double decode_ms = 0, develop_ms = 0;
    uint64_t timed_frames = 0;
};

static void pipeline_release(prr_pipeline *p)
{
    for (prr_context *d : p->dctx)
        if (d)
            (void)hipStreamSynchronize(d->stream);
    if (p->develop)
        (void)hipStreamSynchronize(p->develop);
    for (PipeSet &st : p->sets) {
        for (DevBuf *b : {&st.samples, &st.linear, &st.flags, &st.status})
            if (b->ptr)
                (void)hipFree(b->ptr);
        for (hipEvent_t e : {st.decoded, st.developed, st.t0, st.t1, st.t2, st.t3})
            if (e)
                (void)hipEventDestroy(e);
    }
    if (p->caller)
        (void)hipEventDestroy(p->caller);
    if (p->develop)
        (void)hipStreamDestroy(p->develop);
    for (prr_context *d : p->dctx)
        if (d)
            prr_context_destroy(d);
    delete p;
}

/* A stream on a hardware queue of its own: the runtime shares its few
 * hardware queues among ordinary streams (two pipeline streams can land on
 * one queue and then run one after the other), while a stream created with a
 * CU mask gets a queue to itself. The mask enables every CU. */
static hipError_t dedicated_stream(int device, hipStream_t *out)
{
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess)
        return e;
    std::vector<uint32_t> mask(((size_t)prop.multiProcessorCount + 31) / 32, 0xFFFFFFFFu);
    if (prop.multiProcessorCount % 32)
        mask.back() = (1u << (prop.multiProcessorCount % 32)) - 1u;
    return hipExtStreamCreateWithCUMask(out, (uint32_t)mask.size(), mask.data());
}

extern "C" int prr_pipeline_create(prr_context *ctx, uint32_t depth, prr_pipeline **out,
                                   prr_error *err)
{
    if (!ctx || !out)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (depth == 0)
        depth = 3;
    if (depth < 2 || depth > 8)
        return fail(err, PRR_E_INVALID_ARGUMENT, "pipeline depth", depth);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    prr_pipeline *p = new prr_pipeline();
    p->ctx = ctx;
    p->sets.resize(depth);
    hipError_t e = hipSuccess;
    for (prr_context *&d : p->dctx) {
        int rc = prr_context_create(ctx->device, &d, err);
        if (rc) {
            pipeline_release(p);
            return rc;
        }
        d->launch = ctx->launch;
        d->spw_auto = ctx->spw_auto;
        d->in_flight = PIPE_DECODERS;
        /* Decode and develop streams on hardware queues of their own, so
         * they run side by side. */
        hipStream_t decode_stream = nullptr;
        if (e == hipSuccess)
            e = dedicated_stream(ctx->device, &decode_stream);
        if (e == hipSuccess) {
            (void)hipStreamDestroy(d->stream);
            d->stream = decode_stream;
        }
    }
    if (e == hipSuccess)
        e = dedicated_stream(ctx->device, &p->develop);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&p->caller, hipEventDisableTiming);
    for (PipeSet &st : p->sets) {
        for (hipEvent_t *ev : {&st.decoded, &st.developed})
            if (e == hipSuccess)
                e = hipEventCreateWithFlags(ev, hipEventDisableTiming |
                                                        hipEventDisableSystemFence);
        for (hipEvent_t *ev : {&st.t0, &st.t1, &st.t2, &st.t3})
            if (e == hipSuccess)
                e = hipEventCreate(ev);
    }
    if (e != hipSuccess) {
        pipeline_release(p);
        return device_fail(err, e, "pipeline streams");
    }
    *out = p;
    return PRR_OK;
}

extern "C" void prr_pipeline_destroy(prr_pipeline *p)
{
    if (p)
        pipeline_release(p);
}

/* Averages of the frames timed since timing was switched on. */
static int pipeline_collect(prr_pipeline *p, prr_error *err)
{
    for (PipeSet &st : p->sets) {
        if (!st.timed)
            continue;
        HIP_TRY(hipEventSynchronize(st.t3), "hipEventSynchronize");
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, st.t0, st.t1), "hipEventElapsedTime");
        HIP_TRY(hipEventElapsedTime(&b, st.t2, st.t3), "hipEventElapsedTime");
        p->decode_ms += a;
        p->develop_ms += b;
        p->timed_frames++;
        st.timed = false;
    }
    return PRR_OK;
}

extern "C" int prr_pipeline_submit(prr_pipeline *p, const prr_header *hdr,
                                   const prr_frame_info *info, const uint8_t *d_packet, size_t len,
                                   int idct, int explicit_wb, float wb_red, float wb_blue,
                                   int stage, float *d_r, float *d_g, float *d_b,
                                   prr_device_status *d_status_out, prr_error *err)
{
    if (!p || !hdr || !info || !d_packet || !d_r)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    if ((d_g == nullptr) != (d_b == nullptr))
        return fail(err, PRR_E_INVALID_ARGUMENT, "planes");
    if (stage < PRR_STAGE_DEMOSAICED || stage > PRR_STAGE_REC2020)
        return fail(err, PRR_E_INVALID_ARGUMENT, "stage", (uint64_t)stage);
    /* Every argument the decode and the develop check, checked before
     * either is queued: a refused submit queues nothing. */
    if (((uintptr_t)d_r | (uintptr_t)d_g | (uintptr_t)d_b) & 7)
        return fail(err, PRR_E_INVALID_ARGUMENT, "plane alignment");
    if ((uintptr_t)d_packet & 3)
        return fail(err, PRR_E_INVALID_ARGUMENT, "packet alignment");
    if (idct != PRR_IDCT_APPLE && idct != PRR_IDCT_FFMPEG)
        return fail(err, PRR_E_INVALID_ARGUMENT, "samples or idct");
    HIP_TRY(hipSetDevice(p->ctx->device), "hipSetDevice");
    PipeSet &st = p->sets[p->submitted % p->sets.size()];
    prr_context *dctx = p->dctx[p->submitted % PIPE_DECODERS];
    hipStream_t ds = dctx->stream, vs = p->develop;
    int rc;
    if (p->timing && st.timed && (rc = pipeline_collect(p, err)))
        return rc;
    const size_t coded = (size_t)info->coded_width * info->coded_height;
    const size_t visible = (size_t)hdr->width * hdr->height;
    if (st.samples.bytes < coded * 2 || st.linear.bytes < visible * 2 ||
        st.flags.bytes < std::max<size_t>(info->nb_tiles, 1)) {
        /* Growing a set: its last frame must be done with it. */
        if (st.used)
            HIP_TRY(hipEventSynchronize(st.developed), "hipEventSynchronize");
        if ((rc = ensure(dctx, st.samples, coded * 2, err)) ||
            (rc = ensure(dctx, st.linear, visible * 2, err)) ||
            (rc = ensure(dctx, st.flags, std::max<size_t>(info->nb_tiles, 1), err)))
            return rc;
    }
    if ((rc = ensure(dctx, st.status, sizeof(prr_device_status), err)))
        return rc;
    /* The decode waits (on the device) for the develop that last read this
     * set; the develop waits for this decode. */
    if (st.used)
        HIP_TRY(hipStreamWaitEvent(ds, st.developed, 0), "hipStreamWaitEvent");
    const bool timed = p->timing;
    if (timed)
        HIP_TRY(hipEventRecord(st.t0, ds), "hipEventRecord");
    if ((rc = prr_decode_frame_device(dctx, hdr, info, d_packet, len, idct,
                                      (uint16_t *)st.samples.ptr, (uint16_t *)st.linear.ptr,
                                      (uint8_t *)st.flags.ptr, (prr_device_status *)st.status.ptr,
                                      ds, err)))
        return rc;
    if (timed)
        HIP_TRY(hipEventRecord(st.t1, ds), "hipEventRecord");
    HIP_TRY(hipEventRecord(st.decoded, ds), "hipEventRecord");
    HIP_TRY(hipStreamWaitEvent(vs, st.decoded, 0), "hipStreamWaitEvent");
    if (timed)
        HIP_TRY(hipEventRecord(st.t2, vs), "hipEventRecord");
    if ((rc = develop_linear(dctx, hdr, (const uint16_t *)st.linear.ptr,
                             (const prr_device_status *)st.status.ptr, explicit_wb, wb_red,
                             wb_blue, stage, d_r, d_g, d_b, d_status_out, vs, err)))
        return rc;
    if (timed)
        HIP_TRY(hipEventRecord(st.t3, vs), "hipEventRecord");
    HIP_TRY(hipEventRecord(st.developed, vs), "hipEventRecord");
    st.used = true;
    st.timed = timed;
    p->submitted++;
    return PRR_OK;
}

extern "C" int prr_pipeline_wait(prr_pipeline *p, void *stream, prr_error *err)
{
    if (!p)
        return fail(err, PRR_E_INVALID_ARGUMENT, "pipeline");
    HIP_TRY(hipSetDevice(p->ctx->device), "hipSetDevice");
    /* An event at the end of the caller's stream; both pipeline streams
     * wait for it on the device (the wait takes the event as recorded now,
     * so the one event serves every call). */
    HIP_TRY(hipEventRecord(p->caller, (hipStream_t)stream), "hipEventRecord");
    for (prr_context *d : p->dctx)
        HIP_TRY(hipStreamWaitEvent(d->stream, p->caller, 0), "hipStreamWaitEvent");
    HIP_TRY(hipStreamWaitEvent(p->develop, p->caller, 0), "hipStreamWaitEvent");
    return PRR_OK;
}

extern "C" int prr_pipeline_join(prr_pipeline *p, void *stream, prr_error *err)
{
    if (!p)
        return fail(err, PRR_E_INVALID_ARGUMENT, "pipeline");
    HIP_TRY(hipSetDevice(p->ctx->device), "hipSetDevice");
    for (PipeSet &st : p->sets)
        if (st.used)
            HIP_TRY(hipStreamWaitEvent((hipStream_t)stream, st.developed, 0), "hipStreamWaitEvent");
    return PRR_OK;
}

extern "C" int prr_pipeline_sync(prr_pipeline *p, prr_error *err)
{
    if (!p)
        return fail(err, PRR_E_INVALID_ARGUMENT, "pipeline");
    HIP_TRY(hipSetDevice(p->ctx->device), "hipSetDevice");
    HIP_TRY(hipStreamSynchronize(p->develop), "hipStreamSynchronize");
    for (prr_context *d : p->dctx)
        HIP_TRY(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
    return PRR_OK;
}

extern "C" int prr_pipeline_timing(prr_pipeline *p, int enable, float *decode_ms,
                                   float *develop_ms, uint64_t *frames, prr_error *err)
{
    if (!p)
        return fail(err, PRR_E_INVALID_ARGUMENT, "pipeline");
    int rc = pipeline_collect(p, err);
    if (rc)
        return rc;
    if (decode_ms)
        *decode_ms = p->timed_frames ? (float)(p->decode_ms / (double)p->timed_frames) : 0.0f;
    if (develop_ms)
        *develop_ms = p->timed_frames ? (float)(p->develop_ms / (double)p->timed_frames) : 0.0f;
    if (frames)
        *frames = p->timed_frames;
    p->decode_ms = p->develop_ms = 0;
    p->timed_frames = 0;
    p->timing = enable != 0;
    return PRR_OK;
}

/* ------------------------------------------------------------------ */
/* Lossless JPEG tiles of a DNG (dng.rs:92-128, lj92.rs:14-93). */

static uint64_t lj92_tiles(uint32_t width, uint32_t height)
{
    return (uint64_t)((width + LJ92_TILE - 1) / LJ92_TILE) *
           ((height + LJ92_TILE - 1) / LJ92_TILE);
}

extern "C" size_t prr_dng_lj92_capacity(uint32_t width, uint32_t height)
{
    return (size_t)(lj92_tiles(width, height) * LJ92_TILE_BYTES);
}

static int lj92_check(uint32_t width, uint32_t height, uint32_t stride, prr_error *err)
{
    /* dng.rs's padding repeats the last two rows and columns. */
    if (width < 2 || height < 2 || stride < width)
        return fail(err, PRR_E_INVALID_ARGUMENT, "dng picture size", width, height);
    /* Tile offsets are 32-bit, as TIFF's: the worst case must fit. */
    if (lj92_tiles(width, height) * LJ92_TILE_BYTES > UINT32_MAX)
        return fail(err, PRR_E_INVALID_ARGUMENT, "dng picture too large", width, height);
    return PRR_OK;
}

extern "C" int prr_dng_lj92_tiles_device(prr_context *ctx, const uint16_t *d_values,
                                         uint32_t width, uint32_t height, uint32_t stride,
                                         uint8_t *d_out, size_t out_cap, uint32_t *d_offsets,
                                         uint32_t *d_lengths, void *stream, prr_error *err)
{
    if (!ctx || !d_values || !d_out || !d_offsets || !d_lengths)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    int rc;
    if ((rc = lj92_check(width, height, stride, err)))
        return rc;
    if (out_cap < prr_dng_lj92_capacity(width, height))
        return fail(err, PRR_E_OUTPUT_SIZE, "dng tile streams", prr_dng_lj92_capacity(width, height),
                    out_cap);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    const uint64_t n = lj92_tiles(width, height);
    if ((rc = ensure(ctx, ctx->lj_raw, (size_t)n * LJ92_RAW_WORDS * 4, err)) ||
        (rc = ensure(ctx, ctx->lj_records, (size_t)n * sizeof(Lj92Record), err)))
        return rc;
    HIP_TRY(launch_lj92(d_values, stride, width, height, (uint32_t *)ctx->lj_raw.ptr,
                        (Lj92Record *)ctx->lj_records.ptr, d_out, d_offsets, d_lengths,
                        (hipStream_t)stream),
            "lj92 kernels");
    return PRR_OK;
}

extern "C" int prr_dng_lj92_tiles(prr_context *ctx, const uint16_t *values, uint32_t width,
                                  uint32_t height, uint32_t stride, uint8_t *out, size_t out_cap,
                                  uint32_t *offsets, uint32_t *lengths, size_t n_tiles,
                                  size_t *total, prr_error *err)
{
    if (!ctx || !values || !offsets || !lengths || !total || (!out && out_cap))
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    int rc;
    if ((rc = lj92_check(width, height, stride, err)))
        return rc;
    const uint64_t n = lj92_tiles(width, height);
    if (n_tiles != n)
        return fail(err, PRR_E_OUTPUT_SIZE, "dng tile count", n, n_tiles);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t cap = prr_dng_lj92_capacity(width, height);
    if ((rc = ensure(ctx, ctx->lj_values, (size_t)width * height * 2, err)) ||
        (rc = ensure(ctx, ctx->lj_out, cap, err)) ||
        (rc = ensure(ctx, ctx->lj_meta, (size_t)n * 8, err)))
        return rc;
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpy2DAsync(ctx->lj_values.ptr, (size_t)width * 2, values, (size_t)stride * 2,
                             (size_t)width * 2, height, hipMemcpyHostToDevice, s),
            "hipMemcpy2D values");
    uint32_t *d_off = (uint32_t *)ctx->lj_meta.ptr, *d_len = d_off + n;
    if ((rc = prr_dng_lj92_tiles_device(ctx, (const uint16_t *)ctx->lj_values.ptr, width, height,
                                        width, (uint8_t *)ctx->lj_out.ptr, cap, d_off, d_len, s,
                                        err)))
        return rc;
    HIP_TRY(hipMemcpyAsync(offsets, d_off, (size_t)n * 4, hipMemcpyDeviceToHost, s), "hipMemcpy");
    HIP_TRY(hipMemcpyAsync(lengths, d_len, (size_t)n * 4, hipMemcpyDeviceToHost, s), "hipMemcpy");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    *total = (size_t)offsets[n - 1] + lengths[n - 1];
    if (*total > out_cap)
        return fail(err, PRR_E_OUTPUT_SIZE, "dng tile streams", *total, out_cap);
    HIP_TRY(hipMemcpy(out, ctx->lj_out.ptr, *total, hipMemcpyDeviceToHost), "hipMemcpy out");
    return PRR_OK;
}

/* ------------------------------------------------------------------ */
/* Crop and orientation (mov.rs:121-157, header.rs:98-109, dng.rs:190-196). */

namespace {

struct Cropped {
    uint32_t left, top, cw, ch, ow, oh;
};

/* The crop is applied only when it leaves a picture, as dng.rs:190-196
 * writes DefaultCrop only then. */
bool crop_and_orient(uint32_t width, uint32_t height, const uint8_t *crop, int orientation,
                     Cropped &c)
{
    if (orientation < 1 || orientation > 8 || width == 0 || height == 0)
        return false;
    c.left = c.top = 0;
    c.cw = width;
    c.ch = height;
    if (crop) {
        uint32_t h_margin = (uint32_t)crop[0] + crop[1], v_margin = (uint32_t)crop[2] + crop[3];
        if (width > h_margin && height > v_margin) {
            c.left = crop[0];
            c.top = crop[2];
            c.cw = width - h_margin;
            c.ch = height - v_margin;
        }
    }
    const bool swap = orientation >= 5;
    c.ow = swap ? c.ch : c.cw;
    c.oh = swap ? c.cw : c.ch;
    return true;
}

void source_pixel(int orient, uint32_t cw, uint32_t ch, uint32_t x, uint32_t y, uint32_t &sx,
                  uint32_t &sy)
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

extern "C" int prr_oriented_size(uint32_t width, uint32_t height, const uint8_t crop[4],
                                 int orientation, uint32_t *out_width, uint32_t *out_height)
{
    Cropped c;
    if (!out_width || !out_height || !crop_and_orient(width, height, crop, orientation, c))
        return PRR_E_INVALID_ARGUMENT;
    *out_width = c.ow;
    *out_height = c.oh;
    return PRR_OK;
}

extern "C" int prr_oriented_bayer(int bayer, const uint8_t crop[4], uint32_t width,
                                  uint32_t height, int orientation)
{
    Cropped c;
    if (bayer < 0 || bayer > 3 || !crop_and_orient(width, height, crop, orientation, c) ||
        c.cw < 2 || c.ch < 2)
        return -1;
    /* BayerPattern::red_position (header.rs:66-73) is (code >> 1, code & 1). */
    const uint32_t red_row = (uint32_t)bayer >> 1, red_col = (uint32_t)bayer & 1;
    for (uint32_t y = 0; y < 2; y++)
        for (uint32_t x = 0; x < 2; x++) {
            uint32_t sx, sy;
            source_pixel(orientation, c.cw, c.ch, x, y, sx, sy);
            if (((c.top + sy) & 1) == red_row && ((c.left + sx) & 1) == red_col)
                return (int)(2 * y + x);
        }
    return -1;
}

static int orient_launch(prr_context *ctx, const void *const in[3], void *const out[3],
                         uint32_t planes, int element_bytes, uint32_t width, uint32_t height,
                         uint32_t in_stride, const uint8_t crop[4], int orientation,
                         uint32_t out_stride, void *stream, prr_error *err)
{
    if (!ctx)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    for (uint32_t p = 0; p < planes; p++)
        if (!in[p] || !out[p])
            return fail(err, PRR_E_INVALID_ARGUMENT, "null plane");
    Cropped c;
    if (!crop_and_orient(width, height, crop, orientation, c))
        return fail(err, PRR_E_INVALID_ARGUMENT, "size or orientation", width, (uint64_t)orientation);
    if (in_stride < width || out_stride < c.ow)
        return fail(err, PRR_E_INVALID_ARGUMENT, "stride", in_stride, out_stride);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    OrientDesc d;
    std::memset(&d, 0, sizeof d);
    for (uint32_t p = 0; p < planes; p++) {
        d.in[p] = in[p];
        d.out[p] = out[p];
    }
    d.planes = planes;
    d.orient = (uint32_t)orientation;
    d.cw = c.cw;
    d.ch = c.ch;
    d.ow = c.ow;
    d.oh = c.oh;
    d.left = c.left;
    d.top = c.top;
    d.in_stride = in_stride;
    d.out_stride = out_stride;
    HIP_TRY(launch_orient(d, element_bytes, (hipStream_t)stream), "orient kernel");
    return PRR_OK;
}

extern "C" int prr_orient_planes_device(prr_context *ctx, const float *d_r, const float *d_g,
                                        const float *d_b, uint32_t width, uint32_t height,
                                        uint32_t in_stride, const uint8_t crop[4], int orientation,
                                        float *d_out_r, float *d_out_g, float *d_out_b,
                                        uint32_t out_stride, void *stream, prr_error *err)
{
    const void *in[3] = {d_r, d_g, d_b};
    void *out[3] = {d_out_r, d_out_g, d_out_b};
    return orient_launch(ctx, in, out, 3, 4, width, height, in_stride, crop, orientation,
                         out_stride, stream, err);
}

extern "C" int prr_orient_cfa_device(prr_context *ctx, const uint16_t *d_in, uint32_t width,
                                     uint32_t height, uint32_t in_stride, const uint8_t crop[4],
                                     int orientation, uint16_t *d_out, uint32_t out_stride,
                                     void *stream, prr_error *err)
{
    const void *in[3] = {d_in, nullptr, nullptr};
    void *out[3] = {d_out, nullptr, nullptr};
    return orient_launch(ctx, in, out, 1, 2, width, height, in_stride, crop, orientation,
                         out_stride, stream, err);
}

extern "C" int prr_batch_create(prr_context *ctx, prr_batch **out, prr_error *err)
{
    if (!ctx || !out)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    prr_batch *b = new prr_batch();
    b->ctx = ctx;
    b->spw = 64;
    *out = b;
    return PRR_OK;
}

extern "C" void prr_batch_destroy(prr_batch *b)
{
    if (!b)
        return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (DevBuf *d : {&b->d_frames, &b->d_jobs, &b->d_queue, &b->d_head, &b->d_hist, &b->d_offsets,
                      &b->d_chunks, &b->d_chunk_bytes})
        if (d->ptr)
            (void)hipFree(d->ptr);
    delete b;
}

static uint32_t frame_spw(uint32_t want, uint32_t per_block, uint32_t amax, uint32_t full,
                          uint32_t lds_limit, size_t &lds)
{
    if (full) {
        lds = staged_decode_lds_bytes(amax);
        return want;
    }
    uint32_t spw = want;
    for (;;) {
        lds = decode_lds_bytes(spw, per_block, amax);
        if (lds <= lds_limit || spw <= 4)
            return spw;
        spw >>= 1;
    }
}

/* Tile groups of every frame for the batch's streams per workgroup: each
 * frame's spw (a full decode keeps it, its LDS does not grow with it;
 * a reduced one shrinks until its LDS fits), first queue item and the
 * batch's LDS size. */
static void relayout(prr_batch *b)
{
    b->n_items = 0;
    b->lds = 0;
    for (size_t i = 0; i < b->frames.size(); i++) {
        BatchFrame &bf = b->frames[i];
        const BatchShape &sh = b->shapes[i];
        size_t lds;
        bf.spw = frame_spw(b->spw, sh.per_block, sh.alignment, sh.full,
                           std::min(b->ctx->launch.lds_limit, b->lds_budget), lds);
        const uint32_t T = bf.spw / 4;
        b->n_items += (sh.nb_tiles + T - 1) / T;
        b->lds = std::max(b->lds, lds);
    }
    b->uploaded = false;
}

extern "C" int prr_batch_add(prr_batch *b, const prr_header *hdr, const prr_frame_info *info,
                             const prr_batch_frame *fr, prr_error *err)
{
    if (!b || !hdr || !info || !fr || !fr->d_packet || !fr->d_status)
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    if ((uintptr_t)fr->d_packet & 3)
        return fail(err, PRR_E_INVALID_ARGUMENT, "packet alignment");
    if (b->frames.size() >= 256)
        return fail(err, PRR_E_INVALID_ARGUMENT, "batch holds at most 256 frames");
    BatchFrame bf;
    std::memset(&bf, 0, sizeof bf);
    bf.f = frame_desc(hdr, info, fr->d_packet, fr->len, nullptr);
    uint32_t per_block = 64;
    if (fr->kind == PRR_BATCH_REDUCED) {
        if (fr->scale < PRR_SCALE_HALF || fr->scale > PRR_SCALE_EIGHTH || !fr->d_rgb)
            return fail(err, PRR_E_INVALID_ARGUMENT, "scale or rgb");
        bf.mode = fr->scale == PRR_SCALE_HALF ? BATCH_HALF
                  : fr->scale == PRR_SCALE_QUARTER ? BATCH_QUARTER : BATCH_EIGHTH;
        bf.average = fr->greens == PRR_GREENS_AVERAGE ? 1 : 0;
        bf.rgb = fr->d_rgb;
        per_block = fr->scale == PRR_SCALE_HALF ? 64 : fr->scale == PRR_SCALE_QUARTER ? 16 : 4;
    } else if (fr->kind == PRR_BATCH_DECODE || fr->kind == PRR_BATCH_DEVELOP) {
        if (!fr->d_samples || (fr->idct != PRR_IDCT_APPLE && fr->idct != PRR_IDCT_FFMPEG))
            return fail(err, PRR_E_INVALID_ARGUMENT, "samples or idct");
        if (((uintptr_t)fr->d_samples & 15) || ((uintptr_t)fr->d_linear & 3))
            return fail(err, PRR_E_INVALID_ARGUMENT, "samples or linear alignment");
        bf.mode = fr->idct == PRR_IDCT_FFMPEG ? BATCH_FULL_FFMPEG : BATCH_FULL_APPLE;
        bf.samples = fr->d_samples;
        bf.linear = fr->d_linear;
    } else {
        return fail(err, PRR_E_INVALID_ARGUMENT, "kind", (uint64_t)fr->kind);
    }
    bf.flags = fr->d_tile_flags;
    bf.status = fr->d_status;
    if ((info->nb_tiles + 3) / 4 >= (1u << 20))
        return fail(err, PRR_E_INVALID_ARGUMENT, "too many tile groups in one frame");
    /* Every check comes before the batch changes: a failed add leaves it as
     * it was. */
    if (fr->kind == PRR_BATCH_DEVELOP) {
        if (!fr->d_r || (fr->d_g == nullptr) != (fr->d_b == nullptr))
            return fail(err, PRR_E_INVALID_ARGUMENT, "planes");
        if (fr->stage < PRR_STAGE_DEMOSAICED || fr->stage > PRR_STAGE_REC2020)
            return fail(err, PRR_E_INVALID_ARGUMENT, "stage");
    }
    const uint32_t n_chunks = std::max<uint32_t>((info->nb_tiles + SCAN_CHUNK - 1) / SCAN_CHUNK, 1);
    for (uint32_t c = 0; c < n_chunks; c++)
        b->chunks.push_back(((uint32_t)b->frames.size() << 16) | c);
    b->offset_base.push_back(b->offsets_total);
    b->offsets_total += (size_t)info->nb_tiles + 1;
    b->frames.push_back(bf);
    b->shapes.push_back({per_block, hdr->alignment, info->nb_tiles,
                         fr->kind == PRR_BATCH_REDUCED ? 0u : 1u});
    relayout(b);
    if (fr->kind == PRR_BATCH_DEVELOP) {
        BatchDevelop j;
        j.d = develop_desc_device(hdr, fr->d_samples, fr->d_status, fr->explicit_wb, fr->wb_red,
                                  fr->wb_blue, fr->stage);
        j.r = fr->d_r;
        j.g = fr->d_g;
        j.b = fr->d_b;
        j.tiles_x = ((uint32_t)hdr->width + DEVELOP_TILE_W - 1) / DEVELOP_TILE_W;
        j.first_tile = b->dev_tiles;
        b->dev_tiles += j.tiles_x * (((uint32_t)hdr->height + DEVELOP_TILE_H - 1) / DEVELOP_TILE_H);
        b->jobs.push_back(j);
    }
    b->uploaded = false;
    return PRR_OK;
}

/* Remove every frame and keep the device buffers, so a caller decoding a
 * stream of frame sets (windows of a clip) reuses one batch. Frames added
 * after this are uploaded by the next prr_batch_run, in stream order after
 * the work already enqueued on that stream. */
extern "C" int prr_batch_clear(prr_batch *b, prr_error *err)
{
    if (!b)
        return fail(err, PRR_E_INVALID_ARGUMENT, "batch");
    b->frames.clear();
    b->shapes.clear();
    b->jobs.clear();
    b->offset_base.clear();
    b->offsets_total = 0;
    b->chunks.clear();
    b->n_items = 0;
    b->dev_tiles = 0;
    b->lds = 0;
    b->uploaded = false;
    return PRR_OK;
}

extern "C" int prr_batch_set_config(prr_batch *b, uint32_t streams_per_wg, uint32_t chunk,
                                    uint32_t waves_per_cu, prr_error *err)
{
    if (!b)
        return fail(err, PRR_E_INVALID_ARGUMENT, "batch");
    if (streams_per_wg) {
        if (streams_per_wg != 16 && streams_per_wg != 32 && streams_per_wg != 64)
            return fail(err, PRR_E_INVALID_ARGUMENT, "streams_per_wg", streams_per_wg);
        b->spw = streams_per_wg;
    }
    if (chunk)
        b->chunk = std::min<uint32_t>(chunk, 64);
    b->waves_per_cu = waves_per_cu;
    relayout(b);
    return PRR_OK;
}

/* Enqueue the batch: scans, queue, persistent decode, develop. The first
 * call after adding frames uploads the frame table (a host-to-device copy),
 * so capture a graph only from the second call on. */
extern "C" int prr_batch_run(prr_batch *b, void *stream, prr_error *err)
{
    if (!b)
        return fail(err, PRR_E_INVALID_ARGUMENT, "batch");
    prr_context *ctx = b->ctx;
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    if (b->frames.empty())
        return PRR_OK;
    if (!b->uploaded) {
        int rc;
        if ((rc = ensure(ctx, b->d_offsets, sizeof(uint64_t) * b->offsets_total, err)) ||
            (rc = ensure(ctx, b->d_frames, sizeof(BatchFrame) * b->frames.size(), err)) ||
            (rc = ensure(ctx, b->d_jobs, sizeof(BatchDevelop) * std::max<size_t>(b->jobs.size(), 1), err)) ||
            (rc = ensure(ctx, b->d_queue, sizeof(uint32_t) * std::max<uint32_t>(b->n_items, 1), err)) ||
            (rc = ensure(ctx, b->d_head, 64, err)) ||
            (rc = ensure(ctx, b->d_hist, sizeof(uint32_t) * 64 * b->frames.size(), err)) ||
            (rc = ensure(ctx, b->d_chunks, sizeof(uint32_t) * b->chunks.size(), err)) ||
            (rc = ensure(ctx, b->d_chunk_bytes, sizeof(uint32_t) * b->chunks.size(), err)))
            return rc;
        for (size_t i = 0; i < b->frames.size(); i++)
            b->frames[i].f.offsets = (uint64_t *)b->d_offsets.ptr + b->offset_base[i];
        HIP_TRY(hipMemcpyAsync(b->d_frames.ptr, b->frames.data(), sizeof(BatchFrame) * b->frames.size(),
                               hipMemcpyHostToDevice, s), "hipMemcpy frames");
        HIP_TRY(hipMemcpyAsync(b->d_chunks.ptr, b->chunks.data(), sizeof(uint32_t) * b->chunks.size(),
                               hipMemcpyHostToDevice, s), "hipMemcpy chunks");
        if (!b->jobs.empty())
            HIP_TRY(hipMemcpyAsync(b->d_jobs.ptr, b->jobs.data(), sizeof(BatchDevelop) * b->jobs.size(),
                                   hipMemcpyHostToDevice, s), "hipMemcpy jobs");
        HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, ctx->device), "hipGetDeviceProperties");
        int per_cu = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)batch_decode_kernel(),
                                                             64, b->lds),
                "hipOccupancy");
        if (per_cu < 1)
            per_cu = 1;
        if (b->waves_per_cu)
            per_cu = std::min<int>(per_cu, (int)b->waves_per_cu);
        b->grid = (uint32_t)std::min<uint64_t>((uint64_t)prop.multiProcessorCount * per_cu,
                                               std::max<uint32_t>(b->n_items, 1));
        b->uploaded = true;
    }
    HIP_TRY(launch_batch((const BatchFrame *)b->d_frames.ptr, (uint32_t)b->frames.size(), b->n_items,
                         (uint32_t *)b->d_queue.ptr, (uint32_t *)b->d_head.ptr,
                         (uint32_t *)b->d_hist.ptr, (const uint32_t *)b->d_chunks.ptr,
                         (uint32_t)b->chunks.size(), (uint32_t *)b->d_chunk_bytes.ptr,
                         (const BatchDevelop *)b->d_jobs.ptr,
                         (uint32_t)b->jobs.size(), b->dev_tiles, b->grid, b->chunk, b->lds, s),
            "batch kernels");
    return PRR_OK;
}

extern "C" int prr_batch_info(prr_batch *b, uint32_t *n_items, uint32_t *grid, uint32_t *lds_bytes)
{
    if (!b)
        return PRR_E_INVALID_ARGUMENT;
    if (n_items)
        *n_items = b->n_items;
    if (grid)
        *grid = b->grid;
    if (lds_bytes)
        *lds_bytes = (uint32_t)b->lds;
    return PRR_OK;
}
