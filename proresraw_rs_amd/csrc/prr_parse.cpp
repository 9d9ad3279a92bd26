/* Frame structure on the host (no device work): the packet prefix and
 * header (decoder.rs:45-73, header.rs:156-246), the tile grid (tile.rs:36-72)
 * and the size-table walk (decoder.rs:76-123), the frame quantizer
 * (decoder.rs:33-35, demosaic.rs:50-52), the demosaic parameters
 * (demosaic.rs:80-106) and the output matrix (demosaic.rs:127-153). These
 * read untrusted packets, so they are also built with AddressSanitizer and
 * UndefinedBehaviorSanitizer and fuzzed on the CPU (tests/fuzz_host.cpp).
 * Plain C++: no HIP.
 */
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "prr_fail.h"

/* ------------------------------------------------------------------ */
/* Frame structure. */

namespace {

uint16_t rd16(const uint8_t *p) { return (uint16_t)((p[0] << 8) | p[1]); }
uint32_t rd32(const uint8_t *p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
float rdf32(const uint8_t *p)
{
    uint32_t u = rd32(p);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

constexpr size_t kPrefix = 10;    /* decoder.rs:13 */
constexpr size_t kFixedBody = 70; /* header.rs:11 */
constexpr uint16_t kDefaultCurve[8] = {0, 512, 1024, 2048, 4096, 8192, 16384, 32768};

/* FrameHeader::parse (header.rs:156-246). */
int parse_body(const uint8_t *b, size_t n, prr_header *h, prr_error *err)
{
    if (n < kFixedBody)
        return fail(err, PRR_E_TRUNCATED, "frame header", kFixedBody, n);
    std::memset(h, 0, sizeof *h);
    h->version = b[1];
    if (h->version > 1)
        return fail(err, PRR_E_UNSUPPORTED_VERSION, nullptr, h->version);
    std::memcpy(h->vendor, b + 2, 4);
    h->width = rd16(b + 6);
    h->height = rd16(b + 8);
    if (!h->width || !h->height || (h->width & 1) || (h->height & 1))
        return fail(err, PRR_E_BAD_DIMENSIONS, nullptr, h->width, h->height);
    std::memcpy(h->crop, b + 10, 4);
    h->bayer = (uint8_t)(rd16(b + 14) & 3);
    h->white_level = (uint16_t)(rd16(b + 16) + 256u);
    h->wb_red = rdf32(b + 18);
    h->wb_blue = rdf32(b + 22);
    for (int i = 0; i < 9; i++)
        h->color_matrix[i] = rdf32(b + 26 + 4 * i);
    h->gain = rdf32(b + 62);
    h->wb_cct = rd16(b + 66);
    h->flags = rd16(b + 68);
    h->alignment = (uint8_t)((h->flags >> 1) & 7);
    if (h->alignment > 4)
        return fail(err, PRR_E_BAD_ALIGNMENT, nullptr, h->alignment);
    size_t at = kFixedBody;
    std::memset(h->quant_matrix, 1, sizeof h->quant_matrix);
    if (h->flags & 1) {
        if (at + 64 > n)
            return fail(err, PRR_E_TRUNCATED, "quantization matrix", at + 64, n);
        std::memcpy(h->quant_matrix, b + at, 64);
        at += 64;
        for (int i = 0; i < 64; i++)
            if (h->quant_matrix[i] < 1 || h->quant_matrix[i] > 63)
                return fail(err, PRR_E_BAD_QUANT_MATRIX, nullptr, (uint64_t)i, h->quant_matrix[i]);
    }
    if ((h->flags >> 4) & 1) {
        if (at + 16 > n)
            return fail(err, PRR_E_TRUNCATED, "linearization curve", at + 16, n);
        uint16_t pts[8];
        for (int i = 0; i < 8; i++)
            pts[i] = rd16(b + at + 2 * i);
        if (h->version == 0)
            return fail(err, PRR_E_UNEXPECTED_CURVE);
        for (int i = 0; i + 1 < 8; i++)
            if (pts[i] >= pts[i + 1]) {
                fail(err, PRR_E_BAD_CURVE);
                if (err)
                    std::memcpy(err->points, pts, sizeof pts);
                return PRR_E_BAD_CURVE;
            }
        std::memcpy(h->curve, pts, sizeof pts);
    } else {
        std::memcpy(h->curve, kDefaultCurve, sizeof kDefaultCurve);
    }
    return PRR_OK;
}

uint32_t tiles_per_row_of(uint32_t width, uint32_t alignment)
{
    uint32_t columns = (width + 15) / 16;
    return (columns >> alignment) + (uint32_t)__builtin_popcount(columns & ((1u << alignment) - 1));
}

} // namespace

extern "C" int prr_header_len(const uint8_t *packet, size_t len, size_t *out, prr_error *err)
{
    if (!packet && len)
        return fail(err, PRR_E_INVALID_ARGUMENT, "packet");
    if (len < kPrefix)
        return fail(err, PRR_E_TRUNCATED, "frame prefix", kPrefix, len);
    if (std::memcmp(packet + 4, "prrf", 4) != 0) {
        fail(err, PRR_E_BAD_TAG);
        if (err)
            std::memcpy(err->tag, packet + 4, 4);
        return PRR_E_BAD_TAG;
    }
    uint16_t hs = rd16(packet + 8);
    if (hs < 72)
        return fail(err, PRR_E_HEADER_TOO_SHORT, nullptr, hs);
    *out = kPrefix + hs - 2;
    return PRR_OK;
}

extern "C" int prr_parse_header(const uint8_t *packet, size_t len, prr_header *out,
                                prr_error *err)
{
    size_t end;
    int rc = prr_header_len(packet, len, &end, err);
    if (rc)
        return rc;
    if (end > len)
        return fail(err, PRR_E_TRUNCATED, "frame header", end, len);
    return parse_body(packet + kPrefix, end - kPrefix, out, err);
}

extern "C" int prr_parse_frame_info(const uint8_t *packet, size_t len, prr_header *hdr,
                                    prr_frame_info *info, prr_error *err)
{
    if (!packet && len)
        return fail(err, PRR_E_INVALID_ARGUMENT, "packet");
    if (len < kPrefix)
        return fail(err, PRR_E_TRUNCATED, "frame prefix", kPrefix, len);
    uint32_t declared = rd32(packet);
    if (declared == 0 || (size_t)declared > len)
        return fail(err, PRR_E_FRAME_SIZE_MISMATCH, nullptr, declared, len);
    size_t end;
    int rc = prr_header_len(packet, len, &end, err);
    if (rc)
        return rc;
    rc = prr_parse_header(packet, len, hdr, err);
    if (rc)
        return rc;
    info->coded_width = ((uint32_t)hdr->width + 15) & ~15u;
    info->coded_height = ((uint32_t)hdr->height + 15) & ~15u;
    info->tiles_per_row = tiles_per_row_of(hdr->width, hdr->alignment);
    info->tile_rows = info->coded_height / 16;
    info->nb_tiles = info->tiles_per_row * info->tile_rows;
    info->header_end = end;
    size_t table_end = end + 2 * (size_t)info->nb_tiles;
    if (table_end > len)
        return fail(err, PRR_E_TRUNCATED, "tile size table", table_end, len);
    return PRR_OK;
}

extern "C" uint32_t prr_quantizer(const prr_header *hdr, uint8_t max_qscale)
{
    return (uint32_t)hdr->quant_matrix[0] * (uint32_t)max_qscale;
}

/* parse_frame (decoder.rs:76-123): the size table walked on the host. */
extern "C" int prr_parse_frame(const uint8_t *packet, size_t len, prr_parsed_frame *out,
                               uint64_t *tile_offsets, uint32_t *tile_sizes, size_t cap,
                               prr_error *err)
{
    if (!out || ((tile_offsets == nullptr) != (tile_sizes == nullptr)))
        return fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof *out);
    int rc = prr_parse_frame_info(packet, len, &out->header, &out->info, err);
    if (rc)
        return rc;
    const size_t n = out->info.nb_tiles;
    if (tile_offsets && cap < n)
        return fail(err, PRR_E_OUTPUT_SIZE, "tile capacity", n, cap);
    const uint8_t *table = packet + out->info.header_end;
    uint64_t offset = out->info.header_end + 2 * (uint64_t)n;
    uint8_t mq = 0;
    for (size_t i = 0; i < n; i++) {
        const uint64_t size = rd16(table + 2 * i);
        /* The byte after the tile start, even for a tile too short to hold
         * it (0 past the packet). */
        const uint8_t q = offset + 1 < len ? packet[offset + 1] : 0;
        mq = q > mq ? q : mq;
        if (offset + size > len)
            return fail(err, PRR_E_TILE_OUT_OF_BOUNDS, nullptr, i);
        if (tile_offsets) {
            tile_offsets[i] = offset;
            tile_sizes[i] = (uint32_t)size;
        }
        offset += size;
    }
    out->max_qscale = mq;
    return PRR_OK;
}

extern "C" uint32_t prr_frame_quantizer(const prr_parsed_frame *parsed)
{
    return prr_quantizer(&parsed->header, parsed->max_qscale);
}

/* DemosaicParams::with_white_balance (demosaic.rs:86-106). */
extern "C" void prr_demosaic_params_with_white_balance(const prr_header *hdr, uint32_t quantizer,
                                                       float red, float blue,
                                                       prr_demosaic_params *p)
{
    const uint32_t range = (uint16_t)(hdr->white_level - 256u);
    const double q = (double)quantizer;
    const double base = q * q * 5.0 + 64.0 / 3.0;
    const double range2 = (double)range * (double)range;
    const float rangef = (float)range;
    p->cfa_phase = hdr->bayer;
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

extern "C" void prr_demosaic_params_new(const prr_header *hdr, uint32_t quantizer,
                                        prr_demosaic_params *p)
{
    prr_demosaic_params_with_white_balance(hdr, quantizer, hdr->wb_red, hdr->wb_blue, p);
}

/* output_matrix (demosaic.rs:138-153) with XYZ_TO_REC2020 (demosaic.rs:37-41). */
extern "C" void prr_output_matrix(const prr_header *hdr, int exposure, float m[9])
{
    static const double xyz_to_rec2020[3][3] = {
        {1.716651187971268, -0.355670783776393, -0.25336628137366},
        {-0.666684351832489, 1.616481236634939, 0.015768545813911},
        {0.017639857445311, -0.042770613257809, 0.942103121235474}};
    double gain = (double)hdr->gain;
    if (exposure == PRR_EXPOSURE_NORMALIZED)
        gain = gain * 10.0 / 9.0;
    for (int i = 0; i < 9; i++) {
        const int row = i / 3, col = i % 3;
        const double v = xyz_to_rec2020[row][0] * (double)hdr->color_matrix[col] +
                         xyz_to_rec2020[row][1] * (double)hdr->color_matrix[3 + col] +
                         xyz_to_rec2020[row][2] * (double)hdr->color_matrix[6 + col];
        m[i] = (float)(v * gain);
    }
}

extern "C" const char *prr_version(void) { return "proresraw_amd 0.1.0 (gfx950)"; }
