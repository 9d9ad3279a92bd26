/* CPU fuzz of the host code that reads untrusted input, built with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitized_host.py
 * compiles it with the sources below and runs it):
 *
 *   - packets: prr_header_len, prr_parse_header, prr_parse_frame_info,
 *     prr_parse_frame (tile tables sized exactly as the frame asks),
 *     prr_frame_quantizer, prr_demosaic_params_*, prr_output_matrix on
 *     seeded scene packets and random frames of the synth module, mutated
 *     (bit flips, byte stores, truncation, insertion, deletion, size-table
 *     entries set to extremes) and on random bytes behind a valid prefix;
 *   - movies: prr_mov_read_video_track and the track accessors on the seed
 *     movies given on the command line, mutated the same way (atom sizes
 *     included) and written to a temporary file.
 *
 * Usage: fuzz_host ITERATIONS SEED [movie ...]. A sanitizer report aborts
 * the run (non-zero exit); the last line printed names the case reached. */
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "../include/proresraw_amd.h"
#include "../include/proresraw_mov.h"
#include "../proresraw_rs_amd/csrc/synth.h"

namespace {

uint64_t g_rng = 1;

uint32_t next() { return prs_rng_next(&g_rng); }
uint32_t below(uint32_t n) { return n ? next() % n : 0; }

/* One random edit of `b`. */
void mutate(std::vector<uint8_t> &b)
{
    switch (below(8)) {
    case 0: /* bit flip */
        if (!b.empty())
            b[below((uint32_t)b.size())] ^= (uint8_t)(1u << below(8));
        break;
    case 1: /* byte store, favouring the first 300 bytes (prefix, header, tables) */
        if (!b.empty()) {
            const uint32_t lim = below(2) ? std::min<uint32_t>((uint32_t)b.size(), 300) : (uint32_t)b.size();
            b[below(lim)] = (uint8_t)next();
        }
        break;
    case 2: /* truncate */
        if (!b.empty())
            b.resize(below((uint32_t)b.size() + 1));
        break;
    case 3: { /* insert bytes */
        const size_t at = below((uint32_t)b.size() + 1);
        const uint32_t n = 1 + below(16);
        std::vector<uint8_t> ins(n);
        for (auto &v : ins)
            v = (uint8_t)next();
        b.insert(b.begin() + (long)at, ins.begin(), ins.end());
        break;
    }
    case 4: { /* delete bytes */
        if (b.empty())
            break;
        const size_t at = below((uint32_t)b.size());
        const size_t n = std::min<size_t>(1 + below(16), b.size() - at);
        b.erase(b.begin() + (long)at, b.begin() + (long)(at + n));
        break;
    }
    case 5: /* a big-endian 16-bit value set to an extreme */
        if (b.size() >= 2) {
            const size_t at = below((uint32_t)b.size() - 1);
            const uint16_t v = below(2) ? 0xFFFF : (uint16_t)below(3);
            b[at] = (uint8_t)(v >> 8);
            b[at + 1] = (uint8_t)v;
        }
        break;
    case 6: /* a big-endian 32-bit value set to an extreme (atom sizes) */
        if (b.size() >= 4) {
            const size_t at = below((uint32_t)b.size() - 3);
            const uint32_t pick = below(4);
            const uint32_t v = pick == 0 ? 0xFFFFFFFFu : pick == 1 ? 0u : pick == 2 ? 1u : next();
            for (int k = 0; k < 4; k++)
                b[at + k] = (uint8_t)(v >> (24 - 8 * k));
        }
        break;
    default: /* several edits */
        for (uint32_t k = 0, n = 2 + below(6); k < n; k++) {
            if (!b.empty())
                b[below((uint32_t)b.size())] = (uint8_t)next();
        }
        break;
    }
}

int parse_packet(const std::vector<uint8_t> &pk)
{
    const uint8_t *p = pk.empty() ? nullptr : pk.data();
    /* A heap copy of exactly the packet's length, so reads past its end are
     * seen by the sanitizer. */
    uint8_t *exact = pk.empty() ? nullptr : (uint8_t *)std::malloc(pk.size());
    if (exact)
        std::memcpy(exact, p, pk.size());
    prr_error err;
    size_t hl = 0;
    (void)prr_header_len(exact, pk.size(), &hl, &err);
    prr_header h;
    int ok = 0;
    if (prr_parse_header(exact, pk.size(), &h, &err) == PRR_OK) {
        prr_demosaic_params dp;
        prr_demosaic_params_new(&h, prr_quantizer(&h, (uint8_t)next()), &dp);
        prr_demosaic_params_with_white_balance(&h, next() % 4000, 0.0f, 2.5f, &dp);
        float m[9];
        prr_output_matrix(&h, (int)below(2), m);
    }
    prr_frame_info info;
    if (prr_parse_frame_info(exact, pk.size(), &h, &info, &err) == PRR_OK) {
        prr_parsed_frame pf;
        /* Tables of exactly nb_tiles entries on the heap. */
        const size_t n = info.nb_tiles;
        uint64_t *off = (uint64_t *)std::malloc(sizeof(uint64_t) * (n ? n : 1));
        uint32_t *sz = (uint32_t *)std::malloc(sizeof(uint32_t) * (n ? n : 1));
        if (prr_parse_frame(exact, pk.size(), &pf, off, sz, n, &err) == PRR_OK) {
            (void)prr_frame_quantizer(&pf);
            for (size_t i = 0; i < n; i++)
                if (off[i] + sz[i] > pk.size())
                    std::abort(); /* a tile past the packet must have been refused */
            ok = 1;
        }
        (void)prr_parse_frame(exact, pk.size(), &pf, nullptr, nullptr, 0, &err);
        std::free(off);
        std::free(sz);
    }
    std::free(exact);
    return ok;
}

int accept_all(const uint8_t *, void *) { return 1; }

void read_movie(const std::vector<uint8_t> &bytes)
{
    FILE *f = std::tmpfile();
    if (!f)
        std::abort();
    if (!bytes.empty() && std::fwrite(bytes.data(), 1, bytes.size(), f) != bytes.size())
        std::abort();
    std::fflush(f);
    const int fd = fileno(f);
    char msg[512];
    prr_mov_track *t = nullptr;
    const int accept = (int)below(2);
    if (prr_mov_read_video_track(fd, accept ? accept_all : nullptr, nullptr, &t, msg,
                                 sizeof msg) == PRR_OK) {
        uint8_t cc[4];
        prr_mov_track_fourcc(t, cc);
        uint32_t num, den;
        prr_mov_track_fps(t, &num, &den);
        (void)prr_mov_track_variable_durations(t);
        const prr_mov_sample *s = nullptr;
        const size_t n = prr_mov_track_samples(t, &s);
        uint64_t ct;
        (void)prr_mov_track_creation_time(t, &ct);
        (void)prr_mov_track_orientation(t);
        size_t len;
        for (size_t i = 0; i < prr_mov_track_n_warnings(t); i++)
            (void)prr_mov_track_warning(t, i, &len);
        for (size_t i = 0; i < prr_mov_track_n_metadata(t); i++) {
            (void)prr_mov_track_metadata_key(t, i, &len);
            (void)prr_mov_track_metadata_value(t, i, &len);
        }
        const char *keys[2] = {"com.apple.quicktime.make", "title"};
        (void)prr_mov_track_text(t, keys, 2, &len);
        /* Read a few samples the table names (the reader checks ranges). */
        for (size_t i = 0; i < n && i < 4; i++) {
            std::vector<uint8_t> buf(std::min<size_t>(s[i].size, 1 << 20));
            (void)prr_mov_read_at(fd, s[i].offset, buf.data(), buf.size(), msg, sizeof msg);
        }
        prr_mov_track_free(t);
    }
    std::fclose(f);
}

std::vector<uint8_t> load(const char *path)
{
    std::vector<uint8_t> out;
    FILE *f = std::fopen(path, "rb");
    if (!f)
        return out;
    uint8_t tmp[65536];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0)
        out.insert(out.end(), tmp, tmp + n);
    std::fclose(f);
    return out;
}

} // namespace

int main(int argc, char **argv)
{
    const long iters = argc > 1 ? std::atol(argv[1]) : 1000;
    g_rng = prs_rng_new(argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1);

    /* Seed packets: scenes of a few sizes, alignments and versions, and the
     * synth module's random frames. */
    std::vector<std::vector<uint8_t>> seeds;
    for (int k = 0; k < 6; k++) {
        prs_scene sc;
        std::memset(&sc, 0, sizeof sc);
        prs_spec_default(&sc.spec, (uint16_t)(64 + 32 * k), (uint16_t)(32 + 16 * k), (uint8_t)(k % 4));
        sc.spec.alignment = (uint8_t)(k % 5);
        sc.spec.version = (uint8_t)(k % 2);
        sc.spec.has_qmat = (uint8_t)(k % 3 == 0);
        for (int i = 0; i < 64; i++)
            sc.spec.qmat[i] = (uint8_t)(1 + (i * 7 + k) % 63);
        sc.seed = 100 + (uint64_t)k;
        sc.qscale_min = 15;
        sc.qscale_max = 45;
        sc.n_gradients = 2;
        sc.n_discs = 6;
        sc.corrupt_fraction = k % 2 ? 0.1f : 0.0f;
        sc.dn_per_electron = 0.5f;
        sc.read_sigma = 3.0f;
        sc.threads = 1;
        const size_t cap = prs_encode_scene(&sc, nullptr, 0);
        std::vector<uint8_t> b(cap);
        b.resize(prs_encode_scene(&sc, b.data(), b.size()));
        seeds.push_back(b);
    }
    uint64_t rng = prs_rng_new(7);
    prs_spec specs[20];
    prs_specs(&rng, specs);
    for (int k = 0; k < 20; k += 3) {
        std::vector<uint8_t> b(1 << 20);
        b.resize(prs_random_frame(&specs[k], &rng, b.data(), b.size()));
        seeds.push_back(b);
    }
    std::vector<std::vector<uint8_t>> movies;
    for (int i = 3; i < argc; i++)
        movies.push_back(load(argv[i]));

    long parsed = 0;
    for (long it = 0; it < iters; it++) {
        std::vector<uint8_t> pk = seeds[below((uint32_t)seeds.size())];
        if (below(16) == 0) {
            /* Random bytes behind a valid prefix. */
            pk.resize(10 + below(400));
            for (size_t i = 10; i < pk.size(); i++)
                pk[i] = (uint8_t)next();
        }
        for (uint32_t k = 0, n = 1 + below(4); k < n; k++)
            mutate(pk);
        parsed += parse_packet(pk);
        if (!movies.empty() && it % 4 == 0) {
            std::vector<uint8_t> mv = movies[below((uint32_t)movies.size())];
            for (uint32_t k = 0, n = 1 + below(3); k < n; k++)
                mutate(mv);
            read_movie(mv);
        }
        if ((it + 1) % 2000 == 0) {
            std::printf("iteration %ld\n", it + 1);
            std::fflush(stdout);
        }
    }
    std::printf("done: %ld iterations, %ld packets parsed whole, %zu seed movies\n", iters, parsed,
                movies.size());
    return 0;
}
