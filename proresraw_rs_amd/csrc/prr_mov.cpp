/* QuickTime / ISO base media sample tables: the C-ABI entry points of
 * include/proresraw_mov.h.
 *
 * Restates the crate's movie reader (src/mov.rs) as host C++: the same atoms
 * are read in the same order, with the same acceptance rules and the same
 * messages, so a caller sees identical tracks and errors. Parsing works on
 * spans of the one `moov` body read into memory; failures unwind as a
 * MovFail carrying the message and are turned into PRR_E_MOV at the C
 * boundary. Sample bytes are read with pread (positional, lock-free), and
 * prr_mov_read_samples spreads a list of them over host threads.
 */
#include <errno.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <optional>
#include <sstream>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/proresraw_amd.h"
#include "../../include/proresraw_mov.h"

struct prr_mov_track {
    uint8_t fourcc[4] = {0, 0, 0, 0};
    uint32_t fps_num = 0, fps_den = 1;
    bool variable_durations = false;
    std::vector<prr_mov_sample> samples;
    std::optional<uint64_t> creation_time;
    int orientation = PRR_ORIENT_IDENTITY;
    std::vector<std::string> warnings;
    std::vector<std::pair<std::string, std::string>> metadata;
};

namespace {

struct MovFail {
    std::string msg;
};

template <class... T> std::string cat(const T &...parts)
{
    std::ostringstream s;
    (s << ... << parts);
    return s.str();
}

[[noreturn]] void bail(std::string msg) { throw MovFail{std::move(msg)}; }

/* Larger `moov` atoms are refused (mov.rs:82). */
constexpr uint64_t kMaxMoov = 64ull << 20;
/* More samples than any movie this reader is for (mov.rs:84). */
constexpr uint64_t kMaxSamples = 10000000;

using Kind = std::array<uint8_t, 4>;

/* A fourcc for messages with bytes outside printable ASCII escaped, as
 * Rust's `escape_ascii` writes them (mov.rs:772-775). */
std::string name(const uint8_t *k)
{
    std::string out;
    static const char hex[] = "0123456789abcdef";
    for (int i = 0; i < 4; i++) {
        uint8_t c = k[i];
        switch (c) {
        case '\t': out += "\\t"; break;
        case '\r': out += "\\r"; break;
        case '\n': out += "\\n"; break;
        case '\\': out += "\\\\"; break;
        case '\'': out += "\\'"; break;
        case '"': out += "\\\""; break;
        default:
            if (c >= 0x20 && c < 0x7f) {
                out += (char)c;
            } else {
                out += "\\x";
                out += hex[c >> 4];
                out += hex[c & 15];
            }
        }
    }
    return out;
}
std::string name(const Kind &k) { return name(k.data()); }

/* Shortest decimal that reads back as `v`, without an exponent: how Rust's
 * `{}` prints an f64 (used in two warnings, mov.rs:563-568, 680-686). */
std::string fmt_f64(double v)
{
    if (v == 0)
        return std::signbit(v) ? "-0" : "0";
    char buf[64];
    int digits = 17, exp10 = 0;
    std::string mant;
    for (int p = 1; p <= 17; p++) {
        std::snprintf(buf, sizeof buf, "%.*e", p - 1, v);
        if (std::strtod(buf, nullptr) == v) {
            digits = p;
            break;
        }
    }
    std::snprintf(buf, sizeof buf, "%.*e", digits - 1, v);
    std::string s(buf);
    bool neg = s[0] == '-';
    if (neg)
        s.erase(0, 1);
    size_t e = s.find('e');
    exp10 = std::atoi(s.c_str() + e + 1);
    for (size_t i = 0; i < e; i++)
        if (s[i] != '.')
            mant += s[i];
    while (mant.size() > 1 && mant.back() == '0')
        mant.pop_back();
    /* value = 0.mant x 10^(exp10 + 1) */
    int point = exp10 + 1;
    std::string out;
    if (point <= 0) {
        out = "0." + std::string((size_t)-point, '0') + mant;
    } else if ((size_t)point >= mant.size()) {
        out = mant + std::string((size_t)point - mant.size(), '0');
    } else {
        out = mant.substr(0, (size_t)point) + "." + mant.substr((size_t)point);
    }
    return neg ? "-" + out : out;
}

/* String::from_utf8_lossy: each maximal ill-formed subsequence becomes
 * U+FFFD. */
std::string utf8_lossy(const uint8_t *p, size_t n)
{
    static const char repl[] = "\xEF\xBF\xBD";
    std::string out;
    size_t i = 0;
    while (i < n) {
        uint8_t b = p[i];
        if (b < 0x80) {
            out += (char)b;
            i++;
            continue;
        }
        size_t need;
        uint8_t lo = 0x80, hi = 0xBF;
        if (b >= 0xC2 && b <= 0xDF) {
            need = 1;
        } else if (b >= 0xE0 && b <= 0xEF) {
            need = 2;
            if (b == 0xE0)
                lo = 0xA0;
            if (b == 0xED)
                hi = 0x9F;
        } else if (b >= 0xF0 && b <= 0xF4) {
            need = 3;
            if (b == 0xF0)
                lo = 0x90;
            if (b == 0xF4)
                hi = 0x8F;
        } else {
            out += repl;
            i++;
            continue;
        }
        size_t j = i + 1, got = 0;
        while (got < need && j < n) {
            uint8_t c = p[j];
            uint8_t l = got == 0 ? lo : 0x80, h = got == 0 ? hi : 0xBF;
            if (c < l || c > h)
                break;
            j++;
            got++;
        }
        if (got == need)
            out.append((const char *)p + i, j - i);
        else
            out += repl;
        i = j;
    }
    return out;
}

struct Span {
    const uint8_t *p = nullptr;
    size_t n = 0;
    Span sub(size_t from, size_t to) const { return Span{p + from, to - from}; }
};

uint32_t be32(const uint8_t *b)
{
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
}
uint64_t be64(const uint8_t *b) { return ((uint64_t)be32(b) << 32) | be32(b + 4); }

/* `count` entries of `width` bytes at `at` of `kind`'s body (mov.rs:729-747). */
Span entries(Span body, size_t at, uint64_t count, size_t width, const uint8_t *kind)
{
    unsigned __int128 bytes = (unsigned __int128)count * width;
    if (bytes <= (unsigned __int128)body.n && at <= body.n && bytes <= body.n - at)
        return body.sub(at, at + (size_t)bytes);
    bail(cat("truncated `", name(kind), "` atom: ", count, " entries do not fit in ", body.n,
             " bytes"));
}
Span entries(Span body, size_t at, uint64_t count, size_t width, const char *kind)
{
    return entries(body, at, count, width, (const uint8_t *)kind);
}
uint32_t u32_at(Span body, size_t at, const uint8_t *kind) { return be32(entries(body, at, 1, 4, kind).p); }
uint32_t u32_at(Span body, size_t at, const char *kind) { return u32_at(body, at, (const uint8_t *)kind); }
uint64_t u64_at(Span body, size_t at, const uint8_t *kind) { return be64(entries(body, at, 1, 8, kind).p); }

struct Atom {
    Kind kind;
    Span body;
};

bool is(const Atom &a, const char *k) { return std::memcmp(a.kind.data(), k, 4) == 0; }

/* The child atoms of `data` (mov.rs:694-719). */
std::vector<Atom> children(Span data)
{
    std::vector<Atom> atoms;
    Span rest = data;
    while (rest.n) {
        if (rest.n < 8) {
            /* QuickTime allows a 32-bit zero after the last child. */
            for (size_t i = 0; i < rest.n; i++)
                if (rest.p[i])
                    bail("truncated atom header");
            break;
        }
        Kind kind{rest.p[4], rest.p[5], rest.p[6], rest.p[7]};
        uint64_t header, size;
        switch (be32(rest.p)) {
        case 1:
            header = 16;
            size = u64_at(rest, 8, kind.data());
            break;
        case 0:
            header = 8;
            size = rest.n;
            break;
        default:
            header = 8;
            size = be32(rest.p);
        }
        if (!(header <= size && size <= rest.n))
            bail(cat("truncated `", name(kind), "` atom: it claims ", size, " bytes and ", rest.n,
                     " are left"));
        atoms.push_back(Atom{kind, rest.sub((size_t)header, (size_t)size)});
        rest = rest.sub((size_t)size, rest.n);
    }
    return atoms;
}

const Atom *find(const std::vector<Atom> &atoms, const char *kind)
{
    for (const Atom &a : atoms)
        if (is(a, kind))
            return &a;
    return nullptr;
}

/* mov.rs:721-727. */
Span require(const std::vector<Atom> &atoms, const char *kind, const char *parent)
{
    if (const Atom *a = find(atoms, kind))
        return a->body;
    bail(cat(parent, " has no `", name((const uint8_t *)kind), "` atom"));
}

/* The creation time and timescale of a `mvhd` or `mdhd` (mov.rs:641-649). */
std::pair<uint64_t, uint32_t> times(Span body, const char *kind)
{
    if (!body.n)
        bail(cat("truncated `", kind, "` atom"));
    switch (body.p[0]) {
    case 0:
        return {u32_at(body, 4, kind), u32_at(body, 12, kind)};
    case 1:
        return {u64_at(body, 4, (const uint8_t *)kind), u32_at(body, 20, kind)};
    default:
        bail(cat("`", kind, "` version ", (unsigned)body.p[0], " is not supported"));
    }
}

uint64_t sample_count(uint32_t count)
{
    if (count > kMaxSamples)
        bail(cat("the video track lists ", count, " samples, more than the ", kMaxSamples,
                 " this reader accepts"));
    return count;
}

uint32_t gcd(uint32_t a, uint32_t b) { return b == 0 ? a : gcd(b, a % b); }

/* The orientation of a `tkhd`'s display matrix (mov.rs:650-692): only the
 * signs of a, b, c, d matter (mov.rs:651-692). */
int orientation(Span tkhd, std::vector<std::string> &warnings)
{
    if (!tkhd.n)
        bail("truncated `tkhd` atom");
    size_t at;
    if (tkhd.p[0] == 0)
        at = 40;
    else if (tkhd.p[0] == 1)
        at = 52;
    else
        bail(cat("`tkhd` version ", (unsigned)tkhd.p[0], " is not supported"));
    auto entry = [&](size_t i) { return (double)(int32_t)u32_at(tkhd, at + 4 * i, "tkhd") / 65536.0; };
    double a = entry(0), b = entry(1), c = entry(3), d = entry(4);
    auto sign = [](double v) { return (v > 0) - (v < 0); };
    int sa = sign(a), sb = sign(b), sc = sign(c), sd = sign(d);
    struct Row {
        int a, b, c, d, exif;
    };
    static const Row rows[] = {
        {1, 0, 0, 1, PRR_ORIENT_IDENTITY},    {-1, 0, 0, 1, PRR_ORIENT_MIRROR_HORIZONTAL},
        {-1, 0, 0, -1, PRR_ORIENT_ROTATE_180}, {1, 0, 0, -1, PRR_ORIENT_MIRROR_VERTICAL},
        {0, 1, 1, 0, PRR_ORIENT_TRANSPOSE},    {0, 1, -1, 0, PRR_ORIENT_ROTATE_90},
        {0, -1, -1, 0, PRR_ORIENT_TRANSVERSE}, {0, -1, 1, 0, PRR_ORIENT_ROTATE_270},
    };
    for (const Row &r : rows)
        if (r.a == sa && r.b == sb && r.c == sc && r.d == sd)
            return r.exif;
    warnings.push_back(cat("the video track's display matrix [", fmt_f64(a), " ", fmt_f64(b), "; ",
                           fmt_f64(c), " ", fmt_f64(d),
                           "] is not a quarter turn or mirror, so the frames are shown as stored"));
    return PRR_ORIENT_IDENTITY;
}

/* The most common duration, ignoring the last sample, and whether any other
 * occurs; ties go to the shorter duration (mov.rs:507-521). */
std::pair<uint32_t, bool> common_duration(const std::vector<uint32_t> &durations)
{
    size_t body = std::max<size_t>(durations.size() ? durations.size() - 1 : 0, 1);
    body = std::min(body, durations.size());
    std::map<uint32_t, size_t> counts;
    for (size_t i = 0; i < body; i++)
        counts[durations[i]]++;
    uint32_t mode = 1;
    size_t best = 0;
    for (const auto &kv : counts)
        if (kv.second > best) {
            best = kv.second;
            mode = kv.first;
        }
    return {mode, counts.size() > 1};
}

/* The samples the first edit that shows media presents (mov.rs:522-596). */
std::pair<size_t, size_t> edit(Span elst, const std::vector<uint32_t> &durations, uint32_t timescale,
                               uint32_t movie_timescale, std::vector<std::string> &warnings)
{
    if (!elst.n)
        bail("truncated `elst` atom");
    uint8_t version = elst.p[0];
    uint32_t count = u32_at(elst, 4, "elst");
    struct Edit {
        uint64_t duration;
        int64_t time;
        int32_t rate;
    };
    std::vector<Edit> edits;
    if (version == 0) {
        Span t = entries(elst, 8, count, 12, "elst");
        for (size_t i = 0; i < count; i++) {
            const uint8_t *e = t.p + 12 * i;
            edits.push_back({be32(e), (int64_t)(int32_t)be32(e + 4), (int32_t)be32(e + 8)});
        }
    } else if (version == 1) {
        Span t = entries(elst, 8, count, 20, "elst");
        for (size_t i = 0; i < count; i++) {
            const uint8_t *e = t.p + 20 * i;
            edits.push_back({be64(e), (int64_t)be64(e + 8), (int32_t)be32(e + 16)});
        }
    } else {
        bail(cat("`elst` version ", (unsigned)version, " is not supported"));
    }
    if (edits.size() > 1)
        warnings.push_back(cat("the video track's edit list has ", edits.size(),
                               " edits; only the first that shows media is honoured"));
    if (!edits.empty() && edits[0].time < 0)
        warnings.push_back(
            "the video track's edit list starts with an empty edit (a delay), which is ignored");
    const Edit *shown = nullptr;
    for (const Edit &e : edits)
        if (e.time >= 0) {
            shown = &e;
            break;
        }
    if (!shown) {
        if (!edits.empty())
            warnings.push_back(
                "the video track's edit list shows no media; every sample is presented");
        return {0, durations.size()};
    }
    if (shown->rate != 0x10000)
        warnings.push_back(cat("the video track's edit plays at rate ",
                               fmt_f64((double)shown->rate / 65536.0),
                               "; frames are read at normal speed"));
    /* Compare in the media timescale scaled by the movie's, exactly. */
    using u128 = unsigned __int128;
    u128 start = (u128)(uint64_t)shown->time * movie_timescale;
    u128 end;
    if (shown->duration == 0)
        end = ~(u128)0;
    else if (movie_timescale == 0)
        bail("the movie's timescale is 0");
    else
        end = start + (u128)shown->duration * timescale;
    u128 time = 0;
    size_t first = durations.size(), last = durations.size();
    for (size_t i = 0; i < durations.size(); i++) {
        if (time >= end) {
            last = i;
            break;
        }
        if (time >= start && first == durations.size())
            first = i;
        time += (u128)durations[i] * movie_timescale;
    }
    return {first, std::max(last, first)};
}

/* The UTF-8 items of a `meta` atom whose keys use the `mdta` namespace
 * (mov.rs:597-639). */
std::vector<std::pair<std::string, std::string>> text_metadata(Span meta)
{
    /* QuickTime's `meta` holds its children directly; the ISO form has a
     * version and flags word first. */
    Span body;
    if (meta.n >= 8 && std::memcmp(meta.p + 4, "hdlr", 4) == 0)
        body = meta;
    else if (meta.n >= 4)
        body = meta.sub(4, meta.n);
    else
        bail("truncated `meta` atom");
    std::vector<Atom> atoms = children(body);
    const Atom *keys = find(atoms, "keys"), *ilst = find(atoms, "ilst");
    std::vector<std::pair<std::string, std::string>> items;
    if (!keys || !ilst)
        return items;
    uint32_t count = u32_at(keys->body, 4, "keys");
    std::vector<std::string> names;
    if (keys->body.n < 8)
        bail("truncated `keys` atom");
    Span rest = keys->body.sub(8, keys->body.n);
    for (uint32_t i = 0; i < count; i++) {
        if (rest.n < 4)
            bail("truncated `keys` atom");
        size_t size = be32(rest.p);
        size_t stop = std::max<size_t>(size, 8);
        if (stop > rest.n)
            bail("truncated `keys` atom");
        names.push_back(utf8_lossy(rest.p + 8, stop - 8));
        rest = rest.sub(stop, rest.n);
    }
    /* Each item's type is the 1-based index of its key. */
    for (const Atom &item : children(ilst->body)) {
        size_t index = (size_t)be32(item.kind.data()) - 1;
        if (index >= names.size())
            continue;
        for (const Atom &data : children(item.body)) {
            /* Type indicator (well-known type 1 is UTF-8), locale, value. */
            if (is(data, "data") && data.body.n >= 8 && be32(data.body.p) == 1) {
                std::string value = utf8_lossy(data.body.p + 8, data.body.n - 8);
                while (!value.empty() && value.back() == '\0')
                    value.pop_back();
                items.emplace_back(names[index], value);
            }
        }
    }
    return items;
}

/* An enabled video track's media before its sample table is resolved
 * (mov.rs:294-301). */
struct Video {
    Kind fourcc;
    uint32_t timescale;
    Span tkhd;
    std::optional<Span> elst;
    std::vector<Atom> stbl;

    std::optional<Span> table(const char *kind) const
    {
        if (const Atom *a = find(stbl, kind))
            return a->body;
        return std::nullopt;
    }

    /* Each sample's size from `stsz` or `stz2` (mov.rs:385-419). */
    std::vector<uint32_t> sizes() const
    {
        if (auto stsz = table("stsz")) {
            uint32_t size = u32_at(*stsz, 4, "stsz");
            uint64_t count = sample_count(u32_at(*stsz, 8, "stsz"));
            if (size != 0)
                return std::vector<uint32_t>((size_t)count, size);
            Span t = entries(*stsz, 12, count, 4, "stsz");
            std::vector<uint32_t> out((size_t)count);
            for (size_t i = 0; i < count; i++)
                out[i] = be32(t.p + 4 * i);
            return out;
        }
        auto stz2 = table("stz2");
        if (!stz2)
            bail("the video track has no sample sizes (`stsz` or `stz2`)");
        if (stz2->n < 8)
            bail("truncated `stz2` atom");
        uint8_t field = stz2->p[7];
        uint64_t count = sample_count(u32_at(*stz2, 8, "stz2"));
        std::vector<uint32_t> out;
        out.reserve((size_t)count);
        if (field == 4) {
            Span t = entries(*stz2, 12, (count + 1) / 2, 1, "stz2");
            for (size_t i = 0; i < count; i++)
                out.push_back(i & 1 ? t.p[i / 2] & 15 : t.p[i / 2] >> 4);
        } else if (field == 8) {
            Span t = entries(*stz2, 12, count, 1, "stz2");
            for (size_t i = 0; i < count; i++)
                out.push_back(t.p[i]);
        } else if (field == 16) {
            Span t = entries(*stz2, 12, count, 2, "stz2");
            for (size_t i = 0; i < count; i++)
                out.push_back((uint32_t)(t.p[2 * i] << 8 | t.p[2 * i + 1]));
        } else {
            bail(cat("`stz2` field size ", (unsigned)field, " is not 4, 8 or 16 bits"));
        }
        return out;
    }

    /* Each sample's duration from `stts`, one per sample (mov.rs:420-438). */
    std::vector<uint32_t> durations(size_t samples) const
    {
        Span stts = require(stbl, "stts", "the video track");
        uint32_t count = u32_at(stts, 4, "stts");
        Span t = entries(stts, 8, count, 8, "stts");
        uint64_t total = 0;
        for (size_t i = 0; i < count; i++)
            total += be32(t.p + 8 * i);
        if (total != samples)
            bail(cat("the video track's sample table disagrees: `stts` times ", total,
                     " samples, the sizes list ", samples));
        std::vector<uint32_t> out;
        out.reserve(samples);
        for (size_t i = 0; i < count; i++) {
            uint32_t duration = be32(t.p + 8 * i + 4);
            if (duration == 0)
                bail("the video track has a sample of duration 0");
            out.insert(out.end(), be32(t.p + 8 * i), duration);
        }
        return out;
    }

    /* Each sample's place in the file from the chunk offsets and `stsc`
     * (mov.rs:439-505). */
    std::vector<prr_mov_sample> locate(const std::vector<uint32_t> &sizes, uint64_t length) const
    {
        std::vector<uint64_t> chunks;
        if (auto stco = table("stco")) {
            uint32_t count = u32_at(*stco, 4, "stco");
            Span t = entries(*stco, 8, count, 4, "stco");
            for (size_t i = 0; i < count; i++)
                chunks.push_back(be32(t.p + 4 * i));
        } else {
            auto co64 = table("co64");
            if (!co64)
                bail("the video track has no chunk offsets (`stco` or `co64`)");
            uint32_t count = u32_at(*co64, 4, "co64");
            Span t = entries(*co64, 8, count, 8, "co64");
            for (size_t i = 0; i < count; i++)
                chunks.push_back(be64(t.p + 8 * i));
        }
        Span stsc = require(stbl, "stsc", "the video track");
        uint32_t count = u32_at(stsc, 4, "stsc");
        Span t = entries(stsc, 8, count, 12, "stsc");
        std::vector<std::pair<size_t, size_t>> runs;
        for (size_t i = 0; i < count; i++)
            runs.emplace_back(be32(t.p + 12 * i), be32(t.p + 12 * i + 4));
        bool ordered = !runs.empty() && runs.front().first == 1 && runs.back().first <= chunks.size();
        for (size_t i = 1; ordered && i < runs.size(); i++)
            ordered = runs[i - 1].first < runs[i].first;
        if (!ordered)
            bail(cat("the video track's `stsc` does not list runs of its ", chunks.size(),
                     " chunks from chunk 1 in order"));
        std::vector<prr_mov_sample> samples;
        samples.reserve(sizes.size());
        size_t next_size = 0;
        for (size_t i = 0; i < runs.size(); i++) {
            size_t first = runs[i].first, per_chunk = runs[i].second;
            size_t end = i + 1 < runs.size() ? runs[i + 1].first - 1 : chunks.size();
            for (size_t c = first - 1; c < end; c++) {
                uint64_t offset = chunks[c];
                for (size_t k = 0; k < per_chunk; k++) {
                    if (next_size == sizes.size())
                        bail(cat("the video track's sample table disagrees: its chunks hold more "
                                 "than the ",
                                 sizes.size(), " samples it sizes"));
                    uint32_t size = sizes[next_size++];
                    uint64_t stop = offset + size < offset ? UINT64_MAX : offset + size;
                    if (stop > length)
                        bail(cat("sample ", samples.size(), " (bytes ", offset, "..", stop,
                                 ") lies beyond the end of the file at ", length,
                                 " bytes (a truncated file?)"));
                    samples.push_back(prr_mov_sample{offset, size, 0});
                    offset = stop;
                }
            }
        }
        if (samples.size() != sizes.size())
            bail(cat("the video track's sample table disagrees: its chunks hold ", samples.size(),
                     " of the ", sizes.size(), " samples it sizes"));
        return samples;
    }

    /* The track from the resolved tables (mov.rs:344-378). */
    prr_mov_track read(uint64_t length, uint32_t movie_timescale) const
    {
        std::vector<uint32_t> sz = sizes();
        if (sz.empty())
            bail("the video track has no samples");
        std::vector<uint32_t> dur = durations(sz.size());
        std::vector<prr_mov_sample> samples = locate(sz, length);
        auto [mode, variable] = common_duration(dur);
        uint32_t divisor = gcd(timescale, mode);
        prr_mov_track track;
        track.orientation = orientation(tkhd, track.warnings);
        std::pair<size_t, size_t> presented{0, samples.size()};
        if (elst)
            presented = edit(*elst, dur, timescale, movie_timescale, track.warnings);
        if (presented.first >= presented.second)
            bail(cat("the video track's edit list presents none of its ", samples.size(),
                     " samples"));
        std::memcpy(track.fourcc, fourcc.data(), 4);
        track.fps_num = timescale / divisor;
        track.fps_den = mode / divisor;
        track.variable_durations = variable;
        track.samples.assign(samples.begin() + (ptrdiff_t)presented.first,
                             samples.begin() + (ptrdiff_t)presented.second);
        return track;
    }
};

/* `trak`'s media when it is an enabled video track (mov.rs:303-341). */
std::optional<Video> video_track(Span trak)
{
    std::vector<Atom> atoms = children(trak);
    Span tkhd = require(atoms, "tkhd", "a track");
    if ((u32_at(tkhd, 0, "tkhd") & 1) == 0)
        return std::nullopt;
    std::vector<Atom> mdia = children(require(atoms, "mdia", "a track"));
    /* Version and flags, component type, then the subtype. */
    uint32_t handler = u32_at(require(mdia, "hdlr", "a track's media"), 8, "hdlr");
    if (handler != be32((const uint8_t *)"vide"))
        return std::nullopt;
    uint32_t timescale = times(require(mdia, "mdhd", "the video track"), "mdhd").second;
    if (timescale == 0)
        bail("the video track's timescale is 0");
    std::vector<Atom> minf = children(require(mdia, "minf", "the video track"));
    Video v;
    v.stbl = children(require(minf, "stbl", "the video track"));
    Span stsd = require(v.stbl, "stsd", "the video track");
    if (u32_at(stsd, 4, "stsd") == 0)
        bail("the video track has no sample description");
    /* Version and flags, entry count, then the first entry's size and fourcc. */
    uint32_t fourcc = u32_at(stsd, 12, "stsd");
    v.fourcc = Kind{(uint8_t)(fourcc >> 24), (uint8_t)(fourcc >> 16), (uint8_t)(fourcc >> 8),
                    (uint8_t)fourcc};
    v.timescale = timescale;
    v.tkhd = tkhd;
    if (const Atom *edts = find(atoms, "edts")) {
        for (const Atom &a : children(edts->body))
            if (is(a, "elst")) {
                v.elst = a.body;
                break;
            }
    }
    return v;
}

/* mov.rs:259-292. */
prr_mov_track parse_movie(Span moov, uint64_t length, prr_mov_accept_fn accept, void *user)
{
    std::vector<Atom> atoms = children(moov);
    Span mvhd = require(atoms, "mvhd", "the movie");
    auto [creation_time, movie_timescale] = times(mvhd, "mvhd");
    std::optional<Kind> rejected;
    for (const Atom &trak : atoms) {
        if (!is(trak, "trak"))
            continue;
        std::optional<Video> video = video_track(trak.body);
        if (!video)
            continue;
        if (!accept(video->fourcc.data(), user)) {
            if (!rejected)
                rejected = video->fourcc;
            continue;
        }
        prr_mov_track track = video->read(length, movie_timescale);
        if (creation_time != 0)
            track.creation_time = creation_time;
        if (const Atom *meta = find(atoms, "meta")) {
            /* Metadata is optional: a damaged item costs the tags, not the frames. */
            try {
                track.metadata = text_metadata(meta->body);
            } catch (const MovFail &f) {
                track.warnings.push_back("the movie's metadata is unreadable: " + f.msg);
            }
        }
        return track;
    }
    if (rejected)
        bail(cat("the video track is `", name(*rejected), "`, not a format this reader accepts"));
    bail("the movie has no enabled video track");
}

std::string io_message(int e) { return cat(std::strerror(e), " (os error ", e, ")"); }

/* read_exact_at: the whole range or an error message (mov.rs:183-204). */
bool pread_exact(int fd, uint64_t offset, uint8_t *buf, size_t len, std::string &why)
{
    while (len) {
        ssize_t got = ::pread(fd, buf, len, (off_t)offset);
        if (got < 0) {
            if (errno == EINTR)
                continue;
            why = io_message(errno);
            return false;
        }
        if (got == 0) {
            why = "failed to fill whole buffer";
            return false;
        }
        buf += got;
        len -= (size_t)got;
        offset += (uint64_t)got;
    }
    return true;
}

void read_at(int fd, uint64_t offset, uint8_t *buf, size_t len)
{
    std::string why;
    if (!pread_exact(fd, offset, buf, len, why))
        bail(cat("reading ", len, " bytes at byte ", offset, ": ", why));
}

/* The body of the top-level `moov` atom (mov.rs:205-255). */
std::vector<uint8_t> read_moov(int fd, uint64_t length)
{
    static const char *first_atoms[] = {"ftyp", "moov", "mdat", "wide", "free", "skip", "pnot"};
    uint64_t position = 0;
    while (position < length) {
        uint8_t header[16] = {0};
        size_t available = (size_t)std::min<uint64_t>(length - position, 16);
        if (available < 8)
            bail(cat("truncated atom header at byte ", position));
        read_at(fd, position, header, available);
        const uint8_t *kind = header + 4;
        if (position == 0) {
            bool known = false;
            for (const char *k : first_atoms)
                known |= std::memcmp(kind, k, 4) == 0;
            if (!known)
                bail(cat("not a QuickTime movie: it starts with `", name(kind),
                         "`, not `ftyp` or `moov`"));
        }
        uint64_t header_size, size;
        switch (be32(header)) {
        case 1:
            if (available != 16)
                bail(cat("truncated `", name(kind), "` atom header"));
            header_size = 16;
            size = be64(header + 8);
            break;
        case 0:
            header_size = 8;
            size = length - position;
            break;
        default:
            header_size = 8;
            size = be32(header);
        }
        if (size < header_size)
            bail(cat("invalid `", name(kind), "` atom size ", size, " at byte ", position));
        if (size > length - position)
            bail(cat("truncated `", name(kind), "` atom at byte ", position, ": it claims ", size,
                     " bytes and the file has ", length - position,
                     " left (an interrupted recording?)"));
        if (std::memcmp(kind, "moov", 4) == 0) {
            uint64_t body = size - header_size;
            if (body > kMaxMoov)
                bail(cat("the `moov` atom is ", body, " bytes, more than the ", kMaxMoov,
                         " this reader accepts"));
            std::vector<uint8_t> moov((size_t)body);
            read_at(fd, position + header_size, moov.data(), moov.size());
            return moov;
        }
        position += size;
    }
    if (length == 0)
        bail("not a QuickTime movie: the file is empty");
    bail("the movie has no `moov` atom, so its frames cannot be found (an interrupted "
         "recording?)");
}

int accept_prores_raw(const uint8_t fourcc[4], void *) { return prr_is_prores_raw_fourcc(fourcc); }

int mov_fail(const std::string &msg, char *out, size_t cap)
{
    if (out && cap) {
        size_t n = std::min(msg.size(), cap - 1);
        std::memcpy(out, msg.data(), n);
        out[n] = 0;
    }
    return PRR_E_MOV;
}

} // namespace

extern "C" {

int prr_is_prores_raw_fourcc(const uint8_t fourcc[4])
{
    if (!fourcc)
        return 0;
    char lower[4];
    for (int i = 0; i < 4; i++)
        lower[i] = (char)((fourcc[i] >= 'A' && fourcc[i] <= 'Z') ? fourcc[i] + 32 : fourcc[i]);
    return std::memcmp(lower, "aprn", 4) == 0 || std::memcmp(lower, "aprh", 4) == 0;
}

int prr_mov_read_video_track(int fd, prr_mov_accept_fn accept, void *user, prr_mov_track **out,
                             char *msg, size_t msg_cap)
{
    if (!out)
        return PRR_E_INVALID_ARGUMENT;
    *out = nullptr;
    try {
        struct stat st;
        if (::fstat(fd, &st) != 0)
            bail(cat("reading the file's length: ", io_message(errno)));
        uint64_t length = (uint64_t)st.st_size;
        std::vector<uint8_t> moov = read_moov(fd, length);
        prr_mov_track track = parse_movie(Span{moov.data(), moov.size()}, length,
                                          accept ? accept : accept_prores_raw, user);
        *out = new prr_mov_track(std::move(track));
        return PRR_OK;
    } catch (const MovFail &f) {
        return mov_fail(f.msg, msg, msg_cap);
    } catch (const std::bad_alloc &) {
        return mov_fail("out of memory", msg, msg_cap);
    } catch (const std::exception &e) {
        return mov_fail(e.what(), msg, msg_cap);
    }
}

void prr_mov_track_free(prr_mov_track *track) { delete track; }

void prr_mov_track_fourcc(const prr_mov_track *track, uint8_t out[4]) { std::memcpy(out, track->fourcc, 4); }

void prr_mov_track_fps(const prr_mov_track *track, uint32_t *num, uint32_t *den)
{
    *num = track->fps_num;
    *den = track->fps_den;
}

int prr_mov_track_variable_durations(const prr_mov_track *track) { return track->variable_durations; }

size_t prr_mov_track_samples(const prr_mov_track *track, const prr_mov_sample **samples)
{
    if (samples)
        *samples = track->samples.data();
    return track->samples.size();
}

int prr_mov_track_creation_time(const prr_mov_track *track, uint64_t *out)
{
    if (!track->creation_time)
        return 0;
    if (out)
        *out = *track->creation_time;
    return 1;
}

int prr_mov_track_orientation(const prr_mov_track *track) { return track->orientation; }

size_t prr_mov_track_n_warnings(const prr_mov_track *track) { return track->warnings.size(); }

static const char *out_string(const std::string &s, size_t *len)
{
    if (len)
        *len = s.size();
    return s.c_str();
}

const char *prr_mov_track_warning(const prr_mov_track *track, size_t i, size_t *len)
{
    return i < track->warnings.size() ? out_string(track->warnings[i], len) : nullptr;
}

size_t prr_mov_track_n_metadata(const prr_mov_track *track) { return track->metadata.size(); }

const char *prr_mov_track_metadata_key(const prr_mov_track *track, size_t i, size_t *len)
{
    return i < track->metadata.size() ? out_string(track->metadata[i].first, len) : nullptr;
}

const char *prr_mov_track_metadata_value(const prr_mov_track *track, size_t i, size_t *len)
{
    return i < track->metadata.size() ? out_string(track->metadata[i].second, len) : nullptr;
}

const char *prr_mov_track_text(const prr_mov_track *track, const char *const *keys, size_t n_keys,
                               size_t *len)
{
    for (size_t k = 0; k < n_keys; k++)
        for (const auto &item : track->metadata)
            if (item.first == keys[k])
                return out_string(item.second, len);
    return nullptr;
}

int prr_mov_read_at(int fd, uint64_t offset, uint8_t *buf, size_t len, char *msg, size_t msg_cap)
{
    if (!buf && len)
        return PRR_E_INVALID_ARGUMENT;
    try {
        read_at(fd, offset, buf, len);
        return PRR_OK;
    } catch (const MovFail &f) {
        return mov_fail(f.msg, msg, msg_cap);
    }
}

int prr_mov_read_samples(int fd, const prr_mov_sample *samples, size_t n, uint8_t *dst,
                         const uint64_t *dst_offsets, unsigned threads, char *msg, size_t msg_cap)
{
    if (n && (!samples || !dst || !dst_offsets))
        return PRR_E_INVALID_ARGUMENT;
    if (threads == 0)
        threads = (unsigned)std::min<size_t>(n, 16);
    threads = (unsigned)std::max<size_t>(1, std::min<size_t>(threads, n));
    std::atomic<size_t> next{0};
    std::mutex lock;
    size_t bad = SIZE_MAX;
    std::string why;
    auto worker = [&]() {
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= n)
                return;
            std::string e;
            if (!pread_exact(fd, samples[i].offset, dst + dst_offsets[i], samples[i].size, e)) {
                std::lock_guard<std::mutex> g(lock);
                if (i < bad) {
                    bad = i;
                    why = cat("reading ", samples[i].size, " bytes at byte ", samples[i].offset,
                              ": ", e);
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; t++)
        pool.emplace_back(worker);
    worker();
    for (std::thread &t : pool)
        t.join();
    if (bad != SIZE_MAX)
        return mov_fail(why, msg, msg_cap);
    return PRR_OK;
}

} // extern "C"
