/* Oracle (test infrastructure only): the crate's lossless JPEG tiles of a
 * DNG, restated in C.
 *
 *   orc_dng_tile      - dng.rs:115-128 `tile`: a 256x256 tile of the visible
 *                       picture, edge sensels repeating the nearest sensel of
 *                       the same CFA phase;
 *   orc_lj92_table    - lj92.rs:105-206 `HuffmanTable::optimal`;
 *   orc_lj92_encode   - lj92.rs:14-93 `encode` and `differences` with its
 *                       BitWriter (lj92.rs:208-247);
 *   orc_dng_lj92_tiles- dng.rs:92-113: every tile encoded (threads over
 *                       tiles, as the crate's rayon map), concatenated in
 *                       tile order with their offsets and lengths.
 *
 * Only tests/ and the benchmark's CPU baseline load this code.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TILE 256

/* dng.rs:119-120: the last column or row of the same CFA phase as i, for i
 * past n. */
static size_t clamp_phase(size_t i, size_t n) { return i < n ? i : n - 2 + (i - n) % 2; }

void orc_dng_tile(const uint16_t *values, size_t width, size_t height, size_t left, size_t top,
                  uint16_t *tile)
{
    for (size_t y = 0; y < TILE; y++) {
        const uint16_t *row = values + clamp_phase(top + y, height) * width;
        for (size_t x = 0; x < TILE; x++)
            tile[y * TILE + x] = row[clamp_phase(left + x, width)];
    }
}

/* lj92.rs:90-93: the bit length of |d|. */
static unsigned category(int32_t d)
{
    uint32_t a = d < 0 ? (uint32_t)(-(int64_t)d) : (uint32_t)d;
    return a ? 32u - (unsigned)__builtin_clz(a) : 0u;
}

/* HuffmanTable::optimal (lj92.rs:105-206). counts[16] = BITS, values[] =
 * HUFFVAL (returns how many), code[17] / length[17] per category (length 0
 * when unused). */
int orc_lj92_table(const uint64_t histogram[17], uint8_t counts[16], uint8_t values[17],
                   uint16_t code[17], uint8_t length[17])
{
    enum { RESERVED = 17 };
    uint64_t freq[18];
    size_t size[18];
    int others[18];
    for (int v = 0; v < 17; v++)
        freq[v] = histogram[v];
    freq[RESERVED] = 1;
    for (int v = 0; v < 18; v++) {
        size[v] = 0;
        others[v] = -1;
    }
    for (;;) {
        int least = -1, second = -1;
        for (int v = 0; v < 18; v++) {
            if (freq[v] == 0)
                continue;
            if (least < 0 || freq[v] <= freq[least]) {
                second = least;
                least = v;
            } else if (second < 0 || freq[v] <= freq[second]) {
                second = v;
            }
        }
        if (least < 0 || second < 0)
            break;
        int v1 = least, v2 = second;
        freq[v1] += freq[v2];
        freq[v2] = 0;
        int v = v1;
        for (;;) {
            size[v]++;
            if (others[v] < 0)
                break;
            v = others[v];
        }
        others[v] = v2;
        v = v2;
        for (;;) {
            size[v]++;
            if (others[v] < 0)
                break;
            v = others[v];
        }
    }
    uint32_t bits[33] = {0};
    for (int v = 0; v < 18; v++)
        if (size[v] > 0)
            bits[size[v]]++;
    for (int i = 32; i > 16; i--) {
        while (bits[i] > 0) {
            int j = i - 2;
            while (bits[j] == 0)
                j--;
            bits[i] -= 2;
            bits[i - 1] += 1;
            bits[j + 1] += 2;
            bits[j] -= 1;
        }
    }
    int i = 16;
    while (i > 0 && bits[i] == 0)
        i--;
    if (bits[i])
        bits[i]--;
    int n = 0;
    for (size_t s = 1; s <= 32; s++)
        for (int v = 0; v < 17; v++)
            if (size[v] == s)
                values[n++] = (uint8_t)v;
    for (int k = 0; k < 16; k++)
        counts[k] = (uint8_t)bits[k + 1];
    memset(code, 0, 17 * sizeof code[0]);
    memset(length, 0, 17);
    uint16_t c = 0;
    int k = 0;
    for (int len = 1; len <= 16; len++) {
        for (int m = 0; m < counts[len - 1]; m++) {
            code[values[k]] = c;
            length[values[k]] = (uint8_t)len;
            c++;
            k++;
        }
        c <<= 1;
    }
    return n;
}

typedef struct {
    uint8_t *out;
    size_t len, cap;
    uint64_t buffer;
    uint32_t count;
} Writer;

static void push(Writer *w, uint8_t b)
{
    if (w->len < w->cap)
        w->out[w->len] = b;
    w->len++;
}

/* BitWriter::put (lj92.rs:224-237); finish (lj92.rs:239-246) is the
 * padding at the end of orc_lj92_encode. */
static void put(Writer *w, uint64_t value, unsigned length)
{
    w->buffer = (w->buffer << length) | value;
    w->count += length;
    while (w->count >= 8) {
        w->count -= 8;
        uint8_t byte = (uint8_t)(w->buffer >> w->count);
        push(w, byte);
        if (byte == 0xFF)
            push(w, 0);
    }
    w->buffer &= (1ull << w->count) - 1;
}

static void push16(Writer *w, unsigned v)
{
    push(w, (uint8_t)(v >> 8));
    push(w, (uint8_t)v);
}

/* encode and differences (lj92.rs:14-88): the stream's length (bytes beyond `cap` are
 * counted, not written). */
size_t orc_lj92_encode(const uint16_t *samples, size_t width, size_t height, size_t components,
                       uint8_t *out, size_t cap)
{
    size_t n = width * height;
    int32_t *diff = malloc((n ? n : 1) * sizeof *diff);
    uint64_t histogram[17] = {0};
    for (size_t i = 0; i < n; i++) {
        size_t row = i / width, column = i % width;
        uint16_t pred = column >= components ? samples[i - components]
                        : row > 0            ? samples[i - width]
                                             : (uint16_t)(1u << 15);
        diff[i] = (int16_t)(uint16_t)(samples[i] - pred);
        histogram[category(diff[i])]++;
    }
    uint8_t counts[16], values[17], length[17];
    uint16_t code[17];
    int nv = orc_lj92_table(histogram, counts, values, code, length);
    Writer w = {out, 0, cap, 0, 0};
    push16(&w, 0xFFD8);
    push16(&w, 0xFFC3);
    push16(&w, 8 + 3 * (unsigned)components);
    push(&w, 16);
    push16(&w, (unsigned)height);
    push16(&w, (unsigned)(width / components));
    push(&w, (uint8_t)components);
    for (size_t id = 0; id < components; id++) {
        push(&w, (uint8_t)id);
        push(&w, 0x11);
        push(&w, 0);
    }
    push16(&w, 0xFFC4);
    push16(&w, 2 + 1 + 16 + (unsigned)nv);
    push(&w, 0);
    for (int k = 0; k < 16; k++)
        push(&w, counts[k]);
    for (int k = 0; k < nv; k++)
        push(&w, values[k]);
    push16(&w, 0xFFDA);
    push16(&w, 6 + 2 * (unsigned)components);
    push(&w, (uint8_t)components);
    for (size_t id = 0; id < components; id++) {
        push(&w, (uint8_t)id);
        push(&w, 0);
    }
    push(&w, 1);
    push(&w, 0);
    push(&w, 0);
    for (size_t i = 0; i < n; i++) {
        int32_t d = diff[i];
        unsigned ssss = category(d);
        put(&w, code[ssss], length[ssss]);
        if (ssss > 0 && ssss < 16) {
            int32_t extra = d < 0 ? d - 1 : d;
            put(&w, (uint64_t)(uint32_t)extra & ((1ull << ssss) - 1), ssss);
        }
    }
    if (w.count > 0) {
        unsigned pad = 8 - w.count;
        put(&w, (1ull << pad) - 1, pad);
    }
    push16(&w, 0xFFD9);
    free(diff);
    return w.len;
}

/* The worst-case stream of one 256x256 tile: headers, 31 bits a sample
 * (a 16-bit code and 15 extra bits), every byte stuffed, EOI. */
size_t orc_lj92_tile_capacity(void) { return 128 + 2 * ((size_t)TILE * TILE * 31 / 8 + 1); }

typedef struct {
    const uint16_t *values;
    size_t width, height, across, n_tiles;
    uint8_t *scratch;
    size_t *lens;
    size_t next;
    pthread_mutex_t lock;
} Job;

static void *tiles_worker(void *arg)
{
    Job *j = arg;
    uint16_t *tile = malloc(sizeof(uint16_t) * TILE * TILE);
    for (;;) {
        pthread_mutex_lock(&j->lock);
        size_t t = j->next++;
        pthread_mutex_unlock(&j->lock);
        if (t >= j->n_tiles)
            break;
        orc_dng_tile(j->values, j->width, j->height, t % j->across * TILE, t / j->across * TILE,
                     tile);
        j->lens[t] = orc_lj92_encode(tile, TILE, TILE, 2, j->scratch + t * orc_lj92_tile_capacity(),
                                     orc_lj92_tile_capacity());
    }
    free(tile);
    return NULL;
}

/* dng.rs:92-113: every 256x256 tile of the width x height picture encoded
 * with two components; the streams concatenated in tile order into `out`
 * (at most `cap` bytes) with each one's offset and length. Returns the total
 * bytes (0 when they do not fit). */
size_t orc_dng_lj92_tiles(const uint16_t *values, size_t width, size_t height, uint8_t *out,
                          size_t cap, uint32_t *offsets, uint32_t *lengths, int threads)
{
    size_t across = (width + TILE - 1) / TILE, down = (height + TILE - 1) / TILE;
    Job j = {values, width, height, across, across * down, NULL, NULL, 0,
             PTHREAD_MUTEX_INITIALIZER};
    j.scratch = malloc(j.n_tiles * orc_lj92_tile_capacity());
    j.lens = calloc(j.n_tiles, sizeof *j.lens);
    if (threads < 1)
        threads = 1;
    pthread_t *pool = malloc(sizeof *pool * (size_t)threads);
    for (int t = 0; t < threads; t++)
        pthread_create(&pool[t], NULL, tiles_worker, &j);
    for (int t = 0; t < threads; t++)
        pthread_join(pool[t], NULL);
    free(pool);
    size_t total = 0;
    for (size_t t = 0; t < j.n_tiles; t++) {
        if (total + j.lens[t] <= cap)
            memcpy(out + total, j.scratch + t * orc_lj92_tile_capacity(), j.lens[t]);
        offsets[t] = (uint32_t)total;
        lengths[t] = (uint32_t)j.lens[t];
        total += j.lens[t];
    }
    free(j.scratch);
    free(j.lens);
    return total <= cap ? total : 0;
}
