/* proresraw_mov: the sample table of one video track of a QuickTime (or ISO
 * base media) movie, and positional reads of its samples, behind a C ABI.
 *
 * Replaces the crate's `mov` module (src/mov.rs) and `is_prores_raw_fourcc`
 * (src/lib.rs:68-75). Only the `moov` atom is read into memory; samples are
 * read with positional I/O, so one file descriptor serves every reader thread
 * without a lock. The reader is host code by nature (container parsing);
 * prr_mov_read_samples gathers many samples at once into one caller buffer
 * (for example pinned host memory that is then copied to the device and
 * decoded as a batch).
 *
 * Conventions: functions return 0 (PRR_OK) or PRR_E_MOV (104) with the
 * crate's message for mov::Error (a sentence naming the atom or structure at
 * fault, outermost context first) written NUL-terminated into `msg` (at most
 * `msg_cap` bytes, truncated when longer; `msg` may be NULL), or
 * PRR_E_INVALID_ARGUMENT for a NULL pointer.
 */
#ifndef PRORESRAW_MOV_H
#define PRORESRAW_MOV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PRR_E_MOV = 104 };

/* Sample (mov.rs:91-98): where one sample's bytes are in the file. */
typedef struct {
    uint64_t offset;
    uint32_t size;
    uint32_t reserved;
} prr_mov_sample;

/* Orientation (mov.rs:124-146), as its EXIF / TIFF value 1..8
 * (Orientation::exif, mov.rs:149-152). */
enum prr_orientation {
    PRR_ORIENT_IDENTITY = 1,
    PRR_ORIENT_MIRROR_HORIZONTAL = 2,
    PRR_ORIENT_ROTATE_180 = 3,
    PRR_ORIENT_MIRROR_VERTICAL = 4,
    PRR_ORIENT_TRANSPOSE = 5,
    PRR_ORIENT_ROTATE_90 = 6,
    PRR_ORIENT_TRANSVERSE = 7,
    PRR_ORIENT_ROTATE_270 = 8,
};

/* The `accept` argument of read_video_track (mov.rs:172): non-zero when a
 * sample description fourcc is wanted. */
typedef int (*prr_mov_accept_fn)(const uint8_t fourcc[4], void *user);

/* Track (mov.rs:100-119), opaque; read with the accessors below. */
typedef struct prr_mov_track prr_mov_track;

/* is_prores_raw_fourcc (lib.rs:72-75): `aprn` or `aprh`, any case. */
int prr_is_prores_raw_fourcc(const uint8_t fourcc[4]);

/* read_video_track (mov.rs:171-176): the first enabled video track of the
 * movie open on `fd` whose sample description `accept` takes (NULL accepts
 * prr_is_prores_raw_fourcc), with its samples. Free with prr_mov_track_free. */
int prr_mov_read_video_track(int fd, prr_mov_accept_fn accept, void *user, prr_mov_track **out,
                             char *msg, size_t msg_cap);
void prr_mov_track_free(prr_mov_track *track);

/* Track fields (mov.rs:102-118). */
void prr_mov_track_fourcc(const prr_mov_track *track, uint8_t out[4]);
void prr_mov_track_fps(const prr_mov_track *track, uint32_t *num, uint32_t *den);
int prr_mov_track_variable_durations(const prr_mov_track *track);
/* The presented samples, in order; the array lives as long as the track. */
size_t prr_mov_track_samples(const prr_mov_track *track, const prr_mov_sample **samples);
/* 1 and the creation time (seconds since 1904-01-01 UTC) when recorded, else 0. */
int prr_mov_track_creation_time(const prr_mov_track *track, uint64_t *out);
/* prr_orientation value (Orientation::exif). */
int prr_mov_track_orientation(const prr_mov_track *track);
/* Strings are UTF-8, NUL-terminated, and may also hold NUL bytes (as a
 * Rust String may): `len`, when not NULL, receives the byte length. */
size_t prr_mov_track_n_warnings(const prr_mov_track *track);
const char *prr_mov_track_warning(const prr_mov_track *track, size_t i, size_t *len);
/* The movie's text metadata items (`mdta` keys), in file order. */
size_t prr_mov_track_n_metadata(const prr_mov_track *track);
const char *prr_mov_track_metadata_key(const prr_mov_track *track, size_t i, size_t *len);
const char *prr_mov_track_metadata_value(const prr_mov_track *track, size_t i, size_t *len);
/* Track::text (mov.rs:159-169): the first value stored under any of
 * `keys` (NUL-terminated), in that order, or NULL. */
const char *prr_mov_track_text(const prr_mov_track *track, const char *const *keys,
                               size_t n_keys, size_t *len);

/* read_at (mov.rs:178-182): `len` bytes at byte `offset` of `fd`. */
int prr_mov_read_at(int fd, uint64_t offset, uint8_t *buf, size_t len, char *msg, size_t msg_cap);

/* Many read_at calls at once (no counterpart in the crate, whose callers
 * loop over read_at from rayon threads): sample i is read to
 * dst + dst_offsets[i], by `threads` host threads (0 = one per sample up to
 * 16). On failure the message is that of the lowest-numbered failing
 * sample. */
int prr_mov_read_samples(int fd, const prr_mov_sample *samples, size_t n, uint8_t *dst,
                         const uint64_t *dst_offsets, unsigned threads, char *msg,
                         size_t msg_cap);

#ifdef __cplusplus
}
#endif
#endif
