/* TPGPRVCP version-2 checkpoints for the engine's maps (C17).
 *
 * The on-disk record is the reference's restart format (engine/include/checkpoint_rev.h:7-40,
 * written by engine/src/checkpoint_rev.c), kept byte for byte so that either program resumes
 * the other's runs: the 8-byte magic "TPGPRVCP", version 2, a fixed header of 64-bit
 * little-endian fields, the Hamiltonian terms, the map (key words, then the value, per term),
 * and the SHA-256 of everything before it.  A companion "<path>.manifest.json" carries the
 * payload hash and run metadata.
 *
 * The code is organised around one description of the record: `HEADER` and `TERM` list every
 * serialised field with its C location, its encoding, the largest value its encoding admits
 * and whether it belongs to the run configuration.  The same tables drive the encoder, the
 * decoder (with its range checks) and the configuration comparison; the semantic invariants
 * of a record are checked by one function on save and on load.  Reading maps the file, checks
 * its length and checksum before anything is allocated, then decodes straight from the
 * mapping.  Every file is published by write-to-temporary, fsync, rename and a directory
 * fsync.
 */
#define _POSIX_C_SOURCE 200809L
#include "../../include/tpgp_engine.h"

#include <errno.h>
#include <fcntl.h>
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

_Static_assert(FLT_RADIX == 2 && DBL_MANT_DIG == 53 && sizeof(double) == 8,
               "the record stores IEEE binary64 values");
_Static_assert(sizeof(size_t) == 8, "the record stores sizes as 64-bit words");

/* ---- SHA-256 (FIPS 180-4) ---------------------------------------------------------------- */
typedef struct {
    uint32_t state[8];
    uint64_t length;
    unsigned char buffer[64];
    size_t fill;
} sha_ctx;

static uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static const uint32_t sha_k[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u,
    0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu,
    0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu,
    0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u,
    0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu,
    0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
    0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u,
    0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u,
    0xc67178f2u};

static void sha_compress(sha_ctx *c, const unsigned char *blk) {
    uint32_t w[64];
    for (int t = 0; t < 16; t++)
        w[t] = (uint32_t)blk[4 * t] << 24 | (uint32_t)blk[4 * t + 1] << 16 |
               (uint32_t)blk[4 * t + 2] << 8 | (uint32_t)blk[4 * t + 3];
    for (int t = 16; t < 64; t++) {
        const uint32_t s0 = rotr32(w[t - 15], 7) ^ rotr32(w[t - 15], 18) ^ (w[t - 15] >> 3);
        const uint32_t s1 = rotr32(w[t - 2], 17) ^ rotr32(w[t - 2], 19) ^ (w[t - 2] >> 10);
        w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    uint32_t v[8];
    memcpy(v, c->state, sizeof(v));
    for (int t = 0; t < 64; t++) {
        const uint32_t e = v[4], a = v[0];
        const uint32_t ch = (e & v[5]) ^ (~e & v[6]);
        const uint32_t maj = (a & v[1]) ^ (a & v[2]) ^ (v[1] & v[2]);
        const uint32_t t1 = v[7] + (rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25)) + ch + sha_k[t] +
                            w[t];
        const uint32_t t2 = (rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22)) + maj;
        memmove(v + 1, v, 7 * sizeof(uint32_t));
        v[4] += t1;
        v[0] = t1 + t2;
    }
    for (int i = 0; i < 8; i++) c->state[i] += v[i];
}
static void sha_start(sha_ctx *c) {
    static const uint32_t iv[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    memcpy(c->state, iv, sizeof(iv));
    c->length = 0;
    c->fill = 0;
}
static void sha_feed(sha_ctx *c, const void *data, size_t n) {
    const unsigned char *p = data;
    c->length += n;
    while (n) {
        size_t take = 64 - c->fill < n ? 64 - c->fill : n;
        memcpy(c->buffer + c->fill, p, take);
        c->fill += take;
        p += take;
        n -= take;
        if (c->fill == 64) {
            sha_compress(c, c->buffer);
            c->fill = 0;
        }
    }
}
static void sha_finish(sha_ctx *c, unsigned char out[32]) {
    const uint64_t bits = c->length * 8;
    unsigned char pad = 0x80, zero = 0, len[8];
    sha_feed(c, &pad, 1);
    while (c->fill != 56) sha_feed(c, &zero, 1);
    for (int i = 0; i < 8; i++) len[i] = (unsigned char)(bits >> (56 - 8 * i));
    sha_feed(c, len, 8);
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = (unsigned char)(c->state[i] >> (24 - 8 * j));
}
void tpgp_payload_sha256(const void *data, size_t bytes, unsigned char digest[32]) {
    sha_ctx c;
    sha_start(&c);
    sha_feed(&c, data, bytes);
    sha_finish(&c, digest);
}

/* ---- the record's description ------------------------------------------------------------ */
typedef enum {
    AS_SIZE,   /* size_t, stored as is */
    AS_INT,    /* int (or an int-sized enum), stored as is; negative values are invalid */
    AS_UINT,   /* unsigned */
    AS_DEPTH,  /* int >= -1, stored plus one */
    AS_F64,    /* double, stored as its bit pattern */
    AS_TEXT32  /* char[32], stored as 32 raw bytes */
} encoding;
enum { CONFIG = 1 }; /* the field is part of the run configuration */
typedef struct {
    encoding as;
    size_t at;        /* offset in the owning struct */
    uint64_t top;     /* largest admissible stored value (integer encodings) */
    unsigned flags;
} field;

#define AT(member) offsetof(tpgp_checkpoint, member)
#define ANY UINT64_MAX
/* Header fields, in record order after the magic and the version. */
static const field HEADER[] = {
    {AS_SIZE, AT(map.n_spins), ANY, CONFIG},
    {AS_SIZE, AT(map.len), ANY, 0},
    {AS_SIZE, AT(term_count), ANY, CONFIG},
    {AS_SIZE, AT(step), ANY, 0},
    {AS_SIZE, AT(total_steps), ANY, CONFIG},
    {AS_F64, AT(dt), 0, CONFIG},
    {AS_DEPTH, AT(residual_depth), 11, CONFIG},
    {AS_INT, AT(initial_global), 1, CONFIG},
    {AS_TEXT32, AT(model), 0, CONFIG},
    {AS_F64, AT(j2), 0, CONFIG},
    {AS_SIZE, AT(linear_size), ANY, CONFIG},
    {AS_SIZE, AT(resumed_from_step), ANY, 0},
    {AS_SIZE, AT(prior_total_steps), ANY, 0},
    {AS_SIZE, AT(prior_max_terms), ANY, 0},
    {AS_INT, AT(stop_reason), 6, 0},
    {AS_INT, AT(last_status), TPGP_EBUDGET, 0},
    {AS_F64, AT(options.threshold), 0, CONFIG},
    {AS_SIZE, AT(options.max_terms), ANY, CONFIG},
    {AS_UINT, AT(options.threads), 65536, CONFIG},
    {AS_INT, AT(options.backend), 4, CONFIG},
    {AS_INT, AT(options.validated), 1, CONFIG},
    {AS_F64, AT(options.max_drop_l2), 0, CONFIG},
    {AS_F64, AT(ledger.dropped_l1), 0, 0},
    {AS_F64, AT(ledger.dropped_l2), 0, 0},
    {AS_F64, AT(ledger.dropped_l2_squared), 0, 0},
    {AS_F64, AT(ledger.arithmetic_l2), 0, 0},
    {AS_SIZE, AT(ledger.rotations), ANY, 0},
    {AS_SIZE, AT(ledger.removal_events), ANY, 0},
    {AS_SIZE, AT(ledger.peak_terms), ANY, 0},
    {AS_INT, AT(ledger.arithmetic_enclosed), 1, 0},
    {AS_F64, AT(pf_bound), 0, 0},
    {AS_F64, AT(residual_bound), 0, 0},
    {AS_F64, AT(max_fid_bound), 0, 0},
};
#undef AT
#define AT(member) offsetof(tpgp_term, member)
/* One Hamiltonian term; every field is configuration. */
static const field TERM[] = {
    {AS_SIZE, AT(i), ANY, CONFIG},   {AS_SIZE, AT(j), ANY, CONFIG},
    {AS_UINT, AT(qi), 3, CONFIG},    {AS_UINT, AT(qj), 3, CONFIG},
    {AS_F64, AT(h), 0, CONFIG},      {AS_SIZE, AT(layer), ANY, CONFIG},
};
#undef AT
#define COUNT(t) (sizeof(t) / sizeof((t)[0]))
static const char MAGIC[8] = {'T', 'P', 'G', 'P', 'R', 'V', 'C', 'P'};
enum { VERSION = 2, DIGEST = 32 };

static size_t encoded_size(const field *f) { return f->as == AS_TEXT32 ? 32 : 8; }
static size_t table_size(const field *t, size_t n) {
    size_t bytes = 0;
    for (size_t k = 0; k < n; k++) bytes += encoded_size(t + k);
    return bytes;
}
/* The stored 64-bit word of an integer or floating field. */
static uint64_t stored_word(const field *f, const void *base) {
    const char *p = (const char *)base + f->at;
    switch (f->as) {
    case AS_SIZE: return *(const size_t *)p;
    case AS_INT: return (uint64_t)(int64_t) * (const int *)p;
    case AS_UINT: return *(const unsigned *)p;
    case AS_DEPTH: return (uint64_t)((int64_t) * (const int *)p + 1);
    case AS_F64: {
        uint64_t u;
        memcpy(&u, p, 8);
        return u;
    }
    default: return 0;
    }
}
/* Store a decoded word; 0 when it is outside what the field admits. */
static int load_word(const field *f, void *base, uint64_t w) {
    char *p = (char *)base + f->at;
    if (f->as != AS_F64 && w > f->top) return 0;
    switch (f->as) {
    case AS_SIZE: *(size_t *)p = (size_t)w; break;
    case AS_INT: *(int *)p = (int)w; break;
    case AS_UINT: *(unsigned *)p = (unsigned)w; break;
    case AS_DEPTH: *(int *)p = (int)w - 1; break;
    case AS_F64: memcpy(p, &w, 8); break;
    default: return 0;
    }
    return 1;
}

/* ---- semantic invariants ------------------------------------------------------------------
 * One function for save and load: NULL when the record is well formed, else the first broken
 * rule.  The rules are the format's (checkpoint_rev.c:121-197): geometry, run state, options,
 * ledger, then the map's canonical order and the terms. */
static int mul_fits(size_t a, size_t b) { return b == 0 || a <= SIZE_MAX / b; }
static int nonneg(double x) { return !isnan(x) && x >= 0; }
static int finite_nonneg(double x) { return isfinite(x) && x >= 0; }

static const char *broken_rule(const tpgp_checkpoint *s) {
    if (!s) return "no record";
    const tpgp_map *m = &s->map;
    /* geometry and storage */
    if (!m->n_spins || m->n_spins > SIZE_MAX - 31 || m->n_words != (m->n_spins + 31) / 32)
        return "spin count and key width disagree";
    if (m->len > m->capacity || (m->len && (!m->keys || !m->values)) ||
        (s->term_count && !s->terms))
        return "map or term storage missing";
    if (!mul_fits(m->len, m->n_words) || !mul_fits(m->len * m->n_words, 8) ||
        !mul_fits(m->len, sizeof(double)) || !mul_fits(s->term_count, sizeof(tpgp_term)))
        return "record sizes overflow";
    /* run state */
    if (!isfinite(s->dt) || !(s->dt > 0)) return "time step must be finite and positive";
    if (s->step > s->total_steps || s->resumed_from_step > s->step)
        return "step counters out of order";
    if (s->stop_reason == 1 && s->step != s->total_steps)
        return "a completed run must stand at its last step";
    if (s->residual_depth < -1 || s->residual_depth > 10) return "residual depth out of range";
    if (s->initial_global != 0 && s->initial_global != 1) return "initial-state flag not 0/1";
    if (s->stop_reason < 0 || s->stop_reason > 6) return "unknown stop reason";
    if ((int)s->last_status < TPGP_OK || (int)s->last_status > TPGP_EBUDGET)
        return "unknown status";
    if (!isfinite(s->j2)) return "J2 must be finite";
    const char *end = memchr(s->model, 0, sizeof(s->model));
    if (!end || end == s->model) return "model name empty or unterminated";
    for (const char *c = s->model; c < end; c++)
        if (!((*c >= 'a' && *c <= 'z') || (*c >= '0' && *c <= '9') || *c == '_'))
            return "model name has characters outside [a-z0-9_]";
    /* options */
    const tpgp_options *o = &s->options;
    if (!finite_nonneg(o->threshold) || !finite_nonneg(o->max_drop_l2))
        return "threshold and drop budget must be finite and non-negative";
    if (o->threads < 1 || o->threads > 65536) return "thread count out of range";
    if (o->backend < 0 || o->backend > 4 || (o->validated != 0 && o->validated != 1))
        return "backend or validated flag out of range";
    if (o->max_terms && m->len > o->max_terms) return "map exceeds its term cap";
    /* ledger and bounds */
    const tpgp_ledger *l = &s->ledger;
    if (!nonneg(l->dropped_l1) || !nonneg(l->dropped_l2) || !nonneg(l->dropped_l2_squared) ||
        !nonneg(l->arithmetic_l2) || !nonneg(s->pf_bound))
        return "ledger sums must be non-negative";
    if (l->arithmetic_enclosed != 0 && l->arithmetic_enclosed != 1) return "enclosed flag not 0/1";
    if (l->arithmetic_enclosed && !o->validated) return "arithmetic enclosure without validation";
    if (l->peak_terms < m->len) return "peak term count below the map's length";
    if (!finite_nonneg(s->max_fid_bound)) return "FID accuracy limit must be finite";
    if (s->max_fid_bound > 0 && !o->validated && s->residual_depth < 0)
        return "an accuracy limit needs validation or a residual bound";
    if (s->residual_depth >= 0 ? !nonneg(s->residual_bound) : !isnan(s->residual_bound))
        return "residual bound inconsistent with the residual depth";
    /* the map: finite nonzero values, zero key padding, strictly ascending keys */
    const size_t W = m->n_words, pad = m->n_spins % 32;
    for (size_t k = 0; k < m->len; k++) {
        if (!isfinite(m->values[k]) || m->values[k] == 0) return "map value zero or not finite";
        const uint64_t *key = m->keys + k * W;
        if (pad && key[W - 1] >> (2 * pad)) return "key padding bits set";
        if (k) {
            const uint64_t *prev = key - W;
            size_t w = W;
            while (w > 0 && prev[w - 1] == key[w - 1]) w--;
            if (w == 0 || prev[w - 1] > key[w - 1]) return "map keys not strictly ascending";
        }
    }
    for (size_t k = 0; k < s->term_count; k++) {
        const tpgp_term *t = s->terms + k;
        if (t->i >= m->n_spins || t->j >= m->n_spins || t->i == t->j || t->qi > 3 ||
            t->qj > 3 || !isfinite(t->h))
            return "invalid Hamiltonian term";
    }
    return NULL;
}

void tpgp_checkpoint_free(tpgp_checkpoint *s) {
    if (!s) return;
    tpgp_map_free(&s->map);
    free(s->terms);
    memset(s, 0, sizeof(*s));
}

static tpgp_status fail(char *error, size_t cap, tpgp_status st, const char *what,
                        const char *detail) {
    if (error && cap) {
        if (detail) snprintf(error, cap, "%s: %s", what, detail);
        else snprintf(error, cap, "%s", what);
    }
    return st;
}

/* ---- durable files --------------------------------------------------------------------------
 * A sink buffers bytes for a file descriptor and hashes what it is told to.  durable_write()
 * produces a file's whole content into a fresh temporary beside the target (created with
 * O_EXCL), forces it to disk, renames it over the target and forces the directory entry, so a
 * reader sees the old file or the new one, never a mixture. */
typedef struct {
    int fd;
    int failed;
    size_t used;
    sha_ctx sha;
    unsigned char buf[1 << 16];
} sink;
static void sink_flush(sink *k) {
    const unsigned char *p = k->buf;
    while (!k->failed && k->used) {
        ssize_t n = write(k->fd, p, k->used);
        if (n < 0 && errno == EINTR) continue;
        if (n <= 0) {
            k->failed = 1;
            break;
        }
        p += n;
        k->used -= (size_t)n;
    }
    k->used = 0;
}
static void sink_put(sink *k, const void *data, size_t n, int hashed) {
    const unsigned char *p = data;
    if (hashed) sha_feed(&k->sha, p, n);
    while (n && !k->failed) {
        size_t room = sizeof(k->buf) - k->used, take = n < room ? n : room;
        memcpy(k->buf + k->used, p, take);
        k->used += take;
        p += take;
        n -= take;
        if (k->used == sizeof(k->buf)) sink_flush(k);
    }
}
static void sink_word(sink *k, uint64_t w) {
    unsigned char b[8];
    for (int i = 0; i < 8; i++) b[i] = (unsigned char)(w >> (8 * i));
    sink_put(k, b, 8, 1);
}
static void sink_fields(sink *k, const field *t, size_t n, const void *base) {
    for (size_t i = 0; i < n; i++) {
        if (t[i].as == AS_TEXT32) sink_put(k, (const char *)base + t[i].at, 32, 1);
        else sink_word(k, stored_word(t + i, base));
    }
}

typedef void (*producer)(sink *k, const void *arg);
static int sync_directory_of(const char *path) {
    const char *slash = strrchr(path, '/');
    char *dir;
    if (!slash) dir = strdup(".");
    else if (slash == path) dir = strdup("/");
    else {
        dir = malloc((size_t)(slash - path) + 1);
        if (dir) {
            memcpy(dir, path, (size_t)(slash - path));
            dir[slash - path] = 0;
        }
    }
    if (!dir) return 0;
    int fd = open(dir, O_RDONLY);
    free(dir);
    if (fd < 0) return 0;
    int ok = fsync(fd) == 0;
    close(fd);
    return ok;
}
/* 0 on success, else a static message. */
static const char *durable_write(const char *path, producer make, const void *arg) {
    const size_t n = strlen(path);
    char *tmp = malloc(n + 16);
    sink *k = malloc(sizeof(sink));
    const char *why = NULL;
    if (!tmp || !k) {
        free(tmp);
        free(k);
        return "out of memory";
    }
    memcpy(tmp, path, n);
    memcpy(tmp + n, ".tmp.XXXXXX", 12);
    k->fd = mkstemp(tmp);
    if (k->fd < 0) {
        free(tmp);
        free(k);
        return "cannot create a temporary file beside the target";
    }
    k->failed = 0;
    k->used = 0;
    sha_start(&k->sha);
    make(k, arg);
    sink_flush(k);
    if (k->failed) why = "write failed; the previous file is untouched";
    else if (fsync(k->fd)) why = "fsync of the new file failed";
    if (close(k->fd) && !why) why = "close of the new file failed";
    if (!why && rename(tmp, path)) why = "rename over the target failed";
    if (why) unlink(tmp);
    else if (!sync_directory_of(path)) why = "fsync of the containing directory failed";
    free(tmp);
    free(k);
    return why;
}

/* ---- save ------------------------------------------------------------------------------------ */
typedef struct {
    const tpgp_checkpoint *s;
    unsigned char digest[DIGEST];
} record_job;
static void make_record(sink *k, const void *arg) {
    record_job *job = (record_job *)arg;
    const tpgp_checkpoint *s = job->s;
    sink_put(k, MAGIC, sizeof(MAGIC), 1);
    sink_word(k, VERSION);
    sink_fields(k, HEADER, COUNT(HEADER), s);
    for (size_t t = 0; t < s->term_count; t++) sink_fields(k, TERM, COUNT(TERM), s->terms + t);
    const size_t W = s->map.n_words;
    for (size_t e = 0; e < s->map.len; e++) {
        for (size_t w = 0; w < W; w++) sink_word(k, s->map.keys[e * W + w]);
        uint64_t v;
        memcpy(&v, s->map.values + e, 8);
        sink_word(k, v);
    }
    sha_finish(&k->sha, job->digest);
    sink_put(k, job->digest, DIGEST, 0);
}

/* The manifest: one "key": value line per entry, in the reference's key order
 * (checkpoint_rev.c:360-404).  Numbers in decimal, doubles as C99 hex floats, strings quoted. */
static void put_text(sink *k, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static void put_text(sink *k, const char *fmt, ...) {
    char line[256];
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(line, sizeof(line), fmt, ap);
    va_end(ap);
    if (n < 0 || (size_t)n >= sizeof(line)) {
        k->failed = 1;
        return;
    }
    sink_put(k, line, (size_t)n, 0);
}
static void make_manifest(sink *k, const void *arg) {
    const record_job *job = arg;
    const tpgp_checkpoint *s = job->s;
    static const char *const reason[7] = {"running",          "complete",
                                          "walltime",         "signal",
                                          "resource_failure", "numerical_failure",
                                          "accuracy_limit"};
    char hex[2 * DIGEST + 1];
    for (int b = 0; b < DIGEST; b++) snprintf(hex + 2 * b, 3, "%02x", job->digest[b]);
    put_text(k, "{\n  \"format\": \"TPGPRVCP-2\",\n  \"sha256_payload\": \"%s\",\n", hex);
    put_text(k, "  \"model\": \"%s\",\n  \"n_spins\": %zu,\n", s->model, s->map.n_spins);
    put_text(k, "  \"active_terms\": %zu,\n  \"hamiltonian_terms\": %zu,\n", s->map.len,
             s->term_count);
    put_text(k, "  \"step\": %zu,\n  \"total_steps\": %zu,\n", s->step, s->total_steps);
    put_text(k, "  \"dt_hex\": \"%a\",\n  \"backend\": %d,\n", s->dt, s->options.backend);
    put_text(k, "  \"threads\": %u,\n  \"validated\": %d,\n", s->options.threads,
             s->options.validated);
    put_text(k, "  \"threshold_hex\": \"%a\",\n  \"max_terms\": %zu,\n", s->options.threshold,
             s->options.max_terms);
    put_text(k, "  \"residual_depth\": %d,\n  \"max_fid_bound_hex\": \"%a\",\n",
             s->residual_depth, s->max_fid_bound);
    put_text(k, "  \"resumed_from_step\": %zu,\n  \"prior_total_steps\": %zu,\n",
             s->resumed_from_step, s->prior_total_steps);
    put_text(k, "  \"prior_max_terms\": %zu,\n  \"stop_reason\": \"%s\",\n", s->prior_max_terms,
             reason[s->stop_reason]);
    put_text(k, "  \"status_code\": %d,\n  \"status\": \"%s\"\n}\n", (int)s->last_status,
             tpgp_status_string(s->last_status));
}

tpgp_status tpgp_checkpoint_save(const char *path, const tpgp_checkpoint *s, char *error,
                                 size_t cap) {
    if (!path || !*path) return fail(error, cap, TPGP_EINVAL, "checkpoint save", "no path");
    const char *bad = broken_rule(s);
    if (bad) return fail(error, cap, TPGP_EINVAL, "checkpoint refused", bad);
    record_job job = {s, {0}};
    const char *why = durable_write(path, make_record, &job);
    if (why) return fail(error, cap, TPGP_EINVAL, "checkpoint not written", why);
    const size_t n = strlen(path);
    char *manifest = malloc(n + sizeof(".manifest.json"));
    if (!manifest)
        return fail(error, cap, TPGP_ENOMEM, "checkpoint written, manifest not", "out of memory");
    memcpy(manifest, path, n);
    memcpy(manifest + n, ".manifest.json", sizeof(".manifest.json"));
    why = durable_write(manifest, make_manifest, &job);
    free(manifest);
    if (why) return fail(error, cap, TPGP_EINVAL, "checkpoint written, manifest not", why);
    return fail(error, cap, TPGP_OK, "ok", NULL);
}

/* ---- load ------------------------------------------------------------------------------------ */
typedef struct {
    const unsigned char *p, *end;
} cursor;
static uint64_t take_word(cursor *c) {
    uint64_t w = 0;
    for (int i = 0; i < 8; i++) w |= (uint64_t)c->p[i] << (8 * i);
    c->p += 8;
    return w;
}
/* Decode a table's fields into base; 0 when a stored value is out of range.  The caller has
 * checked that the bytes are there. */
static int take_fields(cursor *c, const field *t, size_t n, void *base) {
    int ok = 1;
    for (size_t i = 0; i < n; i++) {
        if (t[i].as == AS_TEXT32) {
            memcpy((char *)base + t[i].at, c->p, 32);
            c->p += 32;
        } else {
            ok &= load_word(t + i, base, take_word(c));
        }
    }
    return ok;
}

static tpgp_status decode(const unsigned char *data, size_t size, tpgp_checkpoint *s,
                          const char **why) {
    memset(s, 0, sizeof(*s));
    const size_t head = sizeof(MAGIC) + 8 + table_size(HEADER, COUNT(HEADER));
    const size_t term_bytes = table_size(TERM, COUNT(TERM));
    if (size < head + DIGEST) return *why = "file shorter than a record header", TPGP_EINVAL;
    const size_t payload = size - DIGEST;
    unsigned char digest[DIGEST];
    tpgp_payload_sha256(data, payload, digest);
    if (memcmp(digest, data + payload, DIGEST)) return *why = "checksum mismatch", TPGP_EINVAL;
    cursor c = {data, data + payload};
    if (memcmp(c.p, MAGIC, sizeof(MAGIC))) return *why = "not a TPGPRVCP record", TPGP_EINVAL;
    c.p += sizeof(MAGIC);
    if (take_word(&c) != VERSION) return *why = "unsupported record version", TPGP_EINVAL;
    s->options = tpgp_default_options();
    if (!take_fields(&c, HEADER, COUNT(HEADER), s))
        return *why = "header field out of range", TPGP_EINVAL;
    /* the body's length follows from the header's counts */
    const size_t n = s->map.n_spins, len = s->map.len, terms = s->term_count;
    if (n == 0 || n > SIZE_MAX - 31) return *why = "bad spin count", TPGP_EINVAL;
    const size_t W = (n + 31) / 32, rest = (size_t)(c.end - c.p);
    if (!mul_fits(terms, term_bytes) || !mul_fits(len, W + 1) || !mul_fits(len * (W + 1), 8) ||
        terms * term_bytes > rest || len * (W + 1) * 8 != rest - terms * term_bytes)
        return *why = "record length disagrees with its header", TPGP_EINVAL;
    const size_t count = len;
    /* reset the map to empty storage before allocating: a failure below frees only ours */
    if (tpgp_map_init(&s->map, n) != TPGP_OK) return *why = "bad spin count", TPGP_EINVAL;
    if (count) {
        s->map.keys = malloc(count * W * sizeof(uint64_t));
        s->map.values = malloc(count * sizeof(double));
    }
    if (terms) s->terms = malloc(terms * sizeof(tpgp_term));
    if ((count && (!s->map.keys || !s->map.values)) || (terms && !s->terms))
        return *why = "out of memory", TPGP_ENOMEM;
    s->map.len = s->map.capacity = count;
    for (size_t t = 0; t < terms; t++) {
        memset(s->terms + t, 0, sizeof(tpgp_term));
        if (!take_fields(&c, TERM, COUNT(TERM), s->terms + t))
            return *why = "term field out of range", TPGP_EINVAL;
    }
    for (size_t e = 0; e < count; e++) {
        for (size_t w = 0; w < W; w++) s->map.keys[e * W + w] = take_word(&c);
        const uint64_t v = take_word(&c);
        memcpy(s->map.values + e, &v, 8);
    }
    const char *bad = broken_rule(s);
    if (bad) return *why = bad, TPGP_EINVAL;
    return TPGP_OK;
}

tpgp_status tpgp_checkpoint_load(const char *path, tpgp_checkpoint *dst, char *error, size_t cap) {
    if (!path || !dst) return fail(error, cap, TPGP_EINVAL, "checkpoint load", "no path or record");
    int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(error, cap, TPGP_EINVAL, "cannot open checkpoint", strerror(errno));
    struct stat st;
    if (fstat(fd, &st) || !S_ISREG(st.st_mode) || st.st_size <= 0) {
        close(fd);
        return fail(error, cap, TPGP_EINVAL, "checkpoint rejected", "not a non-empty file");
    }
    const size_t size = (size_t)st.st_size;
    void *map = mmap(NULL, size, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED)
        return fail(error, cap, TPGP_EINVAL, "cannot map checkpoint", strerror(errno));
    tpgp_checkpoint s;
    const char *why = NULL;
    tpgp_status status = decode(map, size, &s, &why);
    munmap(map, size);
    if (status != TPGP_OK) {
        tpgp_checkpoint_free(&s);
        return fail(error, cap, status, "checkpoint rejected", why);
    }
    tpgp_checkpoint_free(dst);
    *dst = s;
    return fail(error, cap, TPGP_OK, "ok", NULL);
}

/* ---- configuration --------------------------------------------------------------------------
 * Two records describe the same run when every CONFIG field of the header and every term agree
 * (doubles by bit pattern, the model as a string); state, step counters, bounds and the ledger
 * are not configuration (checkpoint_rev.h:36-40). */
static int same_fields(const field *t, size_t n, const void *a, const void *b) {
    for (size_t i = 0; i < n; i++) {
        if (!(t[i].flags & CONFIG)) continue;
        if (t[i].as == AS_TEXT32) {
            if (strncmp((const char *)a + t[i].at, (const char *)b + t[i].at, 32)) return 0;
        } else if (stored_word(t + i, a) != stored_word(t + i, b)) {
            return 0;
        }
    }
    return 1;
}
int tpgp_checkpoint_same_configuration(const tpgp_checkpoint *a, const tpgp_checkpoint *b) {
    if (!a || !b || !same_fields(HEADER, COUNT(HEADER), a, b)) return 0;
    for (size_t t = 0; t < a->term_count; t++)
        if (!same_fields(TERM, COUNT(TERM), a->terms + t, b->terms + t)) return 0;
    return 1;
}
