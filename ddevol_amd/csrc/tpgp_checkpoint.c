/* TPGPRVCP version-2 checkpoints for the engine's maps (C17).
 *
 * The record layout, its invariants and the commit protocol follow the reference's checkpoint
 * format (engine/src/checkpoint_rev.c:121-604) so that records are interchangeable:
 *   "TPGPRVCP", u64 version 2, then n_spins, len, term_count, step, total_steps, dt,
 *   residual_depth + 1, initial_global, model[32], j2, linear_size, resumed_from_step,
 *   prior_total_steps, prior_max_terms, stop_reason, last_status, threshold, max_terms, threads,
 *   backend, validated, max_drop_l2, the ledger (4 doubles, 3 counts, enclosed flag), pf_bound,
 *   residual_bound, max_fid_bound, the terms (i, j, qi, qj, h, layer), the map (W key words and
 *   the value per term); all little-endian 64-bit; then the SHA-256 of everything before.
 */
#define _POSIX_C_SOURCE 200809L
#include "../../include/tpgp_engine.h"

#include <fcntl.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

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

/* ---- record invariants (checkpoint_rev.c:121-206) ----------------------------------------- */
static void set_error(char *e, size_t cap, const char *m) {
    if (e && cap) snprintf(e, cap, "%s", m);
}
static int mul_ok(size_t a, size_t b, size_t *out) {
    if (b && a > SIZE_MAX / b) return 0;
    *out = a * b;
    return 1;
}
static int bits_equal(double a, double b) { return memcmp(&a, &b, sizeof(double)) == 0; }

static int valid_record(const tpgp_checkpoint *s) {
    if (!s || FLT_RADIX != 2 || DBL_MANT_DIG != 53 || !s->map.n_spins ||
        s->map.n_spins > SIZE_MAX - 31 || s->map.n_words != (s->map.n_spins + 31) / 32 ||
        s->map.len > s->map.capacity || (s->map.len && (!s->map.keys || !s->map.values)) ||
        (s->term_count && !s->terms))
        return 0;
    if (!isfinite(s->dt) || s->dt <= 0 || s->step > s->total_steps || s->residual_depth < -1 ||
        s->residual_depth > 10 || (s->initial_global != 0 && s->initial_global != 1) ||
        !memchr(s->model, 0, sizeof(s->model)) || !s->model[0] || !isfinite(s->j2) ||
        s->resumed_from_step > s->step || s->stop_reason < 0 || s->stop_reason > 6 ||
        s->last_status < TPGP_OK || s->last_status > TPGP_EBUDGET)
        return 0;
    size_t t;
    if (!mul_ok(s->map.len, s->map.n_words, &t) || !mul_ok(t, 8, &t) ||
        !mul_ok(s->map.len, 8, &t) || !mul_ok(s->term_count, sizeof(tpgp_term), &t) ||
        !isfinite(s->max_fid_bound) || s->max_fid_bound < 0 ||
        (s->max_fid_bound > 0 && !s->options.validated && s->residual_depth < 0) ||
        (s->stop_reason == 1 && s->step != s->total_steps))
        return 0;
    for (const char *c = s->model; *c; c++)
        if (!((*c >= 'a' && *c <= 'z') || (*c >= '0' && *c <= '9') || *c == '_')) return 0;
    const tpgp_options *o = &s->options;
    if (!isfinite(o->threshold) || o->threshold < 0 || !o->threads || o->threads > 65536 ||
        o->backend < 0 || o->backend > 4 || (o->validated != 0 && o->validated != 1) ||
        !isfinite(o->max_drop_l2) || o->max_drop_l2 < 0 ||
        (o->max_terms && s->map.len > o->max_terms))
        return 0;
    const double nonneg[5] = {s->ledger.dropped_l1, s->ledger.dropped_l2,
                              s->ledger.dropped_l2_squared, s->ledger.arithmetic_l2, s->pf_bound};
    for (int k = 0; k < 5; k++)
        if (isnan(nonneg[k]) || nonneg[k] < 0) return 0;
    if ((s->ledger.arithmetic_enclosed != 0 && s->ledger.arithmetic_enclosed != 1) ||
        (s->ledger.arithmetic_enclosed && !o->validated) || s->ledger.peak_terms < s->map.len)
        return 0;
    if (s->residual_depth >= 0 ? (isnan(s->residual_bound) || s->residual_bound < 0)
                               : !isnan(s->residual_bound))
        return 0;
    const size_t W = s->map.n_words, rem = s->map.n_spins % 32;
    for (size_t k = 0; k < s->map.len; k++) {  /* finite nonzero values, zero padding, order */
        const double v = s->map.values[k];
        if (!isfinite(v) || v == 0) return 0;
        if (rem && (s->map.keys[(k + 1) * W - 1] >> (2 * rem))) return 0;
        if (k) {
            int less = 0;
            for (size_t w = W; w-- > 0;) {
                const uint64_t a = s->map.keys[(k - 1) * W + w], b = s->map.keys[k * W + w];
                if (a != b) {
                    less = a < b;
                    break;
                }
                if (w == 0) return 0; /* duplicate key */
            }
            if (!less) return 0;
        }
    }
    for (size_t k = 0; k < s->term_count; k++) {
        const tpgp_term *q = s->terms + k;
        if (q->i >= s->map.n_spins || q->j >= s->map.n_spins || q->i == q->j || q->qi > 3 ||
            q->qj > 3 || !isfinite(q->h))
            return 0;
    }
    return 1;
}

void tpgp_checkpoint_free(tpgp_checkpoint *s) {
    if (!s) return;
    tpgp_map_free(&s->map);
    free(s->terms);
    memset(s, 0, sizeof(*s));
}

/* ---- writing: exclusive temporary file, fsync, atomic rename, parent fsync -------------- */
typedef struct {
    FILE *f;
    sha_ctx sha;
    int ok;
} out_stream;
static void out_bytes(out_stream *o, const void *p, size_t n) {
    if (!o->ok) return;
    if (fwrite(p, 1, n, o->f) != n) {
        o->ok = 0;
        return;
    }
    sha_feed(&o->sha, p, n);
}
static void out_u64(out_stream *o, uint64_t x) {
    unsigned char b[8];
    for (int i = 0; i < 8; i++) b[i] = (unsigned char)(x >> (8 * i));
    out_bytes(o, b, 8);
}
static void out_f64(out_stream *o, double x) {
    uint64_t u;
    memcpy(&u, &x, 8);
    out_u64(o, u);
}
static FILE *open_temporary(const char *path, char **tmp_name) {
    const size_t n = strlen(path) + 16;
    *tmp_name = malloc(n);
    if (!*tmp_name) return NULL;
    snprintf(*tmp_name, n, "%s.tmp.XXXXXX", path);
    const int fd = mkstemp(*tmp_name);
    FILE *f = fd >= 0 ? fdopen(fd, "wb") : NULL;
    if (!f) {
        if (fd >= 0) {
            close(fd);
            unlink(*tmp_name);
        }
        free(*tmp_name);
        *tmp_name = NULL;
    }
    return f;
}
static int publish(FILE *f, const char *tmp_name, const char *path) {
    int ok = fflush(f) == 0 && !ferror(f) && fsync(fileno(f)) == 0;
    ok = (fclose(f) == 0) && ok;
    ok = ok && rename(tmp_name, path) == 0;
    if (!ok) {
        unlink(tmp_name);
        return 0;
    }
    char *dir = strdup(path);
    if (!dir) return 0;
    char *slash = strrchr(dir, '/');
    if (!slash)
        strcpy(dir, ".");
    else if (slash == dir)
        slash[1] = 0;
    else
        *slash = 0;
    const int fd = open(dir, O_RDONLY);
    free(dir);
    if (fd < 0) return 0;
    ok = fsync(fd) == 0;
    close(fd);
    return ok;
}

tpgp_status tpgp_checkpoint_save(const char *path, const tpgp_checkpoint *s, char *error,
                                 size_t cap) {
    if (!path || !*path || !valid_record(s)) {
        set_error(error, cap, "invalid checkpoint state");
        return TPGP_EINVAL;
    }
    char *tmp = NULL;
    FILE *f = open_temporary(path, &tmp);
    if (!f) {
        set_error(error, cap, "cannot create checkpoint temporary file");
        return TPGP_EINVAL;
    }
    out_stream o = {f, {{0}}, 1};
    sha_start(&o.sha);
    out_bytes(&o, "TPGPRVCP", 8);
    const uint64_t head[5] = {2, s->map.n_spins, s->map.len, s->term_count, s->step};
    for (int k = 0; k < 5; k++) out_u64(&o, head[k]);
    out_u64(&o, s->total_steps);
    out_f64(&o, s->dt);
    out_u64(&o, (uint64_t)(s->residual_depth + 1));
    out_u64(&o, (uint64_t)s->initial_global);
    out_bytes(&o, s->model, 32);
    out_f64(&o, s->j2);
    out_u64(&o, s->linear_size);
    out_u64(&o, s->resumed_from_step);
    out_u64(&o, s->prior_total_steps);
    out_u64(&o, s->prior_max_terms);
    out_u64(&o, (uint64_t)s->stop_reason);
    out_u64(&o, (uint64_t)s->last_status);
    out_f64(&o, s->options.threshold);
    out_u64(&o, s->options.max_terms);
    out_u64(&o, s->options.threads);
    out_u64(&o, (uint64_t)s->options.backend);
    out_u64(&o, (uint64_t)s->options.validated);
    out_f64(&o, s->options.max_drop_l2);
    out_f64(&o, s->ledger.dropped_l1);
    out_f64(&o, s->ledger.dropped_l2);
    out_f64(&o, s->ledger.dropped_l2_squared);
    out_f64(&o, s->ledger.arithmetic_l2);
    out_u64(&o, s->ledger.rotations);
    out_u64(&o, s->ledger.removal_events);
    out_u64(&o, s->ledger.peak_terms);
    out_u64(&o, (uint64_t)s->ledger.arithmetic_enclosed);
    out_f64(&o, s->pf_bound);
    out_f64(&o, s->residual_bound);
    out_f64(&o, s->max_fid_bound);
    for (size_t k = 0; k < s->term_count; k++) {
        const tpgp_term *q = s->terms + k;
        out_u64(&o, q->i);
        out_u64(&o, q->j);
        out_u64(&o, q->qi);
        out_u64(&o, q->qj);
        out_f64(&o, q->h);
        out_u64(&o, q->layer);
    }
    const size_t W = s->map.n_words;
    for (size_t k = 0; k < s->map.len; k++) {
        for (size_t w = 0; w < W; w++) out_u64(&o, s->map.keys[k * W + w]);
        out_f64(&o, s->map.values[k]);
    }
    unsigned char digest[32];
    sha_finish(&o.sha, digest);
    if (o.ok && fwrite(digest, 1, 32, f) != 32) o.ok = 0;
    if (!o.ok) {
        fclose(f);
        unlink(tmp);
        free(tmp);
        set_error(error, cap, "checkpoint write failed; old record preserved");
        return TPGP_EINVAL;
    }
    if (!publish(f, tmp, path)) {
        free(tmp);
        set_error(error, cap, "checkpoint atomic commit or directory synchronization failed");
        return TPGP_EINVAL;
    }
    free(tmp);
    /* the manifest: payload hash and run metadata (checkpoint_rev.c:360-404) */
    const size_t n = strlen(path) + 20;
    char *mpath = malloc(n);
    if (!mpath) {
        set_error(error, cap, "checkpoint saved but manifest allocation failed");
        return TPGP_ENOMEM;
    }
    snprintf(mpath, n, "%s.manifest.json", path);
    f = open_temporary(mpath, &tmp);
    if (!f) {
        free(mpath);
        set_error(error, cap, "checkpoint saved but manifest creation failed");
        return TPGP_EINVAL;
    }
    char hex[65], dt_hex[40], thr_hex[40], bound_hex[40];
    for (int k = 0; k < 32; k++) snprintf(hex + 2 * k, 3, "%02x", digest[k]);
    snprintf(dt_hex, sizeof(dt_hex), "%a", s->dt);
    snprintf(thr_hex, sizeof(thr_hex), "%a", s->options.threshold);
    snprintf(bound_hex, sizeof(bound_hex), "%a", s->max_fid_bound);
    static const char *reason[7] = {"running",          "complete",          "walltime", "signal",
                                    "resource_failure", "numerical_failure", "accuracy_limit"};
    /* one "key": value line per field; strings quoted, numbers in decimal */
    char num[11][24];
    const unsigned long long counts[10] = {
        s->map.n_spins,        s->map.len,           s->term_count,      s->step,
        s->total_steps,        s->options.max_terms, s->resumed_from_step,
        s->prior_total_steps,  s->prior_max_terms,   s->options.threads};
    for (int k = 0; k < 10; k++) snprintf(num[k], sizeof(num[k]), "%llu", counts[k]);
    snprintf(num[10], sizeof(num[10]), "%d", s->residual_depth);
    char small[3][8];
    snprintf(small[0], sizeof(small[0]), "%d", s->options.backend);
    snprintf(small[1], sizeof(small[1]), "%d", s->options.validated);
    snprintf(small[2], sizeof(small[2]), "%d", (int)s->last_status);
    const struct {
        const char *key, *value;
        int quoted;
    } rows[] = {{"format", "TPGPRVCP-2", 1},       {"sha256_payload", hex, 1},
                {"model", s->model, 1},            {"n_spins", num[0], 0},
                {"active_terms", num[1], 0},       {"hamiltonian_terms", num[2], 0},
                {"step", num[3], 0},               {"total_steps", num[4], 0},
                {"dt_hex", dt_hex, 1},             {"backend", small[0], 0},
                {"threads", num[9], 0},            {"validated", small[1], 0},
                {"threshold_hex", thr_hex, 1},     {"max_terms", num[5], 0},
                {"residual_depth", num[10], 0},    {"max_fid_bound_hex", bound_hex, 1},
                {"resumed_from_step", num[6], 0},  {"prior_total_steps", num[7], 0},
                {"prior_max_terms", num[8], 0},    {"stop_reason", reason[s->stop_reason], 1},
                {"status_code", small[2], 0},      {"status", tpgp_status_string(s->last_status), 1}};
    const size_t n_rows = sizeof(rows) / sizeof(rows[0]);
    fputs("{\n", f);
    for (size_t k = 0; k < n_rows; k++)
        fprintf(f, "  \"%s\": %s%s%s%s\n", rows[k].key, rows[k].quoted ? "\"" : "", rows[k].value,
                rows[k].quoted ? "\"" : "", k + 1 < n_rows ? "," : "");
    fputs("}\n", f);
    const int ok = publish(f, tmp, mpath);
    free(tmp);
    free(mpath);
    if (!ok) {
        set_error(error, cap, "checkpoint saved but manifest commit failed");
        return TPGP_EINVAL;
    }
    set_error(error, cap, "ok");
    return TPGP_OK;
}

/* ---- reading: whole-file checksum first, then a strict parse ------------------------------ */
typedef struct {
    FILE *f;
    uint64_t left;
    int ok;
} in_stream;
static void in_bytes(in_stream *r, void *p, size_t n) {
    if (!r->ok || n > r->left || fread(p, 1, n, r->f) != n) {
        r->ok = 0;
        memset(p, 0, n);
        return;
    }
    r->left -= n;
}
static uint64_t in_u64(in_stream *r) {
    unsigned char b[8];
    in_bytes(r, b, 8);
    uint64_t x = 0;
    for (int i = 0; i < 8; i++) x |= (uint64_t)b[i] << (8 * i);
    return x;
}
static size_t in_size(in_stream *r) {
    const uint64_t x = in_u64(r);
    if (x > SIZE_MAX) r->ok = 0;
    return (size_t)x;
}
static double in_f64(in_stream *r) {
    const uint64_t u = in_u64(r);
    double d;
    memcpy(&d, &u, 8);
    return d;
}

tpgp_status tpgp_checkpoint_load(const char *path, tpgp_checkpoint *dst, char *error, size_t cap) {
    if (!path || !dst || FLT_RADIX != 2 || DBL_MANT_DIG != 53) {
        set_error(error, cap, "invalid checkpoint load arguments");
        return TPGP_EINVAL;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error(error, cap, "cannot open checkpoint");
        return TPGP_EINVAL;
    }
    struct stat st;
    if (fstat(fileno(f), &st) || st.st_size < 328) {
        fclose(f);
        set_error(error, cap, "checkpoint is truncated or has invalid length");
        return TPGP_EINVAL;
    }
    const uint64_t payload = (uint64_t)st.st_size - 32;
    unsigned char chunk[1 << 16], want[32], got[32];
    sha_ctx sha;
    sha_start(&sha);
    int good = 1;
    for (uint64_t left = payload; left && good;) {
        const size_t n = left > sizeof(chunk) ? sizeof(chunk) : (size_t)left;
        good = fread(chunk, 1, n, f) == n;
        if (good) sha_feed(&sha, chunk, n);
        left -= n;
    }
    sha_finish(&sha, got);
    if (!good || fread(want, 1, 32, f) != 32 || memcmp(got, want, 32) || fgetc(f) != EOF) {
        fclose(f);
        set_error(error, cap, "checkpoint checksum/length mismatch");
        return TPGP_EINVAL;
    }
    rewind(f);
    in_stream r = {f, payload, 1};
    tpgp_checkpoint s;
    memset(&s, 0, sizeof(s));
    char magic[8];
    in_bytes(&r, magic, 8);
    const uint64_t version = in_u64(&r);
    const size_t spins = in_size(&r), count = in_size(&r);
    s.term_count = in_size(&r);
    s.step = in_size(&r);
    s.total_steps = in_size(&r);
    s.dt = in_f64(&r);
    const uint64_t depth = in_u64(&r), global = in_u64(&r);
    if (depth > 11 || global > 1) r.ok = 0;
    s.residual_depth = (int)depth - 1;
    s.initial_global = (int)global;
    in_bytes(&r, s.model, 32);
    s.j2 = in_f64(&r);
    s.linear_size = in_size(&r);
    s.resumed_from_step = in_size(&r);
    s.prior_total_steps = in_size(&r);
    s.prior_max_terms = in_size(&r);
    const uint64_t reason = in_u64(&r), last = in_u64(&r);
    if (reason > 6 || last > TPGP_EBUDGET) r.ok = 0;
    s.stop_reason = (int)reason;
    s.last_status = (tpgp_status)last;
    s.options = tpgp_default_options();
    s.options.threshold = in_f64(&r);
    s.options.max_terms = in_size(&r);
    const uint64_t threads = in_u64(&r), backend = in_u64(&r), validated = in_u64(&r);
    if (threads > 65536 || backend > 4 || validated > 1) r.ok = 0;
    s.options.threads = (unsigned)threads;
    s.options.backend = (int)backend;
    s.options.validated = (int)validated;
    s.options.max_drop_l2 = in_f64(&r);
    s.ledger.dropped_l1 = in_f64(&r);
    s.ledger.dropped_l2 = in_f64(&r);
    s.ledger.dropped_l2_squared = in_f64(&r);
    s.ledger.arithmetic_l2 = in_f64(&r);
    s.ledger.rotations = in_size(&r);
    s.ledger.removal_events = in_size(&r);
    s.ledger.peak_terms = in_size(&r);
    const uint64_t enclosed = in_u64(&r);
    if (enclosed > 1) r.ok = 0;
    s.ledger.arithmetic_enclosed = (int)enclosed;
    s.pf_bound = in_f64(&r);
    s.residual_bound = in_f64(&r);
    s.max_fid_bound = in_f64(&r);
    const size_t W = spins && spins <= SIZE_MAX - 31 ? (spins + 31) / 32 : 0;
    size_t keys = 0, key_bytes = 0, term_bytes = 0, disk_terms = 0, disk_map = 0;
    if (!r.ok || memcmp(magic, "TPGPRVCP", 8) || version != 2 || !W ||
        !mul_ok(count, W, &keys) || !mul_ok(keys, 8, &key_bytes) ||
        !mul_ok(s.term_count, sizeof(tpgp_term), &term_bytes) ||
        !mul_ok(s.term_count, 48, &disk_terms) || !mul_ok(count, W + 1, &disk_map) ||
        !mul_ok(disk_map, 8, &disk_map) || disk_terms > UINT64_MAX - disk_map ||
        r.left != (uint64_t)disk_terms + disk_map)
        goto invalid;
    if (tpgp_map_init(&s.map, spins)) goto invalid;
    s.map.len = s.map.capacity = count;
    s.map.keys = count ? malloc(key_bytes) : NULL;
    s.map.values = count ? malloc(count * sizeof(double)) : NULL;
    s.terms = s.term_count ? malloc(term_bytes) : NULL;
    if ((count && (!s.map.keys || !s.map.values)) || (s.term_count && !s.terms)) {
        tpgp_checkpoint_free(&s);
        fclose(f);
        set_error(error, cap, "checkpoint allocation failed");
        return TPGP_ENOMEM;
    }
    for (size_t k = 0; k < s.term_count; k++) {
        tpgp_term *q = s.terms + k;
        q->i = in_size(&r);
        q->j = in_size(&r);
        const uint64_t qi = in_u64(&r), qj = in_u64(&r);
        if (qi > 3 || qj > 3) r.ok = 0;
        q->qi = (unsigned)qi;
        q->qj = (unsigned)qj;
        q->h = in_f64(&r);
        q->layer = in_size(&r);
    }
    for (size_t k = 0; k < count; k++) {
        for (size_t w = 0; w < W; w++) s.map.keys[k * W + w] = in_u64(&r);
        s.map.values[k] = in_f64(&r);
    }
    if (!r.ok || r.left || !valid_record(&s)) goto invalid;
    fclose(f);
    tpgp_checkpoint_free(dst);
    *dst = s;
    set_error(error, cap, "ok");
    return TPGP_OK;
invalid:
    tpgp_checkpoint_free(&s);
    fclose(f);
    set_error(error, cap, "checkpoint format or numerical invariants are invalid");
    return TPGP_EINVAL;
}

/* Same run configuration (checkpoint_rev.c:585-604): everything but state, step and runtime. */
int tpgp_checkpoint_same_configuration(const tpgp_checkpoint *a, const tpgp_checkpoint *b) {
    if (!a || !b || a->map.n_spins != b->map.n_spins || a->term_count != b->term_count ||
        a->total_steps != b->total_steps || !bits_equal(a->dt, b->dt) ||
        a->residual_depth != b->residual_depth || a->initial_global != b->initial_global ||
        strncmp(a->model, b->model, 32) || !bits_equal(a->j2, b->j2) ||
        a->linear_size != b->linear_size)
        return 0;
    const tpgp_options *x = &a->options, *y = &b->options;
    if (!bits_equal(x->threshold, y->threshold) || x->max_terms != y->max_terms ||
        x->threads != y->threads || x->backend != y->backend || x->validated != y->validated ||
        !bits_equal(x->max_drop_l2, y->max_drop_l2))
        return 0;
    for (size_t k = 0; k < a->term_count; k++) {
        const tpgp_term *p = a->terms + k, *q = b->terms + k;
        if (p->i != q->i || p->j != q->j || p->qi != q->qi || p->qj != q->qj ||
            !bits_equal(p->h, q->h) || p->layer != q->layer)
            return 0;
    }
    return 1;
}
