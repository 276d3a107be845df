/* The independent endpoint (residual) bound of one Strang step (C17 + MPFR, with the GPU).
 *
 * tpgp_residual_bound(a, b, ...) bounds the normalized Hilbert-Schmidt error produced over one
 * step of length dt by the cubic Hermite path through the maps a (before) and b (after), as
 * engine/src/residual_rev.c:181-194 defines it:
 *   L = the Liouvillian generator applied to sparse interval vectors (residual_rev.c:60-111);
 *   D1 = 6 b - 6 a - 4 dt L a - 2 dt L b - dt^2 L L a,
 *   D2 = 6 b - 6 a - 2 dt L a - 4 dt L b + dt^2 L L b            (residual_rev.c:187);
 *   the bound is the integral over the Bezier control points of |u D1 + v D2|, subdivided
 *   `depth` times (residual_rev.c:175-180), each norm either from the enclosed Gram matrix of
 *   (D1, D2) (residual_rev.c:116-174) or directly over the sparse supports.
 * D1 and D2 are built on the GPU (tpgp_hip_residual_vectors): four generator applications and
 * two combinations, every entry a chain of correctly rounded 256-bit products and sums, equal to
 * the reference's MPFR entries bit for bit and in its first-insertion order.  The rest is serial
 * in the reference (sums over D1 and D2 in entry order, each addition rounded) and runs here in
 * MPFR at 256 bits with the reference's rounding directions and order of operations, so the
 * rounded-up binary64 result is the reference's bit for bit.
 */
#include "../../include/tpgp_engine.h"
#include "../../include/tpgp_hip.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef TPGP_HAVE_MPFR
#include <mpfr.h>

enum { kBits = 256 };

typedef struct {
    mpfr_t lo, hi;
} interval;

/* A sparse interval vector: entries in first-insertion order, an open-addressing index. */
typedef struct {
    size_t words, len, cap, slots;
    uint64_t *keys;   /* [cap * words] */
    interval *vals;   /* [cap] */
    size_t *index;    /* [slots]: entry + 1, 0 = empty */
} ivec;

static int arithmetic_ok(void) {
    return FLT_RADIX == 2 && DBL_MANT_DIG == 53 && sizeof(size_t) <= 8 &&
           mpfr_get_emin() <= -16384 && mpfr_get_emax() >= 16384;
}

static uint64_t mix(uint64_t x) { /* splitmix64 finaliser */
    x ^= x >> 31;
    x *= 0x7fb5d329728ea185ull;
    x ^= x >> 27;
    x *= 0x81dadef4bc2dd44dull;
    return x ^ (x >> 33);
}
static uint64_t key_hash(const uint64_t *k, size_t words) {
    uint64_t h = 0x243f6a8885a308d3ull;
    for (size_t w = 0; w < words; ++w) h = mix(h + k[w] + w);
    return h;
}

static void ivec_free(ivec *v) {
    for (size_t i = 0; i < v->len; ++i) {
        mpfr_clear(v->vals[i].lo);
        mpfr_clear(v->vals[i].hi);
    }
    free(v->keys);
    free(v->vals);
    free(v->index);
    memset(v, 0, sizeof(*v));
}

static int ivec_reindex(ivec *v, size_t slots) {
    size_t *ix = calloc(slots, sizeof(*ix));
    if (!ix) return 0;
    for (size_t i = 0; i < v->len; ++i) {
        size_t s = key_hash(v->keys + i * v->words, v->words) & (slots - 1);
        while (ix[s]) s = (s + 1) & (slots - 1);
        ix[s] = i + 1;
    }
    free(v->index);
    v->index = ix;
    v->slots = slots;
    return 1;
}

static interval *ivec_find(const ivec *v, const uint64_t *k) {
    if (!v->slots) return NULL;
    size_t s = key_hash(k, v->words) & (v->slots - 1);
    while (v->index[s]) {
        const size_t i = v->index[s] - 1;
        if (!memcmp(v->keys + i * v->words, k, v->words * sizeof(*k))) return &v->vals[i];
        s = (s + 1) & (v->slots - 1);
    }
    return NULL;
}

/* y += s x with outward rounding (residual_rev.c:55). */
static void add_scaled(interval *y, const interval *x, const mpfr_t s, mpfr_t t) {
    const int nonneg = mpfr_sgn(s) >= 0;
    mpfr_mul(t, nonneg ? x->lo : x->hi, s, MPFR_RNDD);
    mpfr_add(y->lo, y->lo, t, MPFR_RNDD);
    mpfr_mul(t, nonneg ? x->hi : x->lo, s, MPFR_RNDU);
    mpfr_add(y->hi, y->hi, t, MPFR_RNDU);
}

/* The engine's process-wide device context (created on first use) and its error note. */
tpgp_hip_context *tpgp_engine_default_context(void);
void tpgp_engine_note(const char *where, const char *what);

/* A vector built on the device, in its entry order, with an index for lookups.  The device
 * endpoints are exact 256-bit values, so the conversion to MPFR at 256 bits is exact. */
static int from_device(ivec *v, const tpgp_hip_interval_vector *d, size_t words) {
    v->words = words;
    v->len = 0;
    v->cap = d->len ? d->len : 1;
    v->keys = malloc(v->cap * words * sizeof(*v->keys));
    v->vals = malloc(v->cap * sizeof(*v->vals));
    if (!v->keys || !v->vals) return 0;
    if (d->len) memcpy(v->keys, d->keys, d->len * words * sizeof(*v->keys));
    mpz_t z;
    mpz_init(z);
    for (size_t i = 0; i < d->len; ++i) {
        const tpgp_hip_f256 *ends[2] = {&d->lower[i], &d->upper[i]};
        mpfr_ptr out[2] = {v->vals[i].lo, v->vals[i].hi};
        mpfr_init2(v->vals[i].lo, kBits);
        mpfr_init2(v->vals[i].hi, kBits);
        v->len = i + 1;
        for (int k = 0; k < 2; ++k) {
            const tpgp_hip_f256 *x = ends[k];
            if (!x->m[3]) {
                mpfr_set_zero(out[k], x->neg ? -1 : 1);
                continue;
            }
            mpz_import(z, 4, -1, sizeof(uint64_t), 0, 0, x->m);
            mpfr_set_z_2exp(out[k], z, (mpfr_exp_t)x->e - kBits, MPFR_RNDN);
            if (x->neg) mpfr_neg(out[k], out[k], MPFR_RNDN);
        }
    }
    mpz_clear(z);
    size_t slots = 128;
    while (slots < 2 * (d->len + 1)) slots *= 2;
    return ivec_reindex(v, slots);
}

/* Upper bound of |u D1 + v D2| over the union of supports (D1's entries, then D2's that D1
 * lacks), each entry's enclosure squared upward (residual_rev.c:116-121). */
static void direct_norm(mpfr_t out, const ivec *d1, const ivec *d2, double u, double v) {
    mpfr_t su, sv, a, c, t;
    mpfr_inits2(kBits, su, sv, a, c, t, (mpfr_ptr)0);
    mpfr_set_d(su, u, MPFR_RNDN);
    mpfr_set_d(sv, v, MPFR_RNDN);
    mpfr_set_zero(out, 1);
    interval x;
    mpfr_init2(x.lo, kBits);
    mpfr_init2(x.hi, kBits);
    for (int pass = 0; pass < 2; ++pass) {
        const ivec *src = pass ? d2 : d1;
        for (size_t i = 0; i < src->len; ++i) {
            const uint64_t *k = src->keys + i * src->words;
            const interval *p1 = ivec_find(d1, k), *p2 = ivec_find(d2, k);
            if (pass && p1) continue;
            mpfr_set_zero(x.lo, 1);
            mpfr_set_zero(x.hi, 1);
            if (p1) add_scaled(&x, p1, su, t);
            if (p2) add_scaled(&x, p2, sv, t);
            mpfr_abs(a, x.lo, MPFR_RNDU);
            mpfr_abs(c, x.hi, MPFR_RNDU);
            if (mpfr_greater_p(c, a)) mpfr_set(a, c, MPFR_RNDU);
            mpfr_sqr(a, a, MPFR_RNDU);
            mpfr_add(out, out, a, MPFR_RNDU);
        }
    }
    mpfr_sqrt(out, out, MPFR_RNDU);
    mpfr_clear(x.lo);
    mpfr_clear(x.hi);
    mpfr_clears(su, sv, a, c, t, (mpfr_ptr)0);
}

/* Enclosures of <D1,D1>, <D1,D2> (signed) and <D2,D2> (residual_rev.c:122-153). */
typedef struct {
    interval aa, ab, bb;
} gram;

static void iv_init(interval *x) {
    mpfr_init2(x->lo, kBits);
    mpfr_init2(x->hi, kBits);
    mpfr_set_zero(x->lo, 1);
    mpfr_set_zero(x->hi, 1);
}
static void iv_clear(interval *x) {
    mpfr_clear(x->lo);
    mpfr_clear(x->hi);
}

static void add_square(interval *sum, const interval *x, mpfr_t p, mpfr_t q) {
    mpfr_sqr(p, x->lo, MPFR_RNDU);
    mpfr_sqr(q, x->hi, MPFR_RNDU);
    if (mpfr_greater_p(q, p)) mpfr_set(p, q, MPFR_RNDU);
    mpfr_add(sum->hi, sum->hi, p, MPFR_RNDU);
    if (mpfr_sgn(x->lo) <= 0 && mpfr_sgn(x->hi) >= 0) {
        mpfr_set_zero(p, 1);
    } else {
        mpfr_sqr(p, x->lo, MPFR_RNDD);
        mpfr_sqr(q, x->hi, MPFR_RNDD);
        if (mpfr_less_p(q, p)) mpfr_set(p, q, MPFR_RNDD);
    }
    mpfr_add(sum->lo, sum->lo, p, MPFR_RNDD);
}

static void add_product(interval *sum, const interval *x, const interval *y, mpfr_t p, mpfr_t q,
                        mpfr_t r) {
    mpfr_mul(p, x->lo, y->lo, MPFR_RNDD);
    mpfr_mul(q, x->lo, y->lo, MPFR_RNDU);
    mpfr_srcptr xs[2] = {x->lo, x->hi}, ys[2] = {y->lo, y->hi};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            if (!i && !j) continue;
            mpfr_mul(r, xs[i], ys[j], MPFR_RNDD);
            if (mpfr_less_p(r, p)) mpfr_set(p, r, MPFR_RNDD);
            mpfr_mul(r, xs[i], ys[j], MPFR_RNDU);
            if (mpfr_greater_p(r, q)) mpfr_set(q, r, MPFR_RNDU);
        }
    mpfr_add(sum->lo, sum->lo, p, MPFR_RNDD);
    mpfr_add(sum->hi, sum->hi, q, MPFR_RNDU);
}

static void gram_of(gram *g, const ivec *d1, const ivec *d2) {
    mpfr_t p, q, r;
    mpfr_inits2(kBits, p, q, r, (mpfr_ptr)0);
    for (size_t i = 0; i < d1->len; ++i) {
        const interval *y = ivec_find(d2, d1->keys + i * d1->words);
        add_square(&g->aa, &d1->vals[i], p, q);
        if (y) add_product(&g->ab, &d1->vals[i], y, p, q, r);
    }
    for (size_t i = 0; i < d2->len; ++i) add_square(&g->bb, &d2->vals[i], p, q);
    mpfr_clears(p, q, r, (mpfr_ptr)0);
}

/* |u D1 + v D2| from the Gram enclosure, upper ends only; when more than half the working bits
 * could cancel, the direct norm instead (residual_rev.c:154-174). */
static void gram_norm(mpfr_t out, const gram *g, const ivec *d1, const ivec *d2, double u,
                      double v) {
    if (u == 0 && v == 0) {
        mpfr_set_zero(out, 1);
        return;
    }
    mpfr_t uu, vv, uv, z, scale, floor_;
    mpfr_inits2(kBits, uu, vv, uv, z, scale, floor_, (mpfr_ptr)0);
    mpfr_set_d(uu, u, MPFR_RNDN);
    mpfr_set_d(vv, v, MPFR_RNDN);
    mpfr_mul(uv, uu, vv, MPFR_RNDN);  /* subdivision weights: all of these are exact */
    mpfr_mul_2ui(uv, uv, 1, MPFR_RNDN);
    mpfr_sqr(uu, uu, MPFR_RNDN);
    mpfr_sqr(vv, vv, MPFR_RNDN);
    mpfr_mul(out, uu, g->aa.hi, MPFR_RNDU);
    mpfr_mul(z, vv, g->bb.hi, MPFR_RNDU);
    mpfr_add(out, out, z, MPFR_RNDU);
    mpfr_set(scale, out, MPFR_RNDU);
    mpfr_mul(z, uv, g->ab.hi, MPFR_RNDU);
    mpfr_add(out, out, z, MPFR_RNDU);
    mpfr_div_2ui(floor_, scale, kBits / 2, MPFR_RNDU);
    if (mpfr_sgn(out) < 0 || mpfr_lessequal_p(out, floor_))
        direct_norm(out, d1, d2, u, v);
    else
        mpfr_sqrt(out, out, MPFR_RNDU);
    mpfr_clears(uu, vv, uv, z, scale, floor_, (mpfr_ptr)0);
}

/* The subdivided integral of the norm along the cubic Bezier control points c (residual_rev.c:
 * 175-180): at the leaves, the four control norms each weighted 1/(12 2^depth); otherwise the
 * de Casteljau halves, left then right. */
static void integrate(mpfr_t sum, const ivec *d1, const ivec *d2, const gram *g,
                      const double c[4][2], unsigned depth, unsigned top) {
    if (!depth) {
        mpfr_t z;
        mpfr_init2(z, kBits);
        for (int i = 0; i < 4; ++i) {
            if (g)
                gram_norm(z, g, d1, d2, c[i][0], c[i][1]);
            else
                direct_norm(z, d1, d2, c[i][0], c[i][1]);
            mpfr_div_2ui(z, z, top, MPFR_RNDU);
            mpfr_div_ui(z, z, 12, MPFR_RNDU);
            mpfr_add(sum, sum, z, MPFR_RNDU);
        }
        mpfr_clear(z);
        return;
    }
    double q[4][2], left[4][2], right[4][2];
    memcpy(q, c, sizeof(q));
    memcpy(left[0], q[0], sizeof(q[0]));
    memcpy(right[3], q[3], sizeof(q[3]));
    for (int level = 1; level <= 3; ++level) {  /* dyadic averages: exact in binary64 */
        for (int i = 0; i + level < 4; ++i)
            for (int k = 0; k < 2; ++k) q[i][k] = (q[i][k] + q[i + 1][k]) * .5;
        memcpy(left[level], q[0], sizeof(q[0]));
        memcpy(right[3 - level], q[3 - level], sizeof(q[0]));
    }
    integrate(sum, d1, d2, g, left, depth - 1, top);
    integrate(sum, d1, d2, g, right, depth - 1, top);
}

/* Maps as the public API builds them (residual_rev.c:21-37): sizes consistent, finite values,
 * zero padding, keys strictly ascending (highest word first). */
static int map_ok(const tpgp_map *m) {
    if (!m || !m->n_spins || m->n_words != m->n_spins / 32 + (m->n_spins % 32 != 0) ||
        m->len > m->capacity || m->n_words > SIZE_MAX / sizeof(uint64_t) ||
        m->capacity > SIZE_MAX / sizeof(double) ||
        m->capacity > SIZE_MAX / sizeof(uint64_t) / m->n_words ||
        (m->capacity && (!m->keys || !m->values)))
        return 0;
    const size_t rem = m->n_spins % 32;
    for (size_t i = 0; i < m->len; ++i) {
        const uint64_t *k = m->keys + i * m->n_words;
        if (!isfinite(m->values[i]) || (rem && (k[m->n_words - 1] >> (2 * rem)))) return 0;
        if (i) {
            const uint64_t *p = k - m->n_words;
            int cmp = 0;
            for (size_t w = m->n_words; w-- > 0 && !cmp;)
                cmp = p[w] < k[w] ? -1 : p[w] > k[w] ? 1 : 0;
            if (cmp >= 0) return 0;
        }
    }
    return 1;
}

static tpgp_status residual(const tpgp_map *a, const tpgp_map *b, const tpgp_term *t, size_t n,
                            double dt, unsigned depth, double *upper, int with_gram) {
    if (!upper || !map_ok(a) || !map_ok(b) || a->n_spins != b->n_spins ||
        a->n_words != b->n_words || !isfinite(dt) || dt <= 0 || depth > 10 || (!t && n) ||
        n > SIZE_MAX / sizeof(*t))
        return TPGP_EINVAL;
    if (!arithmetic_ok()) return TPGP_ENUMERIC;
    for (size_t j = 0; j < n; ++j)
        if (t[j].i >= a->n_spins || t[j].j >= a->n_spins || t[j].i == t[j].j || t[j].qi > 3 ||
            t[j].qj > 3 || !isfinite(t[j].h))
            return TPGP_EINVAL;
    tpgp_hip_context *ctx = tpgp_engine_default_context();
    if (!ctx) return TPGP_EBACKEND;
    /* D1, D2 on the device (tpgp_hip_term has tpgp_term's layout) */
    tpgp_hip_interval_vector dv[2];
    char err[512];
    const int hs = tpgp_hip_residual_vectors(
        ctx, a->keys, a->values, a->len, b->keys, b->values, b->len, a->n_words, a->n_spins,
        (const tpgp_hip_term *)t, n, dt, with_gram ? 1 : 0, dv, err, sizeof(err));
    if (hs) {
        tpgp_engine_note("residual vectors", err);
        return hs == TPGP_HIP_ENOMEM ? TPGP_ENOMEM : hs == TPGP_HIP_EINVAL ? TPGP_EINVAL
                                                                          : TPGP_EBACKEND;
    }
    ivec v[2];
    memset(v, 0, sizeof(v));
    int ok = from_device(&v[0], &dv[0], a->n_words) && from_device(&v[1], &dv[1], a->n_words);
    tpgp_hip_interval_vector_free(&dv[0]);
    tpgp_hip_interval_vector_free(&dv[1]);
    tpgp_status status = ok ? TPGP_OK : TPGP_ENOMEM;
    if (ok) {
        static const double c[4][2] = {{0, 0}, {1, 0}, {0, 1}, {0, 0}};
        mpfr_t sum;
        mpfr_init2(sum, kBits);
        mpfr_set_zero(sum, 1);
        gram g;
        iv_init(&g.aa);
        iv_init(&g.ab);
        iv_init(&g.bb);
        const int use_gram = with_gram && depth > 0;
        if (use_gram) gram_of(&g, &v[0], &v[1]);
        integrate(sum, &v[0], &v[1], use_gram ? &g : NULL, c, depth, depth);
        const double up = mpfr_get_d(sum, MPFR_RNDU);
        iv_clear(&g.aa);
        iv_clear(&g.ab);
        iv_clear(&g.bb);
        mpfr_clear(sum);
        if (!isfinite(up))
            status = TPGP_ENUMERIC;
        else
            *upper = up;
    }
    for (int i = 0; i < 2; ++i) ivec_free(&v[i]);
    return status;
}

tpgp_status tpgp_residual_bound(const tpgp_map *a, const tpgp_map *b, const tpgp_term *terms,
                                size_t count, double dt, unsigned depth, double *upper) {
    return residual(a, b, terms, count, dt, depth, upper, 1);
}
tpgp_status tpgp_residual_bound_reference(const tpgp_map *a, const tpgp_map *b,
                                          const tpgp_term *terms, size_t count, double dt,
                                          unsigned depth, double *upper) {
    return residual(a, b, terms, count, dt, depth, upper, 0);
}

#else /* without MPFR: no enclosure (residual_rev.c:198-199) */

tpgp_status tpgp_residual_bound(const tpgp_map *a, const tpgp_map *b, const tpgp_term *terms,
                                size_t count, double dt, unsigned depth, double *upper) {
    (void)a, (void)b, (void)terms, (void)count, (void)dt, (void)depth, (void)upper;
    return TPGP_EBACKEND;
}
tpgp_status tpgp_residual_bound_reference(const tpgp_map *a, const tpgp_map *b,
                                          const tpgp_term *terms, size_t count, double dt,
                                          unsigned depth, double *upper) {
    (void)a, (void)b, (void)terms, (void)count, (void)dt, (void)depth, (void)upper;
    return TPGP_EBACKEND;
}

#endif
