/* The independent endpoint (residual) bound of one Strang step, host side (C17 + MPFR).
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
 * All arithmetic is MPFR at 256 bits with directed rounding, in the reference's order of
 * operations (vector entries in first-insertion order, generator terms in term order), so the
 * rounded-up binary64 result is the reference's bit for bit.  The reference has no GPU path
 * for it; the generator here runs on the host next to the interval arithmetic it feeds.
 */
#include "../../include/tpgp_engine.h"

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

/* The entry of key k, appended as [+0, +0] when absent (residual_rev.c:44-52). */
static interval *ivec_slot(ivec *v, const uint64_t *k) {
    interval *x = ivec_find(v, k);
    if (x) return x;
    if (v->len == v->cap) {
        const size_t cap = v->cap ? 2 * v->cap : 64;
        uint64_t *keys = realloc(v->keys, cap * v->words * sizeof(*keys));
        if (!keys) return NULL;
        v->keys = keys;
        interval *vals = realloc(v->vals, cap * sizeof(*vals));
        if (!vals) return NULL;
        v->vals = vals;
        v->cap = cap;
    }
    if (2 * (v->len + 1) > v->slots && !ivec_reindex(v, v->slots ? 2 * v->slots : 128))
        return NULL;
    const size_t i = v->len++;
    memcpy(v->keys + i * v->words, k, v->words * sizeof(*k));
    mpfr_init2(v->vals[i].lo, kBits);
    mpfr_init2(v->vals[i].hi, kBits);
    mpfr_set_zero(v->vals[i].lo, 1);
    mpfr_set_zero(v->vals[i].hi, 1);
    size_t s = key_hash(k, v->words) & (v->slots - 1);
    while (v->index[s]) s = (s + 1) & (v->slots - 1);
    v->index[s] = i + 1;
    return &v->vals[i];
}

/* y += s x with outward rounding (residual_rev.c:55). */
static void add_scaled(interval *y, const interval *x, const mpfr_t s, mpfr_t t) {
    const int nonneg = mpfr_sgn(s) >= 0;
    mpfr_mul(t, nonneg ? x->lo : x->hi, s, MPFR_RNDD);
    mpfr_add(y->lo, y->lo, t, MPFR_RNDD);
    mpfr_mul(t, nonneg ? x->hi : x->lo, s, MPFR_RNDU);
    mpfr_add(y->hi, y->hi, t, MPFR_RNDU);
}

/* Point intervals of a map, in map order (residual_rev.c:56). */
static int from_map(ivec *v, const tpgp_map *m) {
    for (size_t i = 0; i < m->len; ++i) {
        interval *x = ivec_slot(v, m->keys + i * m->n_words);
        if (!x || !isfinite(m->values[i])) return 0;
        mpfr_set_d(x->lo, m->values[i], MPFR_RNDN);
        mpfr_set_d(x->hi, m->values[i], MPFR_RNDN);
    }
    return 1;
}

/* y += s x over x's entries in order; nothing (not even new entries) for s = 0
 * (residual_rev.c:57). */
static int add_vec(ivec *y, const ivec *x, const mpfr_t s) {
    if (mpfr_zero_p(s)) return 1;
    mpfr_t t;
    mpfr_init2(t, kBits);
    int ok = 1;
    for (size_t i = 0; i < x->len && ok; ++i) {
        interval *z = ivec_slot(y, x->keys + i * x->words);
        if (!z)
            ok = 0;
        else
            add_scaled(z, &x->vals[i], s, t);
    }
    mpfr_clear(t);
    return ok;
}

/* Exponent of i in the product of Pauli symbols q p (I0 X1 Y2 Z3). */
static unsigned phase(unsigned q, unsigned p) {
    if (!q || !p || q == p) return 0;
    return (p + 3 - q) % 3 == 1 ? 1u : 3u;
}
static unsigned symbol(const uint64_t *k, size_t site) {
    return (unsigned)(k[site / 32] >> (2 * (site % 32))) & 3u;
}

/* The contribution of term t to the generator applied to entry (k, x): for an anticommuting
 * term, out[k ^ t] += s x with s = +-2h (residual_rev.c:60-70, 96-108). */
static int apply_term(ivec *out, const uint64_t *k, const interval *x, const tpgp_term *t,
                      int skip_zero_h, uint64_t *tmp_key, mpfr_t s, mpfr_t tmp) {
    const unsigned p = (phase(t->qi, symbol(k, t->i)) + phase(t->qj, symbol(k, t->j))) & 3u;
    if (!(p & 1u) || (skip_zero_h && t->h == 0)) return 1;
    memcpy(tmp_key, k, out->words * sizeof(*k));
    tmp_key[t->i / 32] ^= (uint64_t)t->qi << (2 * (t->i % 32));
    tmp_key[t->j / 32] ^= (uint64_t)t->qj << (2 * (t->j % 32));
    interval *z = ivec_slot(out, tmp_key);
    if (!z) return 0;
    mpfr_set_d(s, t->h, MPFR_RNDN);
    mpfr_mul_2ui(s, s, 1, MPFR_RNDN);
    if (p == 3) mpfr_neg(s, s, MPFR_RNDN);
    add_scaled(z, x, s, tmp);
    return 1;
}

/* Generator, diagnostic form: every entry against every term in term order
 * (residual_rev.c:60-70). */
static int generator_all(ivec *out, const ivec *in, const tpgp_term *t, size_t n) {
    out->words = in->words;
    uint64_t *key = malloc(in->words * sizeof(*key));
    if (!key) return 0;
    mpfr_t s, tmp;
    mpfr_inits2(kBits, s, tmp, (mpfr_ptr)0);
    int ok = 1;
    for (size_t i = 0; i < in->len && ok; ++i)
        for (size_t j = 0; j < n && ok; ++j)
            ok = apply_term(out, in->keys + i * in->words, &in->vals[i], &t[j], 0, key, s, tmp);
    mpfr_clears(s, tmp, (mpfr_ptr)0);
    free(key);
    return ok;
}

static int cmp_size(const void *a, const void *b) {
    const size_t x = *(const size_t *)a, y = *(const size_t *)b;
    return (x > y) - (x < y);
}

/* Generator: per nonzero entry only the terms incident on one of its non-identity sites, in
 * term order (the others commute); terms with h = 0 add nothing (residual_rev.c:74-111). */
static int generator(ivec *out, const ivec *in, const tpgp_term *t, size_t n) {
    if (!n) return generator_all(out, in, t, n);
    size_t sites = 0;
    for (size_t j = 0; j < n; ++j) {
        if (t[j].i >= sites) sites = t[j].i + 1;
        if (t[j].j >= sites) sites = t[j].j + 1;
    }
    if (n > SIZE_MAX / (2 * sizeof(size_t))) return 0;
    size_t *first = calloc(sites + 1, sizeof(size_t)), *fill = calloc(sites, sizeof(size_t));
    size_t *incident = malloc(2 * n * sizeof(size_t)), *list = malloc(n * sizeof(size_t));
    unsigned char *taken = calloc(n, 1);
    uint64_t *key = malloc(in->words * sizeof(*key));
    int ok = first && fill && incident && list && taken && key;
    if (ok) {
        for (size_t j = 0; j < n; ++j) {  /* terms of each site, ascending */
            first[t[j].i + 1]++;
            first[t[j].j + 1]++;
        }
        for (size_t s = 1; s <= sites; ++s) first[s] += first[s - 1];
        memcpy(fill, first, sites * sizeof(size_t));
        for (size_t j = 0; j < n; ++j) {
            incident[fill[t[j].i]++] = j;
            incident[fill[t[j].j]++] = j;
        }
    }
    out->words = in->words;
    mpfr_t s, tmp;
    mpfr_inits2(kBits, s, tmp, (mpfr_ptr)0);
    for (size_t e = 0; ok && e < in->len; ++e) {
        const interval *x = &in->vals[e];
        if (mpfr_zero_p(x->lo) && mpfr_zero_p(x->hi)) continue;
        const uint64_t *k = in->keys + e * in->words;
        size_t count = 0;
        for (size_t w = 0; w < in->words; ++w) {
            for (uint64_t bits = k[w]; bits;) {
                const unsigned local = (unsigned)__builtin_ctzll(bits) / 2;
                bits &= ~(3ull << (2 * local));
                const size_t site = 32 * w + local;
                if (site >= sites) continue;
                for (size_t q = first[site]; q < first[site + 1]; ++q)
                    if (!taken[incident[q]]) {
                        taken[incident[q]] = 1;
                        list[count++] = incident[q];
                    }
            }
        }
        qsort(list, count, sizeof(size_t), cmp_size);
        for (size_t q = 0; q < count; ++q) {
            taken[list[q]] = 0;
            if (ok) ok = apply_term(out, k, x, &t[list[q]], 1, key, s, tmp);
        }
    }
    mpfr_clears(s, tmp, (mpfr_ptr)0);
    free(first);
    free(fill);
    free(incident);
    free(list);
    free(taken);
    free(key);
    return ok;
}

/* out += factor dt^power in; the scalar is exact at 256 bits (residual_rev.c:112-114). */
static int add_dt(ivec *out, const ivec *in, double dt, int power, int factor) {
    mpfr_t s;
    mpfr_init2(s, kBits);
    if (power == 0)
        mpfr_set_ui(s, 1, MPFR_RNDN);
    else
        mpfr_set_d(s, dt, MPFR_RNDN);
    if (power == 2) mpfr_sqr(s, s, MPFR_RNDN);
    mpfr_mul_si(s, s, factor, MPFR_RNDN);
    const int ok = add_vec(out, in, s);
    mpfr_clear(s);
    return ok;
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
    /* v: a, b, L a, L b, L L a, L L b, D1, D2 */
    ivec v[8];
    memset(v, 0, sizeof(v));
    for (int i = 0; i < 8; ++i) v[i].words = a->n_words;
    int (*gen)(ivec *, const ivec *, const tpgp_term *, size_t) =
        with_gram ? generator : generator_all;
    int ok = from_map(&v[0], a) && from_map(&v[1], b) && gen(&v[2], &v[0], t, n) &&
             gen(&v[3], &v[1], t, n) && gen(&v[4], &v[2], t, n) && gen(&v[5], &v[3], t, n);
    ok = ok && add_dt(&v[6], &v[1], dt, 0, 6) && add_dt(&v[6], &v[0], dt, 0, -6) &&
         add_dt(&v[6], &v[2], dt, 1, -4) && add_dt(&v[6], &v[3], dt, 1, -2) &&
         add_dt(&v[6], &v[4], dt, 2, -1);
    ok = ok && add_dt(&v[7], &v[1], dt, 0, 6) && add_dt(&v[7], &v[0], dt, 0, -6) &&
         add_dt(&v[7], &v[2], dt, 1, -2) && add_dt(&v[7], &v[3], dt, 1, -4) &&
         add_dt(&v[7], &v[5], dt, 2, 1);
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
        if (use_gram) gram_of(&g, &v[6], &v[7]);
        integrate(sum, &v[6], &v[7], use_gram ? &g : NULL, c, depth, depth);
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
    for (int i = 0; i < 8; ++i) ivec_free(&v[i]);
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
