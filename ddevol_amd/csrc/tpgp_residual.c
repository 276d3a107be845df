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
 * the reference's MPFR entries bit for bit and in its first-insertion order; the device also
 * pairs the two vectors' keys and forms every entry's outward-rounded Gram terms (squares and
 * products of its ends).  What is serial in the reference stays here: the sums over D1 and D2 in
 * entry order, each addition rounded, in the same correctly rounded 256-bit arithmetic
 * (tpgp_f256.h, MPFR's results at 256 bits), and the O(1) norm and integral arithmetic in MPFR,
 * with the reference's rounding directions and order of operations, so the rounded-up binary64
 * result is the reference's bit for bit.
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

#include "tpgp_f256.h"

enum { kBits = 256 };
#define NO_PARTNER 0xffffffffu

static int arithmetic_ok(void) {
    return FLT_RADIX == 2 && DBL_MANT_DIG == 53 && sizeof(size_t) <= 8 &&
           mpfr_get_emin() <= -16384 && mpfr_get_emax() >= 16384;
}

/* The engine's process-wide device context (created on first use) and its error note. */
tpgp_hip_context *tpgp_engine_default_context(void);
void tpgp_engine_note(const char *where, const char *what);

/* tpgp_hip_f256 is tpgp_f256.h's f256. */
static f256 at(const tpgp_hip_f256 *v, size_t i) {
    f256 x;
    memcpy(&x, &v[i], sizeof(x));
    return x;
}

/* x exactly at 256 bits (the device values have 256-bit significands). */
static void to_mpfr(mpfr_t out, f256 x) {
    if (f256_is_zero(&x)) {
        mpfr_set_zero(out, x.neg ? -1 : 1);
        return;
    }
    mpz_t z;
    mpz_init(z);
    mpz_import(z, 4, -1, sizeof(uint64_t), 0, 0, x.m);
    mpfr_set_z_2exp(out, z, (mpfr_exp_t)x.e - kBits, MPFR_RNDN);
    if (x.neg) mpfr_neg(out, out, MPFR_RNDN);
    mpz_clear(z);
}

/* y += s x with outward rounding (residual_rev.c:55): f256 operations are MPFR's at 256 bits. */
static void add_scaled(f256 y[2], f256 lo, f256 hi, f256 s) {
    const int nonneg = f256_sgn(&s) >= 0;
    y[0] = f256_add(y[0], f256_mul(nonneg ? lo : hi, s, F256_DOWN), F256_DOWN);
    y[1] = f256_add(y[1], f256_mul(nonneg ? hi : lo, s, F256_UP), F256_UP);
}

/* Upper bound of |u D1 + v D2| over the union of supports (D1's entries, then D2's that D1
 * lacks), each entry's enclosure squared upward and summed upward in that order
 * (residual_rev.c:116-121); u and v are binary64 (at most 31 significant bits). */
static void direct_norm(mpfr_t out, const tpgp_hip_interval_vector *d1,
                        const tpgp_hip_interval_vector *d2, double u, double v) {
    const f256 su = f256_from_double(u), sv = f256_from_double(v);
    f256 sum = f256_zero(0);
    for (int pass = 0; pass < 2; ++pass) {
        const tpgp_hip_interval_vector *src = pass ? d2 : d1;
        for (size_t i = 0; i < src->len; ++i) {
            if (pass && src->partner[i] != NO_PARTNER) continue;  /* counted with D1's entry */
            f256 x[2] = {f256_zero(0), f256_zero(0)};
            if (!pass) {
                add_scaled(x, at(d1->lower, i), at(d1->upper, i), su);
                const uint32_t j = d1->partner[i];
                if (j != NO_PARTNER) add_scaled(x, at(d2->lower, j), at(d2->upper, j), sv);
            } else {
                add_scaled(x, at(d2->lower, i), at(d2->upper, i), sv);
            }
            f256 a = f256_abs(x[0]), c = f256_abs(x[1]);
            if (f256_cmp(&c, &a) > 0) a = c;
            sum = f256_add(sum, f256_mul_full(a, a, F256_UP), F256_UP);
        }
    }
    to_mpfr(out, sum);
    mpfr_sqrt(out, out, MPFR_RNDU);
}

/* Enclosures of <D1,D1>, <D1,D2> (signed) and <D2,D2> (residual_rev.c:122-153): the device's
 * per-entry terms summed in entry order, lower ends rounded down and upper ends up. */
typedef struct {
    mpfr_t aa[2], ab[2], bb[2];
} gram;

static void gram_of(gram *g, const tpgp_hip_interval_vector *d1,
                    const tpgp_hip_interval_vector *d2) {
    f256 aa[2] = {f256_zero(0), f256_zero(0)}, ab[2] = {f256_zero(0), f256_zero(0)};
    f256 bb[2] = {f256_zero(0), f256_zero(0)};
    for (size_t i = 0; i < d1->len; ++i) {
        aa[1] = f256_add(aa[1], at(d1->square[1], i), F256_UP);
        aa[0] = f256_add(aa[0], at(d1->square[0], i), F256_DOWN);
        if (d1->partner[i] != NO_PARTNER) {
            ab[0] = f256_add(ab[0], at(d1->product[0], i), F256_DOWN);
            ab[1] = f256_add(ab[1], at(d1->product[1], i), F256_UP);
        }
    }
    for (size_t i = 0; i < d2->len; ++i) {
        bb[1] = f256_add(bb[1], at(d2->square[1], i), F256_UP);
        bb[0] = f256_add(bb[0], at(d2->square[0], i), F256_DOWN);
    }
    for (int k = 0; k < 2; ++k) {
        mpfr_inits2(kBits, g->aa[k], g->ab[k], g->bb[k], (mpfr_ptr)0);
        to_mpfr(g->aa[k], aa[k]);
        to_mpfr(g->ab[k], ab[k]);
        to_mpfr(g->bb[k], bb[k]);
    }
}
static void gram_clear(gram *g) {
    for (int k = 0; k < 2; ++k) mpfr_clears(g->aa[k], g->ab[k], g->bb[k], (mpfr_ptr)0);
}

/* |u D1 + v D2| from the Gram enclosure, upper ends only; when more than half the working bits
 * could cancel, the direct norm instead (residual_rev.c:154-174). */
static void gram_norm(mpfr_t out, const gram *g, const tpgp_hip_interval_vector *d1,
                      const tpgp_hip_interval_vector *d2, double u, double v) {
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
    mpfr_mul(out, uu, g->aa[1], MPFR_RNDU);
    mpfr_mul(z, vv, g->bb[1], MPFR_RNDU);
    mpfr_add(out, out, z, MPFR_RNDU);
    mpfr_set(scale, out, MPFR_RNDU);
    mpfr_mul(z, uv, g->ab[1], MPFR_RNDU);
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
static void integrate(mpfr_t sum, const tpgp_hip_interval_vector *d1,
                      const tpgp_hip_interval_vector *d2, const gram *g, const double c[4][2],
                      unsigned depth, unsigned top) {
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
    static const double c[4][2] = {{0, 0}, {1, 0}, {0, 1}, {0, 0}};
    tpgp_status status = TPGP_OK;
    mpfr_t sum;
    mpfr_init2(sum, kBits);
    mpfr_set_zero(sum, 1);
    const int use_gram = with_gram && depth > 0;
    gram g;
    if (use_gram) gram_of(&g, &dv[0], &dv[1]);
    integrate(sum, &dv[0], &dv[1], use_gram ? &g : NULL, c, depth, depth);
    const double up = mpfr_get_d(sum, MPFR_RNDU);
    if (use_gram) gram_clear(&g);
    mpfr_clear(sum);
    tpgp_hip_interval_vector_free(&dv[0]);
    tpgp_hip_interval_vector_free(&dv[1]);
    if (!isfinite(up))
        status = TPGP_ENUMERIC;
    else
        *upper = up;
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
