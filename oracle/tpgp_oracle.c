/* CPU oracle for the Pauli-propagation branch step -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates the reference CPU engine (engine/src/tpgp_rev.c, engine/src/residual_rev.c) so the
 * HIP engine can be checked bit for bit on the GPU box, where the reference sources do not exist.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Floating point: IEEE binary64, round-to-nearest, gradual underflow, no contraction
 * (built with -ffp-contract=off -fno-fast-math, as the reference is).
 */
#include "tpgp_oracle.h"
#include <fenv.h>
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#ifdef TPGP_HAVE_MPFR
#include <mpfr.h>
#endif
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- directed-rounding helpers: tpgp_rev.c:25-63 ------------------------------------------ */
/* upadd/upmul: separately rounded result nudged one binary64 step toward +inf, with exact zero
 * short-cuts (tpgp_rev.c:25-36). */
double orc_upadd(double a, double b) {
    if (a == 0) return b;
    if (b == 0) return a;
    return nextafter(a + b, INFINITY);
}
double orc_upmul(double a, double b) {
    if (a == 0 || b == 0) return 0;
    return nextafter(a * b, INFINITY);
}
/* Larger of the two neighbouring gaps around x (tpgp_rev.c:37-42). */
double orc_round_radius(double x) {
    if (!isfinite(x)) return INFINITY;
    double above = fabs(nextafter(x, INFINITY) - x);
    double below = fabs(x - nextafter(x, -INFINITY));
    return fmax(above, below);
}
/* Upper bound of the Euclidean norm: MPFR-192 RNDU accumulation of RNDU squares, RNDU sqrt and
 * RNDU conversion; without MPFR an upadd/upmul chain then one step up (tpgp_rev.c:43-63). */
double orc_upper_l2(const double *v, size_t n) {
#ifdef TPGP_HAVE_MPFR
    mpfr_t acc, sq;
    mpfr_init2(acc, 192);
    mpfr_init2(sq, 192);
    mpfr_set_zero(acc, 0);
    for (size_t k = 0; k < n; k++) {
        mpfr_set_d(sq, v[k], MPFR_RNDN);
        mpfr_sqr(sq, sq, MPFR_RNDU);
        mpfr_add(acc, acc, sq, MPFR_RNDU);
    }
    mpfr_sqrt(acc, acc, MPFR_RNDU);
    double r = mpfr_get_d(acc, MPFR_RNDU);
    mpfr_clear(acc);
    mpfr_clear(sq);
    return r;
#else
    double acc = 0;
    for (size_t k = 0; k < n; k++) acc = orc_upadd(acc, orc_upmul(v[k], v[k]));
    return acc == 0 ? 0 : nextafter(sqrt(acc), INFINITY);
#endif
}

int orc_has_mpfr(void) {
#ifdef TPGP_HAVE_MPFR
    return 1;
#else
    return 0;
#endif
}
int orc_has_openmp(void) {
#ifdef _OPENMP
    return 1;
#else
    return 0;
#endif
}
void orc_ledger_init(orc_ledger *l, int validated) { /* tpgp_rev.c:145-148 */
    memset(l, 0, sizeof(*l));
    l->arithmetic_enclosed = validated && orc_has_mpfr();
}

/* ---- keys and the sorted map: tpgp_rev.c:64-116, 149-259 ---------------------------------- */
/* Order: uint64 words compared from the highest word down (tpgp_rev.c:64-72). */
static int key_order(const uint64_t *a, const uint64_t *b, size_t words) {
    size_t w = words;
    while (w--) {
        if (a[w] != b[w]) return a[w] < b[w] ? -1 : 1;
    }
    return 0;
}
/* Padding bits above the last site must be zero (tpgp_rev.c:73-76). */
static int key_padding_ok(size_t n_spins, size_t words, const uint64_t *key) {
    unsigned used = (unsigned)(n_spins % 32);
    return used == 0 || (key[words - 1] >> (2 * used)) == 0;
}
static unsigned symbol_at(const uint64_t *key, size_t site) { /* tpgp_rev.c:187-189 */
    return (unsigned)(key[site / 32] >> (2 * (site % 32))) & 3u;
}
static void symbol_put(uint64_t *key, size_t site, unsigned p) { /* tpgp_rev.c:190-193 */
    unsigned shift = (unsigned)(2 * (site % 32));
    key[site / 32] = (key[site / 32] & ~((uint64_t)3 << shift)) | ((uint64_t)(p & 3u) << shift);
}
static int mul_ok(size_t a, size_t b, size_t *out) {
    if (b && a > SIZE_MAX / b) return 0;
    *out = a * b;
    return 1;
}
static int map_ok(const orc_map *m) { /* tpgp_rev.c:77-83 */
    size_t nk, bytes;
    return m && m->n_spins && m->n_spins <= SIZE_MAX - 31 && m->n_words == (m->n_spins + 31) / 32 &&
           m->len <= m->capacity && mul_ok(m->capacity, m->n_words, &nk) &&
           mul_ok(nk, sizeof(uint64_t), &bytes) && m->capacity <= SIZE_MAX / sizeof(double) &&
           (!m->capacity || (m->keys && m->values));
}
/* Capacity doubling from 8 (tpgp_rev.c:84-116). */
static int grow(orc_map *m, size_t need) {
    if (need <= m->capacity) return ORC_OK;
    size_t cap = m->capacity ? m->capacity : 8;
    while (cap < need) cap = cap > SIZE_MAX / 2 ? need : cap * 2;
    size_t nk, bytes;
    if (!mul_ok(cap, m->n_words, &nk) || !mul_ok(nk, sizeof(uint64_t), &bytes) ||
        cap > SIZE_MAX / sizeof(double))
        return ORC_ENOMEM;
    uint64_t *k = malloc(bytes);
    double *v = malloc(cap * sizeof(double));
    if (!k || !v) {
        free(k);
        free(v);
        return ORC_ENOMEM;
    }
    if (m->len) {
        memcpy(k, m->keys, m->len * m->n_words * sizeof(uint64_t));
        memcpy(v, m->values, m->len * sizeof(double));
    }
    free(m->keys);
    free(m->values);
    m->keys = k;
    m->values = v;
    m->capacity = cap;
    return ORC_OK;
}
static void map_reset(orc_map *m, size_t n_spins) {
    free(m->keys);
    free(m->values);
    memset(m, 0, sizeof(*m));
    m->n_spins = n_spins;
    m->n_words = (n_spins + 31) / 32;
}
orc_map *orc_map_new(size_t n_spins) { /* tpgp_map_init, tpgp_rev.c:149-156 */
    if (!n_spins || n_spins > SIZE_MAX - 31) return NULL;
    orc_map *m = calloc(1, sizeof(*m));
    if (m) map_reset(m, n_spins);
    return m;
}
void orc_map_delete(orc_map *m) {
    if (!m) return;
    free(m->keys);
    free(m->values);
    free(m);
}
size_t orc_map_len(const orc_map *m) { return m->len; }
size_t orc_map_words(const orc_map *m) { return m->n_words; }
void orc_map_export(const orc_map *m, uint64_t *keys, double *values) {
    if (m->len) {
        memcpy(keys, m->keys, m->len * m->n_words * sizeof(uint64_t));
        memcpy(values, m->values, m->len * sizeof(double));
    }
}
/* Load an already canonical map: strictly ascending keys, zero padding, finite nonzero values. */
int orc_map_assign(orc_map *m, const uint64_t *keys, const double *values, size_t len) {
    if (!map_ok(m) || (len && (!keys || !values))) return ORC_EINVAL;
    for (size_t k = 0; k < len; k++) {
        const uint64_t *key = keys + k * m->n_words;
        if (!key_padding_ok(m->n_spins, m->n_words, key) || !isfinite(values[k]) || values[k] == 0)
            return ORC_EINVAL;
        if (k && key_order(key - m->n_words, key, m->n_words) >= 0) return ORC_EINVAL;
    }
    size_t n = m->n_spins;
    map_reset(m, n);
    int s = grow(m, len);
    if (s) return s;
    orc_map_export(&(orc_map){n, m->n_words, len, len, (uint64_t *)keys, (double *)values}, m->keys,
                   m->values);
    m->len = len;
    return ORC_OK;
}
int orc_map_copy(orc_map *dst, const orc_map *src) { /* tpgp_map_clone, tpgp_rev.c:164-186 */
    if (!dst || !map_ok(src)) return ORC_EINVAL;
    if (dst == src) return ORC_OK;
    orc_map tmp = {0};
    map_reset(&tmp, src->n_spins);
    int s = grow(&tmp, src->len);
    if (s) {
        map_reset(&tmp, 1);
        return s;
    }
    tmp.len = src->len;
    orc_map_export(src, tmp.keys, tmp.values);
    free(dst->keys);
    free(dst->values);
    *dst = tmp;
    return ORC_OK;
}
/* First position whose key is not below `key` (tpgp_rev.c:194-205). */
static size_t first_not_below(const orc_map *m, const uint64_t *key, int *hit) {
    size_t lo = 0, hi = m->len;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (key_order(m->keys + mid * m->n_words, key, m->n_words) < 0)
            lo = mid + 1;
        else
            hi = mid;
    }
    *hit = lo < m->len && key_order(m->keys + lo * m->n_words, key, m->n_words) == 0;
    return lo;
}
/* Insert or accumulate, removing exact cancellations (tpgp_rev.c:206-243). */
int orc_map_add(orc_map *m, const uint64_t *key, double v) {
    if (!map_ok(m) || !key || !isfinite(v) || !key_padding_ok(m->n_spins, m->n_words, key))
        return ORC_EINVAL;
    if (v == 0) return ORC_OK;
    size_t W = m->n_words;
    int hit;
    size_t at = first_not_below(m, key, &hit);
    if (hit) {
        double total = m->values[at] + v;
        if (!isfinite(total)) return ORC_ENUMERIC;
        if (total != 0) {
            m->values[at] = total;
            return ORC_OK;
        }
        size_t tail = m->len - at - 1;
        memmove(m->keys + at * W, m->keys + (at + 1) * W, tail * W * sizeof(uint64_t));
        memmove(m->values + at, m->values + at + 1, tail * sizeof(double));
        m->len--;
        return ORC_OK;
    }
    uint64_t *saved = malloc(W * sizeof(uint64_t)); /* key may alias map storage */
    if (!saved) return ORC_ENOMEM;
    memcpy(saved, key, W * sizeof(uint64_t));
    int s = grow(m, m->len + 1);
    if (s) {
        free(saved);
        return s;
    }
    size_t tail = m->len - at;
    memmove(m->keys + (at + 1) * W, m->keys + at * W, tail * W * sizeof(uint64_t));
    memmove(m->values + at + 1, m->values + at, tail * sizeof(double));
    memcpy(m->keys + at * W, saved, W * sizeof(uint64_t));
    free(saved);
    m->values[at] = v;
    m->len++;
    return ORC_OK;
}
int orc_map_add_single(orc_map *m, size_t site, unsigned p, double v) { /* tpgp_rev.c:244-254 */
    if (!m || site >= m->n_spins || p > 3) return ORC_EINVAL;
    uint64_t *key = calloc(m->n_words, sizeof(uint64_t));
    if (!key) return ORC_ENOMEM;
    symbol_put(key, site, p);
    int s = orc_map_add(m, key, v);
    free(key);
    return s;
}
double orc_map_get(const orc_map *m, const uint64_t *key) { /* tpgp_rev.c:255-259 */
    int hit;
    size_t at = first_not_below(m, key, &hit);
    return hit ? m->values[at] : 0;
}
double orc_map_l2(const orc_map *m) { /* long-double sum of squares, tpgp_rev.c:266-271 */
    long double acc = 0;
    for (size_t k = 0; k < m->len; k++) acc += (long double)m->values[k] * m->values[k];
    return (double)sqrtl(acc);
}
/* Weight and last nonzero symbol of a key (the scan in tpgp_rev.c:316-327). */
static void weight_symbol(const uint64_t *key, size_t words, unsigned *weight, unsigned *symbol) {
    *weight = 0;
    *symbol = 0;
    for (size_t w = 0; w < words; w++)
        for (uint64_t q = key[w]; q; q >>= 2)
            if (q & 3u) {
                ++*weight;
                *symbol = (unsigned)(q & 3u);
            }
}
/* Neumaier-compensated long-double sum over weight-1 keys of symbol p (tpgp_rev.c:311-338). */
double orc_map_single_sum(const orc_map *m, unsigned p) {
    if (p < 1 || p > 3) return NAN;
    long double sum = 0, comp = 0;
    for (size_t k = 0; k < m->len; k++) {
        unsigned weight, symbol;
        weight_symbol(m->keys + k * m->n_words, m->n_words, &weight, &symbol);
        if (weight != 1 || symbol != p) continue;
        long double x = m->values[k], t = sum + x;
        comp += fabsl(sum) >= fabsl(x) ? (sum - t) + x : (x - t) + sum;
        sum = t;
    }
    return (double)(sum + comp);
}
/* MPFR-192 readout: RN midpoint of sum/normalizer plus a radius enclosing its rounding
 * (tpgp_rev.c:339-388). */
int orc_readout(const orc_map *m, unsigned p, size_t normalizer, double *midpoint,
                double *radius) {
    if (!map_ok(m) || !normalizer || p < 1 || p > 3 || !midpoint || !radius) return ORC_EINVAL;
#ifdef TPGP_HAVE_MPFR
    mpfr_t lo, hi, mid, x, den, r;
    mpfr_inits2(192, lo, hi, mid, x, den, r, (mpfr_ptr)0);
    mpfr_set_zero(lo, 0);
    mpfr_set_zero(hi, 0);
    mpfr_set_zero(mid, 0);
    mpfr_set_uj(den, (uintmax_t)normalizer, MPFR_RNDN);
    for (size_t k = 0; k < m->len; k++) {
        unsigned weight, symbol;
        weight_symbol(m->keys + k * m->n_words, m->n_words, &weight, &symbol);
        if (weight != 1 || symbol != p) continue;
        mpfr_set_d(x, m->values[k], MPFR_RNDN);
        mpfr_add(lo, lo, x, MPFR_RNDD);
        mpfr_add(hi, hi, x, MPFR_RNDU);
        mpfr_add(mid, mid, x, MPFR_RNDN);
    }
    mpfr_div(lo, lo, den, MPFR_RNDD);
    mpfr_div(hi, hi, den, MPFR_RNDU);
    mpfr_div(mid, mid, den, MPFR_RNDN);
    double centre = mpfr_get_d(mid, MPFR_RNDN);
    mpfr_set_d(mid, centre, MPFR_RNDN);
    mpfr_sub(r, mid, lo, MPFR_RNDU);
    mpfr_sub(x, hi, mid, MPFR_RNDU);
    mpfr_max(r, r, x, MPFR_RNDU);
    double rad = mpfr_get_d(r, MPFR_RNDU);
    mpfr_clears(lo, hi, mid, x, den, r, (mpfr_ptr)0);
    if (!isfinite(centre) || !isfinite(rad)) return ORC_ENUMERIC;
    *midpoint = centre;
    *radius = rad;
    return ORC_OK;
#else
    return ORC_EBACKEND;
#endif
}

/* ---- Pauli algebra: tpgp_rev.c:414-418 ---------------------------------------------------- */
/* Exponent of i in the product QP, symbols I0 X1 Y2 Z3; 4 flags an invalid symbol. */
unsigned orc_pauli_phase(unsigned q, unsigned p) {
    if (q > 3 || p > 3) return 4;
    if (!q || !p || q == p) return 0;
    /* cyclic X->Y->Z->X gives +i, the reverse order -i */
    return ((p + 3 - q) % 3 == 1) ? 1u : 3u;
}
static int anticommutes(unsigned q, unsigned p) { return q && p && q != p; }

/* ---- trigonometric factors: tpgp_rev.c:484-534 -------------------------------------------- */
static int trig_factors(double h, double dur, int validated, double *c, double *s, double *rc,
                        double *rs) {
    *rc = *rs = 0;
#ifdef TPGP_HAVE_MPFR
    if (validated) {
        mpfr_t x, lo, hi, mid, a, b;
        mpfr_inits2(256, x, lo, hi, mid, a, b, (mpfr_ptr)0);
        mpfr_set_d(x, h, MPFR_RNDN);
        mpfr_set_d(a, dur, MPFR_RNDN);
        mpfr_mul(x, x, a, MPFR_RNDN); /* exact: <=106 significant bits */
        mpfr_mul_2ui(x, x, 1, MPFR_RNDN);
        for (int which = 0; which < 2; which++) {
            int (*f)(mpfr_ptr, mpfr_srcptr, mpfr_rnd_t) = which ? mpfr_sin : mpfr_cos;
            f(lo, x, MPFR_RNDD);
            f(hi, x, MPFR_RNDU);
            f(mid, x, MPFR_RNDN);
            double centre = mpfr_get_d(mid, MPFR_RNDN);
            mpfr_set_d(mid, centre, MPFR_RNDN);
            mpfr_sub(a, mid, lo, MPFR_RNDU);
            mpfr_sub(b, hi, mid, MPFR_RNDU);
            mpfr_max(a, a, b, MPFR_RNDU);
            double rad = mpfr_get_d(a, MPFR_RNDU);
            if (which) {
                *s = centre;
                *rs = rad;
            } else {
                *c = centre;
                *rc = rad;
            }
        }
        mpfr_clears(x, lo, hi, mid, a, b, (mpfr_ptr)0);
        return ORC_OK;
    }
#else
    if (validated) return ORC_EBACKEND;
#endif
    double x = 2 * h * dur;
    if (!isfinite(x)) return ORC_ENUMERIC;
    *c = cos(x);
    *s = sin(x);
    return ORC_OK;
}
int orc_trig(double h, double dur, int validated, double out[4]) {
    return trig_factors(h, dur, validated, out, out + 1, out + 2, out + 3);
}

/* ---- raw branches: tpgp_rev.c:432-467 (identical to engine/src/cuda_rev.cu:78-115) -------- */
typedef struct {
    const uint64_t *keys;
    const double *values;
    size_t words, i, j, begin, end;
    unsigned qi, qj;
    double c, s, rc, rs;
    int enclose;
    uint64_t *out_keys;
    double *out_values, *out_radii;
} branch_range;

static void branch_slots(const branch_range *b) {
    size_t W = b->words;
    for (size_t k = b->begin; k < b->end; k++) {
        const uint64_t *src = b->keys + k * W;
        uint64_t *a_key = b->out_keys + 2 * k * W, *d_key = a_key + W;
        memcpy(a_key, src, W * sizeof(uint64_t));
        memcpy(d_key, src, W * sizeof(uint64_t));
        double v = b->values[k];
        unsigned pi = symbol_at(src, b->i), pj = symbol_at(src, b->j);
        if (!(anticommutes(b->qi, pi) ^ anticommutes(b->qj, pj))) {
            b->out_values[2 * k] = v;
            b->out_values[2 * k + 1] = 0;
            continue;
        }
        unsigned phase = (orc_pauli_phase(b->qi, pi) + orc_pauli_phase(b->qj, pj)) & 3u;
        symbol_put(d_key, b->i, pi ^ b->qi);
        symbol_put(d_key, b->j, pj ^ b->qj);
        double x = v * b->c, y = v * b->s;
        b->out_values[2 * k] = x;
        b->out_values[2 * k + 1] = (phase == 1 ? 1 : -1) * y;
        if (b->enclose) {
            b->out_radii[2 * k] = orc_upadd(orc_upmul(fabs(v), b->rc), orc_round_radius(x));
            b->out_radii[2 * k + 1] = orc_upadd(orc_upmul(fabs(v), b->rs), orc_round_radius(y));
        }
    }
}
static void *branch_thread(void *arg) {
    branch_slots(arg);
    return NULL;
}
void orc_branches(const uint64_t *keys, const double *values, size_t count, size_t words,
                  size_t i, size_t j, unsigned qi, unsigned qj, double c, double s, double rc,
                  double rs, int enclose, uint64_t *out_keys, double *out_values,
                  double *out_radii) {
    for (size_t k = 0; k < 2 * count; k++) out_radii[k] = 0;
    branch_range b = {keys, values, words, i, j, 0, count, qi, qj, c, s, rc, rs,
                      enclose, out_keys, out_values, out_radii};
    branch_slots(&b);
}

/* ---- stable ordering of raw slots: tpgp_rev.c:468-483 ------------------------------------ */
/* The reference uses a top-down stable merge sort; any stable sort yields the same permutation.
 * This is a bottom-up merge sort of slot indices by key. */
static void stable_order(size_t *idx, size_t *tmp, size_t n, const uint64_t *keys, size_t W) {
    size_t *src = idx, *dst = tmp;
    for (size_t width = 1; width < n; width *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * width) {
            size_t mid = lo + width < n ? lo + width : n;
            size_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            size_t a = lo, b = mid, o = lo;
            /* take from the right run only when strictly smaller: equal keys keep slot order */
            while (a < mid && b < hi)
                dst[o++] = key_order(keys + src[b] * W, keys + src[a] * W, W) < 0 ? src[b++]
                                                                                   : src[a++];
            while (a < mid) dst[o++] = src[a++];
            while (b < hi) dst[o++] = src[b++];
        }
        size_t *t = src;
        src = dst;
        dst = t;
    }
    if (src != idx) memcpy(idx, src, n * sizeof(size_t));
}

/* ---- one rotation: tpgp_rev.c:535-771 ----------------------------------------------------- */
static int ledger_ok(const orc_ledger *l) {
    return isfinite(l->dropped_l1) && l->dropped_l1 >= 0 && isfinite(l->dropped_l2) &&
           l->dropped_l2 >= 0 && isfinite(l->dropped_l2_squared) && l->dropped_l2_squared >= 0 &&
           isfinite(l->arithmetic_l2) && l->arithmetic_l2 >= 0;
}
int orc_rotate_product(orc_map *map, size_t i, size_t j, unsigned qi, unsigned qj, double h,
                       double dur, const orc_options *o, orc_ledger *ledger) {
    /* Argument checks in the reference's order and precedence (tpgp_rev.c:538-557). */
    if (!map_ok(map) || !o || !ledger || i >= map->n_spins || j >= map->n_spins || i == j ||
        qi > 3 || qj > 3 || !isfinite(h) || !isfinite(dur) || !isfinite(o->threshold) ||
        o->threshold < 0 || !isfinite(o->max_drop_l2) || o->max_drop_l2 < 0 || o->threads == 0)
        return ORC_EINVAL;
    if (!ledger_ok(ledger)) return ORC_EINVAL;
    if (o->validated) {
        volatile double tiny = DBL_MIN, half = .5;
        if (FLT_RADIX != 2 || DBL_MANT_DIG != 53 || fegetround() != FE_TONEAREST ||
            tiny * half == 0)
            return ORC_ENUMERIC;
    }
    if (o->backend < ORC_SERIAL || o->backend > ORC_OPENMP) return ORC_EBACKEND;
    if (o->backend == ORC_OPENMP && !orc_has_openmp()) return ORC_EBACKEND;
    double c, s, rc, rs;
    int status = trig_factors(h, dur, o->validated, &c, &s, &rc, &rs);
    if (status) return status;
    if (h == 0 || dur == 0 || map->len == 0) return ORC_OK; /* no-op, not counted (562-563) */
    if (map->len > SIZE_MAX / 2) return ORC_ENOMEM;

    const size_t W = map->n_words, n = 2 * map->len;
    uint64_t *rk = NULL;
    double *rv = NULL, *rr = NULL, *removed = NULL, *mrad = NULL;
    size_t *idx = NULL, *tmp = NULL;
    orc_map out = {0};
    map_reset(&out, map->n_spins);
    size_t kb;
    if (!mul_ok(n, W * sizeof(uint64_t), &kb)) return ORC_ENOMEM;
    rk = malloc(kb);
    rv = malloc(n * sizeof(double));
    rr = calloc(n, sizeof(double));
    idx = malloc(n * sizeof(size_t));
    tmp = malloc(n * sizeof(size_t));
    removed = malloc(n * sizeof(double));
    mrad = malloc(n * sizeof(double));
    if (!rk || !rv || !rr || !idx || !tmp || !removed || !mrad) {
        status = ORC_ENOMEM;
        goto done;
    }
    /* Branch generation: serial, pthreads over contiguous ranges, or an OpenMP region
     * (tpgp_rev.c:617-681). */
    branch_range base = {map->keys, map->values, W, i, j, 0, map->len, qi, qj, c, s, rc, rs,
                         o->validated, rk, rv, rr};
    unsigned nt = o->threads > map->len ? (unsigned)map->len : o->threads;
    if (o->backend == ORC_PTHREAD && nt > 1) {
        pthread_t *th = calloc(nt, sizeof(*th));
        branch_range *jobs = calloc(nt, sizeof(*jobs));
        if (!th || !jobs) {
            free(th);
            free(jobs);
            status = ORC_ENOMEM;
            goto done;
        }
        unsigned started = 0;
        for (unsigned t = 0; t < nt; t++) {
            jobs[t] = base;
            jobs[t].begin = map->len * t / nt;
            jobs[t].end = map->len * (t + 1) / nt;
            if (pthread_create(th + t, NULL, branch_thread, jobs + t)) {
                status = ORC_EBACKEND;
                break;
            }
            started++;
        }
        for (unsigned t = 0; t < started; t++) pthread_join(th[t], NULL);
        free(th);
        free(jobs);
        if (status) goto done;
    }
#ifdef _OPENMP
    else if (o->backend == ORC_OPENMP && nt > 1) {
#pragma omp parallel num_threads(nt)
        {
            branch_range job = base;
            size_t t = (size_t)omp_get_thread_num(), T = (size_t)omp_get_num_threads();
            job.begin = map->len * t / T;
            job.end = map->len * (t + 1) / T;
            branch_slots(&job);
        }
    }
#endif
    else
        branch_slots(&base);
    for (size_t k = 0; k < n; k++) { /* tpgp_rev.c:682-688 */
        if (!isfinite(rv[k]) || !isfinite(rr[k])) {
            status = ORC_ENUMERIC;
            goto done;
        }
        idx[k] = k;
    }
    stable_order(idx, tmp, n, rk, W);
    status = grow(&out, n);
    if (status) goto done;
    /* Merge equal keys, drop zeros, truncate non-identity |v| <= threshold (690-729). */
    size_t dropped = 0, nrad = 0;
    double l1 = 0, sq = 0;
    for (size_t k = 0; k < n;) {
        size_t p = idx[k++];
        const uint64_t *key = rk + p * W;
        double v = rv[p], r = rr[p];
        while (k < n && key_order(key, rk + idx[k] * W, W) == 0) {
            size_t q = idx[k++];
            double before = v;
            v += rv[q];
            r = orc_upadd(r, rr[q]);
            if (o->validated && before != 0 && rv[q] != 0) r = orc_upadd(r, orc_round_radius(v));
        }
        if (!isfinite(v) || !isfinite(r)) {
            status = ORC_ENUMERIC;
            goto done;
        }
        if (r) mrad[nrad++] = r;
        int identity = 1;
        for (size_t w = 0; w < W && identity; w++) identity = key[w] == 0;
        if (v == 0) continue;
        if (!identity && fabs(v) <= o->threshold) {
            removed[dropped++] = fabs(v);
            l1 = orc_upadd(l1, fabs(v));
            sq = orc_upadd(sq, orc_upmul(v, v));
            continue;
        }
        memcpy(out.keys + out.len * W, key, W * sizeof(uint64_t));
        out.values[out.len++] = v;
    }
    if (o->max_terms && out.len > o->max_terms) { /* transactional cap (730-733) */
        status = ORC_ECAP;
        goto done;
    }
    orc_ledger next = *ledger; /* ledger update and budget (734-757) */
    double l2 = orc_upper_l2(removed, dropped), ar = orc_upper_l2(mrad, nrad);
    next.dropped_l1 = orc_upadd(next.dropped_l1, l1);
    next.dropped_l2 = orc_upadd(next.dropped_l2, l2);
    next.dropped_l2_squared = orc_upadd(next.dropped_l2_squared, sq);
    next.arithmetic_l2 = orc_upadd(next.arithmetic_l2, ar);
    next.rotations++;
    if (dropped) next.removal_events++;
    if (out.len > next.peak_terms) next.peak_terms = out.len;
    next.arithmetic_enclosed = next.arithmetic_enclosed && o->validated && orc_has_mpfr();
    if (!isfinite(next.dropped_l2) || !isfinite(next.arithmetic_l2)) {
        status = ORC_ENUMERIC;
        goto done;
    }
    if (o->max_drop_l2 && next.dropped_l2 > o->max_drop_l2) {
        status = ORC_EBUDGET;
        goto done;
    }
    free(map->keys);
    free(map->values);
    *map = out;
    memset(&out, 0, sizeof(out));
    *ledger = next;
done:
    free(rk);
    free(rv);
    free(rr);
    free(idx);
    free(tmp);
    free(removed);
    free(mrad);
    free(out.keys);
    free(out.values);
    return status;
}
int orc_rotate(orc_map *m, size_t i, size_t j, unsigned qi, unsigned qj, double theta,
               const orc_options *o, orc_ledger *l) { /* tpgp_rev.c:772-775 */
    return orc_rotate_product(m, i, j, qi, qj, theta, 1.0, o, l);
}

/* ---- second-order Trotter step, transactional: tpgp_rev.c:776-809 ------------------------- */
int orc_strang_step(orc_map *m, const orc_term *t, size_t n, double dt, const orc_options *o,
                    orc_ledger *l) {
    if (!m || (!t && n) || !o || !l || !isfinite(dt)) return ORC_EINVAL;
    double half = dt * 0.5;
    if (half * 2 != dt) return ORC_ENUMERIC;
    orc_map work = {0};
    int s = orc_map_copy(&work, m);
    if (s) return s;
    orc_ledger trial = *l;
    for (size_t step = 0; step < 2 * n && !s; step++) {
        const orc_term *term = step < n ? t + step : t + (2 * n - 1 - step);
        s = orc_rotate_product(&work, term->i, term->j, term->qi, term->qj, term->h, half, o,
                               &trial);
    }
    if (s) {
        free(work.keys);
        free(work.values);
        return s;
    }
    free(m->keys);
    free(m->values);
    *m = work;
    *l = trial;
    return ORC_OK;
}

/* ---- Trotter error bounds: tpgp_rev.c:858-961 --------------------------------------------- */
static int terms_anticommute(const orc_term *a, const orc_term *b) {
    int parity = 0;
    if (a->i == b->i) parity ^= anticommutes(a->qi, b->qi);
    if (a->i == b->j) parity ^= anticommutes(a->qi, b->qj);
    if (a->j == b->i) parity ^= anticommutes(a->qj, b->qi);
    if (a->j == b->j) parity ^= anticommutes(a->qj, b->qj);
    return parity;
}
double orc_strang_bound_taylor(const orc_term *t, size_t n, double dt) {
    if ((!t && n) || !isfinite(dt)) return NAN;
    double lambda = 0;
    for (size_t k = 0; k < n; k++) {
        if (!isfinite(t[k].h) || t[k].i == t[k].j || t[k].qi > 3 || t[k].qj > 3) return NAN;
        lambda = orc_upadd(lambda, fabs(t[k].h));
    }
    double x = orc_upmul(lambda, fabs(dt));
    return orc_upmul(orc_upmul(orc_upmul(x, x), x), nextafter(2.0 / 3.0, INFINITY));
}
/* Nested-commutator bound on the anticommutation graph; adjacency lists are ascending. */
double orc_strang_bound(const orc_term *t, size_t n, double dt) {
    double general = orc_strang_bound_taylor(t, n, dt);
    if (isnan(general)) return NAN;
    size_t *deg = calloc(n ? n : 1, sizeof(size_t));
    size_t **adj = calloc(n ? n : 1, sizeof(size_t *));
    int ok = deg && adj;
    for (size_t a = 0; ok && a < n; a++)
        for (size_t b = a + 1; b < n; b++)
            if (t[a].h && t[b].h && terms_anticommute(t + a, t + b)) {
                size_t *pa = realloc(adj[a], (deg[a] + 1) * sizeof(size_t));
                if (pa) adj[a] = pa;
                size_t *pb = pa ? realloc(adj[b], (deg[b] + 1) * sizeof(size_t)) : NULL;
                if (pb) adj[b] = pb;
                if (!pa || !pb) {
                    ok = 0;
                    break;
                }
                adj[a][deg[a]++] = b;
                adj[b][deg[b]++] = a;
            }
    double C = 0;
    for (size_t a = 0; ok && a < n; a++) {
        double ha = fabs(t[a].h), inner = 0;
        for (size_t k = 0; k < deg[a]; k++)
            if (adj[a][k] > a) inner = orc_upadd(inner, fabs(t[adj[a][k]].h));
        C = orc_upadd(C, orc_upmul(orc_upmul(orc_upmul(ha, ha), inner),
                                   nextafter(1.0 / 6.0, INFINITY)));
        for (size_t k = 0; k < deg[a]; k++) {
            size_t b = adj[a][k];
            if (b <= a) continue;
            /* terms anticommuting with exactly one of a, b: symmetric difference */
            size_t x = 0, y = 0;
            double sum = 0;
            while (x < deg[a] || y < deg[b]) {
                size_t c;
                if (y == deg[b] || (x < deg[a] && adj[a][x] < adj[b][y]))
                    c = adj[a][x++];
                else if (x == deg[a] || adj[b][y] < adj[a][x])
                    c = adj[b][y++];
                else {
                    x++;
                    y++;
                    continue;
                }
                if (c > a) sum = orc_upadd(sum, fabs(t[c].h));
            }
            C = orc_upadd(C, orc_upmul(orc_upmul(orc_upmul(ha, fabs(t[b].h)), sum),
                                       nextafter(1.0 / 3.0, INFINITY)));
        }
    }
    for (size_t k = 0; adj && k < n; k++) free(adj[k]);
    free(adj);
    free(deg);
    if (!ok) return general;
    double T = fabs(dt), nested = orc_upmul(orc_upmul(orc_upmul(orc_upmul(2, C), T), T), T);
    return general < nested ? general : nested;
}

/* ---- FID allowance helpers: residual_rev.c:195-201 ---------------------------------------- */
double orc_bound_add_up(double a, double b) {
    if (isnan(a) || isnan(b) || a < 0 || b < 0) return NAN;
#ifdef TPGP_HAVE_MPFR
    if (isinf(a) || isinf(b)) return INFINITY;
    mpfr_t x, y;
    mpfr_inits2(256, x, y, (mpfr_ptr)0);
    mpfr_set_d(x, a, MPFR_RNDN);
    mpfr_set_d(y, b, MPFR_RNDN);
    mpfr_add(x, x, y, MPFR_RNDU);
    double z = mpfr_get_d(x, MPFR_RNDU);
    mpfr_clears(x, y, (mpfr_ptr)0);
    return z;
#else
    return orc_upadd(a, b);
#endif
}
double orc_bound_fid_up(double a, size_t n) {
#ifdef TPGP_HAVE_MPFR
    if (!n || a < 0 || isnan(a)) return NAN;
    if (isinf(a)) return INFINITY;
    mpfr_t x, y;
    mpfr_inits2(256, x, y, (mpfr_ptr)0);
    mpfr_set_d(x, a, MPFR_RNDN);
    mpfr_set_uj(y, (uintmax_t)n, MPFR_RNDN);
    mpfr_sqrt(y, y, MPFR_RNDD);
    mpfr_div(x, x, y, MPFR_RNDU);
    double z = mpfr_get_d(x, MPFR_RNDU);
    mpfr_clears(x, y, (mpfr_ptr)0);
    return z;
#else
    (void)a;
    (void)n;
    return NAN;
#endif
}

/* ---- Liouvillian -i[H, A]: tpgp_rev.c:810-857 ---------------------------------------------
 * Every term q (sites i, j) and key P that anticommute contribute v = (2 h x)(+-1) to the key
 * q xor P (sign + when the phase of qP is 1), x the coefficient of P.  Restated as: emit all
 * contributions in term-major, key-minor order, stable-sort them by key, and sum each key's run
 * left to right -- the same additions in the same order as the reference's sequential inserts
 * (a run whose sum is exactly zero leaves no key; an intermediate zero restarts from zero
 * exactly as a removal followed by a fresh insert does).  Terms are processed up to the first
 * invalid one; any non-finite contribution or partial sum among them gives ENUMERIC, else the
 * invalid term gives EINVAL -- the reference's status order. */
typedef struct {
    const uint64_t *key; /* into the contribution key array */
    size_t order;
} liou_ref;
static size_t liou_words;
static int liou_cmp(const void *a, const void *b) {
    const liou_ref *x = a, *y = b;
    int c = key_order(x->key, y->key, liou_words);
    if (c) return c;
    return (x->order > y->order) - (x->order < y->order);
}
int orc_liouvillian(orc_map *out, const orc_map *in, const orc_term *terms, size_t n) {
    if (!out || !in || (!terms && n) || !map_ok(in)) return ORC_EINVAL;
    const size_t W = in->n_words, N = in->n_spins;
    size_t valid = 0;
    while (valid < n) {
        const orc_term *q = terms + valid;
        if (q->i >= N || q->j >= N || q->i == q->j || q->qi > 3 || q->qj > 3 || !isfinite(q->h))
            break;
        valid++;
    }
    size_t count = 0;
    for (size_t t = 0; t < valid; t++)
        for (size_t k = 0; k < in->len; k++) {
            const uint64_t *p = in->keys + k * W;
            count += anticommutes(terms[t].qi, symbol_at(p, terms[t].i)) !=
                     anticommutes(terms[t].qj, symbol_at(p, terms[t].j));
        }
    uint64_t *ck = malloc((count ? count : 1) * W * sizeof(uint64_t));
    double *cv = malloc((count ? count : 1) * sizeof(double));
    liou_ref *ref = malloc((count ? count : 1) * sizeof(liou_ref));
    if (!ck || !cv || !ref) {
        free(ck);
        free(cv);
        free(ref);
        return ORC_ENOMEM;
    }
    int numeric_failure = 0;
    size_t at = 0;
    for (size_t t = 0; t < valid; t++) {
        const orc_term *q = terms + t;
        for (size_t k = 0; k < in->len; k++) {
            const uint64_t *p = in->keys + k * W;
            const unsigned a = symbol_at(p, q->i), b = symbol_at(p, q->j);
            if (anticommutes(q->qi, a) == anticommutes(q->qj, b)) continue;
            const unsigned phase = (orc_pauli_phase(q->qi, a) + orc_pauli_phase(q->qj, b)) & 3u;
            const double v = 2 * q->h * in->values[k] * (phase == 1 ? 1 : -1);
            if (!isfinite(v)) numeric_failure = 1;
            uint64_t *dst = ck + at * W;
            memcpy(dst, p, W * sizeof(uint64_t));
            dst[q->i / 32] ^= (uint64_t)q->qi << (2 * (q->i % 32));
            dst[q->j / 32] ^= (uint64_t)q->qj << (2 * (q->j % 32));
            cv[at] = v;
            ref[at].key = dst;
            ref[at].order = at;
            at++;
        }
    }
    liou_words = W;
    qsort(ref, count, sizeof(liou_ref), liou_cmp);
    orc_map *tmp = orc_map_new(N);
    int s = tmp ? ORC_OK : ORC_ENOMEM;
    for (size_t r = 0; !s && !numeric_failure && r < count;) {
        size_t e = r + 1;
        while (e < count && key_order(ref[e].key, ref[r].key, W) == 0) e++;
        double sum = 0;
        for (size_t x = r; x < e; x++) {
            sum = sum + cv[ref[x].order];
            if (!isfinite(sum)) numeric_failure = 1;
        }
        if (!numeric_failure && sum != 0) {
            s = grow(tmp, tmp->len + 1);
            if (!s) {
                memcpy(tmp->keys + tmp->len * W, ref[r].key, W * sizeof(uint64_t));
                tmp->values[tmp->len++] = sum;
            }
        }
        r = e;
    }
    free(ck);
    free(cv);
    free(ref);
    if (!s && numeric_failure) s = ORC_ENUMERIC;
    if (!s && valid < n) s = ORC_EINVAL;
    if (s) {
        orc_map_delete(tmp);
        return s;
    }
    free(out->keys);
    free(out->values);
    *out = *tmp;
    free(tmp);
    return ORC_OK;
}
