/* tpgp_engine: C17 host engine over the tpgp_hip device library (see include/tpgp_engine.h).
 *
 * Host responsibilities, per rotation of engine/src/tpgp_rev.c:535-771: argument validation in
 * the reference's order and precedence, the trigonometric factors (libm, or MPFR-256 enclosures
 * in validated mode), and the ledger update.  Everything that touches the terms -- branching,
 * merging equal keys, truncation, the per-rotation sums, the FID single-site scan -- runs on the
 * GPU through tpgp_hip.  There is no CPU implementation of the propagation in this library.
 *
 * Build: -std=c17 -fno-fast-math -ffp-contract=off, optionally -DTPGP_HAVE_MPFR.
 */
#include "../../include/tpgp_engine.h"
#include "../../include/tpgp_hip.h"

#include <fenv.h>
#include <pthread.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef TPGP_HAVE_MPFR
#include <mpfr.h>
#endif

/* ---- directed rounding (tpgp_rev.c:25-63) ------------------------------------------------- */
static double upadd(double a, double b) {
    if (a == 0) return b;
    if (b == 0) return a;
    return nextafter(a + b, INFINITY);
}
static double upmul(double a, double b) {
    if (a == 0 || b == 0) return 0;
    return nextafter(a * b, INFINITY);
}
static double upper_l2(const double *v, size_t n) {
#ifdef TPGP_HAVE_MPFR
    mpfr_t acc, t;
    mpfr_init2(acc, 192);
    mpfr_init2(t, 192);
    mpfr_set_zero(acc, 0);
    for (size_t k = 0; k < n; k++) {
        mpfr_set_d(t, v[k], MPFR_RNDN);
        mpfr_sqr(t, t, MPFR_RNDU);
        mpfr_add(acc, acc, t, MPFR_RNDU);
    }
    mpfr_sqrt(acc, acc, MPFR_RNDU);
    double out = mpfr_get_d(acc, MPFR_RNDU);
    mpfr_clear(acc);
    mpfr_clear(t);
    return out;
#else
    double acc = 0;
    for (size_t k = 0; k < n; k++) acc = upadd(acc, upmul(v[k], v[k]));
    return acc == 0 ? 0 : nextafter(sqrt(acc), INFINITY);
#endif
}
/* Upper bound of sqrt(S) from an upward-rounded S: one step above the rounded root. */
static double upper_sqrt(double s) { return s == 0 ? 0 : nextafter(sqrt(s), INFINITY); }

/* Message of the last device-side failure in this thread (empty when none). */
static _Thread_local char last_error[512];
const char *tpgp_hip_last_error(void) { return last_error; }
static void note(const char *where, const char *what) {
    snprintf(last_error, sizeof(last_error), "%s: %s", where, what);
}

/* ---- statuses, capabilities, options (tpgp_rev.c:117-148) -------------------------------- */
const char *tpgp_status_string(tpgp_status s) {
    static const char *names[] = {"ok",
                                  "invalid argument",
                                  "allocation failed",
                                  "active-term cap exceeded",
                                  "nonfinite arithmetic or unsupported floating-point environment",
                                  "backend unavailable",
                                  "coefficient-removal budget exceeded"};
    return s >= TPGP_OK && s <= TPGP_EBUDGET ? names[s] : "unknown status";
}
int tpgp_has_mpfr(void) {
#ifdef TPGP_HAVE_MPFR
    return 1;
#else
    return 0;
#endif
}
int tpgp_has_openmp(void) { return 0; }
tpgp_options tpgp_default_options(void) {
    tpgp_options o = {0, 0, 1, TPGP_HIP, 0, 0, NULL, NULL};
    return o;
}
void tpgp_ledger_init(tpgp_ledger *l, int validated) {
    memset(l, 0, sizeof(*l));
    l->arithmetic_enclosed = validated && tpgp_has_mpfr();
}

/* ---- host map container (tpgp_rev.c:64-116, 149-259) -------------------------------------- */
static int mulz(size_t a, size_t b, size_t *o) {
    if (b && a > SIZE_MAX / b) return 0;
    *o = a * b;
    return 1;
}
static int key_cmp(const uint64_t *a, const uint64_t *b, size_t W) {
    for (size_t w = W; w-- > 0;)
        if (a[w] != b[w]) return a[w] < b[w] ? -1 : 1;
    return 0;
}
static int key_padding_ok(const tpgp_map *m, const uint64_t *key) {
    size_t used = m->n_spins % 32;
    return !used || !(key[m->n_words - 1] >> (2 * used));
}
static int map_ok(const tpgp_map *m) {
    size_t nk, bytes;
    return m && m->n_spins && m->n_spins <= SIZE_MAX - 31 && m->n_words == (m->n_spins + 31) / 32 &&
           m->len <= m->capacity && mulz(m->capacity, m->n_words, &nk) &&
           mulz(nk, sizeof(uint64_t), &bytes) && m->capacity <= SIZE_MAX / sizeof(double) &&
           (!m->capacity || (m->keys && m->values));
}
static tpgp_status map_reserve(tpgp_map *m, size_t need) {
    if (need <= m->capacity) return TPGP_OK;
    size_t cap = m->capacity ? m->capacity : 8;
    while (cap < need) cap = cap > SIZE_MAX / 2 ? need : 2 * cap;
    size_t nk, bytes;
    if (!mulz(cap, m->n_words, &nk) || !mulz(nk, sizeof(uint64_t), &bytes) ||
        cap > SIZE_MAX / sizeof(double))
        return TPGP_ENOMEM;
    uint64_t *k = malloc(bytes);
    double *v = malloc(cap * sizeof(double));
    if (!k || !v) {
        free(k);
        free(v);
        return TPGP_ENOMEM;
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
    return TPGP_OK;
}
tpgp_status tpgp_map_init(tpgp_map *m, size_t n) {
    if (!m || !n || n > SIZE_MAX - 31) return TPGP_EINVAL;
    memset(m, 0, sizeof(*m));
    m->n_spins = n;
    m->n_words = (n + 31) / 32;
    return TPGP_OK;
}
void tpgp_map_free(tpgp_map *m) {
    if (!m) return;
    free(m->keys);
    free(m->values);
    memset(m, 0, sizeof(*m));
}
tpgp_status tpgp_map_clone(tpgp_map *dst, const tpgp_map *src) {
    if (!dst || !map_ok(src)) return TPGP_EINVAL;
    if (dst == src) return TPGP_OK;
    tpgp_map t;
    tpgp_status s = tpgp_map_init(&t, src->n_spins);
    if (s) return s;
    if ((s = map_reserve(&t, src->len))) {
        tpgp_map_free(&t);
        return s;
    }
    t.len = src->len;
    if (src->len) {
        memcpy(t.keys, src->keys, src->len * src->n_words * sizeof(uint64_t));
        memcpy(t.values, src->values, src->len * sizeof(double));
    }
    tpgp_map_free(dst);
    *dst = t;
    return TPGP_OK;
}
unsigned tpgp_key_get(const uint64_t *key, size_t site) {
    return (unsigned)(key[site / 32] >> (2 * (site % 32))) & 3u;
}
void tpgp_key_set(uint64_t *key, size_t site, unsigned p) {
    size_t w = site / 32, sh = 2 * (site % 32);
    key[w] = (key[w] & ~((uint64_t)3 << sh)) | ((uint64_t)(p & 3u) << sh);
}
static size_t lower_bound_key(const tpgp_map *m, const uint64_t *key, int *found) {
    size_t lo = 0, hi = m->len;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (key_cmp(m->keys + mid * m->n_words, key, m->n_words) < 0)
            lo = mid + 1;
        else
            hi = mid;
    }
    *found = lo < m->len && key_cmp(m->keys + lo * m->n_words, key, m->n_words) == 0;
    return lo;
}
tpgp_status tpgp_map_add(tpgp_map *m, const uint64_t *key, double v) {
    if (!map_ok(m) || !key || !isfinite(v) || !key_padding_ok(m, key)) return TPGP_EINVAL;
    if (v == 0) return TPGP_OK;
    const size_t W = m->n_words;
    int found;
    size_t at = lower_bound_key(m, key, &found);
    if (found) {
        double sum = m->values[at] + v;
        if (!isfinite(sum)) return TPGP_ENUMERIC;
        if (sum != 0) {
            m->values[at] = sum;
            return TPGP_OK;
        }
        memmove(m->keys + at * W, m->keys + (at + 1) * W, (m->len - at - 1) * W * 8);
        memmove(m->values + at, m->values + at + 1, (m->len - at - 1) * 8);
        m->len--;
        return TPGP_OK;
    }
    uint64_t saved[64];
    uint64_t *copy = W <= 64 ? saved : malloc(W * 8);
    if (!copy) return TPGP_ENOMEM;
    memcpy(copy, key, W * 8);
    tpgp_status s = map_reserve(m, m->len + 1);
    if (!s) {
        memmove(m->keys + (at + 1) * W, m->keys + at * W, (m->len - at) * W * 8);
        memmove(m->values + at + 1, m->values + at, (m->len - at) * 8);
        memcpy(m->keys + at * W, copy, W * 8);
        m->values[at] = v;
        m->len++;
    }
    if (copy != saved) free(copy);
    return s;
}
tpgp_status tpgp_map_add_single(tpgp_map *m, size_t site, unsigned p, double v) {
    if (!m || site >= m->n_spins || p > 3) return TPGP_EINVAL;
    uint64_t *key = calloc(m->n_words, sizeof(uint64_t));
    if (!key) return TPGP_ENOMEM;
    tpgp_key_set(key, site, p);
    tpgp_status s = tpgp_map_add(m, key, v);
    free(key);
    return s;
}
double tpgp_map_get(const tpgp_map *m, const uint64_t *key) {
    int found;
    size_t at = lower_bound_key(m, key, &found);
    return found ? m->values[at] : 0;
}
double tpgp_map_l2(const tpgp_map *m) {
    long double s = 0;
    for (size_t k = 0; k < m->len; k++) s += (long double)m->values[k] * m->values[k];
    return (double)sqrtl(s);
}
static void weight_and_symbol(const uint64_t *key, size_t W, unsigned *weight, unsigned *sym) {
    *weight = 0;
    *sym = 0;
    for (size_t w = 0; w < W; w++)
        for (uint64_t q = key[w]; q; q >>= 2)
            if (q & 3u) {
                ++*weight;
                *sym = (unsigned)(q & 3u);
            }
}
/* Compensated long-double sum of values in the given order (tpgp_rev.c:311-338). */
static double neumaier(const double *v, size_t n) {
    long double sum = 0, comp = 0;
    for (size_t k = 0; k < n; k++) {
        long double x = v[k], t = sum + x;
        comp += fabsl(sum) >= fabsl(x) ? (sum - t) + x : (x - t) + sum;
        sum = t;
    }
    return (double)(sum + comp);
}
/* Weight-1 values of symbol p in map (= site) order. */
static size_t single_values(const tpgp_map *m, unsigned p, double *out) {
    size_t n = 0;
    for (size_t k = 0; k < m->len; k++) {
        unsigned weight, sym;
        weight_and_symbol(m->keys + k * m->n_words, m->n_words, &weight, &sym);
        if (weight == 1 && sym == p) out[n++] = m->values[k];
    }
    return n;
}
double tpgp_map_single_sum(const tpgp_map *m, unsigned p) {
    if (p < 1 || p > 3) return NAN;
    double *v = malloc((m->n_spins + 1) * sizeof(double));
    if (!v) return NAN;
    double r = neumaier(v, single_values(m, p, v));
    free(v);
    return r;
}
/* MPFR-192 readout of values[0..n) / normalizer (tpgp_rev.c:339-388). */
static tpgp_status readout_values(const double *v, size_t n, size_t normalizer, double *midpoint,
                                  double *radius) {
#ifdef TPGP_HAVE_MPFR
    mpfr_t lo, hi, mid, x, den, r;
    mpfr_inits2(192, lo, hi, mid, x, den, r, (mpfr_ptr)0);
    mpfr_set_zero(lo, 0);
    mpfr_set_zero(hi, 0);
    mpfr_set_zero(mid, 0);
    mpfr_set_uj(den, (uintmax_t)normalizer, MPFR_RNDN);
    for (size_t k = 0; k < n; k++) {
        mpfr_set_d(x, v[k], MPFR_RNDN);
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
    if (!isfinite(centre) || !isfinite(rad)) return TPGP_ENUMERIC;
    *midpoint = centre;
    *radius = rad;
    return TPGP_OK;
#else
    (void)v;
    (void)n;
    (void)normalizer;
    (void)midpoint;
    (void)radius;
    return TPGP_EBACKEND;
#endif
}
tpgp_status tpgp_readout(const tpgp_map *m, unsigned p, size_t normalizer, double *midpoint,
                         double *radius) {
    if (!map_ok(m) || !normalizer || p < 1 || p > 3 || !midpoint || !radius) return TPGP_EINVAL;
    double *v = malloc((m->n_spins + 1) * sizeof(double));
    if (!v) return TPGP_ENOMEM;
    tpgp_status s = readout_values(v, single_values(m, p, v), normalizer, midpoint, radius);
    free(v);
    return s;
}
unsigned tpgp_pauli_phase(unsigned q, unsigned p) {
    if (q > 3 || p > 3) return 4;
    if (!q || !p || q == p) return 0;
    return ((p + 3 - q) % 3 == 1) ? 1u : 3u;
}
static int anticomm(unsigned q, unsigned p) { return q && p && q != p; }

/* ---- trigonometric factors (tpgp_rev.c:484-534) ------------------------------------------- */
static tpgp_status trig(double h, double dur, int validated, double *c, double *s, double *rc,
                        double *rs) {
    *rc = *rs = 0;
#ifdef TPGP_HAVE_MPFR
    if (validated) {
        mpfr_t x, lo, hi, mid, a, b;
        mpfr_inits2(256, x, lo, hi, mid, a, b, (mpfr_ptr)0);
        mpfr_set_d(x, h, MPFR_RNDN);
        mpfr_set_d(a, dur, MPFR_RNDN);
        mpfr_mul(x, x, a, MPFR_RNDN);
        mpfr_mul_2ui(x, x, 1, MPFR_RNDN);
        for (int k = 0; k < 2; k++) {
            int (*f)(mpfr_ptr, mpfr_srcptr, mpfr_rnd_t) = k ? mpfr_sin : mpfr_cos;
            f(lo, x, MPFR_RNDD);
            f(hi, x, MPFR_RNDU);
            f(mid, x, MPFR_RNDN);
            double centre = mpfr_get_d(mid, MPFR_RNDN);
            mpfr_set_d(mid, centre, MPFR_RNDN);
            mpfr_sub(a, mid, lo, MPFR_RNDU);
            mpfr_sub(b, hi, mid, MPFR_RNDU);
            mpfr_max(a, a, b, MPFR_RNDU);
            double rad = mpfr_get_d(a, MPFR_RNDU);
            if (k) {
                *s = centre;
                *rs = rad;
            } else {
                *c = centre;
                *rc = rad;
            }
        }
        mpfr_clears(x, lo, hi, mid, a, b, (mpfr_ptr)0);
        return TPGP_OK;
    }
#else
    if (validated) return TPGP_EBACKEND;
#endif
    double x = 2 * h * dur;
    if (!isfinite(x)) return TPGP_ENUMERIC;
    *c = cos(x);
    *s = sin(x);
    return TPGP_OK;
}

/* ---- rotation plans and the ledger fold ---------------------------------------------------- */
typedef struct {
    tpgp_status pre;   /* host-side status of this rotation (validation, trig) */
    int noop;          /* h == 0 or duration == 0: nothing enqueued, nothing counted */
    size_t i, j;
    unsigned qi, qj;
    double c, s, rc, rs;
} rot_plan;

static int ledger_ok(const tpgp_ledger *l) {
    return isfinite(l->dropped_l1) && l->dropped_l1 >= 0 && isfinite(l->dropped_l2) &&
           l->dropped_l2 >= 0 && isfinite(l->dropped_l2_squared) && l->dropped_l2_squared >= 0 &&
           isfinite(l->arithmetic_l2) && l->arithmetic_l2 >= 0;
}
/* Host checks of rotate_product in the reference's order (tpgp_rev.c:538-561), excluding the
 * ledger (checked during the fold, since it evolves) and the map-length no-op. */
static void plan_rotation(rot_plan *p, size_t n_spins, size_t i, size_t j, unsigned qi,
                          unsigned qj, double h, double dur, const tpgp_options *o) {
    memset(p, 0, sizeof(*p));
    p->i = i;
    p->j = j;
    p->qi = qi;
    p->qj = qj;
    if (!o || i >= n_spins || j >= n_spins || i == j || qi > 3 || qj > 3 || !isfinite(h) ||
        !isfinite(dur) || !isfinite(o->threshold) || o->threshold < 0 ||
        !isfinite(o->max_drop_l2) || o->max_drop_l2 < 0 || o->threads == 0) {
        p->pre = TPGP_EINVAL;
        return;
    }
    if (o->validated) {
        volatile double tiny = DBL_MIN, half = .5;
        if (FLT_RADIX != 2 || DBL_MANT_DIG != 53 || fegetround() != FE_TONEAREST ||
            tiny * half == 0) {
            p->pre = TPGP_ENUMERIC;
            return;
        }
    }
    if (o->backend != TPGP_HIP) {
        p->pre = TPGP_EBACKEND;
        return;
    }
    p->pre = trig(h, dur, o->validated, &p->c, &p->s, &p->rc, &p->rs);
    if (p->pre) return;
    p->noop = h == 0 || dur == 0;
}

enum { FOLD_CAPACITY = 100 };

static int cmp_entry(const void *a, const void *b) {
    const tpgp_hip_log_entry *x = a, *y = b;
    if (x->rotation != y->rotation) return x->rotation < y->rotation ? -1 : 1;
    if (x->kind != y->kind) return x->kind < y->kind ? -1 : 1;
    if (x->key[1] != y->key[1]) return x->key[1] < y->key[1] ? -1 : 1;
    if (x->key[0] != y->key[0]) return x->key[0] < y->key[0] ? -1 : 1;
    return 0;
}

/* Apply one executed rotation's record to the trial ledger (tpgp_rev.c:730-753).  entries are
 * this rotation's logged drops (kind 0) and radii (kind 1) in ascending reference key order. */
static int fold_record(const tpgp_hip_record *r, const tpgp_hip_log_entry *ent, size_t n_ent,
                       int exact, const tpgp_options *o, tpgp_ledger *trial, double *scratch) {
    if (r->overflow) return FOLD_CAPACITY;
    if (r->nonfinite) return TPGP_ENUMERIC;
    if (o->max_terms && r->len_out > o->max_terms) return TPGP_ECAP;
    double l1, sq, l2, ar;
    if (exact) {
        size_t nd = 0, nr = 0;
        l1 = sq = 0;
        for (size_t k = 0; k < n_ent; k++)
            if (ent[k].kind == 0) {
                l1 = upadd(l1, ent[k].value);
                sq = upadd(sq, upmul(ent[k].aux, ent[k].aux));
                scratch[nd++] = ent[k].value;
            }
        l2 = upper_l2(scratch, nd);
        for (size_t k = 0; k < n_ent; k++)
            if (ent[k].kind == 1) scratch[nr++] = ent[k].value;
        ar = upper_l2(scratch, nr);
        if (nd != r->dropped || nr != r->radii) return TPGP_EBACKEND;
    } else {
        l1 = r->l1_up;
        sq = r->sq_up;
        l2 = upper_sqrt(r->sq_up);
        ar = upper_sqrt(r->rad_sq_up);
    }
    tpgp_ledger next = *trial;
    next.dropped_l1 = upadd(next.dropped_l1, l1);
    next.dropped_l2 = upadd(next.dropped_l2, l2);
    next.dropped_l2_squared = upadd(next.dropped_l2_squared, sq);
    next.arithmetic_l2 = upadd(next.arithmetic_l2, ar);
    next.rotations++;
    if (r->dropped) next.removal_events++;
    if (r->len_out > next.peak_terms) next.peak_terms = (size_t)r->len_out;
    next.arithmetic_enclosed = next.arithmetic_enclosed && o->validated && tpgp_has_mpfr();
    if (!isfinite(next.dropped_l2) || !isfinite(next.arithmetic_l2)) return TPGP_ENUMERIC;
    if (o->max_drop_l2 && next.dropped_l2 > o->max_drop_l2) return TPGP_EBUDGET;
    *trial = next;
    return TPGP_OK;
}

/* ---- a resident single-map session used by every entry point ----------------------------- */
typedef struct {
    tpgp_hip_context *ctx;
    int own_ctx;
    tpgp_hip_batch *batch;
    size_t cap;
    tpgp_hip_record *rec;
    size_t max_rec;
    tpgp_hip_log_entry *log;
    size_t log_cap;
    double *scratch;
    size_t scratch_cap;
    int no_room; /* a growth for appending rotations was refused: not asked for again */
} session;

static void session_close(session *S) {
    tpgp_hip_batch_destroy(S->batch);
    if (S->own_ctx) tpgp_hip_destroy(S->ctx);
    free(S->rec);
    free(S->log);
    free(S->scratch);
    memset(S, 0, sizeof(*S));
}
static tpgp_status map_status(int hs) {
    return hs == TPGP_HIP_ENOMEM ? TPGP_ENOMEM : hs == TPGP_HIP_EINVAL ? TPGP_EINVAL
                                                                         : TPGP_EBACKEND;
}
static tpgp_status session_open(session *S, const tpgp_map *m, const tpgp_options *o,
                                size_t cap_hint, size_t max_rec) {
    memset(S, 0, sizeof(*S));
    char err[512];
    S->ctx = o->hip_context;
    if (!S->ctx) {
        S->ctx = tpgp_hip_create(0, err, sizeof(err));
        if (!S->ctx) {
            note("context", err);
            return TPGP_EBACKEND;
        }
        S->own_ctx = 1;
    }
    S->cap = cap_hint;
    size_t n = m->n_spins;
    S->batch = tpgp_hip_batch_create(S->ctx, m->n_words, 1, &n, &S->cap, err, sizeof(err));
    if (!S->batch) {
        note("batch", err);
        session_close(S);
        return TPGP_ENOMEM;
    }
    S->max_rec = max_rec;
    S->rec = calloc(max_rec ? max_rec : 1, sizeof(*S->rec));
    if (!S->rec) {
        session_close(S);
        return TPGP_ENOMEM;
    }
    int hs = tpgp_hip_batch_upload(S->batch, 0, m->keys, m->values, m->len, err, sizeof(err));
    if (hs) {
        note("upload", err);
        session_close(S);
        return map_status(hs);
    }
    return TPGP_OK;
}
static tpgp_status session_grow(session *S, size_t need) {
    char err[512];
    size_t cap = S->cap * 2 > need ? S->cap * 2 : need;
    int hs = tpgp_hip_batch_reserve(S->batch, 0, cap, err, sizeof(err));
    if (hs) {
        note("reserve", err);
        return map_status(hs);
    }
    S->cap = cap;
    return TPGP_OK;
}
/* Room for appending rotations (best effort): the library places the classes a rotation
 * rewrites after the slot's used extent and leaves the others in place, which needs free space
 * of about half the map per rotation between partitions.  A slot below three times the map
 * grows to four times it; if the device refuses, the session keeps its slot (rotations then copy
 * every class) and does not ask again. */
static void session_room(session *S, size_t len) {
    char err[512];
    if (S->no_room || len > SIZE_MAX / 5 || 3 * len + 4096 <= S->cap) return;
    const size_t want = 4 * len + 4096;
    if (tpgp_hip_batch_reserve(S->batch, 0, want, err, sizeof(err)) == TPGP_HIP_OK)
        S->cap = want;
    else
        S->no_room = 1;
}
/* Logs of dropped values and radii: device and host buffers hold `need` entries and more. */
static tpgp_status grow_logs(tpgp_hip_batch *batch, tpgp_hip_log_entry **log, size_t *log_cap,
                             size_t need) {
    char err[512];
    size_t cap = need + need / 2 + 4096;
    int hs = tpgp_hip_batch_reserve_log(batch, cap, err, sizeof(err));
    if (hs) {
        note("log reserve", err);
        return map_status(hs);
    }
    tpgp_hip_log_entry *fresh = malloc(cap * sizeof(*fresh));
    if (!fresh) return TPGP_ENOMEM;
    free(*log);
    *log = fresh;
    *log_cap = cap;
    return TPGP_OK;
}
/* A step that reported capacity_needed: 1 = the map outgrew its slot, 2 = logs overflowed. */
static tpgp_status session_regrow(session *S, int kind, size_t need) {
    if (kind == 2) return grow_logs(S->batch, &S->log, &S->log_cap, need);
    return session_grow(S, need + need / 2 + 1024);
}
static tpgp_status session_download(session *S, tpgp_map *m) {
    char err[512];
    size_t len = 0;
    int hs = tpgp_hip_batch_download(S->batch, 0, NULL, NULL, 0, &len, err, sizeof(err));
    if (hs) return map_status(hs);
    tpgp_map t;
    tpgp_status s = tpgp_map_init(&t, m->n_spins);
    if (s) return s;
    if ((s = map_reserve(&t, len))) {
        tpgp_map_free(&t);
        return s;
    }
    hs = tpgp_hip_batch_download(S->batch, 0, t.keys, t.values, len, &t.len, err, sizeof(err));
    if (hs) {
        note("download", err);
        tpgp_map_free(&t);
        return map_status(hs);
    }
    tpgp_map_free(m);
    *m = t;
    return TPGP_OK;
}

/* Enqueue planned rotations [0, n) (the valid prefix, skipping no-ops); returns enqueued count
 * in *enq. */
static tpgp_status enqueue_plans(session *S, const rot_plan *plans, size_t n, int exact,
                                 const tpgp_options *o, uint32_t seq0, size_t *enq) {
    char err[512];
    *enq = 0;
    tpgp_hip_rotation *req = malloc((n + 1) * sizeof(*req));
    uint32_t *seq = malloc((n + 1) * sizeof(*seq));
    if (!req || !seq) {
        free(req);
        free(seq);
        return TPGP_ENOMEM;
    }
    size_t cnt = 0;
    for (size_t k = 0; k < n; k++) {
        if (plans[k].pre) break;
        if (plans[k].noop) continue;
        tpgp_hip_rotation *r = req + cnt;
        memset(r, 0, sizeof(*r));
        r->active = 1;
        r->i = plans[k].i;
        r->j = plans[k].j;
        r->qi = plans[k].qi;
        r->qj = plans[k].qj;
        r->c = plans[k].c;
        r->s = plans[k].s;
        r->rc = plans[k].rc;
        r->rs = plans[k].rs;
        r->threshold = o->threshold;
        r->validated = o->validated;
        r->log = exact;
        seq[cnt++] = seq0 + (uint32_t)k;
    }
    int hs = cnt ? tpgp_hip_batch_rotate_many(S->batch, req, seq, cnt, err, sizeof(err)) : 0;
    free(req);
    free(seq);
    if (hs) {
        note("rotate", err);
        return map_status(hs);
    }
    *enq = cnt;
    return TPGP_OK;
}

/* Wait for the enqueued rotations (and an optional readout), then fold their records into
 * *trial in rotation order.  The reference stops at its first failing rotation; so does this. */
static tpgp_status sync_fold(session *S, const rot_plan *plans, size_t n, size_t enq, int exact,
                             const tpgp_options *o, tpgp_ledger *trial, double *readout,
                             uint64_t *len, int *capacity_needed, size_t *need,
                             uint64_t *updates, uint64_t *in_out, uint64_t *in_place,
                             uint64_t *executed) {
    char err[512];
    *capacity_needed = 0;
    size_t nrec = 0, nlog = 0, nrd = 0;
    if (exact && !S->log) {
        S->log_cap = 1u << 20;
        S->log = malloc(S->log_cap * sizeof(*S->log));
        if (!S->log) return TPGP_ENOMEM;
    }
    int hs = tpgp_hip_batch_sync(S->batch, S->rec, S->max_rec, &nrec, readout, readout ? 2 : 0,
                                 &nrd, len, exact ? S->log : NULL, exact ? S->log_cap : 0, &nlog,
                                 err, sizeof(err));
    if (hs == TPGP_HIP_ECAPACITY && exact) { /* logs too small: grow them and run again */
        *capacity_needed = 2;
        *need = nlog;
        return TPGP_OK;
    }
    if (hs) {
        note("sync", err);
        return map_status(hs);
    }
    if (nrec != enq) {
        note("sync", "record count mismatch");
        return TPGP_EBACKEND;
    }
    if (exact && nlog) qsort(S->log, nlog, sizeof(*S->log), cmp_entry);
    if (S->scratch_cap < nlog + 1) {
        free(S->scratch);
        S->scratch_cap = nlog + 1;
        S->scratch = malloc(S->scratch_cap * sizeof(double));
        if (!S->scratch) return TPGP_ENOMEM;
    }
    size_t r_at = 0, e_at = 0;
    uint64_t up = 0, io = 0, kp = 0, ex = 0;
    for (size_t k = 0; k < n; k++) {
        const rot_plan *p = plans + k;
        if (!ledger_ok(trial)) return TPGP_EINVAL; /* tpgp_rev.c:543-547 */
        if (p->pre) return p->pre;
        if (p->noop) continue;
        const tpgp_hip_record *r = S->rec + r_at++;
        size_t e0 = e_at;
        while (e_at < nlog && S->log[e_at].rotation == r->rotation) e_at++;
        if (!r->executed) continue; /* empty map: no-op (tpgp_rev.c:562-563) */
        int fs = fold_record(r, S->log + e0, e_at - e0, exact, o, trial, S->scratch);
        if (fs == FOLD_CAPACITY) {
            *capacity_needed = 1;
            *need = (size_t)r->len_out;
            return TPGP_OK;
        }
        if (fs) {
            if (fs == TPGP_EBACKEND) note("ledger", "logged entries disagree with the record");
            return (tpgp_status)fs;
        }
        up += r->len_in;
        io += r->len_in + r->len_out;
        kp += 2 * r->in_place;
        ex++;
    }
    if (updates) *updates += up;
    if (in_out) *in_out += io;
    if (in_place) *in_place += kp;
    if (executed) *executed += ex;
    return TPGP_OK;
}

/* Run planned rotations [0, n) on the session's working state; fold them into *trial. */
static tpgp_status run_plans(session *S, const rot_plan *plans, size_t n, int exact,
                             const tpgp_options *o, tpgp_ledger *trial, uint32_t seq0,
                             int *capacity_needed, size_t *need) {
    size_t enq = 0;
    tpgp_status s = enqueue_plans(S, plans, n, exact, o, seq0, &enq);
    if (s) return s;
    return sync_fold(S, plans, n, enq, exact, o, trial, NULL, NULL, capacity_needed, need, NULL,
                     NULL, NULL, NULL);
}

/* Transactional sequence of rotations on a host map (tpgp_rotate / tpgp_strang_step). */
static tpgp_status rotate_sequence(tpgp_map *m, const rot_plan *plans, size_t n,
                                   const tpgp_options *o, tpgp_ledger *l) {
    if (!map_ok(m) || !o || !l) return TPGP_EINVAL;
    int any = 0;
    for (size_t k = 0; k < n && !plans[k].pre; k++) any |= !plans[k].noop;
    if (!any || m->len == 0) {
        /* nothing reaches the device (every rotation is a no-op until the first host-side
         * failure): replay the per-rotation checks in order */
        for (size_t k = 0; k < n; k++) {
            if (!ledger_ok(l)) return TPGP_EINVAL;
            if (plans[k].pre) return plans[k].pre;
        }
        return TPGP_OK;
    }
    for (size_t k = 0; k < n; k++)
        if (!plans[k].pre && !plans[k].noop && m->len > SIZE_MAX / 2) return TPGP_ENOMEM;
    session S;
    size_t cap = m->len * 2 + 1024;
    tpgp_status s = session_open(&S, m, o, cap, n + 1);
    if (s) return s;
    for (;;) {
        tpgp_ledger trial = *l;
        int grow = 0;
        size_t need = 0;
        s = run_plans(&S, plans, n, 1, o, &trial, 0, &grow, &need);
        char err[512];
        if (!s && grow) {
            tpgp_hip_batch_rollback(S.batch, err, sizeof(err));
            if ((s = session_regrow(&S, grow, need))) break;
            continue;
        }
        if (s) {
            tpgp_hip_batch_rollback(S.batch, err, sizeof(err));
            break;
        }
        if (tpgp_hip_batch_commit(S.batch, NULL, err, sizeof(err))) {
            s = TPGP_EBACKEND;
            break;
        }
        if ((s = session_download(&S, m))) break;
        *l = trial;
        break;
    }
    session_close(&S);
    return s;
}

tpgp_status tpgp_rotate(tpgp_map *m, size_t i, size_t j, unsigned qi, unsigned qj, double theta,
                        const tpgp_options *o, tpgp_ledger *l) {
    if (!map_ok(m) || !o || !l) return TPGP_EINVAL;
    rot_plan p;
    plan_rotation(&p, m->n_spins, i, j, qi, qj, theta, 1.0, o);
    return rotate_sequence(m, &p, 1, o, l);
}

tpgp_status tpgp_strang_step(tpgp_map *m, const tpgp_term *t, size_t n, double dt,
                             const tpgp_options *o, tpgp_ledger *l) {
    if (!m || (!t && n) || !o || !l || !isfinite(dt)) return TPGP_EINVAL;
    double half = dt * 0.5;
    if (half * 2 != dt) return TPGP_ENUMERIC; /* tpgp_rev.c:780-784 */
    if (!map_ok(m)) return TPGP_EINVAL;
    rot_plan *plans = malloc((2 * n + 1) * sizeof(*plans));
    if (!plans) return TPGP_ENOMEM;
    for (size_t k = 0; k < 2 * n; k++) {
        const tpgp_term *term = k < n ? t + k : t + (2 * n - 1 - k);
        plan_rotation(plans + k, m->n_spins, term->i, term->j, term->qi, term->qj, term->h, half,
                      o);
    }
    tpgp_status s = rotate_sequence(m, plans, 2 * n, o, l);
    free(plans);
    return s;
}

/* ---- resident evolution ------------------------------------------------------------------- */
struct tpgp_evolution {
    session S;
    rot_plan *plans;
    size_t nplans;
    tpgp_options o;
    tpgp_evolve_options eo;
    size_t n_spins;
    tpgp_evolution_counters c;
    size_t graph_enq; /* rotations a recorded step enqueues */
    uint64_t graphs_recorded, graphs_replayed;
};

tpgp_status tpgp_evolution_open(tpgp_evolution **out, const tpgp_map *m, const tpgp_term *t,
                                size_t n, double dt, const tpgp_options *o,
                                const tpgp_evolve_options *eo) {
    if (!out) return TPGP_EINVAL;
    *out = NULL;
    if (!map_ok(m) || (!t && n) || !o || !eo || !isfinite(dt) || eo->pauli < 1 ||
        eo->pauli > 3 || !eo->normalizer || m->n_spins > 64)
        return TPGP_EINVAL;
    if (o->backend != TPGP_HIP) return TPGP_EBACKEND;
    double half = dt * 0.5;
    if (half * 2 != dt) return TPGP_ENUMERIC; /* tpgp_rev.c:780-784 */
    tpgp_evolution *ev = calloc(1, sizeof(*ev));
    if (!ev) return TPGP_ENOMEM;
    ev->nplans = 2 * n;
    ev->plans = malloc((2 * n + 1) * sizeof(*ev->plans));
    if (!ev->plans) {
        free(ev);
        return TPGP_ENOMEM;
    }
    for (size_t k = 0; k < 2 * n; k++) {
        const tpgp_term *term = k < n ? t + k : t + (2 * n - 1 - k);
        plan_rotation(ev->plans + k, m->n_spins, term->i, term->j, term->qi, term->qj, term->h,
                      half, o);
    }
    ev->o = *o;
    ev->eo = *eo;
    ev->n_spins = m->n_spins;
    size_t cap = eo->initial_capacity ? eo->initial_capacity : 4 * m->len + (1u << 16);
    tpgp_status s = session_open(&ev->S, m, o, cap, 2 * n + 2);
    if (s) {
        free(ev->plans);
        free(ev);
        return s;
    }
    *out = ev;
    return TPGP_OK;
}

/* `steps` Strang steps with a readout before each and after the last: one host sync per step
 * (the readout of the committed state rides with the step's rotations). */
/* Readout of the committed state and, with `more`, the step's rotations.  A step that starts
 * from a state it has started from before is recorded as a HIP graph, and later steps from
 * that state replay it (tpgp_hip_batch_graph_*): the same work in one launch. */
static tpgp_status enqueue_step(tpgp_evolution *ev, int more, size_t *enq) {
    session *S = &ev->S;
    char err[512];
    int replayed = 0, capture = 0, captured = 0;
    *enq = 0;
    if (more) {
        int hs = tpgp_hip_batch_graph_replay(S->batch, 1, &replayed, &capture, err, sizeof(err));
        if (hs) {
            note("graph replay", err);
            return map_status(hs);
        }
        if (replayed) {
            *enq = ev->graph_enq;
            ev->graphs_replayed++;
            return TPGP_OK;
        }
    }
    int recording = capture && !tpgp_hip_batch_graph_begin(S->batch, 1, err, sizeof(err));
    for (;;) {
        tpgp_status s = TPGP_OK;
        int hs = tpgp_hip_batch_readout(S->batch, ev->eo.pauli, NULL, err, sizeof(err));
        if (hs) {
            note("readout", err);
            s = map_status(hs);
        } else if (more) {
            s = enqueue_plans(S, ev->plans, ev->nplans, ev->eo.exact_ledger, &ev->o, 0, enq);
        }
        if (!recording) return s;
        recording = 0;
        hs = tpgp_hip_batch_graph_end(S->batch, !s, &captured, err, sizeof(err));
        if (s) return s;
        if (hs) {
            note("graph", err);
            return map_status(hs);
        }
        if (captured) {
            ev->graph_enq = *enq;
            ev->graphs_recorded++;
            return TPGP_OK;
        }
        *enq = 0; /* not recorded: nothing ran, enqueue the step directly */
    }
}

tpgp_status tpgp_evolution_run(tpgp_evolution *ev, size_t steps, tpgp_ledger *l, double *fid,
                               double *fid_radius, size_t *active, size_t *accepted) {
    if (accepted) *accepted = 0;
    if (!ev || !l) return TPGP_EINVAL;
    session *S = &ev->S;
    const tpgp_options *o = &ev->o;
    const int exact = ev->eo.exact_ledger;
    char err[512];
    double rd[2 * 64];
    tpgp_status s = TPGP_OK;
    for (size_t step = 0;; step++) {
        const int more = step < steps;
        tpgp_ledger trial = *l;
        int grow = 0;
        size_t need = 0;
        uint64_t len = 0, up = 0, io = 0, kp = 0, ex = 0;
        size_t read_len = 0; /* length of the state read out (the committed state) */
        tpgp_hip_batch_info(S->batch, 0, &read_len, NULL);
        if (more && read_len + read_len / 2 + 4096 > S->cap &&
            (s = session_grow(S, 2 * read_len + 4096))) /* room ahead of need */
            return s;
        if (more) session_room(S, read_len);
        for (;;) {
            for (size_t k = 0; k < 2 * 64; k++) rd[k] = NAN; /* filled only by a good sync */
            size_t enq = 0;
            if ((s = enqueue_step(ev, more, &enq))) break;
            trial = *l;
            up = io = kp = ex = 0;
            s = sync_fold(S, ev->plans, more ? ev->nplans : 0, enq, exact, o, &trial, rd, &len,
                          &grow, &need, &up, &io, &kp, &ex);
            if (!s && grow) {
                tpgp_hip_batch_rollback(S->batch, err, sizeof(err));
                if ((s = session_regrow(S, grow, need))) break;
                continue;
            }
            break;
        }
        /* readout of the state before this step (taken before the rotations ran; it is a
         * valid sample even when the step itself fails, as in the reference's step loop) */
        if (!isnan(rd[0])) {
            double mid = 0, rad = NAN;
            tpgp_status rs = TPGP_OK;
            if (tpgp_has_mpfr())
                rs = readout_values(rd, ev->n_spins, ev->eo.normalizer, &mid, &rad);
            else
                mid = neumaier(rd, ev->n_spins) / (double)ev->eo.normalizer;
            if (rs && !s) s = rs;
            if (fid) fid[step] = mid;
            if (fid_radius) fid_radius[step] = rad;
            if (active) active[step] = read_len;
        }
        if (s) {
            tpgp_hip_batch_rollback(S->batch, err, sizeof(err));
            return s;
        }
        if (!more) {
            /* the final readout also committed nothing new: length is the committed length */
            return TPGP_OK;
        }
        if (tpgp_hip_batch_commit(S->batch, NULL, err, sizeof(err))) {
            note("commit", err);
            return TPGP_EBACKEND;
        }
        *l = trial;
        ev->c.steps++;
        ev->c.rotations += ex;
        ev->c.updates += up;
        ev->c.in_out += io;
        ev->c.in_place += kp;
        if (accepted) *accepted = step + 1;
    }
}

void tpgp_evolution_graph_counts(const tpgp_evolution *ev, uint64_t *recorded,
                                 uint64_t *replayed) {
    if (recorded) *recorded = ev ? ev->graphs_recorded : 0;
    if (replayed) *replayed = ev ? ev->graphs_replayed : 0;
}

tpgp_status tpgp_evolution_close(tpgp_evolution *ev, tpgp_map *m) {
    if (!ev) return TPGP_EINVAL;
    tpgp_status s = TPGP_OK;
    if (m) {
        if (m->n_spins != ev->n_spins) s = TPGP_EINVAL;
        else s = session_download(&ev->S, m);
    }
    session_close(&ev->S);
    free(ev->plans);
    free(ev);
    return s;
}

/* The device context of calls without options (tpgp_liouvillian): created on first use on
 * device 0 and kept for the life of the process. */
static pthread_once_t default_once = PTHREAD_ONCE_INIT;
static tpgp_hip_context *default_ctx;
static char default_err[512];
static void default_create(void) {
    default_ctx = tpgp_hip_create(0, default_err, sizeof(default_err));
}
/* For the other units of this library (the residual bound): the same context, or NULL with the
 * reason noted; and the error note. */
__attribute__((visibility("hidden"))) tpgp_hip_context *tpgp_engine_default_context(void) {
    pthread_once(&default_once, default_create);
    if (!default_ctx) note("context", default_err);
    return default_ctx;
}
__attribute__((visibility("hidden"))) void tpgp_engine_note(const char *where, const char *what) {
    note(where, what);
}

tpgp_status tpgp_liouvillian(tpgp_map *out, const tpgp_map *in, const tpgp_term *terms,
                             size_t n) {
    if (!out || !in || (!terms && n) || !in->n_spins || in->n_spins > 64 ||
        in->n_words != (in->n_spins + 31) / 32 || (in->len && (!in->keys || !in->values)))
        return TPGP_EINVAL;
    pthread_once(&default_once, default_create);
    if (!default_ctx) {
        note("context", default_err);
        return TPGP_EBACKEND;
    }
    tpgp_map t;
    tpgp_status s = tpgp_map_init(&t, in->n_spins);
    if (s) return s;
    /* tpgp_hip_term has tpgp_term's layout */
    const tpgp_hip_term *ht = (const tpgp_hip_term *)terms;
    size_t cap = 4 * in->len + 1024, got = 0;
    char err[512];
    for (int attempt = 0; attempt < 2; attempt++) {
        if ((s = map_reserve(&t, cap))) {
            tpgp_map_free(&t);
            return s;
        }
        const int hs = tpgp_hip_liouvillian(default_ctx, in->keys, in->values, in->len,
                                            in->n_words, in->n_spins, ht, n, t.keys, t.values,
                                            t.capacity, &got, err, sizeof(err));
        if (hs == TPGP_HIP_ECAPACITY && attempt == 0) {
            cap = got;
            continue;
        }
        if (hs) {
            note("liouvillian", err);
            tpgp_map_free(&t);
            return hs == TPGP_HIP_ENUMERIC ? TPGP_ENUMERIC : map_status(hs);
        }
        t.len = got;
        tpgp_map_free(out);
        *out = t;
        return TPGP_OK;
    }
    tpgp_map_free(&t);
    return TPGP_EBACKEND;
}

tpgp_status tpgp_evolution_snapshot(tpgp_evolution *ev, tpgp_map *m) {
    if (!ev || !m || m->n_spins != ev->n_spins) return TPGP_EINVAL;
    return session_download(&ev->S, m);
}

tpgp_status tpgp_evolution_l2(tpgp_evolution *ev, double *l2) {
    if (!ev || !l2) return TPGP_EINVAL;
    char err[512];
    double pair[2];
    const int hs = tpgp_hip_batch_sum_squares(ev->S.batch, 0, pair, err, sizeof(err));
    if (hs) {
        note("sum of squares", err);
        return map_status(hs);
    }
    *l2 = (double)sqrtl((long double)pair[0] + (long double)pair[1]);
    return TPGP_OK;
}

void tpgp_evolution_counters_get(const tpgp_evolution *ev, tpgp_evolution_counters *c) {
    if (ev && c) *c = ev->c;
}
int tpgp_evolution_profile(tpgp_evolution *ev, unsigned mask) {
    return ev ? tpgp_hip_batch_profile(ev->S.batch, mask) : TPGP_HIP_EINVAL;
}
int tpgp_evolution_profile_read(const tpgp_evolution *ev, unsigned kind, double *ms,
                                uint64_t *launches) {
    return ev ? tpgp_hip_batch_profile_read(ev->S.batch, kind, ms, launches) : TPGP_HIP_EINVAL;
}
size_t tpgp_evolution_key_words(const tpgp_evolution *ev) {
    return ev ? (ev->n_spins + 31) / 32 : 0;
}

tpgp_status tpgp_hip_evolve(tpgp_map *m, const tpgp_term *t, size_t n, double dt, size_t steps,
                            const tpgp_options *o, const tpgp_evolve_options *eo, tpgp_ledger *l,
                            double *fid, double *fid_radius, size_t *active, size_t *accepted) {
    if (accepted) *accepted = 0;
    if (!l) return TPGP_EINVAL;
    tpgp_evolution *ev = NULL;
    tpgp_status s = tpgp_evolution_open(&ev, m, t, n, dt, o, eo);
    if (s) return s;
    s = tpgp_evolution_run(ev, steps, l, fid, fid_radius, active, accepted);
    tpgp_status cs = tpgp_evolution_close(ev, m);
    return s ? s : cs;
}

/* ---- fleet: many evolutions in one resident batch ----------------------------------------- */
typedef struct {
    tpgp_slot_state state;
    tpgp_status status;
    size_t n_spins, steps, done;
    rot_plan *plans;
    size_t nplans;
    size_t *enq;      /* indices of the plans enqueued in the current step */
    size_t nenq;
    tpgp_options o;
    tpgp_evolve_options eo;
    tpgp_ledger ledger;
    double *fid, *rad;
    size_t *act;
    size_t cap;
    size_t want; /* capacity asked for by an overflowing step */
    uint64_t serial; /* the fleet's admission count when this job was admitted */
} fleet_slot;

struct tpgp_fleet {
    tpgp_hip_context *ctx;
    int own_ctx;
    tpgp_hip_batch *batch;
    size_t W, M;
    fleet_slot *slot;
    tpgp_hip_record *rec;
    size_t max_rec;
    double *rd;
    uint64_t *len;
    tpgp_hip_log_entry *log;
    size_t log_cap;
    double *scratch;
    size_t scratch_cap;
    tpgp_evolution_counters c;
    uint64_t budget; /* device bytes the slots may hold together (0: no limit) */
    size_t cap0;     /* every slot's capacity at open, to which tpgp_fleet_trim returns */
    uint64_t serial; /* admissions so far */
    uint64_t graphs_recorded, graphs_replayed;
};

static void slot_clear(fleet_slot *q) {
    free(q->plans);
    free(q->enq);
    free(q->fid);
    free(q->rad);
    free(q->act);
    size_t cap = q->cap;
    memset(q, 0, sizeof(*q));
    q->cap = cap;
}

void tpgp_fleet_close(tpgp_fleet *f) {
    if (!f) return;
    for (size_t m = 0; m < f->M; m++) slot_clear(&f->slot[m]);
    tpgp_hip_batch_destroy(f->batch);
    if (f->own_ctx) tpgp_hip_destroy(f->ctx);
    free(f->slot);
    free(f->rec);
    free(f->rd);
    free(f->len);
    free(f->log);
    free(f->scratch);
    free(f);
}

tpgp_status tpgp_fleet_open(tpgp_fleet **out, void *context, size_t n_words, size_t n_slots,
                            size_t slot_capacity) {
    if (!out) return TPGP_EINVAL;
    *out = NULL;
    if ((n_words != 1 && n_words != 2) || !n_slots || n_slots > (1u << 16) || !slot_capacity)
        return TPGP_EINVAL;
    tpgp_fleet *f = calloc(1, sizeof(*f));
    if (!f) return TPGP_ENOMEM;
    char err[512];
    f->ctx = context;
    if (!f->ctx) {
        f->ctx = tpgp_hip_create(0, err, sizeof(err));
        if (!f->ctx) {
            note("context", err);
            free(f);
            return TPGP_EBACKEND;
        }
        f->own_ctx = 1;
    }
    f->W = n_words;
    f->M = n_slots;
    f->cap0 = slot_capacity;
    f->slot = calloc(n_slots, sizeof(*f->slot));
    size_t *spins = malloc(n_slots * sizeof(size_t));
    size_t *caps = malloc(n_slots * sizeof(size_t));
    if (!f->slot || !spins || !caps) {
        free(spins);
        free(caps);
        tpgp_fleet_close(f);
        return TPGP_ENOMEM;
    }
    for (size_t m = 0; m < n_slots; m++) {
        spins[m] = n_words == 1 ? 2 : 33; /* placeholder until a job is admitted */
        caps[m] = slot_capacity;
        f->slot[m].cap = slot_capacity;
    }
    f->batch = tpgp_hip_batch_create(f->ctx, n_words, n_slots, spins, caps, err, sizeof(err));
    free(spins);
    free(caps);
    if (!f->batch) {
        note("batch", err);
        tpgp_fleet_close(f);
        return TPGP_ENOMEM;
    }
    f->max_rec = n_slots <= 16 ? 4096 : (65536 / n_slots > 1024 ? 65536 / n_slots : 1024);
    f->rec = malloc(n_slots * f->max_rec * sizeof(*f->rec));
    f->rd = malloc(n_slots * 64 * sizeof(double));
    f->len = malloc(n_slots * sizeof(uint64_t));
    if (!f->rec || !f->rd || !f->len) {
        tpgp_fleet_close(f);
        return TPGP_ENOMEM;
    }
    *out = f;
    return TPGP_OK;
}

tpgp_status tpgp_fleet_admit(tpgp_fleet *f, size_t m, const tpgp_fid_job *job) {
    if (!f || m >= f->M || !job || !job->initial) return TPGP_EINVAL;
    fleet_slot *q = &f->slot[m];
    const tpgp_map *init = job->initial;
    const tpgp_options *o = &job->options;
    const tpgp_evolve_options *eo = &job->evolve;
    if (q->state != TPGP_SLOT_FREE) return TPGP_EINVAL;
    if (!map_ok(init) || init->n_words != f->W || (!job->terms && job->count) ||
        !isfinite(job->dt) || eo->pauli < 1 || eo->pauli > 3 || !eo->normalizer ||
        2 * job->count > f->max_rec)
        return TPGP_EINVAL;
    if (o->backend != TPGP_HIP) return TPGP_EBACKEND;
    double half = job->dt * 0.5;
    if (half * 2 != job->dt) return TPGP_ENUMERIC; /* tpgp_rev.c:780-784 */
    slot_clear(q);
    q->nplans = 2 * job->count;
    q->plans = malloc((q->nplans + 1) * sizeof(*q->plans));
    q->enq = malloc((q->nplans + 1) * sizeof(*q->enq));
    q->fid = malloc((job->steps + 1) * sizeof(double));
    q->rad = malloc((job->steps + 1) * sizeof(double));
    q->act = malloc((job->steps + 1) * sizeof(size_t));
    if (!q->plans || !q->enq || !q->fid || !q->rad || !q->act) {
        slot_clear(q);
        return TPGP_ENOMEM;
    }
    for (size_t k = 0; k < q->nplans; k++) {
        const tpgp_term *term = k < job->count ? job->terms + k : job->terms + (q->nplans - 1 - k);
        plan_rotation(q->plans + k, init->n_spins, term->i, term->j, term->qi, term->qj, term->h,
                      half, o);
    }
    if (init->len > q->cap) {
        char err[512];
        size_t cap = init->len * 2 + 1024;
        int hs = tpgp_hip_batch_reserve(f->batch, m, cap, err, sizeof(err));
        if (hs) {
            note("reserve", err);
            slot_clear(q);
            return map_status(hs);
        }
        q->cap = cap;
    }
    char err[512];
    int hs = tpgp_hip_batch_assign(f->batch, m, init->n_spins, init->keys, init->values,
                                   init->len, err, sizeof(err));
    if (hs) {
        note("assign", err);
        slot_clear(q);
        return map_status(hs);
    }
    q->n_spins = init->n_spins;
    q->steps = job->steps;
    q->serial = ++f->serial;
    q->o = *o;
    q->o.hip_context = f->ctx;
    q->eo = *eo;
    tpgp_ledger_init(&q->ledger, o->validated);
    q->state = TPGP_SLOT_RUNNING;
    return TPGP_OK;
}

/* Fold one slot's records of this step (plans in order, as sync_fold does for one map). */
static tpgp_status fleet_fold(tpgp_fleet *f, size_t m, const tpgp_hip_log_entry *log,
                              size_t nlog, tpgp_ledger *trial, int *grow, size_t *need,
                              uint64_t *up, uint64_t *io, uint64_t *kp, uint64_t *ex) {
    fleet_slot *q = &f->slot[m];
    const int exact = q->eo.exact_ledger;
    size_t r_at = 0, e_at = 0;
    *grow = 0;
    for (size_t k = 0; k < q->nplans; k++) {
        const rot_plan *p = q->plans + k;
        if (!ledger_ok(trial)) return TPGP_EINVAL;
        if (p->pre) return p->pre;
        if (p->noop) continue;
        const tpgp_hip_record *r = f->rec + m * f->max_rec + r_at++;
        size_t e0 = e_at;
        while (e_at < nlog && log[e_at].rotation == r->rotation) e_at++;
        if (!r->executed) continue;
        int fs = fold_record(r, log + e0, e_at - e0, exact, &q->o, trial, f->scratch);
        if (fs == FOLD_CAPACITY) {
            *grow = 1;
            *need = (size_t)r->len_out;
            return TPGP_OK;
        }
        if (fs) return (tpgp_status)fs;
        *up += r->len_in;
        *io += r->len_in + r->len_out;
        *kp += 2 * r->in_place;
        (*ex)++;
    }
    return TPGP_OK;
}

static int cmp_entry_map(const void *a, const void *b) {
    const tpgp_hip_log_entry *x = a, *y = b;
    if (x->map != y->map) return x->map < y->map ? -1 : 1;
    return cmp_entry(a, b);
}

static uint64_t slot_bytes_at(const tpgp_fleet *f, size_t cap) {
    return 3ull * cap * (8 * f->W + 8);
}

/* Grow the slots with want[m] > their capacity, taking them in ascending order of the bytes
 * they add while the fleet stays within its budget.  One re-layout for all the granted slots;
 * if the device cannot hold them together, one slot at a time, smallest first.  A slot that is
 * refused (budget) or cannot be allocated gets refused[m] = 1 and keeps its capacity and its
 * committed state; every other slot is unaffected.  Returns the number of refused slots. */
static size_t grow_slots(tpgp_fleet *f, const size_t *want, int *refused) {
    const size_t M = f->M;
    size_t *order = malloc(M * sizeof(size_t)), n = 0, nref = 0;
    size_t *caps = calloc(M, sizeof(size_t));
    char err[512];
    if (!order || !caps) {
        for (size_t m = 0; m < M; m++)
            if (want[m] > f->slot[m].cap) refused[m] = 1, nref++;
        free(order);
        free(caps);
        return nref;
    }
    uint64_t held = 0; /* every slot's arena share, free ones included: the batch holds it */
    for (size_t m = 0; m < M; m++) {
        held += slot_bytes_at(f, f->slot[m].cap);
        if (want[m] > f->slot[m].cap) order[n++] = m;
    }
    for (size_t a = 1; a < n; a++) /* insertion sort by added bytes (few slots grow at once) */
        for (size_t k = a; k > 0 && want[order[k]] - f->slot[order[k]].cap <
                                        want[order[k - 1]] - f->slot[order[k - 1]].cap; k--) {
            size_t t = order[k];
            order[k] = order[k - 1];
            order[k - 1] = t;
        }
    size_t granted = 0;
    for (size_t a = 0; a < n; a++) {
        const size_t m = order[a];
        const uint64_t add = slot_bytes_at(f, want[m]) - slot_bytes_at(f, f->slot[m].cap);
        if (f->budget && held + add > f->budget) {
            refused[m] = 1;
            nref++;
            continue;
        }
        held += add;
        caps[m] = want[m];
        order[granted++] = m;
    }
    if (granted && tpgp_hip_batch_reserve_many(f->batch, caps, err, sizeof(err))) {
        note("reserve", err);
        /* together they do not fit: one at a time, smallest first */
        for (size_t a = 0; a < granted; a++) {
            const size_t m = order[a];
            memset(caps, 0, M * sizeof(size_t));
            caps[m] = want[m];
            if (tpgp_hip_batch_reserve_many(f->batch, caps, err, sizeof(err))) {
                note("reserve", err);
                refused[m] = 1;
                nref++;
            } else {
                f->slot[m].cap = want[m];
            }
        }
    } else {
        for (size_t a = 0; a < granted; a++) f->slot[order[a]].cap = want[order[a]];
    }
    free(order);
    free(caps);
    return nref;
}

tpgp_status tpgp_fleet_set_budget(tpgp_fleet *f, uint64_t bytes) {
    if (!f) return TPGP_EINVAL;
    f->budget = bytes;
    return TPGP_OK;
}

tpgp_status tpgp_fleet_trim(tpgp_fleet *f) {
    if (!f) return TPGP_EINVAL;
    char err[512];
    size_t *caps = calloc(f->M, sizeof(size_t));
    if (!caps) return TPGP_ENOMEM;
    int any = 0;
    for (size_t m = 0; m < f->M; m++) {
        if (f->slot[m].state != TPGP_SLOT_FREE || f->slot[m].cap <= f->cap0) continue;
        /* a free slot still holds its last map: empty it so the smaller slot can take it */
        int hs = tpgp_hip_batch_upload(f->batch, m, NULL, NULL, 0, err, sizeof(err));
        if (hs) {
            note("trim", err);
            free(caps);
            return map_status(hs);
        }
        caps[m] = f->cap0;
        any = 1;
    }
    int hs = any ? tpgp_hip_batch_resize(f->batch, caps, err, sizeof(err)) : 0;
    if (hs) note("trim", err);
    else
        for (size_t m = 0; m < f->M; m++)
            if (caps[m]) f->slot[m].cap = caps[m];
    free(caps);
    return hs ? map_status(hs) : TPGP_OK;
}

/* One round's device work: readouts of the committed states (one call per readout symbol)
 * and the rotations of every slot with steps left (R requests at most per slot). */
static tpgp_status fleet_enqueue_direct(tpgp_fleet *f, const int *part, int *pmask, size_t R) {
    char err[512];
    const size_t M = f->M;
    for (unsigned pauli = 1; pauli <= 3; pauli++) {
        int need_call = 0;
        for (size_t m = 0; m < M; m++) {
            pmask[m] = part[m] && f->slot[m].eo.pauli == pauli;
            need_call |= pmask[m];
        }
        if (need_call && tpgp_hip_batch_readout(f->batch, pauli, pmask, err, sizeof(err))) {
            note("readout", err);
            return TPGP_EBACKEND;
        }
    }
    if (!R) return TPGP_OK;
    tpgp_hip_rotation *req = calloc(R * M, sizeof(*req));
    uint32_t *seq = malloc(R * sizeof(*seq));
    if (!req || !seq) {
        free(req);
        free(seq);
        return TPGP_ENOMEM;
    }
    for (size_t r = 0; r < R; r++) seq[r] = (uint32_t)r;
    for (size_t m = 0; m < M; m++) {
        fleet_slot *q = &f->slot[m];
        for (size_t r = 0; r < q->nenq; r++) {
            const rot_plan *p = q->plans + q->enq[r];
            tpgp_hip_rotation *t = req + r * M + m;
            t->active = 1;
            t->i = p->i;
            t->j = p->j;
            t->qi = p->qi;
            t->qj = p->qj;
            t->c = p->c;
            t->s = p->s;
            t->rc = p->rc;
            t->rs = p->rs;
            t->threshold = q->o.threshold;
            t->validated = q->o.validated;
            t->log = q->eo.exact_ledger;
        }
    }
    int hs = tpgp_hip_batch_rotate_many(f->batch, req, seq, R, err, sizeof(err));
    free(req);
    free(seq);
    if (hs) {
        note("rotate", err);
        return map_status(hs);
    }
    return TPGP_OK;
}

/* The round's work, replayed from a HIP graph when the same slots (same admissions, same
 * phase: stepping or taking their final readout) start again from a state recorded before;
 * recorded when that state has been seen before (tpgp_hip_batch_graph_*). */
static tpgp_status fleet_enqueue_round(tpgp_fleet *f, const int *part, int *pmask, size_t R) {
    char err[512];
    uint64_t key = 1469598103934665603ull;  /* FNV-1a over the round's slots */
    for (size_t m = 0; m < f->M; m++) {
        const fleet_slot *q = &f->slot[m];
        const uint64_t v = part[m] ? (q->serial << 2) | (q->done >= q->steps ? 2u : 1u) : 0;
        for (int b = 0; b < 64; b += 8) key = (key ^ ((v >> b) & 0xffu)) * 1099511628211ull;
    }
    int replayed = 0, capture = 0, captured = 0;
    int hs = tpgp_hip_batch_graph_replay(f->batch, key, &replayed, &capture, err, sizeof(err));
    if (hs) {
        note("graph replay", err);
        return map_status(hs);
    }
    if (replayed) {
        f->graphs_replayed++;
        return TPGP_OK;
    }
    if (!capture || tpgp_hip_batch_graph_begin(f->batch, key, err, sizeof(err)))
        return fleet_enqueue_direct(f, part, pmask, R);
    tpgp_status s = fleet_enqueue_direct(f, part, pmask, R);
    hs = tpgp_hip_batch_graph_end(f->batch, !s, &captured, err, sizeof(err));
    if (s) return s;
    if (hs) {
        note("graph", err);
        return map_status(hs);
    }
    if (captured) {
        f->graphs_recorded++;
        return TPGP_OK;
    }
    return fleet_enqueue_direct(f, part, pmask, R); /* not recorded: nothing of it ran */
}

tpgp_status tpgp_fleet_step(tpgp_fleet *f, size_t *running) {
    if (!f) return TPGP_EINVAL;
    char err[512];
    const size_t M = f->M;
    int *part = calloc(M, sizeof(int));   /* slots taking part in this round */
    int *mask = calloc(M, sizeof(int));
    int *pmask = calloc(M, sizeof(int));
    if (!part || !mask || !pmask) {
        free(part);
        free(mask);
        free(pmask);
        return TPGP_ENOMEM;
    }
    tpgp_status s = TPGP_OK;
    int any = 0;
    for (size_t m = 0; m < M; m++) any |= part[m] = f->slot[m].state == TPGP_SLOT_RUNNING;
    /* grow slots ahead of need (one re-layout for all of them): a slot whose committed map
     * fills more than half of it gets room for three times that.  Growth ahead is an
     * optimisation: a slot it cannot serve (budget, device memory) keeps its capacity and grows
     * when a step actually overflows it. */
    if (any) {
        size_t *caps = calloc(M, sizeof(size_t));
        int *refused = calloc(M, sizeof(int));
        int grow_any = 0;
        if (!caps || !refused) s = TPGP_ENOMEM;
        for (size_t m = 0; m < M && caps && refused; m++) {
            size_t len = 0;
            if (!part[m] || tpgp_hip_batch_info(f->batch, m, &len, NULL)) continue;
            if (2 * len > f->slot[m].cap) {
                size_t cap = 3 * len + 4096;
                if (cap < 2 * f->slot[m].cap) cap = 2 * f->slot[m].cap;
                caps[m] = cap;
                grow_any = 1;
            }
        }
        if (grow_any && !s) (void)grow_slots(f, caps, refused);
        free(caps);
        free(refused);
    }
    while (any && !s) {
        /* this step's rotations: slot m's r-th enqueued rotation is request r */
        size_t R = 0;
        int exact_any = 0;
        for (size_t m = 0; m < M; m++) {
            fleet_slot *q = &f->slot[m];
            q->nenq = 0;
            if (!part[m] || q->done >= q->steps) continue;
            for (size_t k = 0; k < q->nplans; k++) {
                if (q->plans[k].pre) break;
                if (!q->plans[k].noop) q->enq[q->nenq++] = k;
            }
            if (q->nenq > R) R = q->nenq;
            exact_any |= q->eo.exact_ledger;
        }
        if ((s = fleet_enqueue_round(f, part, pmask, R))) break;
        if (exact_any && !f->log) {
            f->log_cap = 1u << 20;
            f->log = malloc(f->log_cap * sizeof(*f->log));
            if (!f->log) {
                s = TPGP_ENOMEM;
                break;
            }
        }
        size_t nrec = 0, nrd = 0, nlog = 0;
        int hs = tpgp_hip_batch_sync(f->batch, f->rec, f->max_rec, &nrec, f->rd, 1, &nrd, f->len,
                                     exact_any ? f->log : NULL, exact_any ? f->log_cap : 0, &nlog,
                                     err, sizeof(err));
        if (hs == TPGP_HIP_ECAPACITY && exact_any) {
            /* logs too small: discard this round, grow them, run the same slots again */
            tpgp_hip_batch_rollback(f->batch, err, sizeof(err));
            if ((s = grow_logs(f->batch, &f->log, &f->log_cap, nlog))) break;
            continue;
        }
        if (hs) {
            note("sync", err);
            s = hs == TPGP_HIP_ECAPACITY ? TPGP_EBACKEND : map_status(hs);
            break;
        }
        if (nlog) qsort(f->log, nlog, sizeof(*f->log), cmp_entry_map);
        if (f->scratch_cap < nlog + 1) {
            free(f->scratch);
            f->scratch_cap = nlog + 1;
            f->scratch = malloc(f->scratch_cap * sizeof(double));
            if (!f->scratch) {
                s = TPGP_ENOMEM;
                break;
            }
        }
        /* per slot: readout, fold, decision */
        size_t l_at = 0;
        int regrow = 0;
        for (size_t m = 0; m < M; m++) {
            mask[m] = 0;
            size_t l0 = l_at;
            while (l_at < nlog && f->log[l_at].map == m) l_at++;
            if (!part[m]) continue;
            part[m] = 0;
            fleet_slot *q = &f->slot[m];
            double mid = 0, rad = NAN;
            tpgp_status rs = TPGP_OK;
            const double *rd = f->rd + m * 64;
            if (tpgp_has_mpfr())
                rs = readout_values(rd, q->n_spins, q->eo.normalizer, &mid, &rad);
            else
                mid = neumaier(rd, q->n_spins) / (double)q->eo.normalizer;
            size_t read_len = 0; /* the state read out is the committed one */
            tpgp_hip_batch_info(f->batch, m, &read_len, NULL);
            if (q->done >= q->steps) { /* final readout only */
                q->fid[q->done] = mid;
                q->rad[q->done] = rad;
                q->act[q->done] = read_len;
                q->status = rs;
                q->state = rs ? TPGP_SLOT_FAILED : TPGP_SLOT_DONE;
                continue;
            }
            tpgp_ledger trial = q->ledger;
            int grow = 0;
            size_t need = 0;
            uint64_t up = 0, io = 0, kp = 0, ex = 0;
            tpgp_status fs = fleet_fold(f, m, f->log + l0, l_at - l0, &trial, &grow, &need, &up,
                                        &io, &kp, &ex);
            /* the sample of the committed state stands whatever becomes of this step (a slot
             * refused its growth below fails with it as its last sample) */
            q->fid[q->done] = mid;
            q->rad[q->done] = rad;
            q->act[q->done] = read_len;
            if (!fs && grow) {
                size_t cap = need + need / 2 + 1024;
                if (cap < 2 * q->cap) cap = 2 * q->cap;
                q->want = cap;
                part[m] = 1; /* rerun this slot's step after growing */
                regrow = 1;
                continue;
            }
            if (!fs && rs) fs = rs;
            if (fs) {
                q->status = fs;
                q->state = TPGP_SLOT_FAILED;
                continue;
            }
            mask[m] = 1;
            q->ledger = trial;
            q->done++;
            f->c.steps++;
            f->c.rotations += ex;
            f->c.updates += up;
            f->c.in_out += io;
            f->c.in_place += kp;
        }
        if (tpgp_hip_batch_commit(f->batch, mask, err, sizeof(err))) {
            note("commit", err);
            s = TPGP_EBACKEND;
            break;
        }
        any = 0;
        if (regrow) { /* the overflowing slots grow (one re-layout), then run again; a slot
                       * that cannot grow fails alone with ENOMEM at its last accepted step */
            size_t *caps = calloc(M, sizeof(size_t));
            int *refused = calloc(M, sizeof(int));
            if (!caps || !refused) {
                free(caps);
                free(refused);
                s = TPGP_ENOMEM;
                break;
            }
            for (size_t m = 0; m < M; m++)
                if (part[m]) caps[m] = f->slot[m].want;
            (void)grow_slots(f, caps, refused);
            free(caps);
            for (size_t m = 0; m < M; m++) {
                if (!part[m]) continue;
                if (refused[m]) {
                    f->slot[m].status = TPGP_ENOMEM;
                    f->slot[m].state = TPGP_SLOT_FAILED;
                    part[m] = 0;
                }
                any |= part[m];
            }
            free(refused);
        }
    }
    if (running) {
        size_t n = 0;
        for (size_t m = 0; m < M; m++) n += f->slot[m].state == TPGP_SLOT_RUNNING;
        *running = n;
    }
    if (s) {
        tpgp_hip_batch_rollback(f->batch, err, sizeof(err));
    }
    free(part);
    free(mask);
    free(pmask);
    return s;
}

void tpgp_fleet_graph_counts(const tpgp_fleet *f, uint64_t *recorded, uint64_t *replayed) {
    if (recorded) *recorded = f ? f->graphs_recorded : 0;
    if (replayed) *replayed = f ? f->graphs_replayed : 0;
}

tpgp_status tpgp_fleet_slot(const tpgp_fleet *f, size_t m, tpgp_slot_state *state,
                            size_t *steps_done, tpgp_status *status, tpgp_ledger *ledger) {
    if (!f || m >= f->M) return TPGP_EINVAL;
    const fleet_slot *q = &f->slot[m];
    if (state) *state = q->state;
    if (steps_done) *steps_done = q->done;
    if (status) *status = q->status;
    if (ledger) *ledger = q->ledger;
    return TPGP_OK;
}

tpgp_status tpgp_fleet_fid(const tpgp_fleet *f, size_t m, double *fid, double *fid_radius,
                           size_t *active, size_t capacity, size_t *n) {
    if (!f || m >= f->M) return TPGP_EINVAL;
    const fleet_slot *q = &f->slot[m];
    size_t k = q->state == TPGP_SLOT_FREE ? 0 : q->done;
    if (q->state == TPGP_SLOT_DONE || q->state == TPGP_SLOT_FAILED) k = q->done + 1;
    if (k > capacity) k = capacity;
    for (size_t t = 0; t < k; t++) {
        if (fid) fid[t] = q->fid[t];
        if (fid_radius) fid_radius[t] = q->rad[t];
        if (active) active[t] = q->act[t];
    }
    if (n) *n = k;
    return TPGP_OK;
}

tpgp_status tpgp_fleet_release(tpgp_fleet *f, size_t m, tpgp_map *final_map) {
    if (!f || m >= f->M) return TPGP_EINVAL;
    fleet_slot *q = &f->slot[m];
    tpgp_status s = TPGP_OK;
    if (final_map && q->state != TPGP_SLOT_FREE) {
        char err[512];
        size_t len = 0;
        tpgp_map t;
        int hs = tpgp_hip_batch_download(f->batch, m, NULL, NULL, 0, &len, err, sizeof(err));
        if (hs) return map_status(hs);
        if ((s = tpgp_map_init(&t, q->n_spins))) return s;
        if ((s = map_reserve(&t, len))) {
            tpgp_map_free(&t);
            return s;
        }
        hs = tpgp_hip_batch_download(f->batch, m, t.keys, t.values, len, &t.len, err,
                                     sizeof(err));
        if (hs) {
            note("download", err);
            tpgp_map_free(&t);
            return map_status(hs);
        }
        tpgp_map_free(final_map);
        *final_map = t;
    }
    slot_clear(q);
    return s;
}

void tpgp_fleet_counters_get(const tpgp_fleet *f, tpgp_evolution_counters *c) {
    if (f && c) *c = f->c;
}

uint64_t tpgp_fleet_slot_bytes(const tpgp_fleet *f, size_t m) {
    if (!f || m >= f->M) return 0;
    return 3ull * f->slot[m].cap * (8 * f->W + 8);
}

/* ---- Trotter bounds and allowance helpers (tpgp_rev.c:858-961; residual_rev.c:195-201) ----- */
static int terms_anticommute(const tpgp_term *a, const tpgp_term *b) {
    int parity = 0;
    if (a->i == b->i) parity ^= anticomm(a->qi, b->qi);
    if (a->i == b->j) parity ^= anticomm(a->qi, b->qj);
    if (a->j == b->i) parity ^= anticomm(a->qj, b->qi);
    if (a->j == b->j) parity ^= anticomm(a->qj, b->qj);
    return parity;
}
double tpgp_strang_bound_taylor(const tpgp_term *t, size_t n, double dt) {
    if ((!t && n) || !isfinite(dt)) return NAN;
    double lambda = 0;
    for (size_t k = 0; k < n; k++) {
        if (!isfinite(t[k].h) || t[k].i == t[k].j || t[k].qi > 3 || t[k].qj > 3) return NAN;
        lambda = upadd(lambda, fabs(t[k].h));
    }
    double x = upmul(lambda, fabs(dt));
    return upmul(upmul(upmul(x, x), x), nextafter(2.0 / 3.0, INFINITY));
}
double tpgp_strang_bound(const tpgp_term *t, size_t n, double dt) {
    double general = tpgp_strang_bound_taylor(t, n, dt);
    if (isnan(general)) return NAN;
    /* dense adjacency bitmap: fine for the term counts of these models */
    unsigned char *adj = n ? calloc(n * n, 1) : NULL;
    if (n && !adj) return general;
    for (size_t a = 0; a < n; a++)
        for (size_t b = a + 1; b < n; b++)
            if (t[a].h && t[b].h && terms_anticommute(t + a, t + b)) adj[a * n + b] = adj[b * n + a] = 1;
    double C = 0;
    for (size_t a = 0; a < n; a++) {
        double ha = fabs(t[a].h), inner = 0;
        for (size_t c = a + 1; c < n; c++)
            if (adj[a * n + c]) inner = upadd(inner, fabs(t[c].h));
        C = upadd(C, upmul(upmul(upmul(ha, ha), inner), nextafter(1.0 / 6.0, INFINITY)));
        for (size_t b = a + 1; b < n; b++) {
            if (!adj[a * n + b]) continue;
            double sum = 0;
            /* terms anticommuting with exactly one of a, b, in ascending index order */
            for (size_t c = a + 1; c < n; c++)
                if (adj[a * n + c] != adj[b * n + c]) sum = upadd(sum, fabs(t[c].h));
            C = upadd(C, upmul(upmul(upmul(ha, fabs(t[b].h)), sum), nextafter(1.0 / 3.0, INFINITY)));
        }
    }
    free(adj);
    double T = fabs(dt), nested = upmul(upmul(upmul(upmul(2, C), T), T), T);
    return fmin(general, nested);
}
double tpgp_bound_add_up(double a, double b) {
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
    return upadd(a, b);
#endif
}
double tpgp_bound_fid_up(double a, size_t n) {
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
