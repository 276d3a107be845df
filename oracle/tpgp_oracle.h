/* CPU oracle for the Pauli-propagation branch step -- TEST INFRASTRUCTURE ONLY.
 *
 * A C17 restatement of the reference CPU engine's hot path (engine/src/tpgp_rev.c), used by the
 * parity tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  The
 * product library (ddevol_amd) never links, loads or calls anything declared here.
 *
 * Semantics follow the reference line by line (file:line citations in tpgp_oracle.c); status
 * codes, the ledger and the term layout are identical so results compare bitwise.
 */
#ifndef TPGP_ORACLE_H
#define TPGP_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OK = 0, ORC_EINVAL, ORC_ENOMEM, ORC_ECAP, ORC_ENUMERIC, ORC_EBACKEND, ORC_EBUDGET };
enum { ORC_SERIAL = 0, ORC_PTHREAD = 1, ORC_OPENMP = 2 };

typedef struct {
    size_t n_spins, n_words, len, capacity;
    uint64_t *keys; /* keys[k*n_words + w], two bits per site, site s in word s/32 */
    double *values;
} orc_map;
typedef struct {
    size_t i, j;
    unsigned qi, qj;
    double h;
    size_t layer;
} orc_term;
typedef struct {
    double threshold;
    size_t max_terms;
    unsigned threads;
    int backend;
    int validated;
    double max_drop_l2;
} orc_options;
typedef struct {
    double dropped_l1, dropped_l2, dropped_l2_squared;
    double arithmetic_l2;
    size_t rotations, removal_events, peak_terms;
    int arithmetic_enclosed;
} orc_ledger;

int orc_has_mpfr(void);
int orc_has_openmp(void);
void orc_ledger_init(orc_ledger *ledger, int validated);

orc_map *orc_map_new(size_t n_spins);
void orc_map_delete(orc_map *map);
int orc_map_assign(orc_map *map, const uint64_t *keys, const double *values, size_t len);
int orc_map_copy(orc_map *dst, const orc_map *src);
size_t orc_map_len(const orc_map *map);
size_t orc_map_words(const orc_map *map);
void orc_map_export(const orc_map *map, uint64_t *keys, double *values);
int orc_map_add(orc_map *map, const uint64_t *key, double value);
int orc_map_add_single(orc_map *map, size_t site, unsigned pauli, double value);
double orc_map_get(const orc_map *map, const uint64_t *key);
double orc_map_l2(const orc_map *map);
double orc_map_single_sum(const orc_map *map, unsigned pauli);
int orc_readout(const orc_map *map, unsigned pauli, size_t normalizer, double *midpoint,
                double *radius);

unsigned orc_pauli_phase(unsigned q, unsigned p);
int orc_trig(double h, double duration, int validated, double out[4]);
/* Raw branch slots 2k (a) and 2k+1 (d) for given trigonometric factors (scalar spec). */
void orc_branches(const uint64_t *keys, const double *values, size_t count, size_t n_words,
                  size_t i, size_t j, unsigned qi, unsigned qj, double c, double s, double rc,
                  double rs, int enclose, uint64_t *out_keys, double *out_values,
                  double *out_radii);

int orc_rotate_product(orc_map *map, size_t i, size_t j, unsigned qi, unsigned qj, double h,
                       double duration, const orc_options *options, orc_ledger *ledger);
int orc_rotate(orc_map *map, size_t i, size_t j, unsigned qi, unsigned qj, double theta,
               const orc_options *options, orc_ledger *ledger);
int orc_strang_step(orc_map *map, const orc_term *terms, size_t count, double dt,
                    const orc_options *options, orc_ledger *ledger);
double orc_strang_bound_taylor(const orc_term *terms, size_t count, double dt);
double orc_strang_bound(const orc_term *terms, size_t count, double dt);
double orc_bound_add_up(double a, double b);
double orc_bound_fid_up(double a, size_t n);
/* Directed-rounding helpers exposed so tests can check them against their definitions. */
/* -i[H, A] in the Pauli basis (tpgp_rev.c:810-857); replaces *out on success. */
int orc_liouvillian(orc_map *out, const orc_map *in, const orc_term *terms, size_t count);
double orc_upadd(double a, double b);
double orc_upmul(double a, double b);
double orc_round_radius(double x);
double orc_upper_l2(const double *v, size_t n);

#ifdef __cplusplus
}
#endif
#endif
