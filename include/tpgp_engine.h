/* tpgp_engine -- C17 host engine over the tpgp_hip device library.
 *
 * Keeps the reference core's API (engine/include/tpgp_rev.h:9-96) with the same names, types,
 * struct layouts, statuses and transactional behaviour, for the functions on the propagation
 * path.  The only backend is the GPU: options.backend must be TPGP_HIP, which occupies the
 * reference's TPGP_CUDA slot (tpgp_rev.h:23), and options.hip_context occupies its cuda_context
 * slot (tpgp_rev.h:44).  Other backends return TPGP_EBACKEND: there is no CPU fallback.
 */
#ifndef TPGP_ENGINE_H
#define TPGP_ENGINE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* tpgp_rev.h:9-17 */
typedef enum {
    TPGP_OK = 0,
    TPGP_EINVAL,
    TPGP_ENOMEM,
    TPGP_ECAP,
    TPGP_ENUMERIC,
    TPGP_EBACKEND,
    TPGP_EBUDGET
} tpgp_status;
/* tpgp_rev.h:18-24: only the GPU slot is implemented (value 4, the reference's TPGP_CUDA). */
typedef enum { TPGP_HIP = 4 } tpgp_backend;
/* tpgp_rev.h:25-29 */
typedef struct {
    size_t n_spins, n_words, len, capacity;
    uint64_t *keys;
    double *values;
} tpgp_map;
/* tpgp_rev.h:30-35 */
typedef struct {
    size_t i, j;
    unsigned qi, qj;
    double h;
    size_t layer;
} tpgp_term;
/* tpgp_rev.h:36-45 (same layout; metal_context is unused) */
typedef struct {
    double threshold;
    size_t max_terms;
    unsigned threads;
    int backend;
    int validated;
    double max_drop_l2;
    void *metal_context;
    void *hip_context; /* optional reusable tpgp_hip_context, caller-owned */
} tpgp_options;
/* tpgp_rev.h:46-51 */
typedef struct {
    double dropped_l1, dropped_l2, dropped_l2_squared;
    double arithmetic_l2;
    size_t rotations, removal_events, peak_terms;
    int arithmetic_enclosed;
} tpgp_ledger;

/* tpgp_rev.h:53-57 */
const char *tpgp_status_string(tpgp_status status);
int tpgp_has_mpfr(void);
int tpgp_has_openmp(void);
tpgp_options tpgp_default_options(void);
void tpgp_ledger_init(tpgp_ledger *ledger, int validated);
/* tpgp_rev.h:58-73: host containers of the sorted map (zero-initialise before init). */
tpgp_status tpgp_map_init(tpgp_map *map, size_t n_spins);
void tpgp_map_free(tpgp_map *map);
tpgp_status tpgp_map_clone(tpgp_map *dst, const tpgp_map *src);
unsigned tpgp_key_get(const uint64_t *key, size_t site);
void tpgp_key_set(uint64_t *key, size_t site, unsigned pauli);
tpgp_status tpgp_map_add(tpgp_map *map, const uint64_t *key, double value);
tpgp_status tpgp_map_add_single(tpgp_map *map, size_t site, unsigned pauli, double value);
double tpgp_map_get(const tpgp_map *map, const uint64_t *key);
double tpgp_map_l2(const tpgp_map *map);
double tpgp_map_single_sum(const tpgp_map *map, unsigned pauli);
/* tpgp_rev.h:74-78: MPFR-enclosed readout (EBACKEND without MPFR). */
tpgp_status tpgp_readout(const tpgp_map *map, unsigned pauli, size_t normalizer,
                         double *midpoint, double *radius);
/* tpgp_rev.h:81-82 */
unsigned tpgp_pauli_phase(unsigned q, unsigned p);
/* tpgp_rev.h:83-87: transactional rotations on the GPU; ledgers are bit-identical to the
 * reference CPU engine (dropped values are folded on the host in ascending key order). */
tpgp_status tpgp_rotate(tpgp_map *map, size_t i, size_t j, unsigned qi, unsigned qj,
                        double theta, const tpgp_options *options, tpgp_ledger *ledger);
tpgp_status tpgp_strang_step(tpgp_map *map, const tpgp_term *terms, size_t count,
                             double duration, const tpgp_options *options, tpgp_ledger *ledger);
/* tpgp_rev.h:91-96 (harness bounds, evaluated on the host) */
double tpgp_strang_bound(const tpgp_term *terms, size_t count, double duration);
double tpgp_strang_bound_taylor(const tpgp_term *terms, size_t count, double duration);
/* engine/include/residual_rev.h (tpgp_bound_add_up / tpgp_bound_fid_up, used for allowances) */
double tpgp_bound_add_up(double a, double b);
double tpgp_bound_fid_up(double a, size_t n);
/* engine/include/residual_rev.h:1-27 (residual_rev.c:181-194): the independent endpoint bound
 * of one Strang step from the map before (a) and after (b) it, in 256-bit MPFR interval
 * arithmetic; the rounded-up binary64 result equals the reference's.  `_reference` is the
 * diagnostic variant (every term scanned, no Gram enclosure).  EBACKEND without MPFR, EINVAL
 * for malformed maps or terms, ENOMEM, ENUMERIC for a non-finite result; *upper unchanged on
 * failure.  depth <= 10. */
tpgp_status tpgp_residual_bound(const tpgp_map *a, const tpgp_map *b, const tpgp_term *terms,
                                size_t count, double dt, unsigned depth, double *upper);
tpgp_status tpgp_residual_bound_reference(const tpgp_map *a, const tpgp_map *b,
                                          const tpgp_term *terms, size_t count, double dt,
                                          unsigned depth, double *upper);

/* ---- resident evolution: the step loop of engine/src/cli_rev.c:575-670 on one GPU ---------
 * Uploads `map` once, runs `steps` Strang steps with the map resident in HBM, and reads the FID
 * F(k) = sum_s coeff(P_s)/normalizer before every step k = 0..steps (cli_rev.c:576-581).
 * fid[k] is the readout midpoint (MPFR readout when available, else the compensated
 * long-double sum / normalizer); fid_radius[k] its enclosure radius (NaN without MPFR);
 * active[k] the length of the sampled map.  Arrays hold steps+1 entries; any may be NULL.
 * exact_ledger: 1 = bit-identical ledger (host fold of logged drops, slower);
 *               0 = device ledger: rigorous upward-rounded sums in a fixed tree order.
 * On failure the map and ledger keep the last accepted step; *accepted counts steps done and
 * sample *accepted (the state the failing step started from) is still written. */
typedef struct {
    int exact_ledger;
    unsigned pauli;       /* readout symbol (1 = X) */
    size_t normalizer;    /* readout divisor (n_spins for the global initial state) */
    size_t initial_capacity; /* 0: automatic */
} tpgp_evolve_options;
tpgp_status tpgp_hip_evolve(tpgp_map *map, const tpgp_term *terms, size_t count, double dt,
                            size_t steps, const tpgp_options *options,
                            const tpgp_evolve_options *evolve, tpgp_ledger *ledger, double *fid,
                            double *fid_radius, size_t *active, size_t *accepted);

/* The same loop as an incremental session: open uploads the map once; run may be called
 * repeatedly (each call reads the FID before each of its steps and after its last); close
 * downloads the final map into `map` (or discards it for NULL).  Counters: executed rotations,
 * branch-updates (sum of input lengths of executed rotations), sum of input+output lengths, and
 * twice the terms rotations left in place (classes an appending rotation neither read nor
 * wrote: in_out - in_place terms were moved). */
typedef struct tpgp_evolution tpgp_evolution;
typedef struct {
    uint64_t steps, rotations, updates, in_out, in_place;
} tpgp_evolution_counters;
tpgp_status tpgp_evolution_open(tpgp_evolution **evolution, const tpgp_map *map,
                                const tpgp_term *terms, size_t count, double dt,
                                const tpgp_options *options, const tpgp_evolve_options *evolve);
tpgp_status tpgp_evolution_run(tpgp_evolution *evolution, size_t steps, tpgp_ledger *ledger,
                               double *fid, double *fid_radius, size_t *active,
                               size_t *accepted);
tpgp_status tpgp_evolution_close(tpgp_evolution *evolution, tpgp_map *map);
void tpgp_evolution_counters_get(const tpgp_evolution *evolution,
                                 tpgp_evolution_counters *counters);
/* HIP-event timing of the device kernels (kinds as in tpgp_hip.h TPGP_HIP_PROFILE_*). */
int tpgp_evolution_profile(tpgp_evolution *evolution, unsigned mask);
int tpgp_evolution_profile_read(const tpgp_evolution *evolution, unsigned kind,
                                double *milliseconds, uint64_t *launches);
size_t tpgp_evolution_key_words(const tpgp_evolution *evolution);
/* No counterpart in the reference: steps recorded as HIP graphs and steps replayed from them
 * (a step that starts from a state seen before is recorded, later ones replay it; results are
 * those of enqueueing every step directly; TPGP_HIP_GRAPHS=0 turns this off). */
void tpgp_evolution_graph_counts(const tpgp_evolution *evolution, uint64_t *recorded,
                                 uint64_t *replayed);

/* ---- Batched and served FIDs ------------------------------------------------------------
 * A fleet runs many independent FIDs (powder orientations, coupling draws, served requests)
 * as the maps of one resident batch: every device pass works on all running slots at once,
 * whatever their term counts.  Each slot runs exactly the loop of tpgp_hip_evolve -- same
 * rotations, same readouts, same transactional steps, same statuses -- so a FID is bit for bit
 * the same alone and in any mix.  All slots share the key width n_words (1: N <= 32, 2:
 * 33 <= N <= 64).  Slots start with slot_capacity terms and grow on demand. */
typedef struct tpgp_fleet tpgp_fleet;
typedef struct {
    const tpgp_map *initial;     /* canonical initial state (copied at admission) */
    const tpgp_term *terms;      /* copied at admission */
    size_t count;
    double dt;
    size_t steps;                /* Strang steps; steps + 1 FID samples */
    tpgp_options options;        /* backend TPGP_HIP; hip_context ignored (the fleet's is used) */
    tpgp_evolve_options evolve;  /* initial_capacity ignored */
} tpgp_fid_job;
typedef enum {
    TPGP_SLOT_FREE = 0,
    TPGP_SLOT_RUNNING,
    TPGP_SLOT_DONE,
    TPGP_SLOT_FAILED
} tpgp_slot_state;
/* context: caller-owned tpgp_hip_context (NULL: the fleet creates one on device 0). */
tpgp_status tpgp_fleet_open(tpgp_fleet **fleet, void *context, size_t n_words, size_t n_slots,
                            size_t slot_capacity);
void tpgp_fleet_close(tpgp_fleet *fleet);
/* Start a job in a FREE slot. */
tpgp_status tpgp_fleet_admit(tpgp_fleet *fleet, size_t slot, const tpgp_fid_job *job);
/* Advance every RUNNING slot by one Strang step (reading its FID first); a slot whose steps
 * are done takes its last readout and becomes DONE; a failing step leaves the slot FAILED with
 * the map and ledger of its last accepted step.  *running: RUNNING slots after the call. */
tpgp_status tpgp_fleet_step(tpgp_fleet *fleet, size_t *running);
tpgp_status tpgp_fleet_slot(const tpgp_fleet *fleet, size_t slot, tpgp_slot_state *state,
                            size_t *steps_done, tpgp_status *status, tpgp_ledger *ledger);
/* The slot's samples so far (capacity entries at most; returns the count in *n). */
tpgp_status tpgp_fleet_fid(const tpgp_fleet *fleet, size_t slot, double *fid,
                           double *fid_radius, size_t *active, size_t capacity, size_t *n);
/* Free a slot; final_map (optional) receives its last accepted map. */
tpgp_status tpgp_fleet_release(tpgp_fleet *fleet, size_t slot, tpgp_map *final_map);
void tpgp_fleet_counters_get(const tpgp_fleet *fleet, tpgp_evolution_counters *counters);
/* Bytes of device memory the fleet's slots hold (three term buffers per slot). */
uint64_t tpgp_fleet_slot_bytes(const tpgp_fleet *fleet, size_t slot);
/* No counterpart in the reference (its CPU engine runs one FID at a time): a ceiling on the
 * device bytes all slots hold together (0: none).  When slots must grow, growth is granted
 * smallest first while the fleet stays under the ceiling; a slot refused by the ceiling, or by
 * the device's memory, fails alone with TPGP_ENOMEM at its last accepted step, and the other
 * slots run on unaffected. */
tpgp_status tpgp_fleet_set_budget(tpgp_fleet *fleet, uint64_t bytes);
/* No counterpart in the reference: return every FREE slot that has grown to the fleet's
 * opening slot capacity (one re-layout), giving its device memory back; running slots keep
 * theirs.  On TPGP_ENOMEM every slot keeps its capacity. */
tpgp_status tpgp_fleet_trim(tpgp_fleet *fleet);
/* Rounds recorded as HIP graphs and rounds replayed from them, as for an evolution (the same
 * slots with the same jobs, in the same phase, from a state seen before). */
void tpgp_fleet_graph_counts(const tpgp_fleet *fleet, uint64_t *recorded, uint64_t *replayed);

/* -i[H, A] in the Pauli basis, computed on the GPU (default context on device 0), bit for bit
 * as the reference's tpgp_liouvillian (tpgp_rev.h:89, tpgp_rev.c:810-857); replaces *out on
 * success, leaves it unchanged on failure.  At most 64 spins. */
tpgp_status tpgp_liouvillian(tpgp_map *out, const tpgp_map *in, const tpgp_term *terms,
                             size_t count);

/* ---- Checkpoint / restart (the reference's TPGPRVCP version-2 record) --------------------
 * Same record as engine/include/checkpoint_rev.h:7-40 (field order, little-endian encoding,
 * SHA-256 of the payload appended, exclusive temporary file + fsync + atomic rename + parent
 * directory fsync, companion .manifest.json), so a run checkpointed here resumes in the
 * reference's CLI and the other way round.  Context pointers are never serialised. */
typedef struct {
    tpgp_map map;
    tpgp_term *terms;
    size_t term_count;
    tpgp_options options;
    tpgp_ledger ledger;
    double dt, pf_bound, residual_bound, max_fid_bound;
    size_t step, total_steps;
    int residual_depth; /* -1: not requested (residual_bound is then NaN) */
    int initial_global;
    char model[32];
    double j2;
    size_t linear_size;
    size_t resumed_from_step, prior_total_steps, prior_max_terms;
    int stop_reason; /* 0 running, 1 complete, 2 walltime, 3 signal, 4 resource, 5 numerical,
                        6 accuracy */
    tpgp_status last_status;
} tpgp_checkpoint;
void tpgp_checkpoint_free(tpgp_checkpoint *checkpoint);
tpgp_status tpgp_checkpoint_save(const char *path, const tpgp_checkpoint *checkpoint, char *error,
                                 size_t error_capacity);
tpgp_status tpgp_checkpoint_load(const char *path, tpgp_checkpoint *destination, char *error,
                                 size_t error_capacity);
int tpgp_checkpoint_same_configuration(const tpgp_checkpoint *a, const tpgp_checkpoint *b);
/* SHA-256 of a buffer (the checksum the record carries). */
void tpgp_payload_sha256(const void *data, size_t bytes, unsigned char digest[32]);
/* Copy a resident session's last accepted map out (reference layout) without closing it. */
tpgp_status tpgp_evolution_snapshot(tpgp_evolution *evolution, tpgp_map *map);
/* l2 norm of the last accepted map, computed on the device: sqrt of the sum of squares, as
 * tpgp_map_l2 (tpgp_rev.c:266-271) with an accumulation that is exact to ~2^-104 relative. */
tpgp_status tpgp_evolution_l2(tpgp_evolution *evolution, double *l2);

/* Message of the last device-side failure on this thread ("" when none). */
const char *tpgp_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
