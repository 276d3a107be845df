/* tpgp_hip -- MI355X (gfx950) Pauli-propagation branch engine, C ABI.
 *
 * Part 1 replaces the reference's CUDA branch backend one-for-one:
 *   engine/include/tpgp_cuda_rev.h:20-48 (tpgp_cuda_compiled/create/destroy/device_name/branches),
 *   implemented there by engine/src/cuda_rev.cu:117-272.  Same arguments, same byte-identical
 *   output slot order (2k, 2k+1), same statuses, same error-buffer convention, no CPU fallback.
 *
 * Part 2 is the resident extension: a batch of Pauli maps kept in HBM across rotations.  It
 * replaces the per-rotation host work of engine/src/tpgp_rev.c:535-771 (branch, stable sort,
 * merge, truncation, ledger inputs) and the FID scan of tpgp_rev.c:311-388 with device kernels;
 * keys/coefficients go in, merged and truncated maps plus FID samples come out.
 *
 * All pointers are host pointers; sizes are element counts.  No torch types cross this boundary.
 */
#ifndef TPGP_HIP_H
#define TPGP_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ============================ Part 1: branch backend ==================================== */

typedef struct tpgp_hip_context tpgp_hip_context;
/* Values match tpgp_cuda_status (engine/include/tpgp_cuda_rev.h:11-17). */
typedef enum {
    TPGP_HIP_OK = 0,
    TPGP_HIP_EINVAL,
    TPGP_HIP_ENOMEM,
    TPGP_HIP_EDEVICE,
    TPGP_HIP_ENUMERIC
} tpgp_hip_status;

/* Replaces tpgp_cuda_compiled (tpgp_cuda_rev.h:21): build capability, not device presence. */
int tpgp_hip_compiled(void);
/* Replaces tpgp_cuda_create (tpgp_cuda_rev.h:22; cuda_rev.cu:118-157).  Device indices follow
 * HIP_VISIBLE_DEVICES.  NULL plus a message on failure. */
tpgp_hip_context *tpgp_hip_create(int device_index, char *error, size_t error_capacity);
/* Replaces tpgp_cuda_destroy (tpgp_cuda_rev.h:23; cuda_rev.cu:158-169).  NULL-safe. */
void tpgp_hip_destroy(tpgp_hip_context *context);
/* Replaces tpgp_cuda_device_name (tpgp_cuda_rev.h:24): "unavailable" for NULL. */
const char *tpgp_hip_device_name(const tpgp_hip_context *context);

/* Replaces tpgp_cuda_branches (tpgp_cuda_rev.h:43-49; cuda_rev.cu:174-272).
 * For each input term k: output slots 2k and 2k+1.  Commuting P gives (P,v,0),(P,0,0);
 * anticommuting P gives (P, RN(v*c), radius_c) and (key(-iQP), sign*RN(v*s), radius_s).
 * Host-provided c, s, rc, rs; the GPU evaluates no trigonometric function.  Arrays are disjoint
 * host arrays: count*n_words keys, count values, 2*count*n_words output keys, 2*count output
 * values and radii.  EINVAL: bad dimensions/sites/generator/factors, NULL arrays, size overflow,
 * overlapping arrays, nonzero key padding.  ENUMERIC: nonfinite input, product or radius. */
tpgp_hip_status tpgp_hip_branches(tpgp_hip_context *context, const uint64_t *keys,
                                  const double *values, size_t count, size_t n_words,
                                  size_t n_spins, size_t i, size_t j, unsigned qi, unsigned qj,
                                  double c, double s, double rc, double rs, int enclose,
                                  uint64_t *output_keys, double *output_values,
                                  double *output_radii, char *error, size_t error_capacity);

/* No counterpart in the reference: tpgp_hip_branches for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, outputs, statuses and error convention as tpgp_hip_branches, but the device
 * copies of inputs and outputs live in a fixed per-GPU array of TPGP_HIP_NOALLOC_SCRATCH_BYTES
 * bytes, held from the library's load until the process ends, instead of the context's growable
 * scratch; the call performs no hipMalloc, hipFree or other allocating routine.  A call needs
 * count * (24 * n_words + 40) + 256 bytes (one-word keys: up to 524,284 terms); a larger call
 * returns TPGP_HIP_EINVAL with a message.  Every call on one GPU shares that GPU's array, and
 * calls through one context are serialised by its mutex: calls through two contexts on the
 * same GPU that overlap in time race on the array. */
#define TPGP_HIP_NOALLOC_SCRATCH_BYTES (32u << 20)
tpgp_hip_status tpgp_hip_branches_noalloc(tpgp_hip_context *context, const uint64_t *keys,
                                          const double *values, size_t count, size_t n_words,
                                          size_t n_spins, size_t i, size_t j, unsigned qi,
                                          unsigned qj, double c, double s, double rc, double rs,
                                          int enclose, uint64_t *output_keys,
                                          double *output_values, double *output_radii,
                                          char *error, size_t error_capacity);

/* ============================ Part 2: resident batch ==================================== */

/* A batch holds n_maps independent Pauli maps (one per FID: a powder orientation, a coupling
 * draw, a served request) with a common key width n_words (1 for N<=32, 2 for N<=64).  Each
 * map lives in a fixed slot of three HBM arenas: the committed state plus two working buffers.
 * Rotations enqueue device work without host synchronisation; sync() collects per-rotation
 * records; commit()/rollback() make multi-rotation updates transactional per map. */
typedef struct tpgp_hip_batch tpgp_hip_batch;

/* Extended statuses of the resident API (beyond tpgp_hip_status). */
enum { TPGP_HIP_ECAPACITY = 16 /* a map outgrew its slot; roll back and grow */ };

/* Per-map rotation request.  The term (i, j, qi, qj) acts on sites i != j < n_spins of that map;
 * c, s (and rc, rs when validated) are the host's trigonometric factors for 2*h*duration
 * (engine/src/tpgp_rev.c:484-534).  threshold: drop non-identity merged |v| <= threshold
 * (tpgp_rev.c:721).  log: also record dropped values and merged radii (for bit-exact ledgers). */
typedef struct {
    int active;
    unsigned qi, qj;
    int validated;
    int log;
    size_t i, j;
    double c, s, rc, rs;
    double threshold;
} tpgp_hip_rotation;

/* Per-map, per-rotation record filled by sync() (inputs of the ledger of tpgp_rev.c:730-757). */
typedef struct {
    uint64_t len_in, len_out;  /* map length before/after the rotation */
    uint64_t dropped;          /* number of truncated (ledger) terms */
    double l1_up;              /* upward-rounded sum of |dropped| (deterministic tree order) */
    double sq_up;              /* upward-rounded sum of dropped^2 */
    double rad_sq_up;          /* upward-rounded sum of merged radius^2 (validated mode) */
    uint64_t radii;            /* number of nonzero merged radii */
    uint32_t executed;         /* 0: the rotation was a no-op (empty map) */
    uint32_t nonfinite;        /* a raw value, merged value or radius was not finite */
    uint32_t overflow;         /* output exceeded the map's slot capacity */
    uint32_t rotation;         /* caller's rotation sequence number */
    uint64_t in_place;         /* terms of the classes the rotation left where they were in the
                                  slot (neither read nor written; 0 unless it appended) */
} tpgp_hip_record;

/* One dropped coefficient or merged radius of a logged rotation.  key is in the reference
 * layout (site s at bits 2(s%32) of word s/32); kind 0 = dropped |v| (value = |v|, and the
 * signed v in aux), kind 1 = nonzero merged radius. */
typedef struct {
    uint64_t key[2];
    double value;
    double aux;
    uint32_t map, rotation, kind, pad;
} tpgp_hip_log_entry;

tpgp_hip_batch *tpgp_hip_batch_create(tpgp_hip_context *context, size_t n_words, size_t n_maps,
                                      const size_t *n_spins, const size_t *capacities,
                                      char *error, size_t error_capacity);
void tpgp_hip_batch_destroy(tpgp_hip_batch *batch);
/* Replace map `map`'s committed state by a canonical sorted map (reference key order and layout,
 * zero padding, finite nonzero values).  Synchronous. */
int tpgp_hip_batch_upload(tpgp_hip_batch *batch, size_t map, const uint64_t *keys,
                          const double *values, size_t len, char *error, size_t error_capacity);
/* Re-purpose slot `map` for a map of n_spins spins (same key width as the batch) and upload
 * its canonical state as for tpgp_hip_batch_upload.  Pending work of that map is discarded. */
int tpgp_hip_batch_assign(tpgp_hip_batch *batch, size_t map, size_t n_spins,
                          const uint64_t *keys, const double *values, size_t len, char *error,
                          size_t error_capacity);
/* Copy map `map`'s committed state out in reference order.  *len receives the length; arrays
 * must hold `capacity` entries (a NULL array with capacity 0 only queries the length). */
int tpgp_hip_batch_download(tpgp_hip_batch *batch, size_t map, uint64_t *keys, double *values,
                            size_t capacity, size_t *len, char *error, size_t error_capacity);
/* Enqueue one rotation on every map whose request is active (requests[n_maps]).  rotation is a
 * caller sequence number echoed in the records.  Asynchronous. */
int tpgp_hip_batch_rotate(tpgp_hip_batch *batch, const tpgp_hip_rotation *requests,
                          uint32_t rotation, char *error, size_t error_capacity);
/* Enqueue n_rot rotations in order (requests[r*n_maps + m], sequence numbers rotations[r]) with
 * one descriptor upload for the whole sequence.  Same semantics as n_rot rotate() calls. */
int tpgp_hip_batch_rotate_many(tpgp_hip_batch *batch, const tpgp_hip_rotation *requests,
                               const uint32_t *rotations, size_t n_rot, char *error,
                               size_t error_capacity);
/* Enqueue a readout of the n_spins single-site coefficients of symbol `pauli` (1..3) for every
 * map with active[m] != 0 (NULL = all), in the current working state.  Asynchronous. */
int tpgp_hip_batch_readout(tpgp_hip_batch *batch, unsigned pauli, const int *active,
                           char *error, size_t error_capacity);
/* ---- Part 3: the Liouvillian on the device ----------------------------------------------
 * A Hamiltonian term, same layout as tpgp_term (tpgp_rev.h:30-35). */
typedef struct {
    size_t i, j;
    unsigned qi, qj;
    double h;
    size_t layer;
} tpgp_hip_term;
/* -i[H, A] of the map (keys, values, count) with H = sum_t h_t Q_t, bit for bit as
 * tpgp_liouvillian (tpgp_rev.c:810-857): sorted unique keys with nonzero values into
 * out_keys / out_values and the length into *out_count.  n_spins <= 64.  A result longer than
 * out_capacity returns TPGP_HIP_ECAPACITY with the length in *out_count and writes nothing.
 * ENUMERIC: a non-finite contribution or partial sum; EINVAL: an invalid term (whichever the
 * reference meets first in term order).  Synchronous; device memory is allocated per call. */
int tpgp_hip_liouvillian(tpgp_hip_context *context, const uint64_t *keys, const double *values,
                         size_t count, size_t n_words, size_t n_spins,
                         const tpgp_hip_term *terms, size_t n_terms, uint64_t *out_keys,
                         double *out_values, size_t out_capacity, size_t *out_count,
                         char *error, size_t error_capacity);

/* No counterpart in the reference's GPU backend: the interval vectors of the independent endpoint
 * (residual) bound, built on the device.  The reference builds them on its CPU in 256-bit MPFR
 * (engine/src/residual_rev.c:60-114, 181-187):
 *   D1 = 6 b - 6 a - 4 dt L a - 2 dt L b - dt^2 L L a,
 *   D2 = 6 b - 6 a - 2 dt L a - 4 dt L b + dt^2 L L b,
 * L the Liouvillian generator on sparse interval vectors (incident_terms = 1: the form of
 * tpgp_residual_bound, which passes over [0, 0] entries and h = 0 terms; 0: every entry against
 * every term, as tpgp_residual_bound_reference).  Every entry's ends are the reference's bits:
 * the same correctly rounded 256-bit products and sums (down for lower ends, up for upper ends)
 * in the same order, and the entries come in the reference's first-insertion order.
 * With each vector come, per entry, what the reference's Gram enclosure adds up in entry order
 * (residual_rev.c:122-153): the index of the same key in the other vector (partner), the
 * outward-rounded square terms of the entry (the max of the upward squares of its ends, and the
 * min of the downward squares or +0 when the interval holds 0), and for D1 the outward-rounded
 * product terms with its partner in D2 (min and max of the four directed end products).  Only
 * the serial sums remain for the caller.
 * a and b are maps (sorted keys, finite binary64 values) of one spin count; terms as in
 * tpgp_hip_liouvillian; dt finite and positive.  On success out[0] = D1 and out[1] = D2, each
 * with arrays allocated by the library (free with tpgp_hip_interval_vector_free).  EINVAL: bad
 * arguments or an invalid term; ENOMEM: device or host memory, or more than 2^31 contributions
 * in one generator application.  Synchronous; device memory is allocated per call. */
typedef struct {
    uint64_t m[4]; /* significand, m[3] most significant, top bit set unless the value is 0 */
    int32_t e;     /* value = (-1)^neg * (m / 2^256) * 2^e */
    uint32_t neg;
} tpgp_hip_f256;
typedef struct {
    size_t len;
    uint64_t *keys;               /* len * n_words, in first-insertion order */
    tpgp_hip_f256 *lower, *upper; /* the entries' intervals */
    uint32_t *partner;            /* the entry of the same key in the other vector, or UINT32_MAX */
    tpgp_hip_f256 *square[2];     /* Gram terms of the entry's square: [0] lower, [1] upper */
    tpgp_hip_f256 *product[2];    /* D1: terms of the product with its partner (where it has
                                     one); NULL in D2 */
} tpgp_hip_interval_vector;
int tpgp_hip_residual_vectors(tpgp_hip_context *context, const uint64_t *a_keys,
                              const double *a_values, size_t a_len, const uint64_t *b_keys,
                              const double *b_values, size_t b_len, size_t n_words,
                              size_t n_spins, const tpgp_hip_term *terms, size_t n_terms,
                              double dt, int incident_terms, tpgp_hip_interval_vector out[2],
                              char *error, size_t error_capacity);
void tpgp_hip_interval_vector_free(tpgp_hip_interval_vector *vector);

/* Sum of squares of map `map`'s committed values as an unevaluated pair out[0] + out[1]
 * (error-free double-double accumulation, relative error ~len * 2^-104).  Synchronous; the map
 * must have no uncommitted work.  Feeds the l2 norm of tpgp_map_l2 (tpgp_rev.c:266-271). */
int tpgp_hip_batch_sum_squares(tpgp_hip_batch *batch, size_t map, double out[2], char *error,
                               size_t error_capacity);
/* Wait for enqueued work.  Returns the number of records per map since the previous sync in
 * *n_records (records[m*max_records + r], r < n_records), readouts in readouts[m*max_readouts*64
 * + q*64 + s] with *n_readouts, and working lengths in lengths[m].  Any output pointer may be
 * NULL.  Logged entries are returned through log (up to log_capacity) with *n_log. */
int tpgp_hip_batch_sync(tpgp_hip_batch *batch, tpgp_hip_record *records, size_t max_records,
                        size_t *n_records, double *readouts, size_t max_readouts,
                        size_t *n_readouts, uint64_t *lengths, tpgp_hip_log_entry *log,
                        size_t log_capacity, size_t *n_log, char *error, size_t error_capacity);
/* After sync: make the working state of maps with mask[m] != 0 (NULL = all) their committed
 * state; roll back maps with mask[m] == 0. */
int tpgp_hip_batch_commit(tpgp_hip_batch *batch, const int *mask, char *error,
                          size_t error_capacity);
/* Discard all work since the last commit (maps return to their committed state). */
int tpgp_hip_batch_rollback(tpgp_hip_batch *batch, char *error, size_t error_capacity);
/* Grow map `map`'s slot to at least `capacity` terms, keeping its committed state. */
int tpgp_hip_batch_reserve(tpgp_hip_batch *batch, size_t map, size_t capacity, char *error,
                           size_t error_capacity);
/* Grow several slots at once (capacities[m] terms; 0 or less than the current capacity keeps
 * the slot): one re-layout of the arenas for all of them.  Committed states are kept; work
 * since the last commit is discarded. */
int tpgp_hip_batch_reserve_many(tpgp_hip_batch *batch, const size_t *capacities, char *error,
                                size_t error_capacity);
/* No counterpart in the reference (its maps live in host memory): set slot capacities exactly
 * (capacities[m] terms, 0 keeps the slot), shrinking as well as growing, in one re-layout.  A
 * capacity below a map's committed length is rejected (TPGP_HIP_EINVAL) with nothing changed;
 * on TPGP_HIP_ENOMEM every slot keeps its previous capacity.  Committed states are kept; work
 * since the last commit is discarded. */
int tpgp_hip_batch_resize(tpgp_hip_batch *batch, const size_t *capacities, char *error,
                          size_t error_capacity);
/* Grow the device log of dropped values and radii to `entries` (sync reports an overflow as
 * TPGP_HIP_ECAPACITY with the needed count in *n_log: roll back, reserve, run again). */
int tpgp_hip_batch_reserve_log(tpgp_hip_batch *batch, size_t entries, char *error,
                               size_t error_capacity);
/* No counterpart in the reference (its CPU engine launches nothing): step graphs.  A caller
 * that enqueues the same step again and again (readouts and rotate_many calls between two
 * syncs) records it once as a HIP graph and replays it with one launch:
 *   tpgp_hip_batch_graph_replay: if a graph recorded under `key` starts from the batch's
 *     present host state (every map's buffers, digit layout, floor and record counters, and the
 *     same device allocations), launch it and advance that state as the step's planning would
 *     (*replayed = 1).  Otherwise *replayed = 0, and *capture (optional) is 1 when this state
 *     has been seen before under `key`, i.e. a recording is likely to be replayed.
 *   tpgp_hip_batch_graph_begin / _end: record the readouts and rotations enqueued in between
 *     (nothing runs while recording; the context is held by the recording thread); _end with
 *     keep != 0 instantiates the graph and launches it once (*captured = 1).  With keep == 0
 *     (a call failed while recording), or if the graph could not be made, the maps return to
 *     their state at _begin and *captured = 0: nothing of the step ran, enqueue it again
 *     without recording.  Other batch calls are refused while recording.
 * A graph's descriptors live in its own device buffer; graphs are dropped whenever device
 * buffers move (re-layout, log growth).  Results are those of enqueueing the step directly.
 * TPGP_HIP_GRAPHS=0 disables recording and replay; so do validation and profiling. */
int tpgp_hip_batch_graph_replay(tpgp_hip_batch *batch, uint64_t key, int *replayed,
                                int *capture, char *error, size_t error_capacity);
int tpgp_hip_batch_graph_begin(tpgp_hip_batch *batch, uint64_t key, char *error,
                               size_t error_capacity);
int tpgp_hip_batch_graph_end(tpgp_hip_batch *batch, int keep, int *captured, char *error,
                             size_t error_capacity);
/* Committed length and slot capacity of a map. */
int tpgp_hip_batch_info(const tpgp_hip_batch *batch, size_t map, size_t *committed_len,
                        size_t *capacity);
/* Stream the batch enqueues on (a hipStream_t), for callers that time or capture it. */
void *tpgp_hip_batch_stream(const tpgp_hip_batch *batch);
/* Kernel timing with HIP events recorded on the batch stream around every launch of the kinds
 * selected by `mask` (bit k = kind k; 0 disables, TPGP_HIP_PROFILE_ALL selects all).  The call
 * resets the totals; times accumulate at each sync.  Events cost a little time per launch, so
 * a throughput measurement selects only the kinds it reports. */
enum {
    TPGP_HIP_PROFILE_MERGE = 0,     /* rotation write pass: merge and write the outputs */
    TPGP_HIP_PROFILE_SCATTER = 1,   /* partition scatter */
    TPGP_HIP_PROFILE_ROTATION = 2,  /* whole rotation pass */
    TPGP_HIP_PROFILE_PARTITION = 3, /* whole partition pass */
    TPGP_HIP_PROFILE_COUNT = 4,     /* rotation count pass (splits, counts, ledger sums) */
    TPGP_HIP_PROFILE_SCAN = 5,      /* rotation offsets, regions, ledger fold */
    TPGP_HIP_PROFILE_PLAN = 6,      /* rotation tile plan */
    TPGP_HIP_PROFILE_HISTOGRAM = 7, /* partition histogram */
    TPGP_HIP_PROFILE_SPLITS = 8,    /* rotation merge-path splits */
    TPGP_HIP_PROFILE_SMALL = 9,     /* small-map persistent kernel (one chunk of passes) */
    TPGP_HIP_PROFILE_ALL = 0x3ff
};
int tpgp_hip_batch_profile(tpgp_hip_batch *batch, unsigned mask);
int tpgp_hip_batch_profile_read(const tpgp_hip_batch *batch, unsigned kind,
                                double *milliseconds, uint64_t *launches);

/* Launch-shape tuning.  No counterpart in the reference (its CPU path has no launch shapes);
 * the project's autotuner (ddevol_amd/autotune.py) writes these.  Each parameter holds one value
 * per key width (`words` = 1 or 2 key words) and size bucket, bucket = floor(log2(terms the pass
 * reads)) clamped to [0, TPGP_HIP_TUNE_BUCKETS); value 0 restores the built-in default.  Every value gives the same
 * keys, values and ledgers: the parameters only change how a pass is cut into workgroups.
 *   PARTITION_ROUNDS: elements per thread of the partition (the change of active pair) in
 *     coalesced rounds of 256: 4, 8 or 12, i.e. tiles of 1024, 2048 or 3072 terms; default 8.
 *   COUNT_GRID, WRITE_GRID: workgroups of the rotation's count and write passes in sixths of
 *     the estimated tile count (plus 64), 1..24; defaults 3 (a half) and 4 (two thirds).
 * tpgp_hip_create loads a table from the file named by the environment variable
 * TPGP_HIP_TUNING (empty: the built-in defaults), or, when it is unset, from
 * tuning_gfx950.txt in the directory above the library's own (lines "words param bucket value",
 * '#' starts a comment; an invalid line fails the create).  A setting applies to the passes enqueued after it.  EINVAL
 * for a key width, parameter, bucket or value out of range. */
enum {
    TPGP_HIP_TUNE_PARTITION_ROUNDS = 0,
    TPGP_HIP_TUNE_COUNT_GRID = 1,
    TPGP_HIP_TUNE_WRITE_GRID = 2,
    TPGP_HIP_TUNE_PARAMS = 3,
    TPGP_HIP_TUNE_BUCKETS = 48
};
int tpgp_hip_set_tuning(tpgp_hip_context *context, unsigned words, unsigned param,
                        unsigned bucket, unsigned value);
/* The value in force (the default where none is set). */
int tpgp_hip_get_tuning(tpgp_hip_context *context, unsigned words, unsigned param,
                        unsigned bucket, unsigned *value);

#ifdef __cplusplus
}
#endif
#endif
