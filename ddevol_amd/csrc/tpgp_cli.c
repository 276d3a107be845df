/* tpgp_rev: the FID run driver on the MI355X (C17).
 *
 * Same command line, CSV, checkpoint files, stop reasons and exit codes as the reference's
 * driver (engine/src/cli_rev.c), so campaign scripts switch by changing the binary:
 *   models ring | j1j2 | ising | xy | cubic | edges       (cli_rev.c:108-132, 181-246)
 *   initial state sum_s X_s ("global") or X_0 ("local")   (cli_rev.c:514-532)
 *   grid check: END must be an integer multiple of STEP   (cli_rev.c:442-450)
 *   one CSV row per accepted step with the 16 columns     (cli_rev.c:557-608)
 *   checkpoint cadence, walltime/signal pauses (exit 75), accuracy stops (exit 76),
 *   failure checkpoints, restart checks                   (cli_rev.c:27-62, 349-425, 609-676)
 * The map stays resident in HBM for the whole run (tpgp_evolution_*); it is copied to the host
 * only when a checkpoint is written.  The ledger is the exact (bit-identical) one.
 *
 * Differences, all refused loudly rather than approximated:
 *   --backend takes hip (or cuda, the slot the GPU occupies); the CPU and Metal backends do not
 *   exist here.  A restart continues on the GPU whatever backend its record names, and records
 *   written afterwards name the GPU.
 * --residual-depth D adds the independent endpoint bound of every step (tpgp_residual_bound,
 * host MPFR) from the maps before and after it, which are then copied out once per step
 * (cli_rev.c:640-667).
 */
#define _POSIX_C_SOURCE 200809L
#include "../../include/tpgp_engine.h"
#include "../../include/tpgp_hip.h"

#include <errno.h>
#include <fcntl.h>
#include <float.h>
#include <inttypes.h>
#include <math.h>
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/file.h>
#include <sys/resource.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

enum { EXIT_DONE = 0, EXIT_FAILED = 1, EXIT_USAGE = 2, EXIT_PAUSED = 75, EXIT_ACCURACY = 76 };
enum { STOP_NONE, STOP_COMPLETE, STOP_WALLTIME, STOP_SIGNAL, STOP_RESOURCE, STOP_NUMERICAL,
       STOP_ACCURACY };

/* Options whose explicit presence is checked against a restart record, in the bit order of
 * cli_rev.c:63-74. */
static const char *const kChecked[] = {
    "--model",     "--initial",   "--n",        "--linear-size", "--cap",       "--threads",
    "--residual-depth", "--dt",   "--time",     "--j2",          "--threshold", "--drop-budget",
    "--backend",   "--validated", "--edges",    "--max-fid-bound"};
enum { F_MODEL, F_INITIAL, F_N, F_L, F_CAP, F_THREADS, F_RESIDUAL, F_DT, F_TIME, F_J2,
       F_THRESHOLD, F_BUDGET, F_BACKEND, F_VALIDATED, F_EDGES, F_MAX_FID };

typedef struct {
    const char *model, *initial, *output, *edges, *checkpoint, *restart;
    size_t n, L, every;
    double dt, T, j2, checkpoint_seconds, walltime, max_fid_bound;
    int residual_depth, device;
    unsigned given; /* bit F_* set when the option was on the command line */
    tpgp_options o;
} config;

typedef struct {
    tpgp_term *data;
    size_t len, cap;
} term_list;

static volatile sig_atomic_t g_stop = 0;
static void on_signal(int sig) {
    (void)sig;
    g_stop = 1;
}

static double now_seconds(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}
static double peak_rss_mb(void) {
    struct rusage r;
    return getrusage(RUSAGE_SELF, &r) ? NAN : (double)r.ru_maxrss / 1024.0;
}

static void usage(FILE *f) {
    fputs("Usage: tpgp_rev [--model ring|j1j2|ising|xy|cubic|edges] [--n N]\n"
          "  --dt STEP --time END --threshold X --cap TERMS --drop-budget X\n"
          "  --backend hip --device D --threads N --validated\n"
          "  --j2 RATIO --linear-size L --edges PATH --initial global|local\n"
          "  --output FILE (default stdout)\n"
          "  --checkpoint FILE --checkpoint-every STEPS --checkpoint-seconds SECONDS\n"
          "  --restart FILE --walltime SECONDS (stops only at accepted step boundaries)\n"
          "  --max-fid-bound X (stop if the best complete FID bound exceeds X)\n"
          "  A checkpointed pause exits 75; a completed trajectory exits 0.\n"
          "  An accuracy-limit stop exits 76 and requires a checkpoint path.\n"
          "Runs on the MI355X with the map resident in device memory.\n"
          "Hamiltonian: secular d(2ZZ-XX-YY)/4; Ising d ZZ/4; XY d(XX+YY)/4.\n"
          "Cubic is periodic [001], nearest and face-diagonal shells, lattice units.\n"
          "Edge file rows: i j d (zero-based; secular coupling). No hidden fallback.\n",
          f);
}

/* ---- argument parsing -------------------------------------------------------------------
 * One reader for the two kinds of numeric option value: a finite double (strtod, nothing
 * trailing), or a non-negative decimal count that fits a size_t. */
enum { VALUE_REAL, VALUE_COUNT };
static int read_value(const char *text, int kind, void *out) {
    char *end = NULL;
    errno = 0;
    if (kind == VALUE_REAL) {
        const double v = strtod(text, &end);
        if (errno || end == text || *end || !isfinite(v)) return 0;
        *(double *)out = v;
        return 1;
    }
    if (text[0] == '-') return 0; /* strtoumax would wrap a negative count */
    const uintmax_t v = strtoumax(text, &end, 10);
    if (errno || end == text || *end || v > SIZE_MAX) return 0;
    *(size_t *)out = (size_t)v;
    return 1;
}
#define parse_real(text, out) read_value((text), VALUE_REAL, (out))
#define parse_count(text, out) read_value((text), VALUE_COUNT, (out))

/* 0 ok, 1 --help printed, 2 usage error */
static int parse_args(int argc, char **argv, config *c) {
    for (int k = 1; k < argc; k++) {
        const char *a = argv[k];
        for (unsigned f = 0; f < sizeof(kChecked) / sizeof(kChecked[0]); f++)
            if (!strcmp(a, kChecked[f])) c->given |= 1u << f;
        if (!strcmp(a, "--help")) {
            usage(stdout);
            return 1;
        }
        if (!strcmp(a, "--validated")) {
            c->o.validated = 1;
            continue;
        }
        if (k + 1 == argc) {
            usage(stderr);
            return 2;
        }
        const char *v = argv[++k];
        size_t z;
        int ok = 1;
        if (!strcmp(a, "--model")) c->model = v;
        else if (!strcmp(a, "--initial")) c->initial = v;
        else if (!strcmp(a, "--output")) c->output = v;
        else if (!strcmp(a, "--edges")) c->edges = v;
        else if (!strcmp(a, "--checkpoint")) c->checkpoint = v;
        else if (!strcmp(a, "--restart")) c->restart = v;
        else if (!strcmp(a, "--checkpoint-every")) ok = parse_count(v, &c->every) && c->every;
        else if (!strcmp(a, "--checkpoint-seconds"))
            ok = parse_real(v, &c->checkpoint_seconds) && c->checkpoint_seconds >= 0;
        else if (!strcmp(a, "--walltime")) ok = parse_real(v, &c->walltime) && c->walltime >= 0;
        else if (!strcmp(a, "--max-fid-bound"))
            ok = parse_real(v, &c->max_fid_bound) && c->max_fid_bound > 0;
        else if (!strcmp(a, "--n")) ok = parse_count(v, &c->n);
        else if (!strcmp(a, "--linear-size")) ok = parse_count(v, &c->L);
        else if (!strcmp(a, "--cap")) ok = parse_count(v, &c->o.max_terms);
        else if (!strcmp(a, "--threads")) {
            ok = parse_count(v, &z) && z && z <= 65536;
            c->o.threads = (unsigned)z;
        } else if (!strcmp(a, "--residual-depth")) {
            ok = parse_count(v, &z) && z <= 10;
            c->residual_depth = (int)z;
        } else if (!strcmp(a, "--device")) {
            ok = parse_count(v, &z) && z < 64;
            c->device = (int)z;
        } else if (!strcmp(a, "--dt")) ok = parse_real(v, &c->dt);
        else if (!strcmp(a, "--time")) ok = parse_real(v, &c->T);
        else if (!strcmp(a, "--j2")) ok = parse_real(v, &c->j2);
        else if (!strcmp(a, "--threshold")) ok = parse_real(v, &c->o.threshold);
        else if (!strcmp(a, "--drop-budget")) ok = parse_real(v, &c->o.max_drop_l2);
        else if (!strcmp(a, "--backend")) {
            if (strcmp(v, "hip") && strcmp(v, "cuda")) {
                fprintf(stderr, "Backend %s is not available: this driver runs on the MI355X "
                                "(--backend hip).\n", v);
                return 2;
            }
            c->o.backend = TPGP_HIP;
        } else {
            fprintf(stderr, "Unknown option: %s\n", a);
            return 2;
        }
        if (!ok) return 2;
    }
    return 0;
}

/* ---- Hamiltonians (cli_rev.c:108-132, 181-246) ------------------------------------------- */
static int push_term(term_list *l, size_t i, size_t j, unsigned qi, unsigned qj, double h) {
    if (h == 0) return 1; /* zero contributions are never stored */
    if (l->len == l->cap) {
        const size_t cap = l->cap ? 2 * l->cap : 64;
        if (cap > SIZE_MAX / sizeof(tpgp_term)) return 0;
        tpgp_term *p = realloc(l->data, cap * sizeof(tpgp_term));
        if (!p) return 0;
        l->data = p;
        l->cap = cap;
    }
    l->data[l->len] = (tpgp_term){i, j, qi, qj, h, l->len};
    l->len++;
    return 1;
}
/* The two-site Pauli terms of one bond of coupling d, per model, in the order they are applied
 * (cli_rev.c:120-127): Ising d ZZ/4; XY d(XX + YY)/4; otherwise the secular dipolar form
 * d(2ZZ - XX - YY)/4 as ZZ, XX, YY.  Scaling by 1/2 and 1/4 is exact, so d * factor equals the
 * reference's d / 2 and -d / 4 bit for bit. */
typedef struct {
    const char *model;
    unsigned parts;
    struct {
        unsigned q;
        double factor;
    } part[3];
} bond_form;
static const bond_form kBondForms[] = {
    {"ising", 1, {{3, 0.25}}},
    {"xy", 2, {{1, 0.25}, {2, 0.25}}},
    {NULL, 3, {{3, 0.5}, {1, -0.25}, {2, -0.25}}}, /* secular: every other model */
};
static int push_bond(term_list *l, size_t i, size_t j, double d, const char *model) {
    const bond_form *f = kBondForms;
    while (f->model && strcmp(f->model, model)) f++;
    for (unsigned k = 0; k < f->parts; k++)
        if (!push_term(l, i, j, f->part[k].q, f->part[k].q, d * f->part[k].factor)) return 0;
    return 1;
}
static long min_image(long d, size_t L) {
    if (d > (long)L / 2) d -= (long)L;
    if (d < -(long)L / 2) d += (long)L;
    return d;
}
static int cubic_bonds(term_list *l, size_t n, size_t L) {
    if (L < 3 || L > SIZE_MAX / L || L * L > SIZE_MAX / L || L * L * L != n) return 0;
    for (size_t i = 0; i < n; i++)
        for (size_t j = i + 1; j < n; j++) {
            const long dx = min_image((long)(j % L) - (long)(i % L), L);
            const long dy = min_image((long)((j / L) % L) - (long)((i / L) % L), L);
            const long dz = min_image((long)(j / (L * L)) - (long)(i / (L * L)), L);
            const long r2 = dx * dx + dy * dy + dz * dz;
            if (r2 != 1 && r2 != 2) continue;
            const double d = (1 - 3.0 * (double)(dz * dz) / (double)r2) /
                             ((double)r2 * sqrt((double)r2));
            if (!push_bond(l, i, j, d, "secular")) return 0;
        }
    return 1;
}
static int file_bonds(term_list *l, size_t n, const char *path) {
    FILE *f = path ? fopen(path, "r") : NULL;
    if (!f) return 0;
    char line[512];
    int ok = 1;
    while (ok && fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        size_t i, j;
        double d;
        char trailing;
        ok = sscanf(line, "%zu %zu %lf %c", &i, &j, &d, &trailing) == 3 && i < n && j < n &&
             i != j && isfinite(d) && push_bond(l, i, j, d, "secular");
    }
    ok = ok && !ferror(f);
    fclose(f);
    return ok;
}
static int build_terms(term_list *l, const config *c) {
    if (!strcmp(c->model, "cubic")) return cubic_bonds(l, c->n, c->L);
    if (!strcmp(c->model, "edges")) return file_bonds(l, c->n, c->edges);
    const int j1j2 = !strcmp(c->model, "j1j2");
    if (!j1j2 && strcmp(c->model, "ring") && strcmp(c->model, "ising") && strcmp(c->model, "xy"))
        return 0;
    if (c->n < 3 || (j1j2 && c->n < 5)) return 0;
    for (size_t i = 0; i < c->n; i++) {
        if (!push_bond(l, i, (i + 1) % c->n, 1, c->model)) return 0;
        if (j1j2 && !push_bond(l, i, (i + 2) % c->n, c->j2, c->model)) return 0;
    }
    return 1;
}

/* ---- checkpoint lock (cli_rev.c:38-62) ----------------------------------------------------
 * A run that writes checkpoints holds an exclusive advisory lock (flock, as the reference's
 * driver takes it, so the two programs exclude each other) on "<checkpoint>.lock" for its whole
 * life, and never replaces a record that it did not restart from. */
static int lock_beside(const char *path) {
    const size_t n = strlen(path);
    char *lock_path = malloc(n + sizeof(".lock"));
    if (!lock_path) return -1;
    memcpy(lock_path, path, n);
    memcpy(lock_path + n, ".lock", sizeof(".lock"));
    int fd = open(lock_path, O_CREAT | O_RDWR, 0600);
    free(lock_path);
    if (fd >= 0 && flock(fd, LOCK_EX | LOCK_NB) != 0) {
        close(fd);
        fd = -1;
    }
    return fd;
}
static int take_lock(const char *checkpoint, const char *restart) {
    const int fd = lock_beside(checkpoint);
    if (fd < 0) {
        fprintf(stderr, "Checkpoint is locked by another run or cannot be locked.\n");
        return -1;
    }
    struct stat st;
    const int exists = lstat(checkpoint, &st) == 0;
    const int resuming_it = restart && !strcmp(checkpoint, restart);
    if (exists && !resuming_it) {
        fprintf(stderr, "Refusing to overwrite an existing checkpoint without its restart.\n");
        close(fd);
        return -1;
    }
    return fd;
}

/* Explicit options that contradict the restart record (cli_rev.c:366-399). */
static int conflicts_with_record(const config *c, const tpgp_checkpoint *r) {
    const unsigned g = c->given;
    const int global = !strcmp(c->initial, "global");
    return ((g >> F_MODEL & 1) && strcmp(c->model, r->model)) ||
           ((g >> F_INITIAL & 1) && global != r->initial_global) ||
           ((g >> F_N & 1) && c->n != r->map.n_spins) ||
           ((g >> F_L & 1) && c->L != r->linear_size) ||
           ((g >> F_CAP & 1) && c->o.max_terms != 0 &&
            (r->options.max_terms == 0 || c->o.max_terms < r->options.max_terms)) ||
           ((g >> F_THREADS & 1) && c->o.threads != r->options.threads) ||
           ((g >> F_RESIDUAL & 1) && c->residual_depth != r->residual_depth) ||
           ((g >> F_DT & 1) && c->dt != r->dt) ||
           ((g >> F_TIME & 1) && c->T < (double)r->total_steps * r->dt) ||
           ((g >> F_J2 & 1) && c->j2 != r->j2) ||
           ((g >> F_THRESHOLD & 1) && c->o.threshold != r->options.threshold) ||
           ((g >> F_BUDGET & 1) && c->o.max_drop_l2 != r->options.max_drop_l2) ||
           ((g >> F_VALIDATED & 1) && c->o.validated != r->options.validated);
}

/* ---- the run ---------------------------------------------------------------------------- */
typedef struct {
    const config *c;
    const term_list *terms;
    const char *model;
    size_t n, steps, first_step, prior_steps, prior_cap;
    int global;
} run_info;

static tpgp_status write_checkpoint(const run_info *ri, const tpgp_map *map,
                                    const tpgp_options *o, const tpgp_ledger *led, double pf,
                                    double residual, size_t step, int reason, tpgp_status last) {
    tpgp_checkpoint s;
    memset(&s, 0, sizeof(s));
    s.map = *map;
    s.terms = ri->terms->data;
    s.term_count = ri->terms->len;
    s.options = *o;
    s.options.hip_context = NULL;
    s.ledger = *led;
    s.dt = ri->c->dt;
    s.pf_bound = pf;
    s.residual_bound = residual;
    s.max_fid_bound = ri->c->max_fid_bound;
    s.step = step;
    s.total_steps = ri->steps;
    s.residual_depth = ri->c->residual_depth;
    s.initial_global = ri->global;
    snprintf(s.model, sizeof(s.model), "%s", ri->model);
    s.j2 = ri->c->j2;
    s.linear_size = ri->c->L;
    s.resumed_from_step = ri->first_step;
    s.prior_total_steps = ri->prior_steps;
    s.prior_max_terms = ri->prior_cap;
    s.stop_reason = reason;
    s.last_status = last;
    char error[512];
    const tpgp_status st = tpgp_checkpoint_save(ri->c->checkpoint, &s, error, sizeof(error));
    fprintf(stderr, "checkpoint step=%zu/%zu path=%s status=%s stop_reason=%d\n", step,
            ri->steps, ri->c->checkpoint, error, reason);
    return st;
}

/* The FID error allowance of one row (cli_rev.c:584-602): product-formula part and total. */
static void row_bounds(double pf, const tpgp_ledger *led, size_t n, size_t input_terms,
                       double norm0, double radius, double *pf_obs, double *total) {
    const double dual = sqrt((double)n) / (double)input_terms;
    *pf_obs = pf * norm0 * dual;
    *total = *pf_obs + (led->dropped_l2 + led->arithmetic_l2) * dual;
    if (!tpgp_has_mpfr()) return;
    const double loss = tpgp_bound_add_up(led->dropped_l2, led->arithmetic_l2);
    if (input_terms == n) {
        *pf_obs = pf;
        *total = tpgp_bound_add_up(pf, tpgp_bound_fid_up(loss, n));
    } else {
        *pf_obs = pf == 0 ? 0 : nextafter(tpgp_bound_fid_up(pf, n) * (double)n, INFINITY);
        *total = nextafter(tpgp_bound_fid_up(tpgp_bound_add_up(pf, loss), n) * (double)n,
                           INFINITY);
    }
    *total = tpgp_bound_add_up(*total, radius);
}

int main(int argc, char **argv) {
    config c = {.model = "ring", .initial = "global", .n = 6, .L = 3, .every = 100,
                .dt = .01, .T = .1, .j2 = .3, .residual_depth = -1};
    c.o = tpgp_default_options();
    c.o.max_terms = 1000000;
    const int parsed = parse_args(argc, argv, &c);
    if (parsed) return parsed == 1 ? EXIT_DONE : EXIT_USAGE;
    if (c.restart && !c.checkpoint) c.checkpoint = c.restart;
    int lock = -1;
    if (c.checkpoint && (lock = take_lock(c.checkpoint, c.restart)) < 0) return EXIT_USAGE;

    tpgp_checkpoint rec;
    memset(&rec, 0, sizeof(rec));
    size_t prior_steps = 0, prior_cap = 0;
    int code = EXIT_USAGE;
    term_list terms = {0};
    tpgp_map m = {0}, before = {0}, after = {0}; /* before/after: around each step (residual) */
    tpgp_evolution *ev = NULL;
    tpgp_hip_context *ctx = NULL;
    FILE *out = NULL;
    if (c.restart) {
        char error[512];
        if (tpgp_checkpoint_load(c.restart, &rec, error, sizeof(error))) {
            fprintf(stderr, "Restart rejected: %s\n", error);
            goto done;
        }
        if (conflicts_with_record(&c, &rec)) {
            fprintf(stderr, "Restart rejected: explicitly supplied model/numerical options "
                            "differ from the checkpoint.\n");
            goto done;
        }
        /* the record's configuration wins, except explicit horizon / cap increases */
        const size_t cap = c.o.max_terms;
        prior_steps = rec.total_steps;
        prior_cap = rec.options.max_terms;
        c.model = rec.model;
        c.initial = rec.initial_global ? "global" : "local";
        c.n = rec.map.n_spins;
        c.L = rec.linear_size;
        c.dt = rec.dt;
        if (!(c.given >> F_TIME & 1)) c.T = (double)rec.total_steps * c.dt;
        c.j2 = rec.j2;
        c.o = rec.options;
        c.o.backend = TPGP_HIP;
        c.o.hip_context = NULL;
        c.o.metal_context = NULL;
        if (c.given >> F_CAP & 1) c.o.max_terms = cap;
        if (!(c.given >> F_MAX_FID & 1)) c.max_fid_bound = rec.max_fid_bound;
        c.residual_depth = rec.residual_depth;
    }
    if (((c.walltime > 0 || c.max_fid_bound > 0) && !c.checkpoint) ||
        (c.checkpoint && c.output && !strcmp(c.checkpoint, c.output))) {
        fprintf(stderr, "Walltime/accuracy stops require a checkpoint path, distinct from CSV.\n");
        goto done;
    }
    if (!strcmp(c.model, "cubic")) {
        if (c.L < 3 || c.L > 10000) goto done;
        c.n = c.L * c.L * c.L;
    }
    if (!c.n || c.dt <= 0 || c.T < 0 || c.o.threshold < 0 || c.o.max_drop_l2 < 0 ||
        (strcmp(c.initial, "local") && strcmp(c.initial, "global"))) {
        usage(stderr);
        goto done;
    }
    if (c.n > 64) {
        fprintf(stderr, "This driver holds keys of at most 64 spins (two 64-bit words).\n");
        goto done;
    }
    const double ratio = c.T / c.dt, whole = round(ratio);
    if (!isfinite(ratio) || whole >= (double)SIZE_MAX ||
        fabs(ratio - whole) > 32 * DBL_EPSILON * fmax(1, fabs(ratio))) {
        fprintf(stderr,
                "END must be an integer multiple of STEP; no endpoint is silently shortened.\n");
        goto done;
    }
    const size_t steps = c.restart && !(c.given >> F_TIME & 1) ? rec.total_steps : (size_t)whole;
    if (c.o.validated && !tpgp_has_mpfr()) {
        fprintf(stderr, "Validated mode requires MPFR at build time.\n");
        goto done;
    }
    if (c.residual_depth >= 0 && (!tpgp_has_mpfr() || strcmp(c.initial, "global"))) {
        fprintf(stderr, "Residual CLI bounds require MPFR and global initial sum Xi.\n");
        goto done;
    }
    if (c.max_fid_bound > 0 &&
        (!tpgp_has_mpfr() || (!c.o.validated && c.residual_depth < 0))) {
        fprintf(stderr, "An accuracy stop requires a complete MPFR-enclosed error bound.\n");
        goto done;
    }
    {
        char error[512];
        ctx = tpgp_hip_create(c.device, error, sizeof(error));
        if (!ctx) {
            fprintf(stderr, "HIP unavailable: %s\n", error);
            goto done;
        }
        c.o.hip_context = ctx;
        c.o.backend = TPGP_HIP;
    }
    if (c.restart && !c.edges) { /* the saved Hamiltonian */
        terms.data = rec.terms;
        terms.len = terms.cap = rec.term_count;
        rec.terms = NULL;
        rec.term_count = 0;
    } else if (!build_terms(&terms, &c)) {
        fprintf(stderr, "Invalid model geometry/coupling file or allocation failure.\n");
        goto done;
    }
    if (c.restart && c.edges) { /* a coupling file on restart must reproduce the record */
        tpgp_checkpoint asked = rec;
        asked.terms = terms.data;
        asked.term_count = terms.len;
        if (!tpgp_checkpoint_same_configuration(&rec, &asked)) {
            fprintf(stderr, "Restart rejected: coupling file differs from saved Hamiltonian.\n");
            goto done;
        }
    }
    const int global = !strcmp(c.initial, "global");
    const size_t input_terms = global ? c.n : 1;
    tpgp_status status = TPGP_OK;
    if (c.restart) {
        m = rec.map;
        memset(&rec.map, 0, sizeof(rec.map));
    } else {
        status = tpgp_map_init(&m, c.n);
        for (size_t s = 0; s < input_terms && !status; s++)
            status = tpgp_map_add_single(&m, s, 1, 1);
    }
    if (status || (c.o.max_terms && m.len > c.o.max_terms)) {
        fprintf(stderr, "Initial state failed: %s\n",
                status ? tpgp_status_string(status) : "active-term cap exceeded");
        goto done;
    }
    out = c.output ? fopen(c.output, "wx") : stdout;
    if (!out) {
        perror(c.output);
        goto done;
    }
    tpgp_ledger led;
    tpgp_ledger_init(&led, c.o.validated);
    led.peak_terms = m.len;
    if (c.restart) led = rec.ledger;
    const tpgp_evolve_options eo = {.exact_ledger = 1, .pauli = 1, .normalizer = input_terms};
    status = tpgp_evolution_open(&ev, &m, terms.data, terms.len, c.dt, &c.o, &eo);
    if (status) {
        fprintf(stderr, "Device session failed: %s: %s\n", tpgp_status_string(status),
                tpgp_hip_last_error());
        code = EXIT_FAILED;
        goto done;
    }
    const run_info ri = {&c, &terms, c.model, c.n, steps, c.restart ? rec.step : 0,
                         prior_steps, prior_cap, global};
    const double start = now_seconds(), per_step = tpgp_strang_bound(terms.data, terms.len, c.dt),
                 norm0 = sqrt((double)input_terms);
    double pf = c.restart ? rec.pf_bound : 0, last_save_time = start;
    double residual = c.residual_depth < 0 ? NAN : c.restart ? rec.residual_bound : 0;
    tpgp_ledger before_ledger = led;
    int rolled_back = 0, have_after = 0;
    if (c.residual_depth >= 0 &&
        ((status = tpgp_map_init(&before, c.n)) || (status = tpgp_map_init(&after, c.n)))) {
        code = EXIT_FAILED;
        goto done;
    }
    size_t accepted = ri.first_step, last_saved = SIZE_MAX;
    int paused = 0, accuracy = 0;
    signal(SIGINT, on_signal);
    signal(SIGTERM, on_signal);
    fprintf(out,
            "# model=%s,n=%zu,initial=%s,hamiltonian_terms=%zu,dt=%.17g,threshold=%.17g,"
            "backend=%d,threads=%u,validated=%d\n",
            c.model, c.n, c.initial, terms.len, c.dt, c.o.threshold, c.o.backend, c.o.threads,
            c.o.validated);
    fputs("# Initial coefficients are exactly one; normalization is only at readout. Bounds "
          "refer to binary64 Hamiltonian inputs.\n", out);
    fprintf(out,
            "# restart_step=%zu,prior_total_steps=%zu,total_steps=%zu,prior_cap=%zu,cap=%zu,"
            "max_fid_bound=%.17g,prior_max_fid_bound=%.17g\n",
            ri.first_step, prior_steps, steps, prior_cap, c.o.max_terms, c.max_fid_bound,
            c.restart ? rec.max_fid_bound : 0);
    fprintf(out,
            "# Row k represents exact physical time k times the binary64 dt input; the time "
            "column is its rounded display value. dt_hex=%a\n", c.dt);
    fputs("time,fid,active,l2_norm,drop_l1,drop_l2_sum,drop_l2_squared,arithmetic_l2,"
          "pf_observable_bound,total_observable_bound,residual_raw_bound,residual_fid_bound,"
          "readout_radius,arithmetic_enclosed,wall_seconds,peak_rss_mb\n", out);
    /* sample[0]: the readout of the map the next row describes */
    double fid[2], radius[2];
    size_t active[2], done_steps = 0;
    status = tpgp_evolution_run(ev, 0, &led, fid, radius, active, &done_steps);
    for (size_t step = ri.first_step; !status && step <= steps; step++) {
        double l2 = 0;
        if ((status = tpgp_evolution_l2(ev, &l2))) break;
        double pf_obs, total;
        row_bounds(pf, &led, c.n, input_terms, norm0, radius[0], &pf_obs, &total);
        const double residual_fid =
            c.residual_depth >= 0
                ? tpgp_bound_add_up(tpgp_bound_fid_up(residual, c.n), radius[0])
                : NAN;
        fprintf(out,
                "%.17g,%.17g,%zu,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%d,"
                "%.6f,%.3f\n",
                step * c.dt, fid[0], active[0], l2, led.dropped_l1, led.dropped_l2,
                led.dropped_l2_squared, led.arithmetic_l2, pf_obs, total, residual, residual_fid,
                radius[0], led.arithmetic_enclosed, now_seconds() - start, peak_rss_mb());
        if (fflush(out)) {
            status = TPGP_EINVAL;
            break;
        }
        double best = c.o.validated && led.arithmetic_enclosed ? total : INFINITY;
        if (c.residual_depth >= 0 && residual_fid < best) best = residual_fid;
        const int reason = c.max_fid_bound > 0 && best > c.max_fid_bound     ? STOP_ACCURACY
                           : step == steps                                 ? STOP_COMPLETE
                           : g_stop                                        ? STOP_SIGNAL
                           : c.walltime > 0 && now_seconds() - start >= c.walltime ? STOP_WALLTIME
                                                                           : STOP_NONE;
        const int due = reason || step % c.every == 0 ||
                        (c.checkpoint_seconds > 0 &&
                         now_seconds() - last_save_time >= c.checkpoint_seconds);
        if (c.checkpoint && due && last_saved != step) {
            if ((status = tpgp_evolution_snapshot(ev, &m))) break;
            if ((status = write_checkpoint(&ri, &m, &c.o, &led, pf, residual, step, reason,
                                           TPGP_OK)))
                break;
            last_saved = step;
            last_save_time = now_seconds();
        }
        if (reason) {
            paused = reason == STOP_WALLTIME || reason == STOP_SIGNAL;
            accuracy = reason == STOP_ACCURACY;
            break;
        }
        const size_t active_before = active[0];
        if (c.residual_depth >= 0) {
            before_ledger = led;
            if (!have_after && (status = tpgp_evolution_snapshot(ev, &after))) break;
            tpgp_map t = before; /* `before` = the map now (the last step's `after`) */
            before = after;
            after = t;
            have_after = 0;
        }
        status = tpgp_evolution_run(ev, 1, &led, fid, radius, active, &done_steps);
        if (status) {
            fprintf(stderr, "Step %zu rejected: %s; last accepted time %.17g, active=%zu.\n",
                    step + 1, tpgp_status_string(status), step * c.dt, active_before);
            break;
        }
        if (c.residual_depth >= 0) {
            double increment = 0;
            status = tpgp_evolution_snapshot(ev, &after);
            have_after = !status;
            if (!status)
                status = tpgp_residual_bound(&before, &after, terms.data, terms.len, c.dt,
                                             (unsigned)c.residual_depth, &increment);
            if (status) {
                fprintf(stderr, "Independent endpoint bound failed after step %zu: %s.\n",
                        step + 1, tpgp_status_string(status));
                led = before_ledger; /* the step is not accepted */
                rolled_back = 1;
                break;
            }
            residual = tpgp_bound_add_up(residual, increment);
        }
        fid[0] = fid[1];
        radius[0] = radius[1];
        active[0] = active[1];
        pf = tpgp_bound_add_up(pf, per_step);
        accepted = step + 1;
    }
    if (status && c.checkpoint) {
        const int reason = status == TPGP_ECAP || status == TPGP_ENOMEM || status == TPGP_EBUDGET
                               ? STOP_RESOURCE
                               : STOP_NUMERICAL;
        if (rolled_back) /* the map before the step whose bound failed */
            (void)write_checkpoint(&ri, &before, &c.o, &led, pf, residual, accepted, reason,
                                   status);
        else if (!tpgp_evolution_snapshot(ev, &m))
            (void)write_checkpoint(&ri, &m, &c.o, &led, pf, residual, accepted, reason, status);
    }
    fprintf(stderr,
            "status=%s accepted_rotations=%zu peak_terms=%zu elapsed=%.3fs peak_rss=%.3fMiB "
            "arithmetic_enclosed=%d completion=%s step=%zu/%zu\n",
            tpgp_status_string(status), led.rotations, led.peak_terms, now_seconds() - start,
            peak_rss_mb(), led.arithmetic_enclosed,
            status ? "failed" : accuracy ? "accuracy_limit" : paused ? "paused" : "complete",
            accepted, steps);
    if (ferror(out)) status = TPGP_EINVAL;
    code = status ? EXIT_FAILED : accuracy ? EXIT_ACCURACY : paused ? EXIT_PAUSED : EXIT_DONE;
done:
    if (out && out != stdout && fclose(out) && code == EXIT_DONE) code = EXIT_FAILED;
    if (ev) tpgp_evolution_close(ev, NULL);
    if (ctx) tpgp_hip_destroy(ctx);
    free(terms.data);
    tpgp_map_free(&m);
    tpgp_map_free(&before);
    tpgp_map_free(&after);
    tpgp_checkpoint_free(&rec);
    if (lock >= 0) close(lock);
    return code;
}
