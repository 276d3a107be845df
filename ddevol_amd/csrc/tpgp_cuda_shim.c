/* tpgp_cuda_shim -- the reference's CUDA branch-backend symbols, routed to the MI355X library.
 *
 * The reference links its GPU backend at one plug point (engine/cuda/CMakeLists.txt:10-33):
 * either engine/src/cuda_rev.cu or the stub engine/cuda/stub_rev.c, both exporting the five
 * functions of engine/include/tpgp_cuda_rev.h:20-49.  This file exports the same five symbols
 * with the same C signatures and forwards each to its tpgp_hip_* counterpart, so a program
 * built against tpgp_cuda_rev.h (the reference's driver, cli_rev.c:477-485, or its tests) runs
 * on the MI355X by linking libtpgp_cuda_shim.so instead of the stub.
 *
 * The context pointer handed out is the tpgp_hip_context itself: the reference core API of
 * libtpgp_engine.so reads it from the options field at the offset of the reference's
 * `cuda_context` (tpgp_rev.h:44), so a context made here selects the resident GPU engine too.
 *
 * Statuses: tpgp_hip_status has the values of tpgp_cuda_status (OK, EINVAL, ENOMEM, EDEVICE,
 * ENUMERIC = 0..4), so they pass through unchanged.
 */
#include <stddef.h>
#include <stdint.h>

#include "../../include/tpgp_hip.h"

/* The reference's declarations (tpgp_cuda_rev.h:9-17), restated: an opaque context and an
 * int-valued status enum. */
typedef struct tpgp_cuda_context tpgp_cuda_context;
typedef int tpgp_cuda_status;

int tpgp_cuda_compiled(void);
tpgp_cuda_context *tpgp_cuda_create(int device_index, char *error, size_t error_capacity);
void tpgp_cuda_destroy(tpgp_cuda_context *context);
const char *tpgp_cuda_device_name(const tpgp_cuda_context *context);
tpgp_cuda_status tpgp_cuda_branches(tpgp_cuda_context *context, const uint64_t *keys,
                                    const double *values, size_t count, size_t n_words,
                                    size_t n_spins, size_t i, size_t j, unsigned qi, unsigned qj,
                                    double c, double s, double rc, double rs, int enclose,
                                    uint64_t *output_keys, double *output_values,
                                    double *output_radii, char *error, size_t error_capacity);

int tpgp_cuda_compiled(void) { return tpgp_hip_compiled(); }

tpgp_cuda_context *tpgp_cuda_create(int device_index, char *error, size_t error_capacity) {
    return (tpgp_cuda_context *)tpgp_hip_create(device_index, error, error_capacity);
}

void tpgp_cuda_destroy(tpgp_cuda_context *context) {
    tpgp_hip_destroy((tpgp_hip_context *)context);
}

const char *tpgp_cuda_device_name(const tpgp_cuda_context *context) {
    return tpgp_hip_device_name((const tpgp_hip_context *)context);
}

tpgp_cuda_status tpgp_cuda_branches(tpgp_cuda_context *context, const uint64_t *keys,
                                    const double *values, size_t count, size_t n_words,
                                    size_t n_spins, size_t i, size_t j, unsigned qi, unsigned qj,
                                    double c, double s, double rc, double rs, int enclose,
                                    uint64_t *output_keys, double *output_values,
                                    double *output_radii, char *error, size_t error_capacity) {
    return (tpgp_cuda_status)tpgp_hip_branches(
        (tpgp_hip_context *)context, keys, values, count, n_words, n_spins, i, j, qi, qj, c, s,
        rc, rs, enclose, output_keys, output_values, output_radii, error, error_capacity);
}
