// tpgp_generator: the Liouvillian -i[H, A] of a Pauli-basis map on the MI355X (gfx950).
//
// Same result, bit for bit, as the reference's sequential construction (tpgp_rev.c:810-857):
// every term q (sites i, j) and key P that anticommute contribute v = (2 h x)(+-1) to the key
// q xor P (sign + when the phase of qP is 1), x the coefficient of P, and the reference adds the
// contributions into its map in term-major, key-minor order.  Floating-point addition is
// order-dependent, so the device reproduces that order exactly:
//
//   count   per (term, block of 256 keys): the number of anticommuting keys;
//   scan    exclusive prefix over (term, block) in term-major order: every contribution's slot;
//   emit    each contribution at its slot (term-major, key-minor order);
//   sort    stable LSD radix sort by key (64-bit words, lowest word first): contributions of one
//           key stay in emission order;
//   sum     one thread per run of equal keys adds its contributions left to right; a run that
//           sums to exactly zero leaves no key (as the reference's removal on a zero sum; a later
//           insert restarts from zero exactly);
//   compact the kept runs, already in the reference's key order.
//
// Statuses follow the reference's term order: terms are processed up to the first invalid one;
// a non-finite contribution or partial sum among them is ENUMERIC, otherwise the invalid term is
// EINVAL.  Work is O(len * terms) and memory (8W + 8) * 3 bytes per contribution.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/tpgp_hip.h"
#include "tpgp_hip_context.h"

namespace {

constexpr int kGenBlock = 256;

struct GenTerm {  // device copy of a term: sites, symbols and 2h
    uint32_t i, j, qi, qj;
    double h2;
};

__device__ __forceinline__ unsigned gen_symbol(const uint64_t *key, uint32_t site) {
    return (unsigned)(key[site >> 5] >> (2 * (site & 31))) & 3u;
}
__device__ __forceinline__ bool gen_anti(unsigned q, unsigned p) { return q && p && q != p; }
__device__ __forceinline__ unsigned gen_phase(unsigned q, unsigned p) {
    // exponent of i in the product QP (I0 X1 Y2 Z3): X->Y->Z->X cyclic gives +i
    if (!q || !p || q == p) return 0;
    return ((p + 3 - q) % 3 == 1) ? 1u : 3u;
}

// Per (term, block): anticommuting keys of the block.  counts[t * nblocks + block].
template <int W>
__global__ __launch_bounds__(kGenBlock) void k_gen_count(const uint64_t *keys, uint64_t len,
                                                         const GenTerm *terms, uint32_t n_terms,
                                                         uint64_t *counts) {
    const uint64_t k = (uint64_t)blockIdx.x * kGenBlock + threadIdx.x;
    uint64_t key[W] = {};
    if (k < len)
        for (int w = 0; w < W; ++w) key[w] = keys[k * W + w];
    for (uint32_t t = 0; t < n_terms; ++t) {
        const GenTerm q = terms[t];
        const bool flag = k < len && gen_anti(q.qi, gen_symbol(key, q.i)) !=
                                         gen_anti(q.qj, gen_symbol(key, q.j));
        const int c = __syncthreads_count(flag);
        if (threadIdx.x == 0) counts[(uint64_t)t * gridDim.x + blockIdx.x] = (uint64_t)c;
    }
}

// Every contribution at its slot: offsets[t * nblocks + block] plus its rank in the block.
template <int W>
__global__ __launch_bounds__(kGenBlock) void k_gen_emit(const uint64_t *keys, const double *values,
                                                        uint64_t len, const GenTerm *terms,
                                                        uint32_t n_terms, const uint64_t *offsets,
                                                        uint64_t *out_keys, double *out_values,
                                                        uint32_t *nonfinite) {
    __shared__ uint32_t wave_count[kGenBlock / 64];
    const uint64_t k = (uint64_t)blockIdx.x * kGenBlock + threadIdx.x;
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t key[W] = {};
    double x = 0.0;
    if (k < len) {
        for (int w = 0; w < W; ++w) key[w] = keys[k * W + w];
        x = values[k];
    }
    bool bad = false;
    for (uint32_t t = 0; t < n_terms; ++t) {
        const GenTerm q = terms[t];
        const unsigned a = gen_symbol(key, q.i), b = gen_symbol(key, q.j);
        const bool flag = k < len && gen_anti(q.qi, a) != gen_anti(q.qj, b);
        const uint64_t ballot = __ballot(flag);
        if (lane == 0) wave_count[wave] = (uint32_t)__popcll(ballot);
        __syncthreads();
        if (flag) {
            uint64_t at = offsets[(uint64_t)t * gridDim.x + blockIdx.x];
            for (unsigned w = 0; w < wave; ++w) at += wave_count[w];
            at += (uint64_t)__popcll(ballot & ((1ull << lane) - 1ull));
            const unsigned ph = (gen_phase(q.qi, a) + gen_phase(q.qj, b)) & 3u;
            const double v = q.h2 * x * (ph == 1 ? 1.0 : -1.0);
            bad = bad || !__builtin_isfinite(v);
            uint64_t nk[W];
            for (int w = 0; w < W; ++w) nk[w] = key[w];
            nk[q.i >> 5] ^= (uint64_t)q.qi << (2 * (q.i & 31));
            nk[q.j >> 5] ^= (uint64_t)q.qj << (2 * (q.j & 31));
            for (int w = 0; w < W; ++w) out_keys[at * W + w] = nk[w];
            out_values[at] = v;
        }
        __syncthreads();  // wave_count is rewritten for the next term
    }
    if (bad) atomicOr(nonfinite, 1u);
}

// Split interleaved keys (W words each) into one array per word.
template <int W>
__global__ void k_gen_split(const uint64_t *keys, uint64_t n, uint64_t *w0, uint64_t *w1) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        w0[i] = keys[i * W];
        if (W > 1) w1[i] = keys[i * W + 1];
    }
}
__global__ void k_gen_iota(uint32_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}
__global__ void k_gen_gather_u64(const uint64_t *src, const uint32_t *idx, uint64_t *dst,
                                 uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}
__global__ void k_gen_gather_f64(const double *src, const uint32_t *idx, double *dst,
                                 uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

// Run heads of the sorted keys (word arrays w0, w1): 1 where a key differs from its predecessor.
template <int W>
__global__ void k_gen_heads(const uint64_t *w0, const uint64_t *w1, uint64_t n, uint32_t *head) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        head[i] = i == 0 || w0[i] != w0[i - 1] || (W > 1 && w1[i] != w1[i - 1]);
}
// Start of every run (run index = inclusive count of heads - 1).
__global__ void k_gen_starts(const uint32_t *head, const uint64_t *run_of, uint64_t n,
                             uint64_t *start) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (head[i]) start[run_of[i] - 1] = i;
}
// One thread per run: the contributions added left to right in emission order; keep[r] when the
// sum is not exactly zero.
__global__ void k_gen_sum(const double *v, const uint64_t *start, uint64_t runs, uint64_t n,
                          double *sum, uint32_t *keep, uint32_t *nonfinite) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < runs;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e = r + 1 < runs ? start[r + 1] : n;
        double s = 0.0;
        bool bad = false;
        for (uint64_t x = start[r]; x < e; ++x) {
            s = s + v[x];
            bad = bad || !__builtin_isfinite(s);
        }
        sum[r] = s;
        keep[r] = s != 0.0;
        if (bad) atomicOr(nonfinite, 1u);
    }
}
// Kept runs to their output positions (exclusive count of kept runs before them).
template <int W>
__global__ void k_gen_compact(const uint64_t *w0, const uint64_t *w1, const uint64_t *start,
                              const double *sum, const uint32_t *keep, const uint64_t *pos,
                              uint64_t runs, uint64_t *out_keys, double *out_values) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < runs;
         r += (uint64_t)gridDim.x * blockDim.x) {
        if (!keep[r]) continue;
        const uint64_t at = pos[r], i = start[r];
        out_keys[at * W] = w0[i];
        if (W > 1) out_keys[at * W + 1] = w1[i];
        out_values[at] = sum[r];
    }
}

// Device allocations of one call, released together.
struct Arena {
    std::vector<void *> blocks;
    hipError_t err = hipSuccess;
    template <class T> T *get(size_t n) {
        void *p = nullptr;
        if (err == hipSuccess) err = hipMalloc(&p, (n ? n : 1) * sizeof(T));
        if (err != hipSuccess) return nullptr;
        blocks.push_back(p);
        return static_cast<T *>(p);
    }
    ~Arena() {
        for (void *p : blocks) (void)hipFree(p);
    }
};

unsigned grid_for(uint64_t n) {
    const uint64_t g = (n + 255) / 256;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

template <int W>
int liouvillian(tpgp_hip_context *ctx, const uint64_t *keys, const double *values, size_t count,
                size_t n_spins, const GenTerm *h_terms, uint32_t n_valid, uint64_t *out_keys,
                double *out_values, size_t out_capacity, size_t *out_count, int *numeric,
                char *error, size_t cap) {
    hipStream_t st = ctx->stream;
    Arena A;
    hipError_t e;
    const uint64_t nblocks = (count + kGenBlock - 1) / kGenBlock;
    *numeric = 0;
    *out_count = 0;
    if (!count || !n_valid) return TPGP_HIP_OK;
    uint64_t *d_keys = A.get<uint64_t>(count * W);
    double *d_vals = A.get<double>(count);
    GenTerm *d_terms = A.get<GenTerm>(n_valid);
    uint64_t *counts = A.get<uint64_t>(nblocks * n_valid);
    uint64_t *offsets = A.get<uint64_t>(nblocks * n_valid);
    uint32_t *flag = A.get<uint32_t>(2);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    if ((e = hipMemcpyAsync(d_keys, keys, count * W * 8, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_vals, values, count * 8, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_terms, h_terms, n_valid * sizeof(GenTerm), hipMemcpyHostToDevice,
                            st)) ||
        (e = hipMemsetAsync(flag, 0, 2 * sizeof(uint32_t), st)))
        return drain_fail(st, e, "HIP generator upload", error, cap);
    hipLaunchKernelGGL(k_gen_count<W>, dim3((unsigned)nblocks), dim3(kGenBlock), 0, st, d_keys,
                       (uint64_t)count, d_terms, n_valid, counts);
    const uint64_t ncells = nblocks * n_valid;
    if (ncells >= (1ull << 31)) {
        msg(error, cap, "Generator scan too large (%llu count cells).",
            (unsigned long long)ncells);
        return TPGP_HIP_ENOMEM;
    }
    size_t tmp_bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, offsets, (int)ncells, st);
    void *tmp = A.get<unsigned char>(tmp_bytes);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, offsets, (int)ncells, st);
    uint64_t last[2];
    if ((e = hipMemcpyAsync(&last[0], offsets + ncells - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(&last[1], counts + ncells - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return hip_fail(e, "HIP generator count", error, cap);
    const uint64_t C = last[0] + last[1];  // contributions
    if (C == 0) return TPGP_HIP_OK;
    if (C >= (1ull << 31)) {
        msg(error, cap, "Generator scan of %llu contributions exceeds 2^31.",
            (unsigned long long)C);
        return TPGP_HIP_ENOMEM;
    }
    uint64_t *ck = A.get<uint64_t>(C * W);
    double *cv = A.get<double>(C);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipLaunchKernelGGL(k_gen_emit<W>, dim3((unsigned)nblocks), dim3(kGenBlock), 0, st, d_keys,
                       d_vals, (uint64_t)count, d_terms, n_valid, offsets, ck, cv, flag);
    // stable LSD sort by key: the low word, then (two words) the high word -- contributions of
    // one key keep their emission order
    uint64_t *lo = A.get<uint64_t>(C), *hi = W > 1 ? A.get<uint64_t>(C) : nullptr;
    uint64_t *keys_a = A.get<uint64_t>(C), *lo_sorted = A.get<uint64_t>(C);
    uint64_t *hi_sorted = W > 1 ? A.get<uint64_t>(C) : nullptr;
    uint32_t *perm0 = A.get<uint32_t>(C), *perm1 = A.get<uint32_t>(C);
    uint32_t *perm2 = W > 1 ? A.get<uint32_t>(C) : nullptr;
    double *vs = A.get<double>(C);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    const unsigned g = grid_for(C);
    const int n = (int)C;
    hipLaunchKernelGGL(k_gen_split<W>, dim3(g), dim3(256), 0, st, ck, C, lo, hi);
    hipLaunchKernelGGL(k_gen_iota, dim3(g), dim3(256), 0, st, perm0, C);
    const int lo_bits = W == 1 ? (int)(2 * n_spins) : 64;
    const int hi_bits = W == 1 ? 0 : (int)(2 * (n_spins - 32));
    size_t sort_bytes = 0, sort_bytes2 = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, lo, lo_sorted, perm0, perm1, n, 0,
                                       lo_bits, st);
    if (W > 1)
        hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes2, keys_a, hi_sorted, perm1, perm2,
                                           n, 0, hi_bits, st);
    void *stmp = A.get<unsigned char>(sort_bytes > sort_bytes2 ? sort_bytes : sort_bytes2);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipcub::DeviceRadixSort::SortPairs(stmp, sort_bytes, lo, lo_sorted, perm0, perm1, n, 0,
                                       lo_bits, st);
    const uint32_t *order = perm1;  // emission index of every sorted contribution
    if (W > 1) {
        hipLaunchKernelGGL(k_gen_gather_u64, dim3(g), dim3(256), 0, st, hi, perm1, keys_a, C);
        hipcub::DeviceRadixSort::SortPairs(stmp, sort_bytes2, keys_a, hi_sorted, perm1, perm2, n,
                                           0, hi_bits, st);
        hipLaunchKernelGGL(k_gen_gather_u64, dim3(g), dim3(256), 0, st, lo, perm2, lo_sorted, C);
        order = perm2;
    }
    hipLaunchKernelGGL(k_gen_gather_f64, dim3(g), dim3(256), 0, st, cv, order, vs, C);
    const uint64_t *w0s = lo_sorted, *w1s = hi_sorted;
    // sorted words: w0s (low) and, for two words, w1s (high)
    uint32_t *head = A.get<uint32_t>(C);
    uint64_t *run_of = A.get<uint64_t>(C);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipLaunchKernelGGL(k_gen_heads<W>, dim3(g), dim3(256), 0, st, w0s, w1s, C, head);
    size_t scan_bytes = 0;
    hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, head, run_of, n, st);
    void *rtmp = A.get<unsigned char>(scan_bytes);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipcub::DeviceScan::InclusiveSum(rtmp, scan_bytes, head, run_of, n, st);
    uint64_t runs = 0;
    if ((e = hipMemcpyAsync(&runs, run_of + C - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return hip_fail(e, "HIP generator runs", error, cap);
    uint64_t *start = A.get<uint64_t>(runs), *pos = A.get<uint64_t>(runs);
    double *sum = A.get<double>(runs);
    uint32_t *keep = A.get<uint32_t>(runs);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipLaunchKernelGGL(k_gen_starts, dim3(g), dim3(256), 0, st, head, run_of, C, start);
    const unsigned gr = grid_for(runs);
    hipLaunchKernelGGL(k_gen_sum, dim3(gr), dim3(256), 0, st, vs, start, runs, C, sum, keep,
                       flag);
    size_t keep_bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, keep_bytes, keep, pos, (int)runs, st);
    void *ktmp = A.get<unsigned char>(keep_bytes);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipcub::DeviceScan::ExclusiveSum(ktmp, keep_bytes, keep, pos, (int)runs, st);
    uint64_t last_pos = 0;
    uint32_t last_keep = 0, bad = 0;
    if ((e = hipMemcpyAsync(&last_pos, pos + runs - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(&last_keep, keep + runs - 1, 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return hip_fail(e, "HIP generator sums", error, cap);
    if (bad) {
        *numeric = 1;
        return TPGP_HIP_OK;
    }
    const uint64_t kept = last_pos + last_keep;
    *out_count = kept;
    if (kept > out_capacity) {
        msg(error, cap, "The result holds %llu terms; the output buffers hold %zu.",
            (unsigned long long)kept, out_capacity);
        return TPGP_HIP_ECAPACITY;
    }
    uint64_t *ok_keys = A.get<uint64_t>(kept * W);
    double *ok_vals = A.get<double>(kept);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipLaunchKernelGGL(k_gen_compact<W>, dim3(gr), dim3(256), 0, st, w0s, w1s, start, sum, keep,
                       pos, runs, ok_keys, ok_vals);
    if ((e = hipGetLastError()) != hipSuccess)
        return hip_fail(e, "HIP generator launch", error, cap);
    if (kept && ((e = hipMemcpyAsync(out_keys, ok_keys, kept * W * 8, hipMemcpyDeviceToHost,
                                     st)) ||
                 (e = hipMemcpyAsync(out_values, ok_vals, kept * 8, hipMemcpyDeviceToHost, st))))
        return hip_fail(e, "HIP generator download", error, cap);
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "HIP generator", error, cap);
    return TPGP_HIP_OK;
}

}  // namespace

extern "C" int tpgp_hip_liouvillian(tpgp_hip_context *ctx, const uint64_t *keys,
                                    const double *values, size_t count, size_t n_words,
                                    size_t n_spins, const tpgp_hip_term *terms, size_t n_terms,
                                    uint64_t *out_keys, double *out_values, size_t out_capacity,
                                    size_t *out_count, char *error, size_t cap) {
    msg(error, cap, "");
    if (!ctx || !out_count || !n_spins || n_spins > 64 || n_words != (n_spins + 31) / 32 ||
        (count && (!keys || !values)) || (n_terms && !terms) ||
        (out_capacity && (!out_keys || !out_values))) {
        msg(error, cap, "Invalid generator arguments (at most 64 spins).");
        return TPGP_HIP_EINVAL;
    }
    // terms up to the first invalid one (the reference checks each term as it reaches it)
    std::vector<GenTerm> t;
    size_t valid = 0;
    for (; valid < n_terms; ++valid) {
        const tpgp_hip_term &q = terms[valid];
        if (q.i >= n_spins || q.j >= n_spins || q.i == q.j || q.qi > 3 || q.qj > 3 ||
            !std::isfinite(q.h))
            break;
        t.push_back(GenTerm{(uint32_t)q.i, (uint32_t)q.j, q.qi, q.qj, 2 * q.h});
    }
    std::lock_guard<std::recursive_mutex> lock(ctx->mutex);
    DeviceGuard guard(ctx->device);
    if (!guard.ok) {
        msg(error, cap, "HIP device selection failed.");
        return TPGP_HIP_EDEVICE;
    }
    int numeric = 0;
    const int s = n_words == 1
                      ? liouvillian<1>(ctx, keys, values, count, n_spins, t.data(),
                                       (uint32_t)valid, out_keys, out_values, out_capacity,
                                       out_count, &numeric, error, cap)
                      : liouvillian<2>(ctx, keys, values, count, n_spins, t.data(),
                                       (uint32_t)valid, out_keys, out_values, out_capacity,
                                       out_count, &numeric, error, cap);
    if (s) return s;
    if (numeric) {
        msg(error, cap, "Non-finite Liouvillian contribution or sum.");
        *out_count = 0;
        return TPGP_HIP_ENUMERIC;
    }
    if (valid < n_terms) {
        msg(error, cap, "Invalid Hamiltonian term %zu.", valid);
        *out_count = 0;
        return TPGP_HIP_EINVAL;
    }
    return TPGP_HIP_OK;
}
