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
//   sort    stable LSD radix sort by key, 4 bits a pass from the lowest digit of the low word
//           to the highest of the high word: contributions of one key stay in emission order;
//   sum     one thread per run of equal keys adds its contributions left to right; a run that
//           sums to exactly zero leaves no key (as the reference's removal on a zero sum; a later
//           insert restarts from zero exactly);
//   compact the kept runs, already in the reference's key order.
//
// Statuses follow the reference's term order: terms are processed up to the first invalid one;
// a non-finite contribution or partial sum among them is ENUMERIC, otherwise the invalid term is
// EINVAL.  Work is O(len * terms) and memory (8W + 8) * 3 bytes per contribution.
//
// The scans and the sort are this file's own kernels: a tiled three-kernel scan (tile sums,
// one workgroup over the tile sums, tiles with their carry-in) and a stable scatter whose ranks
// come from wave ballots on the digit's bits (lanes with the same digit, lower lanes first),
// wave counts in LDS and a running per-digit base across the tile's rounds.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/tpgp_hip.h"
#include "tpgp_f256.h"
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

// ---- scans: out[i] = sum of in[0..i) (exclusive) or in[0..i] (inclusive), 64-bit sums -------
constexpr int kScanThreads = 256, kScanItems = 8;
constexpr uint64_t kScanTile = (uint64_t)kScanThreads * kScanItems;

// Inclusive scan of one value per thread over the workgroup (wave scans by shuffles, then the
// wave totals); *total receives the workgroup's sum.
__device__ uint64_t block_inclusive(uint64_t x, uint64_t *total) {
    __shared__ uint64_t wsum[kScanThreads / 64];
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (unsigned d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint64_t before = 0, all = 0;
    for (unsigned w = 0; w < kScanThreads / 64; ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();  // wsum is reused by the next call
    *total = all;
    return x + before;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const T *in, uint64_t n,
                                                             uint64_t *tile_sum) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint64_t x = 0;
    for (int r = 0; r < kScanItems; ++r) {
        const uint64_t i = base + (uint64_t)r * kScanThreads + threadIdx.x;
        if (i < n) x += (uint64_t)in[i];
    }
    uint64_t total;
    (void)block_inclusive(x, &total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}
// One workgroup: exclusive scan of the tile sums in place.
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(uint64_t *tile_sum, uint64_t n) {
    uint64_t carry = 0;
    for (uint64_t base = 0; base < n; base += kScanThreads) {
        const uint64_t i = base + threadIdx.x;
        const uint64_t x = i < n ? tile_sum[i] : 0;
        uint64_t total;
        const uint64_t inc = block_inclusive(x, &total);
        if (i < n) tile_sum[i] = carry + inc - x;
        carry += total;
    }
}
// Each tile from its carry-in: element order within the tile is round-major (round r covers
// elements base + r * 256 .. + 255), so a round's scan continues the previous round's.
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const T *in, uint64_t n,
                                                             const uint64_t *tile_sum,
                                                             uint64_t *out, int inclusive) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint64_t carry = tile_sum[blockIdx.x];
    for (int r = 0; r < kScanItems; ++r) {
        const uint64_t i = base + (uint64_t)r * kScanThreads + threadIdx.x;
        const uint64_t x = i < n ? (uint64_t)in[i] : 0;
        uint64_t total;
        const uint64_t inc = block_inclusive(x, &total);
        if (i < n) out[i] = carry + (inclusive ? inc : inc - x);
        carry += total;
    }
}

// ---- stable radix sort of (key, emission index) pairs, 4 bits a pass ------------------------
constexpr int kSortThreads = 256, kSortRounds = 8;
constexpr uint64_t kSortTile = (uint64_t)kSortThreads * kSortRounds;

template <int W>
__device__ __forceinline__ unsigned digit_of(const uint64_t *keys, uint64_t i, unsigned word,
                                             unsigned shift) {
    return (unsigned)(keys[i * W + word] >> shift) & 15u;
}
// Lanes of the wave whose digit equals this lane's (valid lanes only).
__device__ __forceinline__ uint64_t same_digit(unsigned d, bool valid) {
    uint64_t m = __ballot(valid);
    for (unsigned b = 0; b < 4; ++b) {
        const uint64_t set = __ballot(valid && (d >> b & 1u));
        m &= (d >> b & 1u) ? set : ~set;
    }
    return m;
}

// Per tile: how many of its elements have each digit.  hist[d * tiles + tile] (digit-major, so
// one exclusive scan gives every (digit, tile) its first output position).
template <int W>
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint64_t *keys, uint64_t n,
                                                            unsigned word, unsigned shift,
                                                            uint32_t *hist) {
    __shared__ uint32_t cnt[16];
    if (threadIdx.x < 16) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    const unsigned lane = threadIdx.x & 63u;
    for (int r = 0; r < kSortRounds; ++r) {
        const uint64_t i = base + (uint64_t)r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const unsigned d = valid ? digit_of<W>(keys, i, word, shift) : 0u;
        const uint64_t m = same_digit(d, valid);
        // the lowest lane of each digit group counts the group
        if (valid && (m & ((1ull << lane) - 1ull)) == 0) atomicAdd(&cnt[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    if (threadIdx.x < 16) hist[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter: element i goes to first[digit][tile] + (elements of its digit before it in
// the tile).  Within a round, lower waves and lower lanes come first.
template <int W>
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint64_t *keys, const uint32_t *idx, uint64_t n, unsigned word, unsigned shift,
    const uint64_t *first, uint64_t *keys_out, uint32_t *idx_out) {
    constexpr unsigned kWaves = kSortThreads / 64;
    // wave counts by round parity: the digit totals of round r are read (threads 0-15) while
    // other waves may already clear the counts of round r + 1
    __shared__ uint32_t wcnt[2][kWaves][16];
    __shared__ uint64_t run[16];  // elements of each digit placed so far by this tile
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (threadIdx.x < 16) run[threadIdx.x] = first[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        uint32_t (*cnt)[16] = wcnt[r & 1];
        if (lane < 16) cnt[wave][lane] = 0;
        __syncthreads();  // run[] of the previous round is final; this round's counts are clear
        const uint64_t i = base + (uint64_t)r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const unsigned d = valid ? digit_of<W>(keys, i, word, shift) : 0u;
        const uint64_t m = same_digit(d, valid);
        const unsigned below = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        if (valid && below == 0) cnt[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint64_t at = run[d] + below;
            for (unsigned w = 0; w < wave; ++w) at += cnt[w][d];
            for (int k = 0; k < W; ++k) keys_out[at * W + k] = keys[i * W + k];
            idx_out[at] = idx[i];
        }
        __syncthreads();  // every lane has read run[] before it moves on
        if (threadIdx.x < 16) {
            uint64_t add = 0;
            for (unsigned w = 0; w < kWaves; ++w) add += cnt[w][threadIdx.x];
            run[threadIdx.x] += add;
        }
    }
}

__global__ void k_gen_iota(uint32_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}
__global__ void k_gen_gather_f64(const double *src, const uint32_t *idx, double *dst,
                                 uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}
__global__ void k_gen_zero_u32(uint32_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Run heads of the sorted keys (W words each): 1 where a key differs from its predecessor.
template <int W>
__global__ void k_gen_heads(const uint64_t *keys, uint64_t n, uint32_t *head) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        bool h = i == 0;
        for (int k = 0; k < W && !h; ++k) h = keys[i * W + k] != keys[(i - 1) * W + k];
        head[i] = h;
    }
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
__global__ void k_gen_compact(const uint64_t *keys, const uint64_t *start, const double *sum,
                              const uint32_t *keep, const uint64_t *pos, uint64_t runs,
                              uint64_t *out_keys, double *out_values) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < runs;
         r += (uint64_t)gridDim.x * blockDim.x) {
        if (!keep[r]) continue;
        const uint64_t at = pos[r], i = start[r];
        for (int k = 0; k < W; ++k) out_keys[at * W + k] = keys[i * W + k];
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

// out = exclusive (or inclusive) prefix sums of in[0..n), n >= 1.
template <typename T>
hipError_t scan(hipStream_t st, Arena &A, const T *in, uint64_t n, uint64_t *out,
                int inclusive) {
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    uint64_t *sums = A.get<uint64_t>(tiles);
    if (A.err != hipSuccess) return A.err;
    hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, in, n,
                       sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, sums, tiles);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, in, n,
                       sums, out, inclusive);
    return hipGetLastError();
}

// Stable sort of n (key, index) pairs by key: 4-bit passes over the low word's lo_bits, then
// (two words) the high word's hi_bits.  keys/idx hold the input and receive the result (the
// passes alternate with the arena's second buffers).
template <int W>
hipError_t sort_pairs(hipStream_t st, Arena &A, uint64_t *&keys, uint32_t *&idx, uint64_t n,
                      unsigned lo_bits, unsigned hi_bits) {
    const uint64_t tiles = (n + kSortTile - 1) / kSortTile;
    uint32_t *hist = A.get<uint32_t>(16 * tiles);
    uint64_t *first = A.get<uint64_t>(16 * tiles);
    uint64_t *keys2 = A.get<uint64_t>(n * W);
    uint32_t *idx2 = A.get<uint32_t>(n);
    if (A.err != hipSuccess) return A.err;
    for (unsigned word = 0; word < (unsigned)W; ++word) {
        const unsigned bits = word == 0 ? lo_bits : hi_bits;
        for (unsigned shift = 0; shift < bits; shift += 4) {
            hipLaunchKernelGGL(k_sort_hist<W>, dim3((unsigned)tiles), dim3(kSortThreads), 0, st,
                               keys, n, word, shift, hist);
            hipError_t e = scan(st, A, hist, 16 * tiles, first, 0);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_sort_scatter<W>, dim3((unsigned)tiles), dim3(kSortThreads), 0,
                               st, keys, idx, n, word, shift, first, keys2, idx2);
            std::swap(keys, keys2);
            std::swap(idx, idx2);
        }
    }
    return hipGetLastError();
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
                            st)))
        return drain_fail(st, e, "HIP generator upload", error, cap);
    hipLaunchKernelGGL(k_gen_zero_u32, dim3(1), dim3(64), 0, st, flag, 2ull);
    hipLaunchKernelGGL(k_gen_count<W>, dim3((unsigned)nblocks), dim3(kGenBlock), 0, st, d_keys,
                       (uint64_t)count, d_terms, n_valid, counts);
    const uint64_t ncells = nblocks * n_valid;
    if (ncells >= (1ull << 31)) {
        msg(error, cap, "Generator scan too large (%llu count cells).",
            (unsigned long long)ncells);
        return TPGP_HIP_ENOMEM;
    }
    if ((e = scan(st, A, counts, ncells, offsets, 0)) != hipSuccess)
        return drain_fail(st, e, "HIP generator scan", error, cap);
    uint64_t last[2];
    if ((e = hipMemcpyAsync(&last[0], offsets + ncells - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(&last[1], counts + ncells - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return drain_fail(st, e, "HIP generator count", error, cap);
    const uint64_t C = last[0] + last[1];  // contributions
    if (C == 0) return TPGP_HIP_OK;
    if (C >= (1ull << 31)) {
        msg(error, cap, "Generator scan of %llu contributions exceeds 2^31.",
            (unsigned long long)C);
        return TPGP_HIP_ENOMEM;
    }
    uint64_t *ck = A.get<uint64_t>(C * W);
    double *cv = A.get<double>(C);
    uint32_t *order = A.get<uint32_t>(C);  // emission index of every (sorted) contribution
    double *vs = A.get<double>(C);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipLaunchKernelGGL(k_gen_emit<W>, dim3((unsigned)nblocks), dim3(kGenBlock), 0, st, d_keys,
                       d_vals, (uint64_t)count, d_terms, n_valid, offsets, ck, cv, flag);
    const unsigned g = grid_for(C);
    hipLaunchKernelGGL(k_gen_iota, dim3(g), dim3(256), 0, st, order, C);
    // stable by key: contributions of one key keep their emission (term-major) order
    const unsigned lo_bits = W == 1 ? (unsigned)(2 * n_spins) : 64u;
    const unsigned hi_bits = W == 1 ? 0u : (unsigned)(2 * (n_spins - 32));
    if ((e = sort_pairs<W>(st, A, ck, order, C, lo_bits, hi_bits)) != hipSuccess)
        return drain_fail(st, e, "HIP generator sort", error, cap);
    hipLaunchKernelGGL(k_gen_gather_f64, dim3(g), dim3(256), 0, st, cv, order, vs, C);
    uint32_t *head = A.get<uint32_t>(C);
    uint64_t *run_of = A.get<uint64_t>(C);
    if (A.err != hipSuccess) return drain_fail(st, A.err, "HIP generator allocation", error, cap);
    hipLaunchKernelGGL(k_gen_heads<W>, dim3(g), dim3(256), 0, st, ck, C, head);
    if ((e = scan(st, A, head, C, run_of, 1)) != hipSuccess)
        return drain_fail(st, e, "HIP generator scan", error, cap);
    uint64_t runs = 0;
    if ((e = hipMemcpyAsync(&runs, run_of + C - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return drain_fail(st, e, "HIP generator runs", error, cap);
    uint64_t *start = A.get<uint64_t>(runs), *pos = A.get<uint64_t>(runs);
    double *sum = A.get<double>(runs);
    uint32_t *keep = A.get<uint32_t>(runs);
    if (A.err != hipSuccess) return hip_fail(A.err, "HIP generator allocation", error, cap);
    hipLaunchKernelGGL(k_gen_starts, dim3(g), dim3(256), 0, st, head, run_of, C, start);
    const unsigned gr = grid_for(runs);
    hipLaunchKernelGGL(k_gen_sum, dim3(gr), dim3(256), 0, st, vs, start, runs, C, sum, keep,
                       flag);
    if ((e = scan(st, A, keep, runs, pos, 0)) != hipSuccess)
        return drain_fail(st, e, "HIP generator scan", error, cap);
    uint64_t last_pos = 0;
    uint32_t last_keep = 0, bad = 0;
    if ((e = hipMemcpyAsync(&last_pos, pos + runs - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(&last_keep, keep + runs - 1, 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return drain_fail(st, e, "HIP generator sums", error, cap);
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
    hipLaunchKernelGGL(k_gen_compact<W>, dim3(gr), dim3(256), 0, st, ck, start, sum, keep, pos,
                       runs, ok_keys, ok_vals);
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


// ---- the residual bound's interval vectors -----------------------------------------------
//
// tpgp_residual_bound (residual_rev.c:181-194) needs D1 = 6b - 6a - 4dt La - 2dt Lb - dt^2 LLa
// and D2 = 6b - 6a - 2dt La - 4dt Lb + dt^2 LLb as sparse vectors of 256-bit intervals, each
// entry built by a chain of directed-rounded products and sums in a fixed order:
//   L x    (residual_rev.c:60-111): for every entry of x in its first-insertion order, for every
//          anticommuting term in term order, z[key xor q] += s x with s = +-2h, the lower end
//          rounded down and the upper end up (accum, residual_rev.c:55).  The output's entries
//          are in the order the reference first creates them.
//   D1, D2 (residual_rev.c:112-114, 187): y += s_k x_k over x_k = b, a, La, Lb, LLa (or LLb) in
//          that order, each vector's entries in its own order; the scalars s_k are exact.
// Both are one pattern, "emit, stable sort by key, add each run left to right":
//   emit    every contribution (output key, source entry, scalar) at its slot, in the order the
//           reference makes them (entry-major, term-minor; or the vectors one after the other);
//   sort    stable radix sort by key (the Liouvillian's): a key's contributions keep that order;
//   place   a run's output position = the number of runs whose first contribution comes
//           earlier, i.e. an exclusive scan over slots of a flag at each run's first slot;
//   add     one thread per run, chaining the same operations as the reference from [+0, +0].
// The operations are correctly rounded 256-bit products and sums (tpgp_f256.h), so every entry
// is the reference's bit for bit.  What stays on the host is what is serial in the reference:
// the Gram sums and norms over D1 and D2 in their entry order.

struct ResTerm {  // a term as the residual generator reads it
    uint32_t i, j, qi, qj;
    uint32_t live;  // h != 0 (terms with h = 0 add nothing in the incident-terms generator)
    uint32_t pad;
};
static_assert(sizeof(f256) == sizeof(tpgp_hip_f256), "f256 is tpgp_hip_f256's layout");
struct ResScalar {  // s and the source vector its contributions read
    f256 s;
    uint32_t vec, pad;
};
constexpr int kResSources = 5;
struct ResSources {
    const f256 *lo[kResSources];
    const f256 *hi[kResSources];
};

__device__ __forceinline__ unsigned res_phase(const ResTerm &q, const uint64_t *key) {
    return (gen_phase(q.qi, gen_symbol(key, q.i)) + gen_phase(q.qj, gen_symbol(key, q.j))) & 3u;
}
__device__ __forceinline__ bool res_skipped(const f256 *lo, const f256 *hi, uint64_t e,
                                            int incident) {
    // the incident-terms generator passes over entries whose interval is [0, 0]
    return incident && f256_is_zero(&lo[e]) && f256_is_zero(&hi[e]);
}

__global__ void k_res_point(const double *v, uint64_t n, f256 *lo, f256 *hi) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const f256 x = f256_from_double(v[i]);
        lo[i] = x;
        hi[i] = x;
    }
}

// Contributions of every entry of x (residual_rev.c:64-66, 98-100).
template <int W>
__global__ void k_res_count(const uint64_t *keys, const f256 *lo, const f256 *hi, uint64_t n,
                            const ResTerm *terms, uint32_t n_terms, int incident,
                            uint32_t *count) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t c = 0;
        if (!res_skipped(lo, hi, e, incident)) {
            uint64_t key[W];
            for (int w = 0; w < W; ++w) key[w] = keys[e * W + w];
            for (uint32_t t = 0; t < n_terms; ++t) {
                const ResTerm q = terms[t];
                c += (res_phase(q, key) & 1u) && (!incident || q.live);
            }
        }
        count[e] = c;
    }
}

// Every contribution at its slot: output key, source entry, scalar index 2t (+2h) or 2t+1 (-2h).
template <int W>
__global__ void k_res_emit(const uint64_t *keys, const f256 *lo, const f256 *hi, uint64_t n,
                           const ResTerm *terms, uint32_t n_terms, int incident,
                           const uint64_t *offset, uint64_t *out_keys, uint32_t *src,
                           uint32_t *scalar) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x) {
        if (res_skipped(lo, hi, e, incident)) continue;
        uint64_t key[W];
        for (int w = 0; w < W; ++w) key[w] = keys[e * W + w];
        uint64_t at = offset[e];
        for (uint32_t t = 0; t < n_terms; ++t) {
            const ResTerm q = terms[t];
            const unsigned p = res_phase(q, key);
            if (!(p & 1u) || (incident && !q.live)) continue;
            uint64_t nk[W];
            for (int w = 0; w < W; ++w) nk[w] = key[w];
            nk[q.i >> 5] ^= (uint64_t)q.qi << (2 * (q.i & 31));
            nk[q.j >> 5] ^= (uint64_t)q.qj << (2 * (q.j & 31));
            for (int w = 0; w < W; ++w) out_keys[at * W + w] = nk[w];
            src[at] = (uint32_t)e;
            scalar[at] = 2 * t + (p == 3);
            ++at;
        }
    }
}

// One source vector of a combination: its entries at slots base.., all with scalar k.
template <int W>
__global__ void k_res_concat(const uint64_t *keys, uint64_t n, uint64_t base, uint32_t k,
                             uint64_t *out_keys, uint32_t *src, uint32_t *scalar) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x) {
        for (int w = 0; w < W; ++w) out_keys[(base + e) * W + w] = keys[e * W + w];
        src[base + e] = (uint32_t)e;
        scalar[base + e] = k;
    }
}

__global__ void k_res_none(uint32_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0xffffffffu;
}

// Flag the slot of every run's first contribution (the sort kept slots ascending in a run).
__global__ void k_res_first(const uint32_t *order, const uint64_t *start, uint64_t runs,
                            uint32_t *flag) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < runs;
         r += (uint64_t)gridDim.x * blockDim.x)
        flag[order[start[r]]] = 1;
}

// One thread per run: z = [+0, +0], then z += s x for its contributions in slot order, the
// lower end with products and sums rounded down, the upper end up (residual_rev.c:55, 105-106);
// the entry goes to the run's place in first-insertion order.
template <int W>
__global__ void k_res_add(const uint64_t *keys, const uint32_t *order, const uint64_t *start,
                          uint64_t runs, uint64_t n, const uint32_t *src, const uint32_t *scalar,
                          const ResScalar *scalars, ResSources from, const uint64_t *place,
                          uint64_t *out_keys, f256 *out_lo, f256 *out_hi) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < runs;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s0 = start[r], s1 = r + 1 < runs ? start[r + 1] : n;
        f256 zlo = f256_zero(0), zhi = f256_zero(0);
        for (uint64_t p = s0; p < s1; ++p) {
            const uint32_t slot = order[p], e = src[slot];
            const ResScalar c = scalars[scalar[slot]];
            const f256 *lo = from.lo[c.vec], *hi = from.hi[c.vec];
            const bool nonneg = f256_is_zero(&c.s) || !c.s.neg;  // mpfr_sgn(s) >= 0
            zlo = f256_add(zlo, f256_mul(nonneg ? lo[e] : hi[e], c.s, F256_DOWN), F256_DOWN);
            zhi = f256_add(zhi, f256_mul(nonneg ? hi[e] : lo[e], c.s, F256_UP), F256_UP);
        }
        const uint64_t q = place[order[s0]];
        for (int w = 0; w < W; ++w) out_keys[q * W + w] = keys[s0 * W + w];
        out_lo[q] = zlo;
        out_hi[q] = zhi;
    }
}

struct RVec {  // a device interval vector in first-insertion order
    uint64_t *keys = nullptr;
    f256 *lo = nullptr, *hi = nullptr;
    uint64_t len = 0;
};

// Sort the n emitted contributions (keys ck, source entries src, scalar indices sc) by key,
// then add each run into out (allocated from `keep`, which outlives this call's temporaries).
template <int W>
int res_reduce(hipStream_t st, Arena &keep, uint64_t *ck, const uint32_t *src,
               const uint32_t *sc, uint64_t n, Arena &A, const ResScalar *scalars,
               const ResSources &from, unsigned lo_bits, unsigned hi_bits, RVec *out,
               char *error, size_t cap) {
    hipError_t e;
    out->len = 0;
    if (!n) return TPGP_HIP_OK;
    uint32_t *order = A.get<uint32_t>(n), *head = A.get<uint32_t>(n), *flag = A.get<uint32_t>(n);
    uint64_t *run_of = A.get<uint64_t>(n), *place = A.get<uint64_t>(n);
    if (A.err != hipSuccess) return drain_fail(st, A.err, "HIP residual allocation", error, cap);
    const unsigned g = grid_for(n);
    hipLaunchKernelGGL(k_gen_iota, dim3(g), dim3(256), 0, st, order, n);
    if ((e = sort_pairs<W>(st, A, ck, order, n, lo_bits, hi_bits)) != hipSuccess)
        return drain_fail(st, e, "HIP residual sort", error, cap);
    hipLaunchKernelGGL(k_gen_heads<W>, dim3(g), dim3(256), 0, st, ck, n, head);
    if ((e = scan(st, A, head, n, run_of, 1)) != hipSuccess)
        return drain_fail(st, e, "HIP residual scan", error, cap);
    uint64_t runs = 0;
    if ((e = hipMemcpyAsync(&runs, run_of + n - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return drain_fail(st, e, "HIP residual runs", error, cap);
    uint64_t *start = A.get<uint64_t>(runs);
    out->keys = keep.get<uint64_t>(runs * W);
    out->lo = keep.get<f256>(runs);
    out->hi = keep.get<f256>(runs);
    if (A.err != hipSuccess || keep.err != hipSuccess)
        return drain_fail(st, A.err ? A.err : keep.err, "HIP residual allocation", error, cap);
    hipLaunchKernelGGL(k_gen_starts, dim3(g), dim3(256), 0, st, head, run_of, n, start);
    hipLaunchKernelGGL(k_gen_zero_u32, dim3(g), dim3(256), 0, st, flag, n);
    const unsigned gr = grid_for(runs);
    hipLaunchKernelGGL(k_res_first, dim3(gr), dim3(256), 0, st, order, start, runs, flag);
    if ((e = scan(st, A, flag, n, place, 0)) != hipSuccess)
        return drain_fail(st, e, "HIP residual scan", error, cap);
    hipLaunchKernelGGL(k_res_add<W>, dim3(gr), dim3(256), 0, st, ck, order, start, runs, n, src,
                       sc, scalars, from, place, out->keys, out->lo, out->hi);
    if ((e = hipGetLastError()) != hipSuccess)
        return drain_fail(st, e, "HIP residual launch", error, cap);
    out->len = runs;
    return TPGP_HIP_OK;
}

constexpr uint64_t kResMaxSlots = 1ull << 31;  // sort payloads and source entries are 32-bit

// out = L in (the generator of residual_rev.c:60-111; incident = the incident-terms form).
template <int W>
int res_generator(hipStream_t st, Arena &keep, const RVec &in, const ResTerm *d_terms,
                  uint32_t n_terms, const ResScalar *d_scalars, int incident, unsigned lo_bits,
                  unsigned hi_bits, RVec *out, char *error, size_t cap) {
    hipError_t e;
    out->len = 0;
    if (!in.len || !n_terms) return TPGP_HIP_OK;
    Arena A;
    uint32_t *count = A.get<uint32_t>(in.len);
    uint64_t *offset = A.get<uint64_t>(in.len);
    if (A.err != hipSuccess) return drain_fail(st, A.err, "HIP residual allocation", error, cap);
    const unsigned g = grid_for(in.len);
    hipLaunchKernelGGL(k_res_count<W>, dim3(g), dim3(256), 0, st, in.keys, in.lo, in.hi, in.len,
                       d_terms, n_terms, incident, count);
    if ((e = scan(st, A, count, in.len, offset, 0)) != hipSuccess)
        return drain_fail(st, e, "HIP residual scan", error, cap);
    uint64_t last[2] = {0, 0};
    uint32_t last_count = 0;
    if ((e = hipMemcpyAsync(&last[0], offset + in.len - 1, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(&last_count, count + in.len - 1, 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return drain_fail(st, e, "HIP residual count", error, cap);
    const uint64_t n = last[0] + last_count;
    if (n >= kResMaxSlots) {
        msg(error, cap, "Residual generator of %llu contributions exceeds 2^31.",
            (unsigned long long)n);
        return TPGP_HIP_ENOMEM;
    }
    if (!n) return TPGP_HIP_OK;
    uint64_t *ck = A.get<uint64_t>(n * W);
    uint32_t *src = A.get<uint32_t>(n), *sc = A.get<uint32_t>(n);
    if (A.err != hipSuccess) return drain_fail(st, A.err, "HIP residual allocation", error, cap);
    hipLaunchKernelGGL(k_res_emit<W>, dim3(g), dim3(256), 0, st, in.keys, in.lo, in.hi, in.len,
                       d_terms, n_terms, incident, offset, ck, src, sc);
    ResSources from = {};
    for (int k = 0; k < kResSources; ++k) {
        from.lo[k] = in.lo;
        from.hi[k] = in.hi;
    }
    const int s = res_reduce<W>(st, keep, ck, src, sc, n, A, d_scalars, from, lo_bits, hi_bits,
                                out, error, cap);
    if (s) return s;
    // the temporaries are freed when A goes out of scope: finish the work that reads them
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "HIP residual generator", error, cap);
    return TPGP_HIP_OK;
}

// out = sum_k s_k x_k over the five vectors in order (residual_rev.c:57, 112-114).
template <int W>
int res_combine(hipStream_t st, Arena &keep, const RVec *x, const ResScalar *d_scalars,
                unsigned lo_bits, unsigned hi_bits, RVec *out, char *error, size_t cap) {
    hipError_t e;
    uint64_t n = 0;
    for (int k = 0; k < kResSources; ++k) n += x[k].len;
    out->len = 0;
    if (n >= kResMaxSlots) {
        msg(error, cap, "Residual combination of %llu entries exceeds 2^31.",
            (unsigned long long)n);
        return TPGP_HIP_ENOMEM;
    }
    if (!n) return TPGP_HIP_OK;
    Arena A;
    uint64_t *ck = A.get<uint64_t>(n * W);
    uint32_t *src = A.get<uint32_t>(n), *sc = A.get<uint32_t>(n);
    if (A.err != hipSuccess) return drain_fail(st, A.err, "HIP residual allocation", error, cap);
    ResSources from = {};
    uint64_t base = 0;
    for (int k = 0; k < kResSources; ++k) {
        from.lo[k] = x[k].lo;
        from.hi[k] = x[k].hi;
        if (x[k].len)
            hipLaunchKernelGGL(k_res_concat<W>, dim3(grid_for(x[k].len)), dim3(256), 0, st,
                               x[k].keys, x[k].len, base, (uint32_t)k, ck, src, sc);
        base += x[k].len;
    }
    const int s = res_reduce<W>(st, keep, ck, src, sc, n, A, d_scalars, from, lo_bits, hi_bits,
                                out, error, cap);
    if (s) return s;
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "HIP residual combination", error, cap);
    return TPGP_HIP_OK;
}

// The entry of every query key among n sorted keys (their original indices in idx), or
// UINT32_MAX: a binary search per query (keys compare from the highest word down).
template <int W>
__global__ void k_res_join(const uint64_t *query, uint64_t nq, const uint64_t *sorted,
                           const uint32_t *idx, uint64_t n, uint32_t *partner) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t key[W];
        for (int w = 0; w < W; ++w) key[w] = query[i * W + w];
        uint64_t lo = 0, hi = n;  // first sorted key >= key
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            int cmp = 0;
            for (int w = W - 1; w >= 0 && !cmp; --w) {
                const uint64_t a = sorted[mid * W + w];
                cmp = a < key[w] ? -1 : a > key[w] ? 1 : 0;
            }
            if (cmp < 0)
                lo = mid + 1;
            else
                hi = mid;
        }
        bool found = lo < n;
        for (int w = 0; w < W && found; ++w) found = sorted[lo * W + w] == key[w];
        partner[i] = found ? idx[lo] : 0xffffffffu;
    }
}

// The Gram enclosure's terms of every entry (residual_rev.c:122-153), in the reference's
// operations: the square terms [min of the downward squares of the ends, or +0 when the
// interval holds 0; max of the upward squares], and, with a partner y in the other vector, the
// product terms [min of the four downward end products; max of the four upward ones].
__global__ void k_res_gram(const f256 *lo, const f256 *hi, uint64_t n, const uint32_t *partner,
                           const f256 *plo, const f256 *phi, f256 *sq_lo, f256 *sq_hi,
                           f256 *pr_lo, f256 *pr_hi) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const f256 a = lo[i], b = hi[i];
        f256 p = f256_mul_full(a, a, F256_UP), q = f256_mul_full(b, b, F256_UP);
        if (f256_cmp(&q, &p) > 0) p = q;
        sq_hi[i] = p;
        if (f256_sgn(&a) <= 0 && f256_sgn(&b) >= 0) {
            p = f256_zero(0);
        } else {
            p = f256_mul_full(a, a, F256_DOWN);
            q = f256_mul_full(b, b, F256_DOWN);
            if (f256_cmp(&q, &p) < 0) p = q;
        }
        sq_lo[i] = p;
        if (!pr_lo) continue;
        const uint32_t j = partner[i];
        if (j == 0xffffffffu) {
            pr_lo[i] = f256_zero(0);
            pr_hi[i] = f256_zero(0);
            continue;
        }
        const f256 c = plo[j], d = phi[j];
        p = f256_mul_full(a, c, F256_DOWN);
        q = f256_mul_full(a, c, F256_UP);
        auto fold = [&](const f256 &x, const f256 &y) {  // end pairs (lo, hi), (hi, lo), (hi, hi)
            f256 r = f256_mul_full(x, y, F256_DOWN);
            if (f256_cmp(&r, &p) < 0) p = r;
            r = f256_mul_full(x, y, F256_UP);
            if (f256_cmp(&r, &q) > 0) q = r;
        };
        fold(a, d);
        fold(b, c);
        fold(b, d);
        pr_lo[i] = p;
        pr_hi[i] = q;
    }
}

// D1's and D2's partners (the same key in the other vector) and their Gram terms.
template <int W>
int res_gram_terms(hipStream_t st, Arena &keep, const RVec *d, unsigned lo_bits,
                   unsigned hi_bits, uint32_t *partner[2], f256 *sq[2][2], f256 *pr[2],
                   char *error, size_t cap) {
    hipError_t e;
    Arena A;
    for (int k = 0; k < 2; ++k) {
        partner[k] = keep.get<uint32_t>(d[k].len);
        sq[k][0] = keep.get<f256>(d[k].len);
        sq[k][1] = keep.get<f256>(d[k].len);
    }
    pr[0] = keep.get<f256>(d[0].len);
    pr[1] = keep.get<f256>(d[0].len);
    if (keep.err != hipSuccess) return drain_fail(st, keep.err, "HIP residual allocation", error,
                                                  cap);
    for (int k = 0; k < 2; ++k) {  // partner[k]: entries of d[k] found in d[1 - k]
        const RVec &q = d[k], &o = d[1 - k];
        if (!q.len) continue;
        if (!o.len) {
            // no partner for any entry (a fill kernel, not a memset)
            hipLaunchKernelGGL(k_res_none, dim3(grid_for(q.len)), dim3(256), 0, st, partner[k],
                               q.len);
            continue;
        }
        uint64_t *sk = A.get<uint64_t>(o.len * W);
        uint32_t *si = A.get<uint32_t>(o.len);
        if (A.err != hipSuccess) return drain_fail(st, A.err, "HIP residual allocation", error,
                                                   cap);
        if ((e = hipMemcpyAsync(sk, o.keys, o.len * W * 8, hipMemcpyDeviceToDevice, st)))
            return drain_fail(st, e, "HIP residual copy", error, cap);
        hipLaunchKernelGGL(k_gen_iota, dim3(grid_for(o.len)), dim3(256), 0, st, si, o.len);
        if ((e = sort_pairs<W>(st, A, sk, si, o.len, lo_bits, hi_bits)) != hipSuccess)
            return drain_fail(st, e, "HIP residual sort", error, cap);
        hipLaunchKernelGGL(k_res_join<W>, dim3(grid_for(q.len)), dim3(256), 0, st, q.keys,
                           q.len, sk, si, o.len, partner[k]);
    }
    for (int k = 0; k < 2; ++k)
        if (d[k].len)
            hipLaunchKernelGGL(k_res_gram, dim3(grid_for(d[k].len)), dim3(256), 0, st, d[k].lo,
                               d[k].hi, d[k].len, partner[k], d[1].lo, d[1].hi, sq[k][0],
                               sq[k][1], k == 0 ? pr[0] : nullptr, k == 0 ? pr[1] : nullptr);
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "HIP residual Gram terms", error, cap);
    return TPGP_HIP_OK;
}

f256 res_scale(double base_power_dt, int power, int factor) {
    // factor * dt^power exactly (at most 109 significant bits; residual_rev.c:112-114)
    f256 s = f256_from_double(power == 0 ? 1.0 : base_power_dt);
    if (power == 2) s = f256_mul(s, f256_from_double(base_power_dt), F256_DOWN);
    return f256_mul(s, f256_from_double((double)factor), F256_DOWN);
}

template <int W>
int residual_vectors(tpgp_hip_context *ctx, const uint64_t *a_keys, const double *a_values,
                     size_t a_len, const uint64_t *b_keys, const double *b_values, size_t b_len,
                     size_t n_spins, const tpgp_hip_term *terms, size_t n_terms, double dt,
                     int incident, tpgp_hip_interval_vector *out, char *error, size_t cap) {
    hipStream_t st = ctx->stream;
    hipError_t e;
    Arena keep;  // the vectors a, b, La, Lb, LLa, LLb, D1, D2
    RVec v[8];
    // generator terms and their scalars +-2h (residual_rev.c:68, 103: exact in MPFR)
    std::vector<ResTerm> t(n_terms);
    std::vector<ResScalar> gs(2 * n_terms + 1);
    for (size_t k = 0; k < n_terms; ++k) {
        t[k] = ResTerm{(uint32_t)terms[k].i, (uint32_t)terms[k].j, terms[k].qi, terms[k].qj,
                       terms[k].h != 0.0, 0};
        f256 s = f256_from_double(terms[k].h);
        if (!f256_is_zero(&s)) s.e += 1;  // 2h without binary64 overflow
        gs[2 * k] = ResScalar{s, 0, 0};
        gs[2 * k + 1] = ResScalar{f256_neg(s), 0, 0};
    }
    // combination scalars: D1 over (b, a, La, Lb, LLa), D2 over (b, a, La, Lb, LLb)
    const ResScalar cs[2][kResSources] = {
        {{res_scale(dt, 0, 6), 0, 0}, {res_scale(dt, 0, -6), 1, 0}, {res_scale(dt, 1, -4), 2, 0},
         {res_scale(dt, 1, -2), 3, 0}, {res_scale(dt, 2, -1), 4, 0}},
        {{res_scale(dt, 0, 6), 0, 0}, {res_scale(dt, 0, -6), 1, 0}, {res_scale(dt, 1, -2), 2, 0},
         {res_scale(dt, 1, -4), 3, 0}, {res_scale(dt, 2, 1), 4, 0}}};
    ResTerm *d_terms = keep.get<ResTerm>(n_terms);
    ResScalar *d_gs = keep.get<ResScalar>(gs.size());
    ResScalar *d_cs = keep.get<ResScalar>(2 * kResSources);
    v[0].keys = keep.get<uint64_t>(a_len * W);
    v[0].lo = keep.get<f256>(a_len);
    v[0].hi = keep.get<f256>(a_len);
    v[1].keys = keep.get<uint64_t>(b_len * W);
    v[1].lo = keep.get<f256>(b_len);
    v[1].hi = keep.get<f256>(b_len);
    double *d_av = keep.get<double>(a_len), *d_bv = keep.get<double>(b_len);
    if (keep.err != hipSuccess) return hip_fail(keep.err, "HIP residual allocation", error, cap);
    v[0].len = a_len;
    v[1].len = b_len;
    if ((n_terms && (e = hipMemcpyAsync(d_terms, t.data(), n_terms * sizeof(ResTerm),
                                        hipMemcpyHostToDevice, st))) ||
        (e = hipMemcpyAsync(d_gs, gs.data(), gs.size() * sizeof(ResScalar),
                            hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_cs, cs, sizeof(cs), hipMemcpyHostToDevice, st)) ||
        (a_len && (e = hipMemcpyAsync(v[0].keys, a_keys, a_len * W * 8, hipMemcpyHostToDevice,
                                      st))) ||
        (a_len && (e = hipMemcpyAsync(d_av, a_values, a_len * 8, hipMemcpyHostToDevice, st))) ||
        (b_len && (e = hipMemcpyAsync(v[1].keys, b_keys, b_len * W * 8, hipMemcpyHostToDevice,
                                      st))) ||
        (b_len && (e = hipMemcpyAsync(d_bv, b_values, b_len * 8, hipMemcpyHostToDevice, st))))
        return drain_fail(st, e, "HIP residual upload", error, cap);
    if (a_len)
        hipLaunchKernelGGL(k_res_point, dim3(grid_for(a_len)), dim3(256), 0, st, d_av,
                           (uint64_t)a_len, v[0].lo, v[0].hi);
    if (b_len)
        hipLaunchKernelGGL(k_res_point, dim3(grid_for(b_len)), dim3(256), 0, st, d_bv,
                           (uint64_t)b_len, v[1].lo, v[1].hi);
    const unsigned lo_bits = W == 1 ? (unsigned)(2 * n_spins) : 64u;
    const unsigned hi_bits = W == 1 ? 0u : (unsigned)(2 * (n_spins - 32));
    int s = 0;
    // La, Lb, LLa, LLb (residual_rev.c:186)
    for (int k = 0; k < 4 && !s; ++k)
        s = res_generator<W>(st, keep, v[k], d_terms, (uint32_t)n_terms, d_gs, incident, lo_bits,
                             hi_bits, &v[k + 2], error, cap);
    // D1, D2 (residual_rev.c:187)
    for (int d = 0; d < 2 && !s; ++d) {
        const RVec x[kResSources] = {v[1], v[0], v[2], v[3], v[4 + d]};
        s = res_combine<W>(st, keep, x, d_cs + d * kResSources, lo_bits, hi_bits, &v[6 + d],
                           error, cap);
    }
    if (s) return s;
    uint32_t *partner[2];
    f256 *sq[2][2], *pr[2];
    if ((s = res_gram_terms<W>(st, keep, v + 6, lo_bits, hi_bits, partner, sq, pr, error, cap)))
        return s;
    for (int d = 0; d < 2; ++d) {
        const RVec &r = v[6 + d];
        tpgp_hip_interval_vector &o = out[d];
        const size_t n1 = r.len ? r.len : 1;
        o.len = r.len;
        o.keys = (uint64_t *)malloc(n1 * W * 8);
        o.lower = (tpgp_hip_f256 *)malloc(n1 * sizeof(tpgp_hip_f256));
        o.upper = (tpgp_hip_f256 *)malloc(n1 * sizeof(tpgp_hip_f256));
        o.partner = (uint32_t *)malloc(n1 * sizeof(uint32_t));
        for (int k = 0; k < 2; ++k) {
            o.square[k] = (tpgp_hip_f256 *)malloc(n1 * sizeof(tpgp_hip_f256));
            o.product[k] = d == 0 ? (tpgp_hip_f256 *)malloc(n1 * sizeof(tpgp_hip_f256)) : nullptr;
        }
        if (!o.keys || !o.lower || !o.upper || !o.partner || !o.square[0] || !o.square[1] ||
            (d == 0 && (!o.product[0] || !o.product[1]))) {
            (void)hipStreamSynchronize(st);
            msg(error, cap, "Host allocation of the residual vectors failed.");
            return TPGP_HIP_ENOMEM;
        }
        if (!r.len) continue;
        const size_t fb = r.len * sizeof(f256);
        if ((e = hipMemcpyAsync(o.keys, r.keys, r.len * W * 8, hipMemcpyDeviceToHost, st)) ||
            (e = hipMemcpyAsync(o.lower, r.lo, fb, hipMemcpyDeviceToHost, st)) ||
            (e = hipMemcpyAsync(o.upper, r.hi, fb, hipMemcpyDeviceToHost, st)) ||
            (e = hipMemcpyAsync(o.partner, partner[d], r.len * 4, hipMemcpyDeviceToHost, st)) ||
            (e = hipMemcpyAsync(o.square[0], sq[d][0], fb, hipMemcpyDeviceToHost, st)) ||
            (e = hipMemcpyAsync(o.square[1], sq[d][1], fb, hipMemcpyDeviceToHost, st)) ||
            (d == 0 && (e = hipMemcpyAsync(o.product[0], pr[0], fb, hipMemcpyDeviceToHost, st))) ||
            (d == 0 && (e = hipMemcpyAsync(o.product[1], pr[1], fb, hipMemcpyDeviceToHost, st))))
            return drain_fail(st, e, "HIP residual download", error, cap);
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "HIP residual vectors", error, cap);
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


extern "C" void tpgp_hip_interval_vector_free(tpgp_hip_interval_vector *v) {
    if (!v) return;
    free(v->keys);
    free(v->lower);
    free(v->upper);
    free(v->partner);
    for (int k = 0; k < 2; ++k) {
        free(v->square[k]);
        free(v->product[k]);
    }
    *v = tpgp_hip_interval_vector{};
}

extern "C" int tpgp_hip_residual_vectors(tpgp_hip_context *ctx, const uint64_t *a_keys,
                                         const double *a_values, size_t a_len,
                                         const uint64_t *b_keys, const double *b_values,
                                         size_t b_len, size_t n_words, size_t n_spins,
                                         const tpgp_hip_term *terms, size_t n_terms, double dt,
                                         int incident_terms, tpgp_hip_interval_vector out[2],
                                         char *error, size_t cap) {
    msg(error, cap, "");
    if (out)
        for (int d = 0; d < 2; ++d) out[d] = tpgp_hip_interval_vector{};
    if (!ctx || !out || !n_spins || n_spins > 64 || n_words != (n_spins + 31) / 32 ||
        (a_len && (!a_keys || !a_values)) || (b_len && (!b_keys || !b_values)) ||
        (n_terms && !terms) || !std::isfinite(dt) || dt <= 0 ||
        (incident_terms != 0 && incident_terms != 1) || a_len >= kResMaxSlots ||
        b_len >= kResMaxSlots || n_terms > (1u << 30)) {
        msg(error, cap, "Invalid residual-vector arguments (at most 64 spins).");
        return TPGP_HIP_EINVAL;
    }
    for (size_t k = 0; k < n_terms; ++k) {
        const tpgp_hip_term &q = terms[k];
        if (q.i >= n_spins || q.j >= n_spins || q.i == q.j || q.qi > 3 || q.qj > 3 ||
            !std::isfinite(q.h)) {
            msg(error, cap, "Invalid Hamiltonian term %zu.", k);
            return TPGP_HIP_EINVAL;
        }
    }
    std::lock_guard<std::recursive_mutex> lock(ctx->mutex);
    DeviceGuard guard(ctx->device);
    if (!guard.ok) {
        msg(error, cap, "HIP device selection failed.");
        return TPGP_HIP_EDEVICE;
    }
    const int s = n_words == 1
                      ? residual_vectors<1>(ctx, a_keys, a_values, a_len, b_keys, b_values,
                                            b_len, n_spins, terms, n_terms, dt, incident_terms,
                                            out, error, cap)
                      : residual_vectors<2>(ctx, a_keys, a_values, a_len, b_keys, b_values,
                                            b_len, n_spins, terms, n_terms, dt, incident_terms,
                                            out, error, cap);
    if (s)
        for (int d = 0; d < 2; ++d) tpgp_hip_interval_vector_free(&out[d]);
    return s;
}
