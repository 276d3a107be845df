// tpgp_hip: MI355X (gfx950) Pauli-propagation branch engine.
//
// Replaces the reference's CUDA branch backend (engine/src/cuda_rev.cu) and moves the rest of
// the per-rotation work of engine/src/tpgp_rev.c:535-771 (stable sort, merge, truncation, ledger
// inputs) and the FID scan (tpgp_rev.c:311-388) onto the device, with the maps kept resident.
//
// Arithmetic: binary64 round-to-nearest only, no contraction (-ffp-contract=off), FP64
// denormals preserved.  No trigonometric function is evaluated on the device.
//
// Device layout of a resident map.  A map is one sorted array of (key, value) pairs in a slot of
// an HBM arena.  Keys are stored in a *permuted digit layout*: digit position p (bits 2p..2p+1 of
// the 64*W-bit key, word p/32) holds the 2-bit symbol of site sigma[p].  Keys are sorted as
// unsigned integers.  Before a rotation on sites (i, j) the layout is arranged so that i and j
// occupy the two most significant live digits (positions N-1 and N-2).  Then the 16 symbol
// classes (p_top, p_second) are 16 contiguous regions, a rotation pairs region c with region
// c ^ q, and because XOR on the top digits preserves order inside a region, every output region
// is a 2-way merge of two sorted runs.  Changing the active pair is a stable 16-bucket partition
// that moves the new pair's digits to the top.  Coefficients of one output key come from at most
// two raw branches (the key's own a-slot and its partner's d-slot), so the merged value is one
// IEEE addition whose result does not depend on order: outputs are bit-identical to the
// reference's stable merge sort.
//
// Class regions need not be adjacent in the slot.  A rotation whose commuting classes pass
// unchanged may *append*: it leaves those classes where they are and writes the classes it
// rewrites after the slot's used extent, in the same buffer, so that only the anticommuting
// pairs are read and written.  Partitions, uploads and the small-map path write contiguous maps.
// Every reader goes through the class region table (logical start, physical start, extent).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <dlfcn.h>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "../../include/tpgp_hip.h"
#include "tpgp_hip_context.h"

// ------------------------------------------------------------------------------------------
// Scalar device helpers (same semantics as tpgp_rev.c:25-42 and cuda_rev.cu:50-76)
// ------------------------------------------------------------------------------------------
namespace {

constexpr int kBlock = 256;        // 4 waves of 64
constexpr int kMergeItems = 4;     // merge-path elements per thread
constexpr int kMergeTile = kBlock * kMergeItems;
static_assert(kMergeTile <= 65535, "per-thread splits are stored in 16 bits");
constexpr int kChunkTiles = kBlock;  // rotation tiles per scan chunk (one per thread)
constexpr uint32_t kSumBlocks = 512;     // blocks of the sum-of-squares reduction
// PassMap.flags bit: the map's commuting classes cannot lose a key to this rotation's
// truncation (every coefficient there already exceeds the threshold), so they pass unchanged
constexpr uint32_t kPassCommutingKept = 16u;
// PassMap.flags bit (with kPassCommutingKept, in_buf == out_buf): the rotation appends -- its
// unchanged classes stay in place, the classes it rewrites go after the buffer's used extent
constexpr uint32_t kPassAppend = 32u;
// Class region table, kRegStride words per (map, buffer):
//   [0, 17)   logical starts: class c is elements reg[c] .. reg[c+1]-1 of the map in key order;
//   [17, 33)  physical starts: element reg[c] + i lives at slot index reg[kRegPhys + c] + i;
//   [33]      used extent: slot indices from reg[kRegTop] on are free.
// A contiguous map has reg[kRegPhys + c] == reg[c] and reg[kRegTop] == reg[16].
constexpr int kRegStride = 34, kRegPhys = 17, kRegTop = 33;
constexpr int kPartRounds = 8;     // partition: 8 rounds of 256 elements (the default)
constexpr int kPartTile = kBlock * kPartRounds;
constexpr int kPartRoundsMin = 4;  // smallest tunable partition tile: 4 rounds

__device__ __forceinline__ uint64_t dbits(double x) { return __builtin_bit_cast(uint64_t, x); }
__device__ __forceinline__ double dfrom(uint64_t b) { return __builtin_bit_cast(double, b); }

// Adjacent binary64 value toward +inf by bit arithmetic (zero maps to the smallest subnormal).
__device__ __forceinline__ double next_up(double x) {
    if (x != x || x == __builtin_inf()) return x;
    if (x == 0.0) return dfrom(1ull);
    uint64_t b = dbits(x);
    return dfrom(x > 0.0 ? b + 1 : b - 1);
}
__device__ __forceinline__ double next_down(double x) { return -next_up(-x); }
__device__ __forceinline__ double upadd(double a, double b) {
    if (a == 0.0) return b;
    if (b == 0.0) return a;
    return next_up(a + b);
}
__device__ __forceinline__ double upmul(double a, double b) {
    if (a == 0.0 || b == 0.0) return 0.0;
    return next_up(a * b);
}
__device__ __forceinline__ double round_radius(double x) {
    if (!__builtin_isfinite(x)) return __builtin_inf();
    double above = __builtin_fabs(next_up(x) - x);
    double below = __builtin_fabs(x - next_down(x));
    return __builtin_fmax(above, below);
}
__device__ __forceinline__ bool anticomm(unsigned q, unsigned p) { return q && p && q != p; }
// Exponent of i in QP (tpgp_rev.c:414-418): cyclic X->Y->Z gives +i (1), reverse -i (3).
__device__ __forceinline__ unsigned pauli_phase(unsigned q, unsigned p) {
    if (!q || !p || q == p) return 0;
    return ((p + 3 - q) % 3 == 1) ? 1u : 3u;
}

// ------------------------------------------------------------------------------------------
// Keys of W 64-bit words (W=1: N<=32, W=2: N<=64), word W-1 most significant.
// ------------------------------------------------------------------------------------------
template <int W> struct Key { uint64_t w[W]; };

template <int W> __device__ __forceinline__ Key<W> kload(const uint64_t *p) {
    Key<W> k;
#pragma unroll
    for (int i = 0; i < W; ++i) k.w[i] = p[i];
    return k;
}
template <int W> __device__ __forceinline__ void kstore(uint64_t *p, const Key<W> &k) {
#pragma unroll
    for (int i = 0; i < W; ++i) p[i] = k.w[i];
}
template <int W> __device__ __forceinline__ bool klt(const Key<W> &a, const Key<W> &b) {
#pragma unroll
    for (int i = W - 1; i > 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
    return a.w[0] < b.w[0];
}
template <int W> __device__ __forceinline__ bool keq(const Key<W> &a, const Key<W> &b) {
    bool e = true;
#pragma unroll
    for (int i = 0; i < W; ++i) e = e && a.w[i] == b.w[i];
    return e;
}
template <int W> __device__ __forceinline__ Key<W> kxor(const Key<W> &a, const Key<W> &b) {
    Key<W> r;
#pragma unroll
    for (int i = 0; i < W; ++i) r.w[i] = a.w[i] ^ b.w[i];
    return r;
}
template <int W> __device__ __forceinline__ bool kzero(const Key<W> &a) {
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) o |= a.w[i];
    return o == 0;
}
// Word selection by comparison rather than a dynamic array index (keeps keys in registers).
template <int W> __device__ __forceinline__ unsigned kdigit(const Key<W> &a, unsigned p) {
    uint64_t word = a.w[0];
    if constexpr (W == 2) word = (p >> 5) ? a.w[1] : a.w[0];
    return (unsigned)(word >> (2 * (p & 31))) & 3u;
}
template <int W> __device__ __forceinline__ Key<W> kdigit_mask(unsigned p, unsigned d) {
    Key<W> r;
    uint64_t bits = (uint64_t)d << (2 * (p & 31));
    if constexpr (W == 1) {
        r.w[0] = bits;
    } else {
        r.w[0] = (p >> 5) ? 0 : bits;
        r.w[1] = (p >> 5) ? bits : 0;
    }
    return r;
}

// Move the digits at positions pa, pb to positions n-1 (pa's) and n-2 (pb's); the other live
// digits keep their relative order, packed below.  n >= 2, pa != pb, both < n.
__device__ __forceinline__ uint64_t remove_digit64(uint64_t k, unsigned p) {
    uint64_t low = p ? (k & ((1ull << (2 * p)) - 1)) : 0;
    uint64_t high = (k >> 2) & ~(p ? ((1ull << (2 * p)) - 1) : 0ull);
    return low | high;
}
__device__ __forceinline__ unsigned __int128 remove_digit128(unsigned __int128 k, unsigned p) {
    unsigned __int128 mask = p ? ((((unsigned __int128)1) << (2 * p)) - 1) : 0;
    return (k & mask) | ((k >> 2) & ~mask);
}
template <int W>
__device__ __forceinline__ Key<W> move_to_top(const Key<W> &k, unsigned pa, unsigned pb,
                                              unsigned n) {
    unsigned da = kdigit<W>(k, pa), db = kdigit<W>(k, pb);
    unsigned hi = pa > pb ? pa : pb, lo = pa > pb ? pb : pa;
    Key<W> r;
    if constexpr (W == 1) {
        uint64_t x = remove_digit64(remove_digit64(k.w[0], hi), lo);
        x |= (uint64_t)da << (2 * (n - 1));
        x |= (uint64_t)db << (2 * (n - 2));
        r.w[0] = x;
    } else {
        unsigned __int128 x = ((unsigned __int128)k.w[1] << 64) | k.w[0];
        x = remove_digit128(remove_digit128(x, hi), lo);
        x |= (unsigned __int128)da << (2 * (n - 1));
        x |= (unsigned __int128)db << (2 * (n - 2));
        r.w[0] = (uint64_t)x;
        r.w[1] = (uint64_t)(x >> 64);
    }
    return r;
}

// ------------------------------------------------------------------------------------------
// Block-level helpers (256 threads = 4 waves of 64)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned lane_id() { return threadIdx.x & 63u; }
__device__ __forceinline__ unsigned wave_id() { return threadIdx.x >> 6; }

// Exclusive block scan of one unsigned per thread; returns the exclusive prefix, total in *tot.
// Wave-wide inclusive scans by DPP lane moves (all 64 lanes active): row_shr 1, 2, 4, 8 scan
// each row of 16 lanes (lanes before a row's start read 0), then row_bcast 15 and 31 add the
// last lane of each row to the rows after it.  Lane i ends with the sum over lanes <= i, lane 63
// with the total: a fixed addition tree in which every input passes through at most 6
// additions.
template <int CTRL, int ROWS> __device__ __forceinline__ unsigned dpp_u32(unsigned x) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xf, true);
}
template <int CTRL, int ROWS> __device__ __forceinline__ double dpp_f64(double x) {
    const uint64_t b = dbits(x);
    const unsigned lo = dpp_u32<CTRL, ROWS>((unsigned)b);
    const unsigned hi = dpp_u32<CTRL, ROWS>((unsigned)(b >> 32));
    return dfrom(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ unsigned wave_scan_u32(unsigned v) {
    v += dpp_u32<0x111, 0xf>(v);  // row_shr:1
    v += dpp_u32<0x112, 0xf>(v);  // row_shr:2
    v += dpp_u32<0x114, 0xf>(v);  // row_shr:4
    v += dpp_u32<0x118, 0xf>(v);  // row_shr:8
    v += dpp_u32<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_u32<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}
__device__ __forceinline__ double wave_scan_f64(double v) {
    v += dpp_f64<0x111, 0xf>(v);
    v += dpp_f64<0x112, 0xf>(v);
    v += dpp_f64<0x114, 0xf>(v);
    v += dpp_f64<0x118, 0xf>(v);
    v += dpp_f64<0x142, 0xa>(v);
    v += dpp_f64<0x143, 0xc>(v);
    return v;
}

__device__ unsigned block_exclusive_scan(unsigned v, unsigned *scratch /*>= 4*/, unsigned *tot) {
    const unsigned x = wave_scan_u32(v);
    if (lane_id() == 63) scratch[wave_id()] = x;
    __syncthreads();
    unsigned base = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        unsigned s = scratch[w];
        if ((unsigned)w < wave_id()) base += s;
        all += s;
    }
    __syncthreads();
    *tot = all;
    return base + x - v;
}

// ------------------------------------------------------------------------------------------
// Branch kernel (the tpgp_cuda_branches replacement).  Output slot order 2k (a) / 2k+1 (d).
// flag bit 1: nonzero key padding; bit 2: nonfinite input, product or radius.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_branches(
    const uint64_t *__restrict__ keys, const double *__restrict__ values, size_t count,
    size_t words, size_t spins, size_t i, size_t j, unsigned qi, unsigned qj, double c, double s,
    double rc, double rs, int enclose, uint64_t *__restrict__ out_keys,
    double *__restrict__ out_values, double *__restrict__ out_radii, unsigned *flag) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const unsigned rem = (unsigned)(spins % 32);
    unsigned local = 0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
        const uint64_t *src = keys + k * words;
        uint64_t *a = out_keys + 2 * k * words, *d = a + words;
        if (rem && (src[words - 1] >> (2 * rem))) local |= 1u;
        for (size_t w = 0; w < words; ++w) a[w] = d[w] = src[w];
        unsigned pi = (unsigned)(src[i >> 5] >> (2 * (i & 31))) & 3u;
        unsigned pj = (unsigned)(src[j >> 5] >> (2 * (j & 31))) & 3u;
        double v = values[k];
        if (!__builtin_isfinite(v)) local |= 2u;
        out_radii[2 * k] = 0.0;
        out_radii[2 * k + 1] = 0.0;
        if (anticomm(qi, pi) == anticomm(qj, pj)) {
            out_values[2 * k] = v;
            out_values[2 * k + 1] = 0.0;
            continue;
        }
        unsigned ph = (pauli_phase(qi, pi) + pauli_phase(qj, pj)) & 3u;
        d[i >> 5] ^= (uint64_t)qi << (2 * (i & 31));
        d[j >> 5] ^= (uint64_t)qj << (2 * (j & 31));
        double x = v * c, y = v * s;
        out_values[2 * k] = x;
        out_values[2 * k + 1] = ph == 1 ? y : -y;  // exact sign flip, as sign*y
        double ra = 0.0, rd = 0.0;
        if (enclose) {
            ra = upadd(upmul(__builtin_fabs(v), rc), round_radius(x));
            rd = upadd(upmul(__builtin_fabs(v), rs), round_radius(y));
            out_radii[2 * k] = ra;
            out_radii[2 * k + 1] = rd;
        }
        if (!__builtin_isfinite(x) || !__builtin_isfinite(y) || !__builtin_isfinite(ra) ||
            !__builtin_isfinite(rd))
            local |= 2u;
    }
    if (local) atomicOr(flag, local);
}

// ------------------------------------------------------------------------------------------
// Resident batch: device descriptors
// ------------------------------------------------------------------------------------------
struct MapSlot {          // per map, fixed
    uint64_t off;         // element offset of the slot in every arena
    uint64_t cap;         // slot capacity (elements)
    uint32_t n_spins;
    uint32_t pad;
};
struct PassMap {          // per map, per pass (staged from the host)
    uint32_t active;
    uint32_t in_buf, out_buf;
    uint32_t pa, pb;      // partition: positions moved to the top
    uint32_t qt, qs;      // rotation: generator symbols on the top / second digit
    uint32_t flags;       // bit0 validated, bit1 log
    uint32_t rec_slot;    // record index (rotation) or readout index
    uint32_t rot_seq;
    double c, s, rc, rs, threshold;
};
struct ReadoutMap {       // per map, per readout (staged from the host)
    uint32_t active, in_buf, rec_slot, pad;
    uint32_t pos[64];     // digit position of every site
};
struct TileLedger {
    uint64_t dropped, radii;
    double l1, sq, rsq;
    uint32_t nonfinite, pad;
};
struct LogDev {
    uint64_t key[2];
    double value, aux;
    uint32_t map, rotation, kind, pad;
};

struct Arena {
    uint64_t *keys;
    double *vals;
};

struct BatchDev {
    Arena arena[3];
    const MapSlot *slot;      // [M]
    uint64_t *len;            // [M*3]
    uint64_t *regions;        // [M*3*kRegStride] class region table
    uint32_t n_maps;
    uint32_t max_rec, max_rd;
    // scratch
    uint64_t *tstart;         // partition [M+1] / rotation [M*16+1]
    uint32_t *tile_hist;      // [maxT*16] rotation: per tile; partition: class-major (class c
    uint64_t *tile_base;      // [maxT*16]   of tile g at c * ntiles + g)
    uint64_t *map_hist;       // [M*16]
    uint64_t *splits;         // [maxT+1]
    struct TileDesc *tdesc;   // [maxT] rotation tile descriptors (written by the split pass)
    uint64_t *work;           // [maxT] rotation tiles the count / write passes visit
    uint16_t *tsplit;         // [maxT*kBlock] per-thread merge-path splits (count -> write)
    uint64_t *wstart;         // [M*16+2] first work-list slot per (map, class); [M*16] count
                              // tiles, [M*16+1] all work tiles
    TileLedger *tile_led;     // [maxT]
    uint64_t *chunk_start;    // [M+1] first scan chunk of every map (rotation)
    uint64_t *chunk_base;     // [maxC] slot index of every chunk's first output (its offset
                              // in the map's written output plus the rotation's output base)
    uint32_t *chunk_sum;      // [maxC] kept outputs of every chunk
    TileLedger *chunk_led;    // [maxC] ledger partials of every chunk
    uint32_t *counter;        // [4]
    uint32_t *overflow;       // [M]
    uint32_t *small_resume;   // [M] first pass the global small-map kernel runs (the LDS kernel
                              // ran the ones before it)
    tpgp_hip_record *rec;     // [M*max_rec]
    double *rd;               // [M*max_rd*64]
    LogDev *log;              // [log_cap]
    uint64_t *log_count;      // [1]
    uint64_t log_cap;
    uint64_t *debug;          // [16] validation report (debug[0] != 0: violation found)
    double *sumsq;            // [kSumBlocks * 2] per-block double-double partial sums
};

__host__ __device__ __forceinline__ uint64_t *region_row(uint64_t *regions, uint32_t m,
                                                         uint32_t buf) {
    return regions + ((uint64_t)m * 3 + buf) * kRegStride;
}

// Logical element index -> slot index, for indices from e0 on: the class holding e0 is looked up
// once; later classes (met only past the end of that class) by a short forward scan.
struct Phys {
    const uint64_t *reg;
    uint64_t hi;     // logical end of the first class
    uint64_t delta;  // slot index - logical index inside it (mod 2^64)
    unsigned c0;
};
__device__ __forceinline__ Phys phys_from(const uint64_t *reg, uint64_t e0) {
    unsigned c = 0;  // the largest class c with reg[c] <= e0 (e0 < reg[16])
#pragma unroll
    for (unsigned k = 1; k < 16; ++k) c = reg[k] <= e0 ? k : c;
    Phys p;
    p.reg = reg;
    p.c0 = c;
    p.hi = reg[c + 1];
    p.delta = reg[kRegPhys + c] - reg[c];
    return p;
}
__device__ __forceinline__ uint64_t phys_at(const Phys &p, uint64_t e) {
    if (e < p.hi) return e + p.delta;
    unsigned c = p.c0 + 1;
    while (c < 15 && e >= p.reg[c + 1]) ++c;
    return e - p.reg[c] + p.reg[kRegPhys + c];
}
// phys_at for a run of indices below e1 that `one` (uniform) says lie in the first class: the
// common case of a tile inside one class costs one add per element.
__device__ __forceinline__ uint64_t phys_in(const Phys &p, bool one, uint64_t e) {
    return one ? e + p.delta : phys_at(p, e);
}
// The region row of a contiguous map of `len` elements whose logical class starts are in
// reg[0..16] (reg[16] == len): physical starts equal to them, extent len.
__device__ __forceinline__ void contiguous_row(uint64_t *reg, unsigned c, uint64_t start) {
    reg[c] = start;
    if (c < 16) reg[kRegPhys + c] = start;
    else reg[kRegTop] = start;
}

__device__ __forceinline__ uint32_t find_bucket(const uint64_t *start, uint32_t n, uint64_t g) {
    // largest b in [0, n) with start[b] <= g  (start is nondecreasing, start[0] == 0)
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------------------------------------
// Partition (change of active pair): plan -> count -> scan -> scatter
// ------------------------------------------------------------------------------------------
// Stable scatter of one partition tile (elements [e0, e1), at most kPartTile): every element
// goes to run[class] + (number of elements of its class before it in the tile), with digits
// pa, pb moved to the top of its key; run[] (LDS) then advances by the tile's class counts.
// Ranks come from per-round, per-wave ballots; three barriers per tile.
struct ScatterShared {
    unsigned cnt[kPartRounds][kBlock / 64][16];
    uint64_t base[kPartRounds][kBlock / 64][16];
    uint64_t run[16];
};
template <int W>
__device__ __forceinline__ void scatter_tile(const uint64_t *keys, const double *vals,
                                             const uint64_t *reg, uint64_t e0, uint64_t e1,
                                             uint32_t pa, uint32_t pb, uint32_t n,
                                             uint64_t *okeys, double *ovals, ScatterShared &sh) {
    Key<W> kk[kPartRounds];
    double vv[kPartRounds];
    unsigned tag[kPartRounds];  // class (bits 0-4, 16 = none) | rank in wave (bits 8-14)
    const Phys ph = phys_from(reg, e0);  // input element e lives at slot index phys_at(ph, e)
#pragma unroll
    for (int r = 0; r < kPartRounds; ++r) {
        const uint64_t e = e0 + threadIdx.x + (uint64_t)r * kBlock;
        if (e < e1) {
            const uint64_t at = phys_at(ph, e);
            kk[r] = kload<W>(keys + at * W);
            vv[r] = vals[at];
        }
    }
    const uint64_t lanemask = (1ull << lane_id()) - 1ull;
#pragma unroll
    for (int r = 0; r < kPartRounds; ++r) {
        const uint64_t e = e0 + threadIdx.x + (uint64_t)r * kBlock;
        const unsigned cls = e < e1 ? (kdigit<W>(kk[r], pa) << 2) | kdigit<W>(kk[r], pb) : 16u;
        unsigned below = 0;
        for (unsigned c = 0; c < 16; ++c) {
            const uint64_t bal = __ballot(cls == c);
            if (cls == c) below = (unsigned)__popcll(bal & lanemask);
            if (lane_id() == 0) sh.cnt[r][wave_id()][c] = (unsigned)__popcll(bal);
        }
        tag[r] = cls | (below << 8);
    }
    __syncthreads();
    if (threadIdx.x < 16) {  // class c: offsets of every (round, wave) group, in element order
        const unsigned c = threadIdx.x;
        uint64_t at = sh.run[c];
        for (int r = 0; r < kPartRounds; ++r)
            for (int w = 0; w < kBlock / 64; ++w) {
                sh.base[r][w][c] = at;
                at += sh.cnt[r][w][c];
            }
        sh.run[c] = at;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPartRounds; ++r) {
        const unsigned cls = tag[r] & 31u;
        if (cls < 16) {
            const uint64_t dst = sh.base[r][wave_id()][cls] + (tag[r] >> 8);
            kstore<W>(okeys + dst * W, move_to_top<W>(kk[r], pa, pb, n));
            ovals[dst] = vv[r];
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_part_plan(BatchDev b, const PassMap *pm,
                                                      uint32_t tile) {
    __shared__ unsigned scratch[8];
    const uint32_t M = b.n_maps;
    // tiles per map, exclusive scan over maps in chunks of 256
    uint64_t carry = 0;
    for (uint32_t base = 0; base < M; base += kBlock) {
        uint32_t m = base + threadIdx.x;
        unsigned n = 0;
        if (m < M && pm[m].active) {
            uint64_t len = b.len[m * 3 + pm[m].in_buf];
            n = (unsigned)((len + tile - 1) / tile);
            if (len == 0) {  // no tile will describe the (empty) output: do it here
                uint64_t *reg = region_row(b.regions, m, pm[m].out_buf);
                for (int c = 0; c < kRegStride; ++c) reg[c] = 0;
                b.len[m * 3 + pm[m].out_buf] = 0;
            }
        }
        unsigned tot;
        unsigned ex = block_exclusive_scan(n, scratch, &tot);
        if (m < M) b.tstart[m] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) b.tstart[M] = carry;
}

// Per tile of kBlock * R elements: the histogram of the 16 new classes (digits pa, pb).  Each
// thread counts its R keys (coalesced rounds of 256) in 4-bit counters packed in one register: no LDS
// atomics, which the sorted keys would serialize (the new digits come in long runs, so a wave's
// atomics hit one bin).  The counters are widened to 16-bit halves of 8 dwords, summed across
// the wave by DPP scans and across the waves in LDS.
template <int W, int R>
__global__ __launch_bounds__(kBlock) void k_part_count(BatchDev b, const PassMap *__restrict__ pm) {
    static_assert(R <= 15, "4-bit per-thread counters");
    constexpr uint64_t kTile = (uint64_t)kBlock * R;
    __shared__ unsigned whist[kBlock / 64][8];
    const uint32_t M = b.n_maps;
    const uint64_t ntiles = b.tstart[M];
    for (uint64_t g = blockIdx.x; g < ntiles; g += gridDim.x) {
        uint32_t m = find_bucket(b.tstart, M, g);
        const PassMap &P = pm[m];
        uint64_t len = b.len[m * 3 + P.in_buf];
        uint64_t e0 = (g - b.tstart[m]) * kTile;
        uint64_t e1 = e0 + kTile < len ? e0 + kTile : len;
        const uint64_t *keys = b.arena[P.in_buf].keys + b.slot[m].off * W;
        const uint32_t pa = P.pa, pb = P.pb;  // read once
        const Phys ph = phys_from(region_row(b.regions, m, P.in_buf), e0);
        const bool one = e1 <= ph.hi;  // the tile lies in one class
        Key<W> k[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {  // branch-free: past the end, reload the last
            const uint64_t e = e0 + threadIdx.x + (uint64_t)r * kBlock;
            k[r] = kload<W>(keys + phys_in(ph, one, e < e1 ? e : e1 - 1) * W);
        }
        uint64_t cnt = 0;  // 16 bins x 4 bits
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint64_t e = e0 + threadIdx.x + (uint64_t)r * kBlock;
            const unsigned bin = (kdigit<W>(k[r], pa) << 2) | kdigit<W>(k[r], pb);
            cnt += e < e1 ? 1ull << (4 * bin) : 0ull;
        }
        unsigned h[8];  // dword j: bin 2j in bits 0-15, bin 2j+1 in bits 16-31
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint64_t two = cnt >> (8 * j);
            h[j] = (unsigned)(two & 0xfu) | ((unsigned)((two >> 4) & 0xfu) << 16);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) h[j] = wave_scan_u32(h[j]);  // <= 64 * R per half
        if (lane_id() == 63) {
#pragma unroll
            for (int j = 0; j < 8; ++j) whist[wave_id()][j] = h[j];
        }
        __syncthreads();
        if (threadIdx.x < 16) {
            const unsigned j = threadIdx.x >> 1, sh = 16 * (threadIdx.x & 1);
            unsigned c = 0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) c += (whist[w][j] >> sh) & 0xffffu;
            b.tile_hist[(uint64_t)threadIdx.x * ntiles + g] = c;  // class-major
        }
        __syncthreads();
    }
}

// One block per (map, class): exclusive scan of that class's tile counts over the map's tiles
// (offsets within the class) and the class total.  Each thread takes kScanPer consecutive tiles,
// so one block scan covers kBlock * kScanPer tiles (a map of 2e7 terms in two rounds).
__global__ __launch_bounds__(kBlock) void k_part_scan(BatchDev b, const PassMap *pm) {
    constexpr int kScanPer = 16;
    __shared__ unsigned scratch[8];
    const uint32_t m = blockIdx.x >> 4, cls = blockIdx.x & 15;
    const PassMap &P = pm[m];
    if (!P.active) return;
    const uint64_t t0 = b.tstart[m], t1 = b.tstart[m + 1], ntiles = b.tstart[b.n_maps];
    // this class's counts and offsets over all tiles (class-major: contiguous per class)
    const unsigned *row_hist = b.tile_hist + (uint64_t)cls * ntiles;
    uint64_t *row_base = b.tile_base + (uint64_t)cls * ntiles;
    uint64_t carry = 0;
    for (uint64_t base = t0; base < t1; base += (uint64_t)kBlock * kScanPer) {
        const uint64_t g0 = base + (uint64_t)threadIdx.x * kScanPer;
        unsigned v[kScanPer], sum = 0;
#pragma unroll
        for (int i = 0; i < kScanPer; ++i) {
            v[i] = g0 + i < t1 ? row_hist[g0 + i] : 0u;
            sum += v[i];
        }
        unsigned tot;
        unsigned ex = block_exclusive_scan(sum, scratch, &tot);
#pragma unroll
        for (int i = 0; i < kScanPer; ++i) {
            if (g0 + i < t1) row_base[g0 + i] = carry + ex;
            ex += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) b.map_hist[m * 16 + cls] = carry;
}

// Per tile: move every element to its class's run (stable), with digits pa, pb moved to the
// top of the key; one round of 256 elements at a time (ranks from per-wave ballots), which keeps
// register use low for the many-tile launches.  The per-round counts and runs are double
// buffered, so one barrier per round suffices.  The map's first tile also writes the output
// regions and length.
template <int W, int R>
__global__ __launch_bounds__(kBlock) void k_part_scatter(BatchDev b,
                                                         const PassMap *__restrict__ pm) {
    constexpr uint64_t kTile = (uint64_t)kBlock * R;
    __shared__ unsigned wave_cnt[2][kBlock / 64][16];
    __shared__ uint64_t run[2][16];
    __shared__ uint64_t class_total[16];
    const uint32_t M = b.n_maps;
    const uint64_t ntiles = b.tstart[M];
    for (uint64_t g = blockIdx.x; g < ntiles; g += gridDim.x) {
        uint32_t m = find_bucket(b.tstart, M, g);
        const PassMap &P = pm[m];
        const uint32_t n = b.slot[m].n_spins;
        // read once: the rounds' global stores could alias the descriptor for the compiler
        const uint32_t pa = P.pa, pb = P.pb;
        uint64_t len = b.len[m * 3 + P.in_buf];
        uint64_t e0 = (g - b.tstart[m]) * kTile;
        uint64_t e1 = e0 + kTile < len ? e0 + kTile : len;
        const uint64_t *keys = b.arena[P.in_buf].keys + b.slot[m].off * W;
        const double *vals = b.arena[P.in_buf].vals + b.slot[m].off;
        uint64_t *okeys = b.arena[P.out_buf].keys + b.slot[m].off * W;
        double *ovals = b.arena[P.out_buf].vals + b.slot[m].off;
        // the class totals and this tile's offsets first, so their loads fly with the tile's
        uint64_t tile_off = 0;
        if (threadIdx.x < 16) {
            class_total[threadIdx.x] = b.map_hist[m * 16 + threadIdx.x];
            tile_off = b.tile_base[(uint64_t)threadIdx.x * ntiles + g];
        }
        const Phys ph = phys_from(region_row(b.regions, m, P.in_buf), e0);
        const bool one = e1 <= ph.hi;  // the tile lies in one class
        Key<W> kk[R];
        double vv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {  // branch-free: past the end, reload the last
            const uint64_t e = e0 + threadIdx.x + (uint64_t)r * kBlock;
            const uint64_t ec = phys_in(ph, one, e < e1 ? e : e1 - 1);
            kk[r] = kload<W>(keys + ec * W);
            vv[r] = vals[ec];
        }
        __syncthreads();
        if (threadIdx.x < 16) {
            uint64_t base = 0;
            for (unsigned c = 0; c < threadIdx.x; ++c) base += class_total[c];
            run[0][threadIdx.x] = base + tile_off;
            if (g == b.tstart[m]) {  // the output is contiguous
                uint64_t *reg = region_row(b.regions, m, P.out_buf);
                contiguous_row(reg, threadIdx.x, base);
                if (threadIdx.x == 15) {
                    const uint64_t total = base + class_total[15];
                    contiguous_row(reg, 16, total);
                    b.len[m * 3 + P.out_buf] = total;
                }
            }
        }
        const uint64_t lanemask = (1ull << lane_id()) - 1ull;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int cur = r & 1;
            const uint64_t e = e0 + threadIdx.x + (uint64_t)r * kBlock;
            const bool live = e < e1;
            const unsigned cls = live ? (kdigit<W>(kk[r], pa) << 2) | kdigit<W>(kk[r], pb) : 0u;
            // the live lanes of my class: four bit ballots combined; my rank among them, and
            // their count written by the first of them (the wave's other entries stay 0)
            uint64_t same = __ballot(live);
#pragma unroll
            for (int bit = 0; bit < 4; ++bit) {
                const uint64_t bb = __ballot((cls >> bit) & 1u);
                same &= ((cls >> bit) & 1u) ? bb : ~bb;
            }
            const unsigned below = (unsigned)__popcll(same & lanemask);
            if (lane_id() < 16) wave_cnt[cur][wave_id()][lane_id()] = 0;
            if (live && below == 0) wave_cnt[cur][wave_id()][cls] = (unsigned)__popcll(same);
            // publishes this round's counts and the runs the previous round advanced; the other
            // buffers were last read before this barrier's predecessor
            __syncthreads();
            if (live) {
                uint64_t dst = run[cur][cls] + below;
                for (unsigned w = 0; w < wave_id(); ++w) dst += wave_cnt[cur][w][cls];
                kstore<W>(okeys + dst * W, move_to_top<W>(kk[r], pa, pb, n));
                ovals[dst] = vv[r];
            }
            if (threadIdx.x < 16) {
                unsigned add = 0;
                for (unsigned w = 0; w < kBlock / 64; ++w) add += wave_cnt[cur][w][threadIdx.x];
                run[cur ^ 1][threadIdx.x] = run[cur][threadIdx.x] + add;
            }
        }
        __syncthreads();  // LDS is reused by the next tile
    }
}

// Class regions of a canonical map whose top two digits are positions n-1, n-2 (after upload).
template <int W>
__global__ void k_regions(BatchDev b, uint32_t m, uint32_t buf) {
    const uint64_t len = b.len[m * 3 + buf];
    const uint32_t n = b.slot[m].n_spins;
    const uint64_t *keys = b.arena[buf].keys + b.slot[m].off * W;
    unsigned c = threadIdx.x;
    if (c > 16) return;
    uint64_t lo = 0, hi = len;  // first index whose class >= c (the map is contiguous)
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        Key<W> k = kload<W>(keys + mid * W);
        unsigned cls = (kdigit<W>(k, n - 1) << 2) | kdigit<W>(k, n - 2);
        if (cls < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    contiguous_row(region_row(b.regions, m, buf), c, c == 16 ? len : lo);
}

// ------------------------------------------------------------------------------------------
// Rotation: plan -> splits -> merge -> finish
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool class_anti(unsigned c, unsigned qt, unsigned qs) {
    return anticomm(qt, c >> 2) != anticomm(qs, c & 3u);
}

// Merge tiles of (map, class) t = 16 m + c of a rotation: ceil((|c| + |c ^ q|) / kMergeTile)
// for an anticommuting class, ceil(|c| / kMergeTile) otherwise.
__device__ __forceinline__ unsigned class_tiles(const BatchDev &b, const PassMap *pm, uint32_t t) {
    const uint32_t m = t >> 4, c = t & 15;
    const PassMap &P = pm[m];
    if (!P.active) return 0;
    const uint64_t *reg = region_row(b.regions, m, P.in_buf);
    const uint64_t na = reg[c + 1] - reg[c];
    uint64_t nd = 0;
    if (class_anti(c, P.qt, P.qs)) {
        const unsigned cq = c ^ ((P.qt << 2) | P.qs);
        nd = reg[cq + 1] - reg[cq];
    }
    return (unsigned)((na + nd + kMergeTile - 1) / kMergeTile);
}

// What the passes do with the tiles of class c: 1 = merged by the count and write passes (the
// lower class of an anticommuting pair, or a commuting class truncation may change); 2 = copied
// by the write pass only (a commuting class truncation cannot change: the split pass records
// its counts); 3 = left in place (such a class of an appending rotation: the split pass records
// no output for it); 0 = nothing (the higher class of a pair is handled with its partner's
// tiles).  The work list holds the kind-1 tiles in (map, class) order, then the kind-2 tiles.
// Kind-3 classes keep their (empty) tiles so that the ledger partials fold in the same chunks as
// when the class is copied: the upward-rounded ledger sums do not depend on whether a rotation
// appends.
__device__ __forceinline__ unsigned tile_kind(uint32_t qt, uint32_t qs, uint32_t flags,
                                              uint32_t c) {
    if (class_anti(c, qt, qs)) return (c ^ ((qt << 2) | qs)) > c ? 1u : 0u;
    if (!(flags & kPassCommutingKept)) return 1u;
    return (flags & kPassAppend) ? 3u : 2u;
}
__device__ __forceinline__ unsigned tile_kind(const PassMap &P, uint32_t c) {
    return tile_kind(P.qt, P.qs, P.flags, c);
}

__global__ __launch_bounds__(kBlock) void k_rot_plan(BatchDev b, const PassMap *pm) {
    __shared__ unsigned scratch[8];
    const uint32_t M = b.n_maps, n = M * 16;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < n; base += kBlock) {
        uint32_t t = base + threadIdx.x;
        const unsigned tiles = t < n ? class_tiles(b, pm, t) : 0u;
        unsigned tot;
        unsigned ex = block_exclusive_scan(tiles, scratch, &tot);
        if (t < n) b.tstart[t] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) b.tstart[n] = carry;
    __syncthreads();
    // work-list offsets: kind-1 tiles first, then kind-2
    uint64_t wcarry = 0, n1 = 0;
    for (unsigned want = 1; want <= 2; ++want) {
        for (uint32_t base = 0; base < n; base += kBlock) {
            const uint32_t t = base + threadIdx.x;
            const unsigned kind = t < n ? tile_kind(pm[t >> 4], t & 15) : 0u;
            const unsigned tiles = t < n && kind == want ? class_tiles(b, pm, t) : 0u;
            unsigned tot;
            const unsigned ex = block_exclusive_scan(tiles, scratch, &tot);
            if (t < n && kind == want) b.wstart[t] = wcarry + ex;
            wcarry += tot;
        }
        if (want == 1) n1 = wcarry;
    }
    if (threadIdx.x == 0) {
        b.wstart[n] = n1;
        b.wstart[n + 1] = wcarry;
    }
    __syncthreads();
    // scan chunks of every map: ceil(tiles / kChunkTiles), exclusive prefix over maps
    uint64_t ccarry = 0;
    for (uint32_t base = 0; base < M; base += kBlock) {
        const uint32_t m = base + threadIdx.x;
        unsigned chunks = 0;
        if (m < M) {
            const uint64_t tiles = b.tstart[m * 16 + 16] - b.tstart[m * 16];
            chunks = (unsigned)((tiles + kChunkTiles - 1) / kChunkTiles);
        }
        unsigned tot;
        unsigned ex = block_exclusive_scan(chunks, scratch, &tot);
        if (m < M) b.chunk_start[m] = ccarry + ex;
        ccarry += tot;
    }
    if (threadIdx.x == 0) b.chunk_start[M] = ccarry;
}

struct MergeCtx {
    uint32_t m, c, cq;
    bool anti;
    uint64_t A0, nA, D0, nD, k;
    uint64_t off;     // the map's slot offset in the arena
    uint32_t in_buf;  // the arena buffer holding the rotation's input
};
// Where a rotation tile lies: its map and class, the class regions it merges, its index in the
// class; flags bit0: anticommuting class, bit1: last tile of its class.
struct TileDesc {
    uint64_t A0, nA, D0, nD, k, off;
    uint32_t m, c, cq, flags, in_buf, pad;
};

__device__ __forceinline__ void locate(const BatchDev &b, const PassMap *pm,
                                       const uint64_t *tstart, uint64_t g, MergeCtx &x) {
    uint32_t t = find_bucket(tstart, b.n_maps * 16, g);
    x.m = t >> 4;
    x.c = t & 15;
    const PassMap &P = pm[x.m];
    const uint64_t *reg = region_row(b.regions, x.m, P.in_buf);
    x.A0 = reg[kRegPhys + x.c];
    x.nA = reg[x.c + 1] - reg[x.c];
    x.anti = class_anti(x.c, P.qt, P.qs);
    x.cq = x.c ^ ((P.qt << 2) | P.qs);
    x.D0 = reg[kRegPhys + x.cq];
    x.nD = x.anti ? reg[x.cq + 1] - reg[x.cq] : 0;
    x.k = g - tstart[t];
    x.off = b.slot[x.m].off;
    x.in_buf = P.in_buf;
}

// What the split pass keeps in LDS per map while planning.
struct PlanMap {
    uint64_t off;
    uint32_t in_buf, qt, qs, flags, n_spins, pad;
};

// locate from LDS alone: the tile offsets, the classes' physical starts and lengths (per (map,
// class): phys, clen) and the maps' pass fields, all read while planning.
__device__ __forceinline__ void locate_planned(const BatchDev &b, const uint64_t *tstart,
                                               const uint64_t *phys, const uint64_t *clen,
                                               const PlanMap *maps, uint64_t g, MergeCtx &x) {
    uint32_t t = find_bucket(tstart, b.n_maps * 16, g);
    x.m = t >> 4;
    x.c = t & 15;
    const PlanMap &Q = maps[x.m];
    x.anti = class_anti(x.c, Q.qt, Q.qs);
    x.cq = x.c ^ ((Q.qt << 2) | Q.qs);
    const uint32_t tq = (t & ~15u) | x.cq;
    x.A0 = phys[t];
    x.nA = clen[t];
    x.D0 = phys[tq];
    x.nD = x.anti ? clen[tq] : 0;
    x.k = g - tstart[t];
    x.off = Q.off;
    x.in_buf = Q.in_buf;
}

// Tile inputs in LDS: A at [0, na), T(D) at [na, na+nd), the D element after the tile at na+nd
// (pairs with the tile's last A element), the A element before the tile at kMergeTile+1 (its
// pair, if any, was emitted by the previous tile).
template <int W> struct TileIn {
    uint64_t keys[(kMergeTile + 2) * W];
    double vals[kMergeTile + 2];
};
struct CountMisc {
    double red[6][kBlock / 64];     // per wave: l1, sq, rsq of class c, then of its partner
    unsigned redn[3][kBlock / 64];  // per wave: kept, drops, radii (class c in the low 16 bits)
    unsigned nonfinite;
};

struct TileGeom {
    uint64_t d0, d1, a0, a1, b0, b1;
    uint32_t na, nd;
    bool has_next, has_prev;
};

// Range checks of a tile's splits; a violation means an internal inconsistency (reported).
__device__ __forceinline__ bool tile_ok(const MergeCtx &x, const TileGeom &t) {
    return !(t.a0 > t.d0 || t.a1 > t.d1 || t.a1 < t.a0 || t.a0 > x.nA || t.a1 > x.nA ||
             t.d0 - t.a0 > x.nD || t.d1 - t.a1 > x.nD || t.d1 - t.a1 < t.d0 - t.a0 ||
             t.d1 - t.d0 > (uint64_t)kMergeTile);
}

// Stage a tile's inputs in LDS.  Every thread issues all of its global loads (A elements, then
// T(D) elements, in one index space) before its first LDS store, so the tile costs one memory
// latency rather than one per loop trip.
template <int W>
__device__ void load_tile(const BatchDev &b, const PassMap &P, const MergeCtx &x,
                          const Key<W> &q, TileGeom &t, TileIn<W> &in) {
    t.b0 = t.d0 - t.a0;
    t.b1 = t.d1 - t.a1;
    t.na = (uint32_t)(t.a1 - t.a0);
    t.nd = (uint32_t)(t.b1 - t.b0);
    t.has_next = x.anti && t.b1 < x.nD;
    t.has_prev = x.anti && t.a0 > 0;
    // addresses from the tile's own descriptor: no load of the map's descriptors on the way
    const uint64_t *ikeys = b.arena[x.in_buf].keys + x.off * W;
    const double *ivals = b.arena[x.in_buf].vals + x.off;
    constexpr int R = kMergeTile / kBlock;
    const uint32_t na = t.na, tot = t.na + t.nd;  // tot <= kMergeTile (checked by tile_ok)
    const uint64_t abase = x.A0 + t.a0, dbase = x.D0 + t.b0 - na;
    Key<W> kk[R];
    double vv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {  // branch-free (tot >= 1): past the end, reload the last
        const uint32_t e0 = threadIdx.x + r * kBlock, e = e0 < tot ? e0 : tot - 1;
        const uint64_t src = e < na ? abase + e : dbase + e;
        kk[r] = kload<W>(ikeys + src * W);
        vv[r] = ivals[src];
    }
    Key<W> knext, kprev;
    double vnext = 0.0;
    if (threadIdx.x == 0) {
        if (t.has_next) {
            knext = kload<W>(ikeys + (x.D0 + t.b1) * W);
            vnext = ivals[x.D0 + t.b1];
        }
        if (t.has_prev) kprev = kload<W>(ikeys + (x.A0 + t.a0 - 1) * W);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t e = threadIdx.x + r * kBlock;
        if (e < tot) {
            kstore<W>(in.keys + e * W, e < na ? kk[r] : kxor<W>(kk[r], q));
            in.vals[e] = vv[r];
        }
    }
    if (threadIdx.x == 0) {
        if (t.has_next) {
            kstore<W>(in.keys + tot * W, kxor<W>(knext, q));
            in.vals[tot] = vnext;
        }
        if (t.has_prev) kstore<W>(in.keys + (kMergeTile + 1) * W, kprev);
    }
}

// One output key of the rotation: value and merged radius from its a-slot (own key) and/or its
// d-slot (partner), exactly as the reference merges two raw branches (tpgp_rev.c:458-464,
// 695-706); the merged value is one IEEE addition, independent of order.  Radii = false drops
// the radius arithmetic at compile time (the value is the same).
struct Coef {
    double c, s, rc, rs, thr;
    bool anti, validated, dneg;
};
template <bool Radii = true>
__device__ __forceinline__ void merged(const Coef &k, bool hasA, double va, bool hasD, double vd,
                                       double &v, double &r, bool &bad) {
    r = 0.0;
    if (!k.anti) {
        v = va;
        return;
    }
    double xa = 0.0, yd = 0.0, ra = 0.0, rd = 0.0;
    if (hasA) {
        xa = va * k.c;
        if ((Radii && k.validated)) ra = upadd(upmul(__builtin_fabs(va), k.rc), round_radius(xa));
    }
    if (hasD) {
        double y = vd * k.s;
        yd = k.dneg ? -y : y;
        if ((Radii && k.validated)) rd = upadd(upmul(__builtin_fabs(vd), k.rs), round_radius(y));
        if (!__builtin_isfinite(y)) bad = true;
    }
    if (!__builtin_isfinite(xa) || !__builtin_isfinite(ra) || !__builtin_isfinite(rd)) bad = true;
    if (hasA && hasD) {
        v = xa + yd;
        r = upadd(ra, rd);
        if ((Radii && k.validated) && xa != 0.0 && yd != 0.0) r = upadd(r, round_radius(v));
    } else if (hasA) {
        v = xa;
        r = ra;
    } else {
        v = yd;
        r = rd;
    }
    if (!__builtin_isfinite(v) || !__builtin_isfinite(r)) bad = true;
}

// This thread's merge-path split of the staged tile: the number of A elements among the first
// kMergeItems * threadIdx.x positions of merge(A, T(D)) (0 past the tile's end).
template <int W>
__device__ __forceinline__ uint32_t local_split(const TileIn<W> &in, const TileGeom &t) {
    const uint32_t nloc = (uint32_t)(t.d1 - t.d0);
    const uint32_t dt = threadIdx.x * kMergeItems;
    if (dt >= nloc) return 0;
    const uint32_t na = t.na, nd = t.nd;
    uint32_t lo = dt > nd ? dt - nd : 0, hi = dt < na ? dt : na;
    while (lo < hi) {  // first i with A[i] > D[dt-i-1]
        uint32_t mid = (lo + hi) >> 1;
        if (klt<W>(kload<W>(in.keys + (na + dt - mid - 1) * W), kload<W>(in.keys + mid * W)))
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// Walk this thread's kMergeItems merge-path positions of the tile from its split `lo`, calling
// f(key, hasA, va, hasD, vd, i, j) once per output key in ascending order (i, j: the LDS
// indices of its A and T(D) elements, meaningful where hasA / hasD; an absent side's value is
// passed as 0.0).
// Branch-free steps: the current A and T(D) keys stay in registers (read at clamped indices
// inside the staged tile when a run is exhausted, and then ignored); each step chooses the A
// element (with its equal T(D) partner, which may be the element after the tile) or the T(D)
// element alone by selects, so f is inlined once and the wave does not diverge on the kind of
// step.
template <int W, typename F>
__device__ __forceinline__ void walk_from(const TileIn<W> &in, const TileGeom &t, uint32_t lo,
                                          F &&f) {
    const uint32_t nloc = (uint32_t)(t.d1 - t.d0);
    const uint32_t dt = threadIdx.x * kMergeItems;
    if (dt >= nloc) return;
    const uint32_t na = t.na, nd = t.nd, nd_look = nd + (t.has_next ? 1u : 0u);
    auto Akey = [&](uint32_t i) { return kload<W>(in.keys + i * W); };
    auto Dkey = [&](uint32_t j) { return kload<W>(in.keys + (na + j) * W); };
    const uint32_t limit = (nloc - dt) < (uint32_t)kMergeItems ? (nloc - dt) : (uint32_t)kMergeItems;
    uint32_t i = lo, j = dt - lo, used = 0;
    // a D element equal to the A element just before this split was consumed with it
    if (j < nd && (i > 0 || t.has_prev)) {
        Key<W> pa = i > 0 ? Akey(i - 1) : kload<W>(in.keys + (kMergeTile + 1) * W);
        if (keq<W>(pa, Dkey(j))) {
            ++j;
            ++used;
        }
    }
    const uint32_t alast = na ? na - 1 : 0, dlast = nd_look ? nd_look - 1 : 0;
    Key<W> x = Akey(i < alast ? i : alast), y = Dkey(j < dlast ? j : dlast);
#pragma unroll
    for (int step = 0; step < kMergeItems; ++step) {
        if (used >= limit) break;
        const bool hasA = i < na, hasD = j < nd;
        const bool takeA = hasA && (!hasD || !klt<W>(y, x));  // A[i] <= D[j]
        const bool takeD = takeA ? (j < nd_look && keq<W>(y, x)) : true;
        const uint32_t ia = i < alast ? i : alast, jd = j < dlast ? j : dlast;
        const double va = takeA ? in.vals[ia] : 0.0, vd = takeD ? in.vals[na + jd] : 0.0;
        f(takeA ? x : y, takeA, va, takeD, vd, i, j);
        i += takeA ? 1u : 0u;
        j += takeD ? 1u : 0u;
        used += (takeA ? 1u : 0u) + (takeD ? 1u : 0u);
        x = Akey(i < alast ? i : alast);
        y = Dkey(j < dlast ? j : dlast);
    }
}

// walk_from with a branch per kind of step (A with its partner, A alone, T(D) alone): f is
// inlined three times.  The write pass keeps this form: it fits its seven-wave register budget
// without spills, which the select form does not.
template <int W, typename F>
__device__ __forceinline__ void walk_branchy(const TileIn<W> &in, const TileGeom &t, uint32_t lo,
                                          F &&f) {
    const uint32_t nloc = (uint32_t)(t.d1 - t.d0);
    const uint32_t dt = threadIdx.x * kMergeItems;
    if (dt >= nloc) return;
    const uint32_t na = t.na, nd = t.nd, nd_look = nd + (t.has_next ? 1u : 0u);
    auto Akey = [&](uint32_t i) { return kload<W>(in.keys + i * W); };
    auto Dkey = [&](uint32_t j) { return kload<W>(in.keys + (na + j) * W); };
    const uint32_t limit = (nloc - dt) < (uint32_t)kMergeItems ? (nloc - dt) : (uint32_t)kMergeItems;
    uint32_t i = lo, j = dt - lo, used = 0;
    // a D element equal to the A element just before this split was consumed with it
    if (j < nd && (i > 0 || t.has_prev)) {
        Key<W> pa = i > 0 ? Akey(i - 1) : kload<W>(in.keys + (kMergeTile + 1) * W);
        if (keq<W>(pa, Dkey(j))) {
            ++j;
            ++used;
        }
    }
    while (used < limit) {
        bool takeA;
        if (i >= na)
            takeA = false;
        else if (j >= nd)
            takeA = true;
        else
            takeA = !klt<W>(Dkey(j), Akey(i));  // A[i] <= D[j]
        if (takeA) {
            Key<W> ka = Akey(i);
            if (j < nd_look && keq<W>(ka, Dkey(j))) {
                f(ka, true, in.vals[i], true, in.vals[na + j], i, j);
                ++i;
                ++j;
                used += 2;
            } else {
                f(ka, true, in.vals[i], false, 0.0, i, j);
                ++i;
                ++used;
            }
        } else {
            f(Dkey(j), false, 0.0, true, in.vals[na + j], i, j);
            ++j;
            ++used;
        }
    }
}

template <int W, typename F>
__device__ __forceinline__ void walk_tile(const TileIn<W> &in, const TileGeom &t, F &&f) {
    walk_from<W>(in, t, local_split<W>(in, t),
                 [&](const Key<W> &key, bool hasA, double va, bool hasD, double vd, uint32_t,
                     uint32_t) { f(key, hasA, va, hasD, vd); });
}

__device__ __forceinline__ Coef coef_of(const PassMap &P, const MergeCtx &x) {
    Coef k;
    k.c = P.c;
    k.s = P.s;
    k.rc = P.rc;
    k.rs = P.rs;
    k.thr = P.threshold;
    k.anti = x.anti;
    k.validated = P.flags & 1u;
    const unsigned ph = (pauli_phase(P.qt, x.cq >> 2) + pauli_phase(P.qs, x.cq & 3u)) & 3u;
    k.dneg = ph != 1;
    return k;
}

__device__ __forceinline__ void log_entry(const BatchDev &b, uint64_t k0, uint64_t k1, double v,
                                          double aux, uint32_t m, uint32_t rot, uint32_t kind) {
    unsigned long long at = atomicAdd((unsigned long long *)b.log_count, 1ull);
    if (at < b.log_cap) {
        LogDev &L = b.log[at];
        L.key[0] = k0;
        L.key[1] = k1;
        L.value = v;
        L.aux = aux;
        L.map = m;
        L.rotation = rot;
        L.kind = kind;
    }
}

// Merge-path split by 8 cooperating lanes (lanes 8*grp .. 8*grp+7 of the wave, `sub` = lane
// within the group): the number of A elements among the first d elements of merge(A, T(D))
// (ties: A first).  P(x) = A[x] > T(D[d-x-1]) is monotone; the split is the first x in
// [lo, hi] with P(x) true (P(hi) counts as true), found with 8-ary rounds.  All 8 lanes return
// the same value.  Lanes of one group stay in lockstep; groups may take different numbers of
// rounds (the ballot of a round only holds the lanes still searching).
template <int W>
__device__ uint64_t split_search8(const uint64_t *keys, const MergeCtx &x, const Key<W> &q,
                                  uint64_t d, unsigned grp, unsigned sub) {
    uint64_t lo = d > x.nD ? d - x.nD : 0, hi = d < x.nA ? d : x.nA;
    auto pred = [&](uint64_t xi) -> bool {
        Key<W> a = kload<W>(keys + (x.A0 + xi) * W);
        Key<W> dk = kxor<W>(kload<W>(keys + (x.D0 + d - xi - 1) * W), q);
        return klt<W>(dk, a);
    };
    while (hi > lo) {
        const uint64_t span = hi - lo;
        if (span < 8) {  // lane `span` maps to hi (P(hi) is true): the group's byte is never 0
            const uint64_t xi = lo + sub;
            const bool p = xi < hi ? pred(xi) : true;
            const unsigned byte = (unsigned)(__ballot(p) >> (8 * grp)) & 0xffu;
            lo = lo + (uint64_t)(__ffs(byte) - 1);
            break;
        }
        const uint64_t xi = lo + ((span * sub) >> 3);
        const unsigned byte = (unsigned)(__ballot(pred(xi)) >> (8 * grp)) & 0xffu;
        if (byte == 0) {
            lo = lo + ((span * 7) >> 3) + 1;
        } else {
            const int f = __ffs(byte) - 1;
            if (f == 0) break;  // P(lo) true: the split is lo
            const uint64_t xf = lo + ((span * (uint64_t)f) >> 3);
            const uint64_t xp = lo + ((span * (uint64_t)(f - 1)) >> 3);
            lo = xp + 1;
            hi = xf;
        }
    }
    return lo;
}

// Pass 0 of a rotation: every tile's descriptor and merge-path start split, 8 lanes per tile
// (32 tiles per workgroup, all searching at once so the dependent loads overlap).
// With `plan` (at most kPlanMaps maps), every workgroup first derives the tile offsets of all
// (map, class) pairs itself -- what k_rot_plan computes -- and workgroup 0 publishes them and
// the scan chunks for the later passes: one launch less per rotation.
constexpr uint32_t kPlanMaps = kBlock / 16;
template <int W>
__global__ __launch_bounds__(kBlock) void k_rot_splits(BatchDev b, const PassMap *__restrict__ pm,
                                                       int plan) {
    __shared__ uint64_t s_tstart[kBlock + 1];
    __shared__ uint64_t s_wstart[kBlock];
    __shared__ uint64_t s_phys[kBlock], s_clen[kBlock];  // per (map, class), when planning
    __shared__ PlanMap s_map[kPlanMaps];
    __shared__ unsigned scratch[8];
    const uint64_t *tstart = b.tstart, *wstart = b.wstart;
    if (plan) {
        const uint32_t M = b.n_maps, n = M * 16;  // n <= kBlock
        // what class_tiles, tile_kind and locate read, in two rounds of loads (the pass fields
        // and slots, then the region table); kept in LDS for the tiles
        unsigned tiles = 0, kind = 0;
        if (threadIdx.x < n) {
            const uint32_t m = threadIdx.x >> 4, c = threadIdx.x & 15;
            const PassMap P = pm[m];
            const uint64_t off = b.slot[m].off;
            const uint32_t n_spins = b.slot[m].n_spins;
            const unsigned cq = c ^ ((P.qt << 2) | P.qs);
            const bool anti = class_anti(c, P.qt, P.qs);
            const uint64_t *reg = region_row(b.regions, m, P.in_buf);
            const uint64_t r0 = reg[c], r1 = reg[c + 1], q0 = reg[cq], q1 = reg[cq + 1];
            s_phys[threadIdx.x] = reg[kRegPhys + c];
            s_clen[threadIdx.x] = r1 - r0;
            // the map's 16 threads store the same record (no branch for the loads to sink into)
            s_map[m] = PlanMap{off, P.in_buf, P.qt, P.qs, P.flags, n_spins, 0};
            // class_tiles (lengths masked, not branched on: their loads stay in the second round)
            const uint64_t nd = (q1 - q0) & (anti ? ~0ull : 0ull);
            const uint64_t len = (r1 - r0 + nd) & (P.active ? ~0ull : 0ull);
            tiles = (unsigned)((len + kMergeTile - 1) / kMergeTile);
            kind = tile_kind(P, c);
        }
        unsigned tot, t1, t2;
        const unsigned ex = block_exclusive_scan(tiles, scratch, &tot);
        const unsigned x1 = block_exclusive_scan(kind == 1 ? tiles : 0u, scratch, &t1);
        const unsigned x2 = block_exclusive_scan(kind == 2 ? tiles : 0u, scratch, &t2);
        if (threadIdx.x < n) {
            s_tstart[threadIdx.x] = ex;
            s_wstart[threadIdx.x] = kind == 2 ? t1 + x2 : x1;
        }
        if (threadIdx.x == 0) s_tstart[n] = tot;
        __syncthreads();
        if (blockIdx.x == 0) {
            if (threadIdx.x < n) b.tstart[threadIdx.x] = s_tstart[threadIdx.x];
            if (threadIdx.x == 0) {
                b.tstart[n] = s_tstart[n];
                b.wstart[n] = t1;
                b.wstart[n + 1] = t1 + t2;
            }
            const unsigned chunks =
                threadIdx.x < M ? (unsigned)((s_tstart[threadIdx.x * 16 + 16] -
                                              s_tstart[threadIdx.x * 16] + kChunkTiles - 1) /
                                             kChunkTiles)
                                : 0u;
            unsigned ctot;
            const unsigned cex = block_exclusive_scan(chunks, scratch, &ctot);
            if (threadIdx.x < M) b.chunk_start[threadIdx.x] = cex;
            if (threadIdx.x == 0) b.chunk_start[M] = ctot;
        }
        tstart = s_tstart;
        wstart = s_wstart;
    }
    const uint64_t ntiles = tstart[b.n_maps * 16];
    const unsigned grp = lane_id() >> 3, sub = lane_id() & 7u;
    const uint64_t per_block = kBlock / 8;
    for (uint64_t base = (uint64_t)blockIdx.x * per_block; base < ntiles;
         base += (uint64_t)gridDim.x * per_block) {
        const uint64_t g = base + threadIdx.x / 8;
        if (g >= ntiles) continue;  // whole 8-lane groups leave together
        MergeCtx x;
        uint32_t n_sp, qt, qs, flags;
        if (plan) {
            locate_planned(b, tstart, s_phys, s_clen, s_map, g, x);
            const PlanMap &Q = s_map[x.m];
            n_sp = Q.n_spins;
            qt = Q.qt;
            qs = Q.qs;
            flags = Q.flags;
        } else {
            locate(b, pm, tstart, g, x);
            const PassMap &P = pm[x.m];
            n_sp = b.slot[x.m].n_spins;
            qt = P.qt;
            qs = P.qs;
            flags = P.flags;
        }
        const uint64_t d0 = x.k * kMergeTile;
        uint64_t a0 = d0;
        // the higher class of an anticommuting pair is counted and written with its partner's
        // tiles and splits: its own are never read
        if (x.anti && x.c < x.cq) {
            const uint64_t *ikeys = b.arena[x.in_buf].keys + x.off * W;
            const Key<W> q = kxor<W>(kdigit_mask<W>(n_sp - 1, qt), kdigit_mask<W>(n_sp - 2, qs));
            a0 = split_search8<W>(ikeys, x, q, d0, grp, sub);
        }
        if (sub == 0) {
            b.splits[g] = a0;
            TileDesc t;
            t.A0 = x.A0;
            t.nA = x.nA;
            t.D0 = x.D0;
            t.nD = x.nD;
            t.k = x.k;
            t.m = x.m;
            t.c = x.c;
            t.cq = x.cq;
            t.off = x.off;
            t.in_buf = x.in_buf;
            t.pad = 0;
            const bool last = g + 1 >= tstart[x.m * 16 + x.c + 1];
            t.flags = (x.anti ? 1u : 0u) | (last ? 2u : 0u);
            b.tdesc[g] = t;
            const unsigned kind = tile_kind(qt, qs, flags, x.c);
            if (kind == 1 || kind == 2) b.work[wstart[x.m * 16 + x.c] + x.k] = g;
            if (kind >= 2) {  // unchanged keys: the count pass does not visit the tile; copied
                              // (kind 2) or left in place with no output of its own (kind 3)
                const uint64_t d1 = d0 + kMergeTile < x.nA ? d0 + kMergeTile : x.nA;
                b.tile_hist[g] = kind == 2 ? (unsigned)(d1 - d0) : 0u;
                b.tile_led[g] = TileLedger{};
            }
        }
    }
}

// A tile's context and splits from its descriptor (uniform loads: no bucket search).
__device__ __forceinline__ void tile_from_desc(const BatchDev &b, uint64_t g, MergeCtx &x,
                                               uint64_t &a0, uint64_t &a1) {
    const TileDesc t = b.tdesc[g];
    x.A0 = t.A0;
    x.nA = t.nA;
    x.D0 = t.D0;
    x.nD = t.nD;
    x.k = t.k;
    x.m = t.m;
    x.c = t.c;
    x.cq = t.cq;
    x.anti = (t.flags & 1u) != 0;
    x.off = t.off;
    x.in_buf = t.in_buf;
    a0 = b.splits[g];
    const uint64_t next = b.splits[g + 1];  // unconditional (splits has maxT + 1 entries): the
    a1 = (t.flags & 2u) ? t.nA : next;      // load does not wait for the descriptor
}

// The coefficients of the partner class c ^ q of an anticommuting tile: its own entries are
// the tile's D elements, its partner's the tile's A elements.
__device__ __forceinline__ Coef partner_coef(const PassMap &P, const MergeCtx &x) {
    Coef k = coef_of(P, x);
    const unsigned ph = (pauli_phase(P.qt, x.c >> 2) + pauli_phase(P.qs, x.c & 3u)) & 3u;
    k.dneg = ph != 1;
    return k;
}

// Per-class tallies of one thread's outputs in the count pass.
struct TileAcc {
    unsigned cnt;  // kept (bits 0-7), dropped (8-15), nonzero radii (16-23): each <= 4
    double l1, sq, rsq;
};
// One output key of class `k`: kept, dropped into the ledger or cancelled (tpgp_rev.c:690-729),
// with the exact-ledger log entries when asked.
template <int W, bool V>
__device__ __forceinline__ void account(const BatchDev &b, uint32_t m, uint32_t rot, bool want_log,
                                        const Coef &k, const Key<W> &key, bool hasA, double va,
                                        bool hasD, double vd, TileAcc &a, bool &bad) {
    double v, r;
    merged<V>(k, hasA, va, hasD, vd, v, r, bad);
    const uint64_t k0 = key.w[0], k1 = W > 1 ? key.w[W - 1] : 0;
    if (V && r != 0.0) {
        a.cnt += 1u << 16;
        a.rsq += upmul(r, r);
        if (want_log) log_entry(b, k0, k1, r, 0.0, m, rot, 1);
    }
    // tallies by selects: adding +0.0 to a non-negative partial sum leaves it unchanged, so the
    // sums are those of the dropped outputs alone
    const bool kept_or_dropped = v != 0.0;  // an exact cancellation is dropped silently
    const bool drop = kept_or_dropped && !kzero<W>(key) && __builtin_fabs(v) <= k.thr;
    a.cnt += drop ? (1u << 8) : (kept_or_dropped ? 1u : 0u);
    a.l1 += drop ? __builtin_fabs(v) : 0.0;
    a.sq += drop ? upmul(v, v) : 0.0;
    if (want_log && drop) log_entry(b, k0, k1, __builtin_fabs(v), v, m, rot, 0);
}

// Pass 1 of a rotation: per tile, the merge-path splits, the number of kept outputs, the
// ledger partial sums and, when asked, the log.  An anticommuting class c and its partner
// c ^ q share their tiles (tile k of both covers the same merge-path range, as in the write
// pass): the lower class's workgroup stages the pair once and one walk tallies both classes;
// the partner's workgroup returns at once.
//
// Ledger partials.  Every summand is >= 0 (|v|, and squares rounded up), so the sums use plain
// round-to-nearest additions in a fixed tree (deterministic) and are made upper bounds once per
// tile: at most kMergeItems - 1 = 3 additions per thread in sequence, 6 levels across the wave
// and 3 in sequence across the waves give a depth of d = 12, so the computed sum S' >=
// (1 - gamma_12) S with gamma_d = d u / (1 - d u), u = 2^-53, and next_up(S' (1 + 24 u)) >= S.
// V: some map of the launch is validated (radii and their ledger sums); without it the radius
// arithmetic is compiled out and one-word keys fit 80 VGPRs: six waves per SIMD.
template <int W, bool V>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W == 1 && !V ? 6 : 1, 8)))
void k_rot_count(BatchDev b, const PassMap *__restrict__ pm) {
    __shared__ TileIn<W> in;
    __shared__ CountMisc misc;
    const uint64_t nwork = b.wstart[b.n_maps * 16];  // kind-1 tiles
    uint64_t gn = blockIdx.x < nwork ? b.work[blockIdx.x] : 0;
    for (uint64_t w = blockIdx.x; w < nwork; w += gridDim.x) {
        const uint64_t g = gn;
        gn = w + gridDim.x < nwork ? b.work[w + gridDim.x] : 0;  // the next tile, early
        MergeCtx x;
        TileGeom t;
        tile_from_desc(b, g, x, t.a0, t.a1);
        if (x.anti && x.cq < x.c) continue;  // tallied with its partner's tile
        const PassMap &P = pm[x.m];
        const uint32_t n = b.slot[x.m].n_spins;
        const Key<W> q = kxor<W>(kdigit_mask<W>(n - 1, P.qt), kdigit_mask<W>(n - 2, P.qs));
        const uint64_t total = x.nA + x.nD;
        t.d0 = x.k * kMergeTile;
        t.d1 = t.d0 + kMergeTile < total ? t.d0 + kMergeTile : total;
        if (!x.anti && (P.flags & kPassCommutingKept)) {  // every key kept as it is
            if (threadIdx.x == 0) {
                b.tile_hist[g] = (unsigned)(t.d1 - t.d0);
                b.tile_led[g] = TileLedger{};
            }
            continue;
        }
        const bool anti = x.anti;
        const uint64_t g2 = anti ? b.tstart[x.m * 16 + x.cq] + x.k : g;
        if (threadIdx.x == 0) misc.nonfinite = 0;
        if (!tile_ok(x, t)) {
            if (threadIdx.x == 0) {
                atomicOr(&b.counter[1], 1u);
                b.tile_hist[g] = 0;
                b.tile_led[g] = TileLedger{};
                b.tile_hist[g2] = 0;
                b.tile_led[g2] = TileLedger{};
            }
            __syncthreads();
            continue;
        }
        load_tile<W>(b, P, x, q, t, in);
        __syncthreads();
        const Coef k0 = coef_of(P, x);
        const Coef k1 = anti ? partner_coef(P, x) : k0;
        const bool want_log = (P.flags & 2u) != 0;
        TileAcc a0{}, a1{};
        bool bad = false;
        const uint32_t lo = local_split<W>(in, t);
        b.tsplit[g * kBlock + threadIdx.x] = (uint16_t)lo;  // the write pass walks from it too
        walk_from<W>(in, t, lo, [&](const Key<W> &key, bool hasA, double va, bool hasD,
                                    double vd, uint32_t, uint32_t) {
            account<W, V>(b, x.m, P.rot_seq, want_log, k0, key, hasA, va, hasD, vd, a0, bad);
            if (anti)
                account<W, V>(b, x.m, P.rot_seq, want_log, k1, kxor<W>(key, q), hasD, vd, hasA,
                              va, a1, bad);
        });
        if (bad) atomicOr(&misc.nonfinite, 1u);
        // wave sums (lane 63) of both classes' tallies, counts packed in 16-bit halves (a tile
        // keeps and drops <= kMergeTile per class)
        const bool radii = V && k0.validated;  // uniform: without enclosures every radius is 0
        const unsigned nk = wave_scan_u32((a0.cnt & 0xffu) | ((a1.cnt & 0xffu) << 16));
        const unsigned nd =
            wave_scan_u32(((a0.cnt >> 8) & 0xffu) | (((a1.cnt >> 8) & 0xffu) << 16));
        a0.l1 = wave_scan_f64(a0.l1);
        a0.sq = wave_scan_f64(a0.sq);
        if (anti) {
            a1.l1 = wave_scan_f64(a1.l1);
            a1.sq = wave_scan_f64(a1.sq);
        }
        unsigned nr = 0;
        if (radii) {
            nr = wave_scan_u32(((a0.cnt >> 16) & 0xffu) | (((a1.cnt >> 16) & 0xffu) << 16));
            a0.rsq = wave_scan_f64(a0.rsq);
            a1.rsq = wave_scan_f64(a1.rsq);
        }
        if (lane_id() == 63) {
            const unsigned w = wave_id();
            misc.red[0][w] = a0.l1;
            misc.red[1][w] = a0.sq;
            misc.red[2][w] = radii ? a0.rsq : 0.0;
            misc.red[3][w] = anti ? a1.l1 : 0.0;
            misc.red[4][w] = anti ? a1.sq : 0.0;
            misc.red[5][w] = anti && radii ? a1.rsq : 0.0;
            misc.redn[0][w] = nk;
            misc.redn[1][w] = nd;
            misc.redn[2][w] = radii ? nr : 0u;
        }
        __syncthreads();
        if (threadIdx.x < 2 && (threadIdx.x == 0 || anti)) {  // thread h writes class h's tile
            constexpr int kDepth = (kMergeItems - 1) + 6 + (kBlock / 64 - 1);
            static_assert(kDepth == 12, "the tree depth the inflation is derived for");
            constexpr double kInflate = 1.0 + 2.0 * kDepth / 9007199254740992.0;  // exact
            const unsigned h = threadIdx.x, sh = 16 * h;
            double s1 = misc.red[3 * h][0], s2 = misc.red[3 * h + 1][0],
                   s3 = misc.red[3 * h + 2][0];
            unsigned c0 = (misc.redn[0][0] >> sh) & 0xffffu;
            unsigned c1 = (misc.redn[1][0] >> sh) & 0xffffu, c2 = (misc.redn[2][0] >> sh) & 0xffffu;
            for (int w = 1; w < kBlock / 64; ++w) {
                s1 += misc.red[3 * h][w];
                s2 += misc.red[3 * h + 1][w];
                s3 += misc.red[3 * h + 2][w];
                c0 += (misc.redn[0][w] >> sh) & 0xffffu;
                c1 += (misc.redn[1][w] >> sh) & 0xffffu;
                c2 += (misc.redn[2][w] >> sh) & 0xffffu;
            }
            TileLedger tl;
            tl.l1 = upmul(s1, kInflate);
            tl.sq = upmul(s2, kInflate);
            tl.rsq = upmul(s3, kInflate);
            tl.dropped = c1;
            tl.radii = c2;
            tl.nonfinite = misc.nonfinite;
            tl.pad = 0;
            b.tile_led[h ? g2 : g] = tl;
            b.tile_hist[h ? g2 : g] = c0;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void ledger_fold(TileLedger &a, const TileLedger &x) {
    a.l1 = upadd(a.l1, x.l1);
    a.sq = upadd(a.sq, x.sq);
    a.rsq = upadd(a.rsq, x.rsq);
    a.dropped += x.dropped;
    a.radii += x.radii;
    a.nonfinite |= x.nonfinite;
}
// Fixed-shape block reduction of per-thread ledgers by wave shuffles and one fold across the
// waves (deterministic for a given input); every thread gets the result.
__device__ TileLedger block_ledger(TileLedger a, TileLedger *sh /*[kBlock / 64]*/) {
    for (int d = 1; d < 64; d <<= 1) {
        TileLedger o;
        o.l1 = __shfl_down(a.l1, d, 64);
        o.sq = __shfl_down(a.sq, d, 64);
        o.rsq = __shfl_down(a.rsq, d, 64);
        o.dropped = __shfl_down(a.dropped, d, 64);
        o.radii = __shfl_down(a.radii, d, 64);
        o.nonfinite = __shfl_down(a.nonfinite, d, 64);
        if (lane_id() + d < 64) ledger_fold(a, o);
    }
    if (lane_id() == 0) sh[wave_id()] = a;
    __syncthreads();
    TileLedger r = sh[0];
    for (int w = 1; w < kBlock / 64; ++w) ledger_fold(r, sh[w]);
    __syncthreads();
    return r;
}

// Sum of one uint64_t per thread over the block; every thread gets it.
__device__ uint64_t block_sum_u64(uint64_t v, uint64_t *sh /*[kBlock / 64]*/) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (lane_id() == 0) sh[wave_id()] = v;
    __syncthreads();
    uint64_t s = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w];
    __syncthreads();
    return s;
}

// Scan level 1, one block per chunk of kChunkTiles tiles of one map: tile offsets within the
// chunk (exclusive scan of the kept counts), the chunk total and the chunk's ledger partials.
// direct: the block also sums the kept counts of all the map's tiles before its chunk and writes
// the chunk's output offset (chunk_base, with the rotation's output base) itself, so the write
// pass can start without the map-level scan; the rest of that scan (regions, ledger, record)
// runs inside the write pass (k_rot_write with finish).  Every block reads the counts of the
// map's earlier tiles (4 KB per earlier chunk, in the same round of loads as its own tiles), so
// the form is used for maps of at most kDirectScanChunks chunks.
constexpr uint64_t kDirectScanChunks = 128;
__global__ __launch_bounds__(kBlock) void k_rot_scan1(BatchDev b, const PassMap *pm, int direct) {
    static_assert(kChunkTiles == kBlock, "one tile per thread");
    __shared__ unsigned scratch[8];
    __shared__ TileLedger sh[kBlock / 64];
    __shared__ uint64_t sh64[kBlock / 64];
    const uint32_t M = b.n_maps;
    const uint64_t nchunks = b.chunk_start[M];
    for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const uint32_t m = find_bucket(b.chunk_start, M, ch);
        // Loads in two dependent rounds, no branch on a loaded value before the second: the
        // map's tile range and pass descriptor, then the extent, the chunk's counts and ledgers
        // and (direct) the earlier tiles' counts.  The ledger fold waits until all are issued.
        const uint64_t t1 = b.tstart[m * 16 + 16], tm = b.tstart[m * 16];
        const uint32_t flags = pm[m].flags, in_buf = pm[m].in_buf;
        const uint64_t g0 = tm + (ch - b.chunk_start[m]) * kChunkTiles;
        const uint64_t g = g0 + threadIdx.x;
        // outputs start at the input's used extent when appending (the row is rewritten only by
        // the write pass, after this launch)
        const uint64_t top = region_row(b.regions, m, in_buf)[kRegTop];
        const unsigned cnt = g < t1 ? b.tile_hist[g] : 0u;
        TileLedger led{};
        if (g < t1) led = b.tile_led[g];
        uint64_t pre = 0;
        const uint64_t obase = (flags & kPassAppend) ? top : 0;
        if (direct) {
            // counts of tiles tm .. g0 - 1: the 16-byte words qa .. qz - 1 that cover them are
            // summed whole, and thread 0 takes off the at most 3 entries before tm and 3 from g0
            // on (tile_hist holds 16 entries per tile of the allocation, so the words and entries
            // up to g0 + 2 are inside it)
            const uint4 *h4 = reinterpret_cast<const uint4 *>(b.tile_hist);
            const uint64_t qa = tm >> 2, qz = (g0 + 3) >> 2;
            unsigned edge = 0;
            if (threadIdx.x == 0) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const uint64_t jh = qa * 4 + k, jt = g0 + k;
                    edge += (jh < tm ? b.tile_hist[jh] : 0u) + (jt < qz * 4 ? b.tile_hist[jt] : 0u);
                }
            }
            constexpr int R = 16;  // one round of loads for maps of up to 64 chunks
            for (uint64_t qb = qa; qb < qz; qb += (uint64_t)R * kBlock) {
                uint4 v[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint64_t q = qb + (uint64_t)r * kBlock + threadIdx.x;
                    v[r] = q < qz ? h4[q] : make_uint4(0u, 0u, 0u, 0u);
                }
                unsigned s32 = 0;  // <= 16 words of 4 counts <= 2048
#pragma unroll
                for (int r = 0; r < R; ++r) s32 += v[r].x + v[r].y + v[r].z + v[r].w;
                pre += s32;
            }
            pre -= edge;  // modulo 2^64: the block sum is exact
        }
        TileLedger L{};
        if (g < t1) ledger_fold(L, led);
        unsigned tot;
        const unsigned at = block_exclusive_scan(cnt, scratch, &tot);
        if (g < t1) b.tile_base[g] = at;
        const TileLedger C = block_ledger(L, sh);
        if (threadIdx.x == 0) {
            b.chunk_sum[ch] = tot;
            b.chunk_led[ch] = C;
        }
        if (direct) {
            pre = block_sum_u64(pre, sh64);
            if (threadIdx.x == 0) b.chunk_base[ch] = obase + pre;
        }
    }
}

// The end of a map's rotation scan (whole block, chunk_base and tile_base of the map written and
// published): output regions and length, overflow, and the record with the map's ledger T.
// base: the slot index of the first written output (the input's used extent when appending,
// else 0), already included in chunk_base.  Lanes 0-15 of the first wave take one class each.
__device__ void finish_map_scan(const BatchDev &b, const PassMap &P, uint32_t m, uint64_t acc,
                                uint64_t base, const TileLedger &T) {
    if (threadIdx.x < 64) {
        const unsigned c = lane_id();
        const bool mine = c < 16;
        const uint64_t *rin = region_row(b.regions, m, P.in_buf);
        uint64_t *rout = region_row(b.regions, m, P.out_buf);
        const bool append = (P.flags & kPassAppend) != 0;
        // the input row, read before any of it is rewritten (appending: the same words)
        uint64_t lo = 0, hi = 0, ph = 0, ws = ~0ull;
        if (mine) {
            lo = rin[c];
            hi = rin[c + 1];
            ph = rin[kRegPhys + c];
            const uint64_t tm0 = b.tstart[m * 16], ch0 = b.chunk_start[m];
            const uint64_t t = b.tstart[m * 16 + c], te = b.tstart[m * 16 + c + 1];
            if (t < te)  // where class c's first tile writes, within the written output
                ws = b.chunk_base[ch0 + (t - tm0) / kChunkTiles] + b.tile_base[t] - base;
        }
        const uint64_t len_in = rin[16];
        // written output of class c: positions wr .. wr_next of the acc written terms; a class
        // without tiles starts where the next class with tiles does (offsets never decrease),
        // so wr is the minimum over classes c..15 (acc where none)
        uint64_t wr = ws < acc ? ws : acc;
        for (unsigned d = 1; d < 16; d <<= 1) {
            const uint64_t o = __shfl_down(wr, d, 64);
            if (c + d < 16) wr = o < wr ? o : wr;
        }
        uint64_t wr_next = __shfl_down(wr, 1, 64);
        if (c == 15) wr_next = acc;
        const bool in_place = mine && append && tile_kind(P, c) == 3u;
        const uint64_t n = !mine ? 0 : in_place ? hi - lo : wr_next - wr;
        uint64_t inc = n, kept = in_place ? n : 0;  // kept: terms of the classes left in place
        for (unsigned d = 1; d < 16; d <<= 1) {
            const uint64_t o = __shfl_up(inc, d, 64);
            if (c >= d) inc += o;
        }
        for (unsigned d = 8; d >= 1; d >>= 1) kept += __shfl_xor(kept, d, 64);
        const uint64_t at = inc - n, total = __shfl(inc, 15, 64), end = base + acc;
        // A map that outgrew its slot is emptied: later passes never read past the slot; its
        // record carries the overflow and the host rolls the map back and grows the slot.
        const bool over = b.overflow[m] || end > b.slot[m].cap;
        if (mine) {
            rout[c] = over ? 0 : at;
            rout[kRegPhys + c] = over ? 0 : in_place ? ph : base + wr;
        }
        if (c == 0) {
            rout[16] = over ? 0 : total;
            rout[kRegTop] = over ? 0 : end;
            b.len[m * 3 + P.out_buf] = over ? 0 : total;
            if (over) b.overflow[m] = 1;
            tpgp_hip_record r;
            r.len_in = len_in;
            r.len_out = total;
            r.dropped = T.dropped;
            r.l1_up = T.l1;
            r.sq_up = T.sq;
            r.rad_sq_up = T.rsq;
            r.radii = T.radii;
            r.executed = len_in > 0;
            r.nonfinite = T.nonfinite;
            r.overflow = over ? 1u : 0u;
            r.rotation = P.rot_seq;
            r.in_place = kept;
            b.rec[(uint64_t)m * b.max_rec + P.rec_slot] = r;
        }
    }
    __syncthreads();
}

// Scan level 2, one block per map: chunk offsets, output regions and length, the ledger fold
// in chunk order, and the record.
__global__ __launch_bounds__(kBlock) void k_rot_scan2(BatchDev b, const PassMap *pm) {
    const uint32_t m = blockIdx.x;
    const PassMap &P = pm[m];
    if (!P.active) return;
    __shared__ unsigned scratch[8];
    __shared__ TileLedger sh[kBlock / 64];
    const uint64_t c0 = b.chunk_start[m], c1 = b.chunk_start[m + 1];
    // the outputs start at the input's used extent when appending (read before
    // finish_map_scan rewrites the row, which is the input's own when appending)
    const uint64_t obase =
        (P.flags & kPassAppend) ? region_row(b.regions, m, P.in_buf)[kRegTop] : 0;
    uint64_t carry = 0;
    TileLedger L{};
    for (uint64_t base = c0; base < c1; base += kBlock) {
        const uint64_t ch = base + threadIdx.x;
        // a chunk keeps < 2^20 terms, so kBlock of them fit 32 bits
        const unsigned v = ch < c1 ? b.chunk_sum[ch] : 0u;
        if (ch < c1) ledger_fold(L, b.chunk_led[ch]);
        unsigned tot;
        const unsigned ex = block_exclusive_scan(v, scratch, &tot);
        if (ch < c1) b.chunk_base[ch] = obase + carry + ex;
        carry += tot;
    }
    const TileLedger T = block_ledger(L, sh);  // (its barriers publish chunk_base too)
    finish_map_scan(b, P, m, carry, obase, T);
}

// The map-level end of a direct rotation scan (k_rot_scan1 with direct has written every
// chunk's offset): the written total and the ledger folded from the chunk partials in
// k_rot_scan2's order (the same bits), then finish_map_scan.  Run by the write pass's first
// blocks (block m for map m) after their tiles: no block of that launch reads what it writes
// (the region rows, lengths and records), and a map found to overflow here has had its tiles
// written only where they fit the slot; it is emptied like any overflowed map.
__device__ void finish_direct(const BatchDev &b, const PassMap *pm, uint64_t *sh64,
                              TileLedger *sh) {
    for (uint32_t m = blockIdx.x; m < b.n_maps; m += gridDim.x) {
        const PassMap &P = pm[m];
        if (!P.active) continue;
        const uint64_t c0 = b.chunk_start[m], c1 = b.chunk_start[m + 1];
        // read before finish_map_scan rewrites the row (the input's own when appending)
        const uint64_t obase =
            (P.flags & kPassAppend) ? region_row(b.regions, m, P.in_buf)[kRegTop] : 0;
        uint64_t acc = 0;
        TileLedger L{};
        for (uint64_t ch = c0 + threadIdx.x; ch < c1; ch += kBlock) {
            acc += b.chunk_sum[ch];
            ledger_fold(L, b.chunk_led[ch]);
        }
        acc = block_sum_u64(acc, sh64);
        const TileLedger T = block_ledger(L, sh);
        finish_map_scan(b, P, m, acc, obase, T);
    }
}

// The whole rotation scan of small maps in one launch, one block per map: its chunks one after
// another (tile offsets within the chunk, chunk offsets), the ledger folded per thread across
// the chunks, then the map's regions and record.  Correct for any size; used when the launch's
// maps are small, where the two-level scan's second launch costs more than the serial chunks.
__global__ __launch_bounds__(kBlock) void k_rot_scan_maps(BatchDev b, const PassMap *pm) {
    __shared__ unsigned scratch[8];
    __shared__ TileLedger sh[kBlock / 64];
    for (uint32_t m = blockIdx.x; m < b.n_maps; m += gridDim.x) {
        const PassMap &P = pm[m];
        if (!P.active) continue;
        const uint64_t t0 = b.tstart[m * 16], t1 = b.tstart[m * 16 + 16];
        // outputs start at the input's used extent when appending (read before the row is
        // rewritten)
        const uint64_t obase =
            (P.flags & kPassAppend) ? region_row(b.regions, m, P.in_buf)[kRegTop] : 0;
        uint64_t carry = 0, ch = b.chunk_start[m];
        TileLedger L{};
        for (uint64_t g0 = t0; g0 < t1; g0 += kChunkTiles, ++ch) {  // one tile per thread
            const uint64_t g = g0 + threadIdx.x;
            const unsigned cnt = g < t1 ? b.tile_hist[g] : 0u;
            unsigned tot;
            const unsigned ex = block_exclusive_scan(cnt, scratch, &tot);
            if (g < t1) {
                b.tile_base[g] = ex;
                ledger_fold(L, b.tile_led[g]);
            }
            if (threadIdx.x == 0) b.chunk_base[ch] = obase + carry;
            carry += tot;
        }
        const TileLedger T = block_ledger(L, sh);  // (its barriers publish the offsets too)
        finish_map_scan(b, P, m, carry, obase, T);
    }
}

// Pass 2 of a rotation: per tile, merge again and write the kept outputs at the tile's offset.
// Paired classes: an anticommuting class c and its partner c ^ q merge the same two sorted runs
// (c's keys, and c ^ q's with the top digits xor-ed), so tile k of c and tile k of c ^ q cover
// the same merge-path range and emit the same lower digits thread by thread.  The lower class's
// workgroup stages the pair once and walks it once for both classes; the higher class's tiles
// are not visited.  Each thread's kept outputs (at most kMergeItems per class, consecutive in
// the output) are held as LDS indices and go straight to their places: no LDS output staging.
// One-word keys: at most 72 VGPRs, so seven waves per SIMD fit without spills (unconstrained
// it takes 74 and six; eight would spill); two-word keys: at most 96 and five waves (it takes
// 98 and four unconstrained).
template <int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W == 1 ? 7 : 5, 8)))
void k_rot_write(BatchDev b, const PassMap *__restrict__ pm, int finish) {
    __shared__ TileIn<W> in;
    __shared__ unsigned scratch[8];
    const uint64_t nwork = b.wstart[b.n_maps * 16 + 1];  // kind-1 and kind-2 tiles
    uint64_t gn = blockIdx.x < nwork ? b.work[blockIdx.x] : 0;
    for (uint64_t w = blockIdx.x; w < nwork; w += gridDim.x) {
        const uint64_t g = gn;
        gn = w + gridDim.x < nwork ? b.work[w + gridDim.x] : 0;  // the next tile, early
        MergeCtx x;
        TileGeom t;
        tile_from_desc(b, g, x, t.a0, t.a1);
        if (x.anti && x.cq < x.c) continue;  // written with its partner's tile
        const PassMap &P = pm[x.m];
        const uint32_t n = b.slot[x.m].n_spins;
        const Key<W> q = kxor<W>(kdigit_mask<W>(n - 1, P.qt), kdigit_mask<W>(n - 2, P.qs));
        const uint64_t total = x.nA + x.nD;
        t.d0 = x.k * kMergeTile;
        t.d1 = t.d0 + kMergeTile < total ? t.d0 + kMergeTile : total;
        const uint64_t tm0 = b.tstart[x.m * 16], c0 = b.chunk_start[x.m];
        // slot index of the tile's first output (chunk_base includes the rotation's output base)
        const uint64_t out_pos = b.chunk_base[c0 + (g - tm0) / kChunkTiles] + b.tile_base[g];
        const uint64_t count = b.tile_hist[g];
        const uint64_t cap = b.slot[x.m].cap;
        // overflowed maps and inconsistent tiles write nothing (both are reported at sync)
        if (!tile_ok(x, t) || b.overflow[x.m]) continue;
        uint64_t *gkeys = b.arena[P.out_buf].keys + b.slot[x.m].off * W;
        double *gvals = b.arena[P.out_buf].vals + b.slot[x.m].off;
        if (!x.anti && (P.flags & kPassCommutingKept)) {  // unchanged keys: a straight copy
            if (out_pos + count > cap) continue;
            const uint64_t *sk = b.arena[x.in_buf].keys + (x.off + x.A0 + t.d0) * W;
            const double *sv = b.arena[x.in_buf].vals + x.off + x.A0 + t.d0;
            uint64_t *gk = gkeys + out_pos * W;
            double *gv = gvals + out_pos;
            constexpr int R = kMergeTile / kBlock;  // count <= kMergeTile: loads first, then stores
            Key<W> kk[R];
            double vv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {  // branch-free (count >= 1)
                const uint64_t i0 = threadIdx.x + (uint64_t)r * kBlock;
                const uint64_t i = i0 < count ? i0 : count - 1;
                kk[r] = kload<W>(sk + i * W);
                vv[r] = sv[i];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint64_t i = threadIdx.x + (uint64_t)r * kBlock;
                if (i < count) {
                    kstore<W>(gk + i * W, kk[r]);
                    gv[i] = vv[r];
                }
            }
            continue;
        }
        uint64_t out_pos2 = 0, count2 = 0;
        if (x.anti) {
            const uint64_t g2 = b.tstart[x.m * 16 + x.cq] + x.k;
            out_pos2 = b.chunk_base[c0 + (g2 - tm0) / kChunkTiles] + b.tile_base[g2];
            count2 = b.tile_hist[g2];
        }
        const uint32_t lo = b.tsplit[g * kBlock + threadIdx.x];  // the count pass's split
        load_tile<W>(b, P, x, q, t, in);
        __syncthreads();
        // One walk serves both classes of a pair: each merge position gives class c's output
        // (key) and class c ^ q's (key ^ q, roles of the two slots swapped); the two kept
        // counts share one block scan (packed in 16-bit halves: a tile keeps <= kMergeTile).
        const bool anti = x.anti;
        const Coef k0 = coef_of(P, x);
        const Coef k1 = anti ? partner_coef(P, x) : k0;
        // A kept output is held as its LDS indices and flags (i | j << 11 | hasA << 22 |
        // hasD << 23: indices <= kMergeTile + 1) and rebuilt from the staged tile when stored:
        // four registers per thread instead of four keys and eight values.
        static_assert(kMergeTile + 1 < 2048, "11-bit LDS indices");
        uint32_t code[kMergeItems];
        unsigned keep0 = 0, keep1 = 0, ncall = 0;
        bool bad = false;
        walk_branchy<W>(in, t, lo, [&](const Key<W> &key, bool hasA, double va, bool hasD,
                                    double vd, uint32_t i, uint32_t j) {
            double v, w = 0.0, r;
            merged<false>(k0, hasA, va, hasD, vd, v, r, bad);
            const bool kept0 = v != 0.0 && (kzero<W>(key) || __builtin_fabs(v) > k0.thr);
            bool kept1 = false;
            if (anti) {
                merged<false>(k1, hasD, vd, hasA, va, w, r, bad);
                kept1 = w != 0.0 && (kzero<W>(kxor<W>(key, q)) || __builtin_fabs(w) > k1.thr);
            }
            const uint32_t c = i | (j << 11) | ((uint32_t)hasA << 22) | ((uint32_t)hasD << 23);
#pragma unroll
            for (int o = 0; o < kMergeItems; ++o) {  // static indices: stays in registers
                if (ncall == (unsigned)o) code[o] = c;
            }
            keep0 |= (unsigned)kept0 << ncall;
            keep1 |= (unsigned)kept1 << ncall;
            ncall++;
        });
        unsigned tot;
        const unsigned myoff = block_exclusive_scan(
            (unsigned)__builtin_popcount(keep0) | ((unsigned)__builtin_popcount(keep1) << 16),
            scratch, &tot);
        const bool fits0 = out_pos + count <= cap, fits1 = anti && out_pos2 + count2 <= cap;
        uint64_t *gk0 = gkeys + (out_pos + (myoff & 0xffffu)) * W;
        double *gv0 = gvals + out_pos + (myoff & 0xffffu);
        uint64_t *gk1 = gkeys + (out_pos2 + (myoff >> 16)) * W;
        double *gv1 = gvals + out_pos2 + (myoff >> 16);
        const uint32_t na = t.na;
#pragma unroll
        for (int o = 0; o < kMergeItems; ++o) {
            const bool s0 = fits0 && ((keep0 >> o) & 1u), s1 = fits1 && ((keep1 >> o) & 1u);
            if (!(s0 || s1)) continue;
            const uint32_t c = code[o], i = c & 0x7ffu, j = (c >> 11) & 0x7ffu;
            const bool hasA = (c >> 22) & 1u, hasD = (c >> 23) & 1u;
            const Key<W> key = kload<W>(in.keys + (hasA ? i : na + j) * W);
            const double va = hasA ? in.vals[i] : 0.0, vd = hasD ? in.vals[na + j] : 0.0;
            double v, r;
            if (s0) {  // the walk's arithmetic again: the same value, bit for bit
                merged<false>(k0, hasA, va, hasD, vd, v, r, bad);
                const unsigned at = (unsigned)__builtin_popcount(keep0 & ((1u << o) - 1u));
                kstore<W>(gk0 + (uint64_t)at * W, key);
                gv0[at] = v;
            }
            if (s1) {
                merged<false>(k1, hasD, vd, hasA, va, v, r, bad);
                const unsigned at = (unsigned)__builtin_popcount(keep1 & ((1u << o) - 1u));
                kstore<W>(gk1 + (uint64_t)at * W, kxor<W>(key, q));
                gv1[at] = v;
            }
        }
        __syncthreads();
    }
    // The map-level end of a direct k_rot_scan1, by the first blocks after their tiles (the
    // tiles do not wait for it); the staged tile's LDS is free again for its reductions.
    if (finish) {
        static_assert(sizeof(TileIn<W>) >= (kBlock / 64) * (sizeof(TileLedger) + 8), "room");
        TileLedger *sh = reinterpret_cast<TileLedger *>(&in);
        uint64_t *sh64 = reinterpret_cast<uint64_t *>(sh + kBlock / 64);
        finish_direct(b, pm, sh64, sh);
    }
}

// ------------------------------------------------------------------------------------------
// Small maps: one workgroup per map runs a whole sequence of passes (partitions and rotations)
// in one launch.  It walks the tiles of every class in order, so output offsets are a running
// count: no split, count or scan kernels, and no launch between passes.
// ------------------------------------------------------------------------------------------
constexpr uint32_t kPassSmall = 4u;      // PassMap.flags: the map runs on the small-map path
constexpr uint32_t kPassPartition = 8u;  // PassMap.flags: the descriptor is a partition

// Merge-path split at diagonal d, searched by one wave over [lo, hi] (64-ary rounds; P(hi)
// counts as true).  All lanes return the same value.
template <int W>
__device__ uint64_t split_wave(const uint64_t *keys, const MergeCtx &x, const Key<W> &q,
                               uint64_t d, uint64_t lo, uint64_t hi) {
    auto pred = [&](uint64_t xi) -> bool {
        Key<W> a = kload<W>(keys + (x.A0 + xi) * W);
        Key<W> dk = kxor<W>(kload<W>(keys + (x.D0 + d - xi - 1) * W), q);
        return klt<W>(dk, a);
    };
    while (hi > lo) {
        const uint64_t span = hi - lo;
        if (span < 64) {  // lane `span` maps to hi (true): the ballot is never 0
            const uint64_t xi = lo + lane_id();
            const bool p = xi < hi ? pred(xi) : true;
            lo = lo + (uint64_t)(__ffsll((unsigned long long)__ballot(p)) - 1);
            break;
        }
        const uint64_t xi = lo + ((span * lane_id()) >> 6);
        const uint64_t bal = __ballot(pred(xi));
        if (bal == 0) {
            lo = lo + ((span * 63) >> 6) + 1;
        } else {
            const int f = __ffsll((unsigned long long)bal) - 1;
            if (f == 0) break;
            const uint64_t xf = lo + ((span * (uint64_t)f) >> 6);
            const uint64_t xp = lo + ((span * (uint64_t)(f - 1)) >> 6);
            lo = xp + 1;
            hi = xf;
        }
    }
    return lo;
}

template <int W> struct SmallShared {
    TileIn<W> in;
    uint64_t okeys[kMergeTile * W];
    double ovals[kMergeTile];
    unsigned scratch[8];
    uint64_t split;
    uint64_t rout[17];
    unsigned hist[16];
    TileLedger red[kBlock / 64];
    unsigned nonfinite;
    ScatterShared sc;
};

template <int W>
__device__ __forceinline__ void small_rotation(const BatchDev &b, const PassMap &P, uint32_t m,
                               SmallShared<W> &sh) {
    const uint32_t n = b.slot[m].n_spins;
    const uint64_t off = b.slot[m].off, cap = b.slot[m].cap;
    const uint64_t *ikeys = b.arena[P.in_buf].keys + off * W;
    uint64_t *gk = b.arena[P.out_buf].keys + off * W;
    double *gv = b.arena[P.out_buf].vals + off;
    const uint64_t *rin = region_row(b.regions, m, P.in_buf);
    uint64_t *rout = region_row(b.regions, m, P.out_buf);
    const Key<W> q = kxor<W>(kdigit_mask<W>(n - 1, P.qt), kdigit_mask<W>(n - 2, P.qs));
    const bool want_log = (P.flags & 2u) != 0;
    bool over = b.overflow[m] != 0;
    uint64_t out_at = 0;
    double l1 = 0.0, sq = 0.0, rsq = 0.0;
    uint64_t n_drop = 0, n_rad = 0;
    bool bad = false;
    if (threadIdx.x == 0) sh.nonfinite = 0;
    for (uint32_t c = 0; c < 16; ++c) {
        if (threadIdx.x == 0) sh.rout[c] = out_at;
        MergeCtx x;
        x.m = m;
        x.c = c;
        x.A0 = rin[kRegPhys + c];
        x.nA = rin[c + 1] - rin[c];
        x.anti = class_anti(c, P.qt, P.qs);
        x.cq = c ^ ((P.qt << 2) | P.qs);
        x.D0 = rin[kRegPhys + x.cq];
        x.nD = x.anti ? rin[x.cq + 1] - rin[x.cq] : 0;
        x.k = 0;
        x.off = b.slot[m].off;
        x.in_buf = P.in_buf;
        const Coef k = coef_of(P, x);
        const uint64_t total = x.nA + x.nD;
        uint64_t a0 = 0;
        for (uint64_t d0 = 0; d0 < total; d0 += kMergeTile) {
            TileGeom t;
            t.d0 = d0;
            t.d1 = d0 + kMergeTile < total ? d0 + kMergeTile : total;
            t.a0 = a0;
            if (!x.anti) {
                t.a1 = t.d1;
            } else if (t.d1 == total) {
                t.a1 = x.nA;
            } else {
                if (wave_id() == 0) {
                    uint64_t lo = t.d1 > x.nD ? t.d1 - x.nD : 0;
                    if (lo < a0) lo = a0;
                    uint64_t hi = a0 + (t.d1 - t.d0);
                    if (hi > x.nA) hi = x.nA;
                    const uint64_t a1 = split_wave<W>(ikeys, x, q, t.d1, lo, hi);
                    if (lane_id() == 0) sh.split = a1;
                }
                __syncthreads();
                t.a1 = sh.split;
            }
            if (!tile_ok(x, t)) {  // internal inconsistency: reported at sync, nothing read
                if (threadIdx.x == 0) atomicOr(&b.counter[1], 1u);
                over = true;
                break;
            }
            load_tile<W>(b, P, x, q, t, sh.in);
            __syncthreads();
            const uint32_t obase = threadIdx.x * kMergeItems;
            unsigned nout = 0;
            walk_tile<W>(sh.in, t, [&](const Key<W> &key, bool hasA, double va, bool hasD,
                                       double vd) {
                double v, r;
                merged(k, hasA, va, hasD, vd, v, r, bad);
                const uint64_t k0 = key.w[0], k1 = W > 1 ? key.w[W - 1] : 0;
                if (r != 0.0) {
                    n_rad++;
                    rsq = upadd(rsq, upmul(r, r));
                    if (want_log) log_entry(b, k0, k1, r, 0.0, m, P.rot_seq, 1);
                }
                if (v == 0.0) return;  // exact cancellation: dropped silently
                if (!kzero<W>(key) && __builtin_fabs(v) <= k.thr) {
                    n_drop++;
                    l1 = upadd(l1, __builtin_fabs(v));
                    sq = upadd(sq, upmul(v, v));
                    if (want_log) log_entry(b, k0, k1, __builtin_fabs(v), v, m, P.rot_seq, 0);
                    return;
                }
                kstore<W>(sh.okeys + (obase + nout) * W, key);
                sh.ovals[obase + nout] = v;
                nout++;
            });
            unsigned tot;
            const unsigned myoff = block_exclusive_scan(nout, sh.scratch, &tot);
            for (unsigned i = 0; i < nout; ++i) {
                kstore<W>(sh.in.keys + (myoff + i) * W, kload<W>(sh.okeys + (obase + i) * W));
                sh.in.vals[myoff + i] = sh.ovals[obase + i];
            }
            __syncthreads();
            if (!over && out_at + tot > cap) over = true;
            if (!over) {
                for (unsigned i = threadIdx.x; i < tot; i += kBlock) {
                    kstore<W>(gk + (out_at + i) * W, kload<W>(sh.in.keys + i * W));
                    gv[out_at + i] = sh.in.vals[i];
                }
            }
            out_at += tot;
            a0 = t.a1;
            __syncthreads();
        }
    }
    if (bad) atomicOr(&sh.nonfinite, 1u);
    // ledger partials: per-thread chains in tile order, then a fixed tree over the block
    for (int d = 1; d < 64; d <<= 1) {
        const double o1 = __shfl_down(l1, d, 64), o2 = __shfl_down(sq, d, 64),
                     o3 = __shfl_down(rsq, d, 64);
        const uint64_t c1 = __shfl_down(n_drop, d, 64), c2 = __shfl_down(n_rad, d, 64);
        if (lane_id() + d < 64) {
            l1 = upadd(l1, o1);
            sq = upadd(sq, o2);
            rsq = upadd(rsq, o3);
            n_drop += c1;
            n_rad += c2;
        }
    }
    if (lane_id() == 0) {
        TileLedger L;
        L.l1 = l1;
        L.sq = sq;
        L.rsq = rsq;
        L.dropped = n_drop;
        L.radii = n_rad;
        L.nonfinite = 0;
        L.pad = 0;
        sh.red[wave_id()] = L;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        TileLedger T = sh.red[0];
        for (int w = 1; w < kBlock / 64; ++w) ledger_fold(T, sh.red[w]);
        for (int c = 0; c < 16; ++c) contiguous_row(rout, c, over ? 0 : sh.rout[c]);
        contiguous_row(rout, 16, over ? 0 : out_at);
        b.len[m * 3 + P.out_buf] = over ? 0 : out_at;
        if (over) b.overflow[m] = 1;
        tpgp_hip_record r;
        r.len_in = rin[16];
        r.len_out = out_at;
        r.dropped = T.dropped;
        r.l1_up = T.l1;
        r.sq_up = T.sq;
        r.rad_sq_up = T.rsq;
        r.radii = T.radii;
        r.executed = rin[16] > 0;
        r.nonfinite = sh.nonfinite;
        r.overflow = over ? 1u : 0u;
        r.rotation = P.rot_seq;
        r.in_place = 0;
        b.rec[(uint64_t)m * b.max_rec + P.rec_slot] = r;
    }
    __syncthreads();
}

template <int W>
__device__ __forceinline__ void small_partition(const BatchDev &b, const PassMap &P, uint32_t m,
                                                SmallShared<W> &sh) {
    const uint32_t n = b.slot[m].n_spins;
    const uint64_t off = b.slot[m].off;
    const uint64_t len = b.len[m * 3 + P.in_buf];
    const uint64_t *keys = b.arena[P.in_buf].keys + off * W;
    const double *vals = b.arena[P.in_buf].vals + off;
    uint64_t *okeys = b.arena[P.out_buf].keys + off * W;
    double *ovals = b.arena[P.out_buf].vals + off;
    const uint64_t *rin = region_row(b.regions, m, P.in_buf);
    if (threadIdx.x < 16) sh.hist[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t e0 = 0; e0 < len; e0 += kPartTile) {  // histogram of the new classes
        Key<W> k[kPartRounds];
        const Phys ph = phys_from(rin, e0);
#pragma unroll
        for (int r = 0; r < kPartRounds; ++r) {
            const uint64_t e = e0 + threadIdx.x + (uint64_t)r * kBlock;
            if (e < len) k[r] = kload<W>(keys + phys_at(ph, e) * W);
        }
#pragma unroll
        for (int r = 0; r < kPartRounds; ++r) {
            const uint64_t e = e0 + threadIdx.x + (uint64_t)r * kBlock;
            if (e < len) atomicAdd(&sh.hist[(kdigit<W>(k[r], P.pa) << 2) | kdigit<W>(k[r], P.pb)], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t *reg = region_row(b.regions, m, P.out_buf);
        uint64_t at = 0;
        for (int c = 0; c < 16; ++c) {
            contiguous_row(reg, c, at);
            sh.sc.run[c] = at;
            at += sh.hist[c];
        }
        contiguous_row(reg, 16, at);
        b.len[m * 3 + P.out_buf] = at;
    }
    __syncthreads();
    for (uint64_t e0 = 0; e0 < len; e0 += kPartTile) {  // stable scatter, tile by tile
        const uint64_t e1 = e0 + kPartTile < len ? e0 + kPartTile : len;
        scatter_tile<W>(keys, vals, rin, e0, e1, P.pa, P.pb, n, okeys, ovals, sh.sc);
    }
}

// ------------------------------------------------------------------------------------------
// Small maps resident in LDS.  A map of at most kLdsCap<W> terms is loaded into LDS once per
// launch and every pass of the chunk runs there, ping-ponging between two LDS buffers; only the
// per-rotation records, the ledger log and the final state go to global memory.  Both passes
// are two walks over the thread's share of the work with a scan between them: the first walk
// counts each thread's kept outputs per output class (class-major table cnt[class][thread]),
// the scan turns the table into output offsets, the second walk writes.
//   Rotation: the work of a commuting class is its terms; the work of an anticommuting pair
//   (c, c ^ q) is the merge path of the two classes (keys of c ^ q xor-ed into c's top digits).
//   Every merge-path position gives one output key in each class of the pair -- the key with
//   its own a-branch and, when the partner class holds the same lower digits, the partner's
//   d-branch -- so both classes' outputs are written in key order by the same walk.
//   Partition: a stable 16-way split by the digits at positions pa, pb, moved to the top.
// A map that starts too large, or whose next output would outgrow LDS or its slot, leaves the
// rest of the chunk to k_small (global memory), which the host launches right after: the LDS
// kernel writes the input of the pass it stops at to that pass's global buffer and records the
// pass in small_resume.
// ------------------------------------------------------------------------------------------
constexpr int kLdsThreads = 512;  // 8 waves: one workgroup per CU (its LDS holds the map)
// A rotation's work is cut into equal shares of weighted merge positions: a position of an
// anticommuting pair (a merge step, two outputs) weighs four of a commuting class (one
// output).  Equal shares of plain positions left the pairs' threads slower than the rest at
// every barrier.
constexpr uint32_t kAntiWeight = 4;
template <int W> struct LdsCap;
template <> struct LdsCap<1> { static constexpr uint32_t value = 4096; };
template <> struct LdsCap<2> { static constexpr uint32_t value = 2816; };

template <int W> struct LdsMap {
    static constexpr uint32_t kCap = LdsCap<W>::value;
    uint64_t keys[2][kCap * W];
    double vals[2][kCap];
    uint16_t cnt[16][kLdsThreads];  // per (output class, thread): kept count, then offset
    uint32_t reg[2][17];            // class regions of each buffer
    uint32_t len[2];
    uint32_t seg[17];               // rotation work: start of the work of the k-th segment in
                                    // weighted positions
    uint32_t sord[16];              // class of the k-th segment: the commuting classes, then the
                                    // lower classes of the anticommuting pairs, then the upper
                                    // ones (no work of their own), each in class order
    uint32_t total;
    unsigned scratch[kLdsThreads / 64];
    TileLedger red[kLdsThreads / 64];
};

__device__ unsigned lds_exclusive_scan(unsigned v, unsigned *scratch, unsigned *tot) {
    const unsigned x = wave_scan_u32(v);
    if (lane_id() == 63) scratch[wave_id()] = x;
    __syncthreads();
    unsigned base = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kLdsThreads / 64; ++w) {
        const unsigned s = scratch[w];
        if ((unsigned)w < wave_id()) base += s;
        all += s;
    }
    __syncthreads();
    *tot = all;
    return base + x - v;
}

// Per-thread ledger partials (upward-rounded chains, so upper bounds) summed over the block for
// thread 0: a DPP tree per wave (depth 6) and a fold across the waves in wave order (depth
// kLdsThreads / 64 - 1 more), in round-to-nearest, then made an upper bound once.  Every
// summand is non-negative, so a sum S' of depth d satisfies S' >= (1 - u)^d S, and
// next_up(S' (1 + 2 d u)) >= S (u = 2^-53).  Counts are exact.  Fixed order: deterministic.
__device__ TileLedger lds_ledger(TileLedger a, TileLedger *sh) {
    const double l1 = wave_scan_f64(a.l1), sq = wave_scan_f64(a.sq), rsq = wave_scan_f64(a.rsq);
    const unsigned nd = wave_scan_u32((unsigned)a.dropped), nr = wave_scan_u32((unsigned)a.radii);
    const bool bad = __ballot(a.nonfinite != 0) != 0;
    if (lane_id() == 63) {
        TileLedger w;
        w.l1 = l1;
        w.sq = sq;
        w.rsq = rsq;
        w.dropped = nd;
        w.radii = nr;
        w.nonfinite = bad ? 1u : 0u;
        w.pad = 0;
        sh[wave_id()] = w;
    }
    __syncthreads();
    TileLedger r = sh[0];
    if (threadIdx.x == 0) {
        for (int w = 1; w < kLdsThreads / 64; ++w) {
            r.l1 += sh[w].l1;
            r.sq += sh[w].sq;
            r.rsq += sh[w].rsq;
            r.dropped += sh[w].dropped;
            r.radii += sh[w].radii;
            r.nonfinite |= sh[w].nonfinite;
        }
        constexpr int kDepth = 6 + (kLdsThreads / 64 - 1);
        constexpr double kInflate = 1.0 + 2.0 * kDepth / 9007199254740992.0;  // exact
        r.l1 = upmul(r.l1, kInflate);
        r.sq = upmul(r.sq, kInflate);
        r.rsq = upmul(r.rsq, kInflate);
    }
    return r;
}

// cnt[c][t] (this thread's kept count per class) -> exclusive output offsets in class-major,
// thread-minor order; reg[c] = the offset of class c, reg[16] = the total (also sh.total).
// Each thread scans 16 consecutive entries of one row.
template <int W>
__device__ void lds_offsets(LdsMap<W> &sh, uint32_t *reg) {
    constexpr unsigned kPerRow = kLdsThreads / 16;  // threads per row of the table
    const unsigned row = threadIdx.x / kPerRow, col0 = (threadIdx.x % kPerRow) * 16;
    unsigned sum = 0;
    for (int k = 0; k < 16; ++k) sum += sh.cnt[row][col0 + k];
    unsigned tot;
    unsigned at = lds_exclusive_scan(sum, sh.scratch, &tot);
    if (col0 == 0) reg[row] = at;
    for (int k = 0; k < 16; ++k) {
        const unsigned here = sh.cnt[row][col0 + k];
        sh.cnt[row][col0 + k] = (uint16_t)at;
        at += here;
    }
    if (threadIdx.x == 0) {
        reg[16] = tot;
        sh.total = tot;
    }
    __syncthreads();
}

// Arena buffer `buf` selected by comparisons: indexing the kernel-argument array with a runtime
// value would make the compiler copy the whole argument block to scratch memory.
__device__ __forceinline__ Arena arena_at(const BatchDev &b, uint32_t buf) {
    return buf == 0 ? b.arena[0] : buf == 1 ? b.arena[1] : b.arena[2];
}

template <int W>
__device__ bool lds_load(const BatchDev &b, uint32_t buf, uint32_t m, LdsMap<W> &sh) {
    const uint64_t len = b.len[m * 3 + buf];
    if (len > LdsMap<W>::kCap) return false;  // uniform: every thread reads the same word
    const uint64_t off = b.slot[m].off;
    const Arena a = arena_at(b, buf);
    const uint64_t *gk = a.keys + off * W;
    const double *gv = a.vals + off;
    const uint64_t *reg = region_row(b.regions, m, buf);
    unsigned c = 0;  // class of this thread's element e (e < len = reg[16], so c <= 15)
    for (uint32_t e = threadIdx.x; e < len; e += kLdsThreads) {
        while (e >= reg[c + 1]) ++c;
        const uint64_t at = e - reg[c] + reg[kRegPhys + c];
        kstore<W>(sh.keys[0] + e * W, kload<W>(gk + at * W));
        sh.vals[0][e] = gv[at];
    }
    if (threadIdx.x < 17) sh.reg[0][threadIdx.x] = (uint32_t)reg[threadIdx.x];
    if (threadIdx.x == 0) sh.len[0] = (uint32_t)len;
    __syncthreads();
    return true;
}

template <int W>
__device__ void lds_store(const BatchDev &b, uint32_t buf, uint32_t m, const LdsMap<W> &sh,
                          int cur) {
    const uint32_t len = sh.len[cur];
    const uint64_t off = b.slot[m].off;
    const Arena a = arena_at(b, buf);
    uint64_t *gk = a.keys + off * W;
    double *gv = a.vals + off;
    for (uint32_t e = threadIdx.x; e < len; e += kLdsThreads) {
        kstore<W>(gk + e * W, kload<W>(sh.keys[cur] + e * W));
        gv[e] = sh.vals[cur][e];
    }
    if (threadIdx.x < 17)
        contiguous_row(region_row(b.regions, m, buf), threadIdx.x, sh.reg[cur][threadIdx.x]);
    if (threadIdx.x == 0) b.len[m * 3 + buf] = len;
}

template <int W>
__device__ void lds_partition(const PassMap &P, uint32_t n, LdsMap<W> &sh, int cur) {
    const uint32_t len = sh.len[cur], per = (len + kLdsThreads - 1) / kLdsThreads;
    const uint32_t e0 = min(len, threadIdx.x * per), e1 = min(len, e0 + per);
    const uint64_t *keys = sh.keys[cur];
#pragma unroll
    for (int c = 0; c < 16; ++c) sh.cnt[c][threadIdx.x] = 0;
    for (uint32_t e = e0; e < e1; ++e) {
        const Key<W> k = kload<W>(keys + e * W);
        sh.cnt[(kdigit<W>(k, P.pa) << 2) | kdigit<W>(k, P.pb)][threadIdx.x]++;
    }
    __syncthreads();
    lds_offsets<W>(sh, sh.reg[cur ^ 1]);
    uint64_t *okeys = sh.keys[cur ^ 1];
    double *ovals = sh.vals[cur ^ 1];
    for (uint32_t e = e0; e < e1; ++e) {
        const Key<W> k = kload<W>(keys + e * W);
        const unsigned cls = (kdigit<W>(k, P.pa) << 2) | kdigit<W>(k, P.pb);
        const uint32_t at = sh.cnt[cls][threadIdx.x]++;
        kstore<W>(okeys + at * W, move_to_top<W>(k, P.pa, P.pb, n));
        ovals[at] = sh.vals[cur][e];
    }
    if (threadIdx.x == 0) sh.len[cur ^ 1] = len;
    __syncthreads();
}

// The coefficient rules of the 16 output classes: shared factors, and per class (bit c) whether
// it anticommutes and whether its partner's d-branch is negated.
struct ClassCoefs {
    Coef base;
    unsigned anti, dneg;
    __device__ __forceinline__ Coef of(unsigned c) const {
        Coef k = base;
        k.anti = (anti >> c) & 1u;
        k.dneg = (dneg >> c) & 1u;
        return k;
    }
};
// The merged coefficient alone (merged() without radii or flags), branch-free: one product per
// present branch and one addition.  An absent branch contributes +0, which changes nothing but
// the sign of a zero result, and zero results are dropped.
__device__ __forceinline__ double merged_value(const Coef &k, bool hasA, double va, bool hasD,
                                               double vd) {
    if (!k.anti) return va;
    const double y = vd * k.s;
    return (hasA ? va * k.c : 0.0) + (hasD ? (k.dneg ? -y : y) : 0.0);
}

// Lane l of every wave evaluates class l % 16; two ballots gather the masks.
__device__ __forceinline__ ClassCoefs class_coefs(const PassMap &P) {
    ClassCoefs r;
    r.base.c = P.c;
    r.base.s = P.s;
    r.base.rc = P.rc;
    r.base.rs = P.rs;
    r.base.thr = P.threshold;
    r.base.validated = P.flags & 1u;
    r.base.anti = false;
    r.base.dneg = false;
    const unsigned c = lane_id() & 15u, cq = c ^ ((P.qt << 2) | P.qs);
    const bool a = class_anti(c, P.qt, P.qs);
    const bool neg = ((pauli_phase(P.qt, cq >> 2) + pauli_phase(P.qs, cq & 3u)) & 3u) != 1;
    r.anti = (unsigned)__ballot(a) & 0xffffu;
    r.dneg = (unsigned)__ballot(neg) & 0xffffu;
    return r;
}

// Walk this thread's work [w0, w1) of a rotation: emit(c, key, hasA, va, hasD, vd) once per
// output key, in ascending key order within every output class c.
// Each segment the thread works on is announced first, begin(c, partner) (partner == c for a
// commuting class), and its outputs then come as emit(side, c, key, hasA, va, hasD, vd), side 0
// for class c and 1 for the partner.  `split` is the merge-path split of the first (partial)
// segment, ~0u if not searched yet; the walk returns it, so the second walk need not search.
template <int W, typename B, typename F>
__device__ __forceinline__ uint32_t lds_rotation_walk(const LdsMap<W> &sh, int cur,
                                                      unsigned anti, unsigned qq,
                                                      const Key<W> &q, uint32_t w0, uint32_t w1,
                                                      uint32_t split, B &&begin, F &&emit) {
    const uint64_t *keys = sh.keys[cur];
    const double *vals = sh.vals[cur];
    const uint32_t *reg = sh.reg[cur];
    uint32_t seg[17];
    unsigned sord[16];
#pragma unroll
    for (int k = 0; k < 17; ++k) seg[k] = sh.seg[k];
#pragma unroll
    for (int k = 0; k < 16; ++k) sord[k] = sh.sord[k];
    uint32_t g = w0;
    while (g < w1) {
        unsigned s = 0;  // the segment holding g: the number of segment ends at or below g
#pragma unroll
        for (int k = 1; k < 16; ++k) s += seg[k] <= g ? 1u : 0u;
        const unsigned c = sord[s];
        // this thread's share of the segment, in weighted units, then in merge positions (an
        // anticommuting pair's positions weigh kAntiWeight each)
        const uint32_t wt = ((anti >> c) & 1u) ? kAntiWeight : 1u;
        const uint32_t o1 = min(w1, seg[s + 1]) - seg[s];
        const uint32_t d0 = (g - seg[s] + wt - 1) / wt, d1 = (o1 + wt - 1) / wt;
        const uint32_t A0 = reg[c], nA = reg[c + 1] - reg[c];
        if (!((anti >> c) & 1u)) {
            begin(c, c);
            for (uint32_t i = d0; i < d1; ++i)
                emit(0, c, kload<W>(keys + (A0 + i) * W), true, vals[A0 + i], false, 0.0);
        } else {
            const unsigned cq = c ^ qq;
            begin(c, cq);
            const uint32_t B0 = reg[cq], nB = reg[cq + 1] - reg[cq];
            auto Akey = [&](uint32_t i) { return kload<W>(keys + (A0 + i) * W); };
            auto Bkey = [&](uint32_t j) { return kxor<W>(kload<W>(keys + (B0 + j) * W), q); };
            // split at diagonal d0: the number of A terms among the first d0 positions of the
            // merge (ties: A first), i.e. the first i with A[i] > T(B[d0 - i - 1])
            uint32_t lo = d0 > nB ? d0 - nB : 0, hi = d0 < nA ? d0 : nA;
            if (g == w0 && split != ~0u) {
                lo = split;
            } else {
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (klt<W>(Bkey(d0 - mid - 1), Akey(mid)))
                        hi = mid;
                    else
                        lo = mid + 1;
                }
                if (g == w0) split = lo;
            }
            uint32_t i = lo, j = d0 - lo, used = 0;
            const uint32_t limit = d1 - d0;
            // a B term equal to the A term just before the split went out with it
            if (i > 0 && j < nB && keq<W>(Akey(i - 1), Bkey(j))) {
                ++j;
                ++used;
            }
            // one step per output key, without branches on the step's kind: the A term (key x,
            // its partner's d-branch if B holds the same lower digits) or the B term alone
            // The current A and B terms stay in registers; each step loads the next of both
            // (an exhausted run is read at a clamped index inside the LDS map and ignored)
            // while it emits, and keeps whichever its run advanced to.
            const uint32_t alast = nA ? nA - 1 : 0, blast = nB ? nB - 1 : 0;
            Key<W> x = Akey(min(i, alast)), y = Bkey(min(j, blast));
            double vx = vals[A0 + min(i, alast)], vy = vals[B0 + min(j, blast)];
            while (used < limit) {
                const bool hasA = i < nA, hasB = j < nB;
                const bool takeA = hasA && (!hasB || !klt<W>(y, x));
                const bool takeB = !takeA || (hasB && keq<W>(y, x));
                const uint32_t in = min(i + 1, alast), jn = min(j + 1, blast);
                const Key<W> xn = Akey(in), yn = Bkey(jn);
                const double vxn = vals[A0 + in], vyn = vals[B0 + jn];
                const Key<W> key = takeA ? x : y;
                emit(0, c, key, takeA, vx, takeB, vy);
                emit(1, cq, kxor<W>(key, q), takeB, vy, takeA, vx);
                if (takeA) {
                    x = xn;
                    vx = vxn;
                    ++i;
                }
                if (takeB) {
                    y = yn;
                    vy = vyn;
                    ++j;
                }
                used += (takeA ? 1u : 0u) + (takeB ? 1u : 0u);
            }
        }
        g = seg[s] + o1;
    }
    return split;
}

// Where a rotation in LDS reports: the ledger log and this map's records and slot capacity
// (copied out of the kernel arguments, so the walks' closures hold a few registers).
struct LdsSink {
    LogDev *log;
    uint64_t *log_count;
    uint64_t log_cap;
    tpgp_hip_record *rec;  // this map's records
    uint64_t cap;          // this map's slot capacity
    __device__ __forceinline__ void entry(uint64_t k0, uint64_t k1, double v, double aux,
                                          uint32_t m, uint32_t rot, uint32_t kind) const {
        unsigned long long at = atomicAdd((unsigned long long *)log_count, 1ull);
        if (at < log_cap) {
            LogDev &L = log[at];
            L.key[0] = k0;
            L.key[1] = k1;
            L.value = v;
            L.aux = aux;
            L.map = m;
            L.rotation = rot;
            L.kind = kind;
        }
    }
};

// One rotation in LDS; false (nothing written anywhere, no record) when its output does not fit.
// V: the rotation computes validated radii.
template <int W, bool V>
__device__ bool lds_rotation(const LdsSink &out_to, const PassMap &P, uint32_t m, uint32_t n,
                             LdsMap<W> &sh, int cur) {
    const Key<W> q = kxor<W>(kdigit_mask<W>(n - 1, P.qt), kdigit_mask<W>(n - 2, P.qs));
    const unsigned qq = (P.qt << 2) | P.qs;
    const ClassCoefs kc = class_coefs(P);
    const uint32_t *rin = sh.reg[cur];
    if (threadIdx.x < 64) {  // work segments: sizes on lanes 0-15, ordered, one wave-wide scan
        const unsigned c = threadIdx.x;
        uint32_t size = 0;
        unsigned kind = 3;  // 0 commuting, 1 lower class of a pair, 2 upper class, 3 no class
        if (c < 16) {
            const uint32_t nc = rin[c + 1] - rin[c];
            if (!((kc.anti >> c) & 1u)) {
                size = nc;
                kind = 0;
            } else if ((c ^ qq) > c) {
                size = (nc + (rin[(c ^ qq) + 1] - rin[c ^ qq])) * kAntiWeight;
                kind = 1;
            } else {
                kind = 2;
            }
        }
        // a wave's lanes walking both kinds of segment run both loops one after the other, so
        // the kinds are not interleaved: the commuting classes first, then the pairs
        const uint64_t m0 = __ballot(kind == 0), m1 = __ballot(kind == 1), m2 = __ballot(kind == 2);
        const uint64_t below = (1ull << c) - 1ull;
        const unsigned n0 = (unsigned)__popcll(m0), n1 = (unsigned)__popcll(m1);
        const unsigned rank = kind == 0 ? (unsigned)__popcll(m0 & below)
                              : kind == 1 ? n0 + (unsigned)__popcll(m1 & below)
                                          : n0 + n1 + (unsigned)__popcll(m2 & below);
        // lane r receives the size and class of the segment of rank r
        const int to = (int)(c < 16 ? rank : c) * 4;
        const uint32_t rsize = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)size);
        const unsigned rclass = (unsigned)__builtin_amdgcn_ds_permute(to, (int)c);
        const uint32_t incl = wave_scan_u32(c < 16 ? rsize : 0u);
        if (c < 16) {
            sh.seg[c + 1] = incl;
            sh.sord[c] = rclass;
        }
        if (c == 0) sh.seg[0] = 0;
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) sh.cnt[c][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t work = sh.seg[16], per = (work + kLdsThreads - 1) / kLdsThreads;
    const uint32_t w0 = min(work, threadIdx.x * per), w1 = min(work, w0 + per);
    const bool want_log = (P.flags & 2u) != 0;
    // The walks keep the segment's two classes' counts (then offsets) in registers: each class
    // is in one segment, so a thread visits it in one stretch.
    // walk 1: kept outputs per class (values only; nothing leaves the workgroup, so a rotation
    // that turns out not to fit has no side effect)
    unsigned ca = 16, cb = 16;  // the segment's classes (16: none yet)
    uint32_t na = 0, nb = 0;
    auto flush = [&]() {
        if (ca < 16) sh.cnt[ca][threadIdx.x] = (uint16_t)na;
        if (cb < 16 && cb != ca) sh.cnt[cb][threadIdx.x] = (uint16_t)nb;
    };
    const uint32_t split = lds_rotation_walk<W>(
        sh, cur, kc.anti, qq, q, w0, w1, ~0u,
        [&](unsigned c, unsigned partner) {
            flush();
            ca = c;
            cb = partner;
            na = nb = 0;
        },
        [&](unsigned side, unsigned c, const Key<W> &key, bool hasA, double va, bool hasD,
            double vd) {
            const double v = merged_value(kc.of(c), hasA, va, hasD, vd);
            if (v == 0.0) return;  // exact cancellation: dropped silently
            if (!kzero<W>(key) && __builtin_fabs(v) <= kc.base.thr) return;
            if (side) ++nb;
            else ++na;
        });
    flush();
    __syncthreads();
    uint32_t *rout = sh.reg[cur ^ 1];
    lds_offsets<W>(sh, rout);
    const uint32_t out = sh.total;
    if (out > LdsMap<W>::kCap || out > out_to.cap) return false;  // uniform
    // walk 2: values with radii, ledger partials and log; the kept outputs written in order
    uint64_t *okeys = sh.keys[cur ^ 1];
    double *ovals = sh.vals[cur ^ 1];
    TileLedger L{};
    bool bad = false;
    (void)lds_rotation_walk<W>(
        sh, cur, kc.anti, qq, q, w0, w1, split,
        [&](unsigned c, unsigned partner) {
            na = sh.cnt[c][threadIdx.x];
            nb = sh.cnt[partner][threadIdx.x];
        },
        [&](unsigned side, unsigned c, const Key<W> &key, bool hasA, double va, bool hasD,
            double vd) {
            const Coef k = kc.of(c);
            double v;
            const uint64_t k0 = key.w[0], k1 = W > 1 ? key.w[W - 1] : 0;
            if constexpr (V) {
                double r;
                merged(k, hasA, va, hasD, vd, v, r, bad);
                if (r != 0.0) {
                    L.radii++;
                    L.rsq = upadd(L.rsq, upmul(r, r));
                    if (want_log) out_to.entry(k0, k1, r, 0.0, m, P.rot_seq, 1);
                }
            } else {
                // without radii merged() flags exactly a non-finite result of an
                // anticommuting class: a non-finite product makes the sum non-finite too
                v = merged_value(k, hasA, va, hasD, vd);
                bad |= k.anti && !__builtin_isfinite(v);
            }
            if (v == 0.0) return;
            if (!kzero<W>(key) && __builtin_fabs(v) <= k.thr) {
                L.dropped++;
                L.l1 = upadd(L.l1, __builtin_fabs(v));
                L.sq = upadd(L.sq, upmul(v, v));
                if (want_log) out_to.entry(k0, k1, __builtin_fabs(v), v, m, P.rot_seq, 0);
                return;
            }
            const uint32_t at = side ? nb++ : na++;
            kstore<W>(okeys + at * W, key);
            ovals[at] = v;
        });
    L.nonfinite = bad ? 1u : 0u;
    const TileLedger T = lds_ledger(L, sh.red);
    if (threadIdx.x == 0) {
        sh.len[cur ^ 1] = out;
        tpgp_hip_record rec;
        rec.len_in = rin[16];
        rec.len_out = out;
        rec.dropped = T.dropped;
        rec.l1_up = T.l1;
        rec.sq_up = T.sq;
        rec.rad_sq_up = T.rsq;
        rec.radii = T.radii;
        rec.executed = rin[16] > 0;
        rec.nonfinite = T.nonfinite;
        rec.overflow = 0;
        rec.rotation = P.rot_seq;
        rec.in_place = 0;
        out_to.rec[P.rec_slot] = rec;
    }
    __syncthreads();
    return true;
}

// One workgroup per small map (maps[blockIdx.x]) runs the chunk's passes with the map in LDS:
// descriptor k of map m is ops[k * M + m]; descriptors without kPassSmall belong to other maps'
// paths.  small_resume[m] receives the first pass left to k_small (n_ops when none is).
template <int W>
__global__ __launch_bounds__(kLdsThreads) void k_small_lds(BatchDev b, const PassMap *ops,
                                                           uint32_t n_ops, const uint32_t *maps,
                                                           uint32_t n_small) {
    __shared__ LdsMap<W> sh;
    if (blockIdx.x >= n_small) return;
    const uint32_t m = maps[blockIdx.x], n = b.slot[m].n_spins;
    uint32_t k = 0;
    while (k < n_ops && !(ops[(uint64_t)k * b.n_maps + m].flags & kPassSmall)) ++k;
    if (k == n_ops || b.overflow[m] != 0 ||
        !lds_load<W>(b, ops[(uint64_t)k * b.n_maps + m].in_buf, m, sh)) {
        if (threadIdx.x == 0) b.small_resume[m] = k;
        return;
    }
    const LdsSink out_to{b.log, b.log_count, b.log_cap, b.rec + (uint64_t)m * b.max_rec,
                         b.slot[m].cap};
    int cur = 0;
    uint32_t last_out = 0;
    for (; k < n_ops; ++k) {
        // a private copy: the passes read its fields inside their loops, and a reference to
        // global memory would be reloaded there (LDS stores and barriers stop the hoisting)
        const PassMap P = ops[(uint64_t)k * b.n_maps + m];
        if (!(P.flags & kPassSmall)) continue;
        bool done = true;
        if (P.flags & kPassPartition)
            lds_partition<W>(P, n, sh, cur);
        else if (P.flags & 1u)
            done = lds_rotation<W, true>(out_to, P, m, n, sh, cur);
        else
            done = lds_rotation<W, false>(out_to, P, m, n, sh, cur);
        if (!done) {  // the pass's input, for the global kernel
            lds_store<W>(b, P.in_buf, m, sh, cur);
            if (threadIdx.x == 0) b.small_resume[m] = k;
            return;
        }
        cur ^= 1;
        last_out = P.out_buf;
    }
    lds_store<W>(b, last_out, m, sh, cur);
    if (threadIdx.x == 0) b.small_resume[m] = n_ops;
}

// The global-memory small-map kernel: one workgroup per small map runs the passes from
// small_resume[m] on (the LDS kernel, launched just before, ran the ones before it).
template <int W>
__global__ __launch_bounds__(kBlock) void k_small(BatchDev b, const PassMap *ops, uint32_t n_ops,
                                                  const uint32_t *maps, uint32_t n_small) {
    __shared__ SmallShared<W> sh;
    if (blockIdx.x >= n_small) return;
    const uint32_t m = maps[blockIdx.x];
    for (uint32_t k = b.small_resume[m]; k < n_ops; ++k) {
        const PassMap &P = ops[(uint64_t)k * b.n_maps + m];
        if (!(P.flags & kPassSmall)) continue;
        if (P.flags & kPassPartition)
            small_partition<W>(b, P, m, sh);
        else
            small_rotation<W>(b, P, m, sh);
    }
}

// One wave per map: binary search of the n single-site keys (symbol `pauli` at site s), each in
// the class region its top two digits name.
template <int W>
__global__ void k_readout(BatchDev b, const ReadoutMap *pm, unsigned pauli) {
    const uint32_t m = blockIdx.x;
    const ReadoutMap &P = pm[m];
    if (!P.active) return;
    const uint32_t n = b.slot[m].n_spins, s = threadIdx.x;
    if (s >= 64) return;
    double out = 0.0;
    if (s < n) {
        const uint64_t *reg = region_row(b.regions, m, P.in_buf);
        const Key<W> target = kdigit_mask<W>(P.pos[s], pauli);
        const unsigned cls = (kdigit<W>(target, n - 1) << 2) | kdigit<W>(target, n - 2);
        const uint64_t base = reg[kRegPhys + cls], len = reg[cls + 1] - reg[cls];
        const uint64_t *keys = b.arena[P.in_buf].keys + (b.slot[m].off + base) * W;
        const double *vals = b.arena[P.in_buf].vals + b.slot[m].off + base;
        uint64_t lo = 0, hi = len;
        while (lo < hi) {
            uint64_t mid = (lo + hi) >> 1;
            if (klt<W>(kload<W>(keys + mid * W), target))
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < len && keq<W>(kload<W>(keys + lo * W), target)) out = vals[lo];
    }
    b.rd[((uint64_t)m * b.max_rd + P.rec_slot) * 64 + s] = out;
}

// Sum of squares of one map's values as an unevaluated double-double (hi, lo) per block.  Every
// square splits exactly (x*x = p + e, e = fma(x, x, -p)) and partial sums combine with
// error-free two-sums, so the total is within a relative ~len * 2^-104 of the exact sum: the
// map's l2 norm comes out correctly rounded (but for ties) whatever the summation order.
__host__ __device__ __forceinline__ void dd_add(double &hi, double &lo, double b_hi,
                                                double b_lo) {
    const double s = hi + b_hi;
    const double bb = s - hi;
    double err = (hi - (s - bb)) + (b_hi - bb);
    err += lo + b_lo;
    hi = s + err;
    lo = err - (hi - s);
}
// Values in the map's key order (logical index k at its slot index through the region row).
__global__ void __launch_bounds__(256) k_sumsq(const double *vals, const uint64_t *reg,
                                               uint64_t len, double *out) {
    double hi = 0.0, lo = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned c = 0;  // class of this thread's k (k < len = reg[16], so c <= 15)
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < len; k += stride) {
        while (k >= reg[c + 1]) ++c;
        const double x = vals[k - reg[c] + reg[kRegPhys + c]];
        const double p = x * x;
        dd_add(hi, lo, p, fma(x, x, -p));
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double oh = __shfl_down(hi, off, 64), ol = __shfl_down(lo, off, 64);
        dd_add(hi, lo, oh, ol);
    }
    __shared__ double part[8];
    const unsigned wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[2 * wave] = hi;
        part[2 * wave + 1] = lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < (blockDim.x >> 6); ++w) dd_add(hi, lo, part[2 * w], part[2 * w + 1]);
        out[2 * blockIdx.x] = hi;
        out[2 * blockIdx.x + 1] = lo;
    }
}

// Validation (TPGP_HIP_VALIDATE=1): every active map's working buffer must be strictly sorted
// and its class regions must match the classes of its keys.  The first violation is recorded.
template <int W>
__global__ void k_validate(BatchDev b, const PassMap *pm, uint32_t tag) {
    const uint32_t M = b.n_maps;
    for (uint32_t m = 0; m < M; ++m) {
        const PassMap &P = pm[m];
        if (!P.active) continue;
        const uint32_t buf = P.out_buf, n = b.slot[m].n_spins;
        const uint64_t len = b.len[m * 3 + buf];
        const uint64_t *reg = region_row(b.regions, m, buf);
        const uint64_t *keys = b.arena[buf].keys + b.slot[m].off * W;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            // logical starts nondecreasing from 0 to len; every class inside the used extent,
            // which is inside the slot; no two nonempty classes overlap
            bool bad = reg[0] != 0 || reg[16] != len || reg[kRegTop] > b.slot[m].cap;
            for (int c = 0; c < 16; ++c) {
                bad = bad || reg[c] > reg[c + 1] ||
                      reg[kRegPhys + c] + (reg[c + 1] - reg[c]) > reg[kRegTop];
                for (int d = 0; d < c; ++d) {
                    const uint64_t nc = reg[c + 1] - reg[c], nd = reg[d + 1] - reg[d];
                    const uint64_t pc = reg[kRegPhys + c], pd = reg[kRegPhys + d];
                    bad = bad || (nc && nd && pc < pd + nd && pd < pc + nc);
                }
            }
            if (bad && atomicCAS((unsigned long long *)&b.debug[0], 0ull,
                                 (unsigned long long)tag) == 0ull) {
                b.debug[1] = m;
                b.debug[2] = ~0ull;
                b.debug[3] = len;
                b.debug[4] = reg[16];
            }
        }
        for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < len;
             e += (uint64_t)gridDim.x * blockDim.x) {
            const Phys ph = phys_from(reg, e);
            Key<W> k = kload<W>(keys + phys_at(ph, e) * W);
            unsigned cls = (kdigit<W>(k, n - 1) << 2) | kdigit<W>(k, n - 2);
            bool bad = e < reg[cls] || e >= reg[cls + 1];
            if (e > 0) {
                const Phys pp = phys_from(reg, e - 1);
                bad = bad || !klt<W>(kload<W>(keys + phys_at(pp, e - 1) * W), k);
            }
            if (bad && atomicCAS((unsigned long long *)&b.debug[0], 0ull,
                                 (unsigned long long)tag) == 0ull) {
                b.debug[1] = m;
                b.debug[2] = e;
                b.debug[3] = len;
                b.debug[4] = cls;
                b.debug[5] = reg[cls];
                b.debug[6] = reg[cls + 1];
                b.debug[7] = k.w[0];
                b.debug[8] = e > 0 ? keys[phys_at(phys_from(reg, e - 1), e - 1) * W] : 0;
            }
        }
    }
}

__global__ void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x)
        p[t] = v;
}
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x)
        p[t] = v;
}
__global__ void k_set_len(uint64_t *len, uint64_t idx, uint64_t v) { len[idx] = v; }

}  // namespace

// ==========================================================================================
// Host side
// ==========================================================================================
namespace {
bool mulz(size_t a, size_t b, size_t *o) {
    if (b && a > SIZE_MAX / b) return false;
    *o = a * b;
    return true;
}
bool addz(size_t a, size_t b, size_t *o) {
    if (a > SIZE_MAX - b) return false;
    *o = a + b;
    return true;
}
bool overlap(const void *a, size_t as, const void *b, size_t bs) {
    uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    if (as > UINTPTR_MAX - x || bs > UINTPTR_MAX - y) return true;
    return x < y + bs && y < x + as;
}
}  // namespace

extern "C" int tpgp_hip_compiled(void) { return 1; }

namespace {
bool tuning_value_ok(unsigned param, unsigned value) {
    if (value == 0) return true;  // the default
    if (param == TPGP_HIP_TUNE_PARTITION_ROUNDS) return value == 4 || value == 8 || value == 12;
    return value <= 24;  // grids, in sixths of the tile estimate
}
const unsigned kTuneDefault[TPGP_HIP_TUNE_PARAMS] = {kPartRounds, 3, 4};

// Lines "words param bucket value" ('#' starts a comment); false with a message on the first bad
// line.
bool load_tuning(tpgp_hip_context *ctx, const char *path, char *error, size_t cap) {
    FILE *f = fopen(path, "r");
    if (!f) {
        msg(error, cap, "TPGP_HIP_TUNING: cannot open %s", path);
        return false;
    }
    char line[256];
    unsigned lineno = 0;
    bool ok = true;
    while (ok && fgets(line, sizeof line, f)) {
        ++lineno;
        char *hash = strchr(line, '#');
        if (hash) *hash = 0;
        unsigned words, p, bkt, v;
        char extra;
        const int got = sscanf(line, "%u %u %u %u %c", &words, &p, &bkt, &v, &extra);
        if (got == EOF || got <= 0) {
            bool blank = true;
            for (const char *c = line; *c; ++c)
                if (!isspace((unsigned char)*c)) blank = false;
            if (blank) continue;
        }
        if (got != 4 || words < 1 || words > 2 || p >= TPGP_HIP_TUNE_PARAMS ||
            bkt >= TPGP_HIP_TUNE_BUCKETS || !tuning_value_ok(p, v)) {
            msg(error, cap,
                "TPGP_HIP_TUNING: %s line %u is not a valid \"words param bucket value\"", path,
                lineno);
            ok = false;
        } else {
            ctx->tune[words - 1][p][bkt] = (unsigned char)v;
        }
    }
    fclose(f);
    return ok;
}
}  // namespace

namespace {
hipError_t resolve_scratch(int device);  // the allocation-free entry point's fixed scratch
}  // namespace

extern "C" tpgp_hip_context *tpgp_hip_create(int index, char *error, size_t cap) {
    msg(error, cap, "");
    int count = 0, previous = 0;
    hipError_t code = hipGetDeviceCount(&count);
    if (code != hipSuccess) {
        hip_fail(code, "HIP device discovery", error, cap);
        return nullptr;
    }
    if (index < 0 || index >= count) {
        msg(error, cap, "Requested HIP device index is unavailable.");
        return nullptr;
    }
    if ((code = hipGetDevice(&previous)) != hipSuccess) {
        hip_fail(code, "HIP current-device query", error, cap);
        return nullptr;
    }
    if ((code = hipSetDevice(index)) != hipSuccess) {
        hip_fail(code, "HIP device selection", error, cap);
        return nullptr;
    }
    hipDeviceProp_t prop;
    code = hipGetDeviceProperties(&prop, index);
    if (code != hipSuccess) {
        hip_fail(code, "HIP device properties", error, cap);
        (void)hipSetDevice(previous);
        return nullptr;
    }
    tpgp_hip_context *ctx = new (std::nothrow) tpgp_hip_context();
    if (!ctx) {
        msg(error, cap, "Could not allocate the HIP host context.");
        (void)hipSetDevice(previous);
        return nullptr;
    }
    ctx->device = index;
    ctx->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 1;
    ctx->storage = nullptr;
    ctx->capacity = 0;
    snprintf(ctx->name, sizeof(ctx->name), "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
             ctx->cus);
    code = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (code != hipSuccess) {
        hip_fail(code, "HIP stream creation", error, cap);
        delete ctx;
        (void)hipSetDevice(previous);
        return nullptr;
    }
    // this device's copy of the allocation-free entry point's fixed scratch
    if ((code = resolve_scratch(index)) != hipSuccess) {
        hip_fail(code, "HIP fixed-scratch address", error, cap);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        (void)hipSetDevice(previous);
        return nullptr;
    }
    if ((code = hipSetDevice(previous)) != hipSuccess) {
        hip_fail(code, "HIP previous-device restoration", error, cap);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return nullptr;
    }
    // the table: TPGP_HIP_TUNING names it (empty: the built-in defaults); unset, the library
    // looks for tuning_gfx950.txt in the directory above its own (the package directory)
    const char *tuning = getenv("TPGP_HIP_TUNING");
    char beside[4096] = {0};
    if (!tuning) {
        Dl_info info;
        if (dladdr((void *)&tpgp_hip_create, &info) && info.dli_fname &&
            strlen(info.dli_fname) < sizeof(beside) - 32) {
            strcpy(beside, info.dli_fname);
            char *slash = strrchr(beside, '/');
            if (slash) *slash = 0;  // .../lib
            slash = strrchr(beside, '/');
            if (slash) {
                strcpy(slash, "/tuning_gfx950.txt");
                FILE *probe = fopen(beside, "r");
                if (probe) {
                    fclose(probe);
                    tuning = beside;
                }
            }
        }
    }
    if (tuning && *tuning && !load_tuning(ctx, tuning, error, cap)) {
        tpgp_hip_destroy(ctx);
        return nullptr;
    }
    return ctx;
}

extern "C" void tpgp_hip_destroy(tpgp_hip_context *ctx) {
    if (!ctx) return;
    {
        DeviceGuard guard(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->storage) (void)hipFree(ctx->storage);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

extern "C" const char *tpgp_hip_device_name(const tpgp_hip_context *ctx) {
    return ctx ? ctx->name : "unavailable";
}

extern "C" int tpgp_hip_set_tuning(tpgp_hip_context *ctx, unsigned words, unsigned param,
                                   unsigned bucket, unsigned value) {
    if (!ctx || words < 1 || words > 2 || param >= TPGP_HIP_TUNE_PARAMS ||
        bucket >= TPGP_HIP_TUNE_BUCKETS || !tuning_value_ok(param, value))
        return TPGP_HIP_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(ctx->mutex);
    ctx->tune[words - 1][param][bucket] = (unsigned char)value;
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_get_tuning(tpgp_hip_context *ctx, unsigned words, unsigned param,
                                   unsigned bucket, unsigned *value) {
    if (!ctx || !value || words < 1 || words > 2 || param >= TPGP_HIP_TUNE_PARAMS ||
        bucket >= TPGP_HIP_TUNE_BUCKETS)
        return TPGP_HIP_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(ctx->mutex);
    const unsigned v = ctx->tune[words - 1][param][bucket];
    *value = v ? v : kTuneDefault[param];
    return TPGP_HIP_OK;
}

namespace {
// Validated sizes of one branch call (tpgp_hip_branches' contract).
struct BranchSizes {
    size_t key_bytes, value_bytes, okey_bytes, oval_bytes, total;
};
tpgp_hip_status check_branches(tpgp_hip_context *ctx, const uint64_t *keys,
                               const double *values, size_t count, size_t words, size_t spins,
                               size_t i, size_t j, unsigned qi, unsigned qj, double c, double s,
                               double rc, double rs, int enclose, const uint64_t *out_keys,
                               const double *out_values, const double *out_radii,
                               BranchSizes *z, char *error, size_t cap) {
    if (!ctx || spins < 2 || words != spins / 32 + (spins % 32 != 0) || i >= spins ||
        j >= spins || i == j || qi > 3 || qj > 3 || !std::isfinite(c) || !std::isfinite(s) ||
        !std::isfinite(rc) || !std::isfinite(rs) || rc < 0 || rs < 0 ||
        (enclose != 0 && enclose != 1)) {
        msg(error, cap,
            "Invalid HIP branch dimensions, sites, generator, or trigonometric enclosure.");
        return TPGP_HIP_EINVAL;
    }
    if (!count) return TPGP_HIP_OK;
    size_t t;
    if (!keys || !values || !out_keys || !out_values || !out_radii ||
        !mulz(count, words, &t) || !mulz(t, sizeof(uint64_t), &z->key_bytes) ||
        !mulz(count, sizeof(double), &z->value_bytes) || !mulz(z->key_bytes, 2, &z->okey_bytes) ||
        !mulz(z->value_bytes, 2, &z->oval_bytes) || !mulz(z->key_bytes, 3, &z->total) ||
        !mulz(z->value_bytes, 5, &t) || !addz(z->total, t, &z->total) ||
        !addz(z->total, 256, &z->total)) {
        msg(error, cap, "Null HIP array or packed-branch allocation size overflow.");
        return TPGP_HIP_EINVAL;
    }
    const void *arrays[] = {keys, values, out_keys, out_values, out_radii};
    const size_t lengths[] = {z->key_bytes, z->value_bytes, z->okey_bytes, z->oval_bytes,
                              z->oval_bytes};
    for (int a = 0; a < 5; ++a)
        for (int b2 = a + 1; b2 < 5; ++b2)
            if (overlap(arrays[a], lengths[a], arrays[b2], lengths[b2])) {
                msg(error, cap, "HIP input/output arrays must be disjoint.");
                return TPGP_HIP_EINVAL;
            }
    return TPGP_HIP_OK;
}

// One branch call on a device buffer of at least z.total bytes (the caller holds the context's
// mutex and has made its device current): inputs in, k_branches, status word and outputs out.
// The status word is cleared by a fill kernel, not a memset, so the sequence holds no memset
// node if a caller captures the stream.
tpgp_hip_status run_branches(tpgp_hip_context *ctx, unsigned char *base, const BranchSizes &z,
                             const uint64_t *keys, const double *values, size_t count,
                             size_t words, size_t spins, size_t i, size_t j, unsigned qi,
                             unsigned qj, double c, double s, double rc, double rs, int enclose,
                             uint64_t *out_keys, double *out_values, double *out_radii,
                             char *error, size_t cap) {
    tpgp_hip_status result = TPGP_HIP_OK;
    hipError_t code;
    uint64_t *dk = (uint64_t *)base;
    double *dv = (double *)(base + z.key_bytes);
    uint64_t *dok = (uint64_t *)(base + z.key_bytes + z.value_bytes);
    double *dov = (double *)(base + z.key_bytes + z.value_bytes + z.okey_bytes);
    double *dor = dov + 2 * count;
    unsigned *dflag = (unsigned *)(dor + 2 * count);
    unsigned flag = 0;
#define HIPCALL(call, what)                                                   \
    do {                                                                      \
        if (result == TPGP_HIP_OK) {                                          \
            code = (call);                                                    \
            if (code != hipSuccess) result = hip_fail(code, what, error, cap); \
        }                                                                     \
    } while (0)
    HIPCALL(hipMemcpyAsync(dk, keys, z.key_bytes, hipMemcpyHostToDevice, ctx->stream),
            "HIP input-key transfer");
    HIPCALL(hipMemcpyAsync(dv, values, z.value_bytes, hipMemcpyHostToDevice, ctx->stream),
            "HIP input-coefficient transfer");
    if (result == TPGP_HIP_OK) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, dflag, 1ull, 0u);
        HIPCALL(hipGetLastError(), "HIP status reset");
    }
    if (result == TPGP_HIP_OK) {
        size_t want = (count + kBlock - 1) / kBlock;
        size_t maxb = (size_t)ctx->cus * 16;
        unsigned blocks = (unsigned)(want < maxb ? want : maxb);
        hipLaunchKernelGGL(k_branches, dim3(blocks), dim3(kBlock), 0, ctx->stream, dk, dv, count,
                           words, spins, i, j, qi, qj, c, s, rc, rs, enclose, dok, dov, dor,
                           dflag);
        HIPCALL(hipGetLastError(), "HIP branch-kernel launch");
    }
    HIPCALL(hipMemcpyAsync(&flag, dflag, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream),
            "HIP status transfer");
    HIPCALL(hipStreamSynchronize(ctx->stream), "HIP branch-kernel execution");
    if (result == TPGP_HIP_OK && flag) {
        result = (flag & 1u) ? TPGP_HIP_EINVAL : TPGP_HIP_ENUMERIC;
        msg(error, cap, (flag & 1u) ? "HIP input keys contain nonzero padding bits."
                                    : "HIP branch arithmetic or coefficients are nonfinite.");
    }
    HIPCALL(hipMemcpyAsync(out_keys, dok, z.okey_bytes, hipMemcpyDeviceToHost, ctx->stream),
            "HIP output-key transfer");
    HIPCALL(hipMemcpyAsync(out_values, dov, z.oval_bytes, hipMemcpyDeviceToHost, ctx->stream),
            "HIP output-coefficient transfer");
    HIPCALL(hipMemcpyAsync(out_radii, dor, z.oval_bytes, hipMemcpyDeviceToHost, ctx->stream),
            "HIP output-radius transfer");
    HIPCALL(hipStreamSynchronize(ctx->stream), "HIP output completion");
#undef HIPCALL
    // drain anything queued before an earlier failure; the caller may release its arrays
    if (result != TPGP_HIP_OK) (void)hipStreamSynchronize(ctx->stream);
    return result;
}
}  // namespace

extern "C" tpgp_hip_status tpgp_hip_branches(
    tpgp_hip_context *ctx, const uint64_t *keys, const double *values, size_t count,
    size_t words, size_t spins, size_t i, size_t j, unsigned qi, unsigned qj, double c, double s,
    double rc, double rs, int enclose, uint64_t *out_keys, double *out_values, double *out_radii,
    char *error, size_t cap) {
    msg(error, cap, "");
    BranchSizes z{};
    tpgp_hip_status st = check_branches(ctx, keys, values, count, words, spins, i, j, qi, qj, c,
                                        s, rc, rs, enclose, out_keys, out_values, out_radii, &z,
                                        error, cap);
    if (st != TPGP_HIP_OK || !count) return st;
    std::lock_guard<std::recursive_mutex> lock(ctx->mutex);
    DeviceGuard guard(ctx->device);
    if (!guard.ok) {
        msg(error, cap, "HIP device selection failed.");
        return TPGP_HIP_EDEVICE;
    }
    hipError_t code;
    if (z.total > ctx->capacity) {
        // release first: avoids a transient double allocation as a trajectory grows
        if (ctx->storage) {
            code = hipFree(ctx->storage);
            ctx->storage = nullptr;
            ctx->capacity = 0;
            if (code != hipSuccess) return hip_fail(code, "HIP scratch release", error, cap);
        }
        code = hipMalloc(&ctx->storage, z.total);
        if (code != hipSuccess) {
            ctx->storage = nullptr;
            return hip_fail(code, "HIP persistent scratch allocation", error, cap);
        }
        ctx->capacity = z.total;
    }
    return run_branches(ctx, (unsigned char *)ctx->storage, z, keys, values, count, words, spins,
                        i, j, qi, qj, c, s, rc, rs, enclose, out_keys, out_values, out_radii,
                        error, cap);
}

// ---- the allocation-free branch call ------------------------------------------------------
// A fixed device array per GPU replaces the context's growable scratch.  In host code the name
// `scratch` is only a stand-in for the symbol: its device address is taken once per device by
// hipGetSymbolAddress (in tpgp_hip_create, with that device current) and only that address is
// handed to copies and kernels.
namespace {
__device__ __align__(256) unsigned char scratch[TPGP_HIP_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch_ptr[kMaxDevices] = {};
// With `device` current: resolve this device's copy of `scratch` once.  0 on success.
hipError_t resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    hipError_t code = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (code == hipSuccess) g_scratch_ptr[device] = (unsigned char *)ptr;
    return code;
}
unsigned char *scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

extern "C" tpgp_hip_status tpgp_hip_branches_noalloc(
    tpgp_hip_context *ctx, const uint64_t *keys, const double *values, size_t count,
    size_t words, size_t spins, size_t i, size_t j, unsigned qi, unsigned qj, double c, double s,
    double rc, double rs, int enclose, uint64_t *out_keys, double *out_values, double *out_radii,
    char *error, size_t cap) {
    msg(error, cap, "");
    BranchSizes z{};
    tpgp_hip_status st = check_branches(ctx, keys, values, count, words, spins, i, j, qi, qj, c,
                                        s, rc, rs, enclose, out_keys, out_values, out_radii, &z,
                                        error, cap);
    if (st != TPGP_HIP_OK || !count) return st;
    if (z.total > (size_t)TPGP_HIP_NOALLOC_SCRATCH_BYTES) {
        msg(error, cap,
            "HIP allocation-free branch call needs %zu bytes of device scratch; its fixed "
            "scratch holds %zu (at most %zu terms of %zu key words).",
            z.total, (size_t)TPGP_HIP_NOALLOC_SCRATCH_BYTES,
            ((size_t)TPGP_HIP_NOALLOC_SCRATCH_BYTES - 256) / (24 * words + 40), words);
        return TPGP_HIP_EINVAL;
    }
    std::lock_guard<std::recursive_mutex> lock(ctx->mutex);
    DeviceGuard guard(ctx->device);
    if (!guard.ok) {
        msg(error, cap, "HIP device selection failed.");
        return TPGP_HIP_EDEVICE;
    }
    unsigned char *base = scratch_of(ctx->device);
    if (!base) {
        msg(error, cap, "HIP fixed scratch of device %d was not resolved.", ctx->device);
        return TPGP_HIP_EDEVICE;
    }
    return run_branches(ctx, base, z, keys, values, count, words, spins, i, j, qi, qj, c, s, rc,
                        rs, enclose, out_keys, out_values, out_radii, error, cap);
}

// ==========================================================================================
// Resident batch
// ==========================================================================================
namespace {
struct HostMap {
    size_t n_spins = 0;
    uint64_t off = 0, cap = 0;
    int kc = 0, kcur = 0;
    std::vector<uint8_t> sig_c, sig;  // sigma[p] = site at digit position p
    uint32_t nrec = 0, nrd = 0;
    uint64_t committed_len = 0;
    uint64_t est_len = 0;  // last known working length (sizes launch grids)
    // every non-identity coefficient of the working / committed state has |v| >= floor
    // (0: nothing known).  A rotation whose threshold is below it cannot truncate a key it
    // leaves unchanged.
    double floor = 0.0, floor_c = 0.0;
    // Appending rotations: the working / committed state may have classes out of key order in
    // its slot (gap), the working buffer's used extent as planned (est_top, in terms), and the
    // planning bound on what one rotation appends, as a fraction of the map (app_level indexes
    // kAppendLevels; learned from the appended counts at sync).
    bool gap = false, gap_c = false;
    uint64_t est_top = 0;
    uint32_t app_level = 4;
    std::vector<uint32_t> app_recs;  // record slots of the rotations planned to append
    // growth of the map over the last synced work (peak output / first input, 1 to 4): a
    // growing map is planned at the size it will reach
    double growth = 1.0;
};
// Fractions of the map one appending rotation is planned to write after the extent.
constexpr double kAppendLevels[] = {0.25, 0.375, 0.5, 0.625, 0.75, 1.0, 1.25, 1.5, 2.0, 2.5};
constexpr uint32_t kAppendTop = sizeof(kAppendLevels) / sizeof(kAppendLevels[0]) - 1;
struct LoggedLayout {
    uint32_t map, rotation;
    std::vector<uint8_t> sigma;
};
// A recorded step (tpgp_hip_batch_graph_*): the HIP graph of the readouts and rotations one
// step enqueues, with the descriptors they read in a device buffer of their own (the same on
// every replay), and the host state of every map before and after the step's planning.  It
// replays only from the exact state it was recorded in (same buffers, digit layouts, floors,
// record counters) and the same device allocations (gen).
struct StepGraph {
    uint64_t key = 0, gen = 0, last_use = 0;
    std::vector<uint8_t> before, after;
    std::vector<LoggedLayout> logged;  // exact-ledger layouts the step registers
    std::vector<uint64_t> est;         // working lengths its launch shapes were planned for
    hipGraphExec_t exec = nullptr;
    uint8_t *desc = nullptr;           // device descriptors
};
}  // namespace

struct tpgp_hip_batch {
    tpgp_hip_context *ctx = nullptr;
    int W = 1;
    uint32_t M = 0;
    std::vector<HostMap> maps;
    uint64_t total_cap = 0;
    uint64_t max_tiles = 0;
    BatchDev dev{};
    MapSlot *d_slot = nullptr;
    // staging ring for per-pass descriptors (pinned host bytes mirrored on the device); a
    // region is reused only after the event recorded behind its upload has completed
    uint8_t *h_stage = nullptr;
    uint8_t *d_stage = nullptr;
    size_t stage_cap = 0, stage_used = 0;  // bytes
    struct StageRegion {
        size_t begin, end;
        hipEvent_t ev;
    };
    std::vector<StageRegion> inflight;
    unsigned grid = 1024;
    uint32_t max_rec = 0, max_rd = 0;
    std::vector<LoggedLayout> logged;
    std::vector<void *> allocations;
    bool poisoned = false;
    bool validate = false;  // TPGP_HIP_VALIDATE=1: check order and regions after every pass
    uint64_t log_entries = 1u << 20;  // device log capacity (grown by reserve_log)
    // path choice (per rotate_many call): modelled costs of one rotation, or a fixed bound
    double cost_launch_us = 45.0;  // multi-kernel chain of one rotation (launches, latency)
    double cost_multi_ns = 0.03;   // multi-kernel streaming cost per term
    double cost_small_ns = 5.0;    // one persistent workgroup on global memory, per term of its
                                   // map (W=1; twice that for two-word keys)
    double cost_lds_us = 3.0;      // one persistent workgroup with the map in LDS: per rotation
    double cost_lds_ns = 1.0;      // and per term
    bool small_fixed = false;      // TPGP_HIP_SMALL_LIMIT given: length bound instead
    uint64_t small_limit = 0;
    std::vector<uint8_t> small;  // per map, decided at each rotate_many call
    uint32_t pass_count = 0;
    // kernel timing with HIP events on the batch stream (kinds: TPGP_HIP_PROFILE_*)
    unsigned profiling = 0;  // bit k: bracket launches of profile kind k with events
    std::vector<hipEvent_t> event_pool;
    struct Pending {
        unsigned kind;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending;
    double prof_ms[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t prof_count[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    // step graphs (TPGP_HIP_GRAPHS=0 disables them)
    bool graphs_on = true;
    // appending rotations (TPGP_HIP_APPEND=0 copies every class instead)
    bool append_on = true;
    // the rotation scan's form by size (0), or one form for every rotation (TPGP_HIP_SCAN =
    // maps | levels | direct: 1, 2, 3), for tests
    int scan_form = 0;
    bool capturing = false, capture_spilled = false;
    std::vector<uint8_t> cap_host;            // descriptors staged while capturing
    std::vector<std::vector<uint8_t>> spill;  // staging beyond cap_host (capture is dropped)
    size_t cap_used = 0, graph_desc_cap = 0;
    StepGraph recording;
    size_t logged_mark = 0;
    std::vector<StepGraph> graphs;
    std::vector<uint64_t> seen;  // fingerprints of recent pre-step states (capture on recurrence)
    uint64_t gen = 0, use_clock = 0;
};

namespace {

template <typename T> hipError_t dalloc(tpgp_hip_batch *b, T **p, size_t count) {
    size_t bytes = count * sizeof(T);
    if (!bytes) bytes = sizeof(T);
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e == hipSuccess) {
        b->allocations.push_back(q);
        *p = (T *)q;
    }
    return e;
}

void collect_profile(tpgp_hip_batch *b);
void drop_graph(StepGraph &g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.desc) (void)hipFree(g.desc);
    g.exec = nullptr;
    g.desc = nullptr;
}
// Device buffers are about to move: recorded steps point at them.
void drop_graphs(tpgp_hip_batch *b) {
    for (StepGraph &g : b->graphs) drop_graph(g);
    b->graphs.clear();
    b->seen.clear();
    ++b->gen;
}
void free_all(tpgp_hip_batch *b) {
    drop_graphs(b);
    collect_profile(b);
    for (auto &R : b->inflight) b->event_pool.push_back(R.ev);
    b->inflight.clear();
    b->stage_used = 0;
    for (hipEvent_t e : b->event_pool) (void)hipEventDestroy(e);
    b->event_pool.clear();
    for (void *p : b->allocations) (void)hipFree(p);
    b->allocations.clear();
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    b->h_stage = nullptr;
}

// Fit all device buffers to the current slot layout (maps' off/cap already assigned).
int allocate_device(tpgp_hip_batch *b, char *error, size_t cap) {
    hipError_t e;
    const uint32_t M = b->M;
    const uint64_t T = b->total_cap;
    b->max_tiles = (2 * T) / kMergeTile + 16ull * M + (T / (kBlock * kPartRoundsMin)) + M + 64;
    BatchDev &d = b->dev;
    d = BatchDev{};
    d.n_maps = M;
    d.max_rec = b->max_rec;
    d.max_rd = b->max_rd;
#define ALLOC(ptr, n)                                                                    \
    do {                                                                                 \
        if ((e = dalloc(b, &(ptr), (n))) != hipSuccess) {                                \
            hip_fail(e, "HIP resident-batch allocation", error, cap);                    \
            return e == hipErrorOutOfMemory ? TPGP_HIP_ENOMEM : TPGP_HIP_EDEVICE;        \
        }                                                                                \
    } while (0)
    for (int k = 0; k < 3; ++k) {
        ALLOC(d.arena[k].keys, T * b->W + 4);
        ALLOC(d.arena[k].vals, T + 4);
    }
    ALLOC(b->d_slot, M);
    d.slot = b->d_slot;
    ALLOC(d.len, (size_t)M * 3);
    ALLOC(d.regions, (size_t)M * 3 * kRegStride);
    ALLOC(d.tstart, (size_t)M * 16 + 2);
    ALLOC(d.tile_hist, b->max_tiles * 16);
    ALLOC(d.tile_base, b->max_tiles * 16);
    ALLOC(d.map_hist, (size_t)M * 16);
    ALLOC(d.splits, b->max_tiles + 1);
    ALLOC(d.tdesc, b->max_tiles);
    ALLOC(d.work, b->max_tiles);
    ALLOC(d.tsplit, b->max_tiles * kBlock);
    ALLOC(d.wstart, (size_t)M * 16 + 2);
    ALLOC(d.tile_led, b->max_tiles);
    ALLOC(d.chunk_start, (size_t)M + 1);
    ALLOC(d.chunk_base, b->max_tiles / kChunkTiles + M + 1);
    ALLOC(d.chunk_sum, b->max_tiles / kChunkTiles + M + 1);
    ALLOC(d.chunk_led, b->max_tiles / kChunkTiles + M + 1);
    ALLOC(d.counter, 4);
    ALLOC(d.overflow, M);
    ALLOC(d.small_resume, M);
    ALLOC(d.rec, (size_t)M * b->max_rec);
    ALLOC(d.rd, (size_t)M * b->max_rd * 64);
    d.log_cap = b->log_entries;
    ALLOC(d.log, d.log_cap);
    ALLOC(d.log_count, 1);
    ALLOC(d.debug, 16);
    ALLOC(d.sumsq, 2 * kSumBlocks);
    ALLOC(b->d_stage, b->stage_cap);
#undef ALLOC
    e = hipHostMalloc((void **)&b->h_stage, b->stage_cap, hipHostMallocDefault);
    if (e != hipSuccess) {
        b->h_stage = nullptr;
        hip_fail(e, "HIP pinned staging allocation", error, cap);
        return TPGP_HIP_ENOMEM;
    }
    hipStream_t st = b->ctx->stream;
    std::vector<MapSlot> slots(M);
    for (uint32_t m = 0; m < M; ++m)
        slots[m] = MapSlot{b->maps[m].off, b->maps[m].cap, (uint32_t)b->maps[m].n_spins, 0};
    if ((e = hipMemcpyAsync(b->d_slot, slots.data(), M * sizeof(MapSlot), hipMemcpyHostToDevice,
                            st)) != hipSuccess)
        return drain_fail(st, e, "HIP slot upload", error, cap);
    hipLaunchKernelGGL(k_fill_u64, dim3(8), dim3(256), 0, st, d.len, (uint64_t)M * 3, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(8), dim3(256), 0, st, d.regions,
                       (uint64_t)M * 3 * kRegStride, 0ull);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, d.overflow, (uint64_t)M, 0u);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, d.log_count, 1ull, 0ull);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, d.counter, 4ull, 0u);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, d.debug, 16ull, 0ull);
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "HIP batch initialisation", error, cap);
    b->stage_used = 0;
    return TPGP_HIP_OK;
}

// Reserve room for n descriptors of type T in the staging ring (zeroed).  Waits only for the
// uploads of earlier regions that the new one overlaps.
void release_stage(tpgp_hip_batch *b);
template <class T> T *stage(tpgp_hip_batch *b, size_t n, T **dev) {
    const size_t bytes = (n * sizeof(T) + 255) & ~(size_t)255;
    if (b->capturing) {  // the step graph's own descriptor buffer, uploaded once at its end
        uint8_t *h;
        if (b->cap_used + bytes <= b->cap_host.size()) {
            h = b->cap_host.data() + b->cap_used;
            *dev = reinterpret_cast<T *>(b->recording.desc + b->cap_used);
            b->cap_used += bytes;
        } else {  // too large for the graph's buffer: finish planning, drop the capture
            b->capture_spilled = true;
            b->spill.emplace_back(bytes);
            h = b->spill.back().data();
            *dev = reinterpret_cast<T *>(b->recording.desc);
        }
        std::memset((void *)h, 0, bytes);
        return reinterpret_cast<T *>(h);
    }
    if (b->stage_used + bytes > b->stage_cap) b->stage_used = 0;
    const size_t lo = b->stage_used, hi = lo + bytes;
    size_t keep = 0;
    for (size_t r = 0; r < b->inflight.size(); ++r) {
        auto &R = b->inflight[r];
        if (R.begin < hi && lo < R.end) {
            (void)hipEventSynchronize(R.ev);
            b->event_pool.push_back(R.ev);
        } else {
            b->inflight[keep++] = R;
        }
    }
    b->inflight.resize(keep);
    T *h = reinterpret_cast<T *>(b->h_stage + lo);
    *dev = reinterpret_cast<T *>(b->d_stage + lo);
    std::memset((void *)h, 0, bytes);
    b->stage_used = hi;
    return h;
}
hipEvent_t pooled_event(tpgp_hip_batch *b);
template <class T> hipError_t push_stage(tpgp_hip_batch *b, T *h, T *d, size_t n) {
    if (b->capturing) return hipSuccess;  // uploaded with the whole graph
    const size_t bytes = n * sizeof(T);
    hipError_t e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, b->ctx->stream);
    if (e != hipSuccess) return e;
    hipEvent_t ev = pooled_event(b);
    if (!ev) return hipStreamSynchronize(b->ctx->stream);  // no event: finish the upload now
    const size_t lo = (size_t)(reinterpret_cast<uint8_t *>(h) - b->h_stage);
    e = hipEventRecord(ev, b->ctx->stream);
    b->inflight.push_back({lo, lo + ((bytes + 255) & ~(size_t)255), ev});
    return e;
}
// After the stream has drained: every staged region is free again.
void release_stage(tpgp_hip_batch *b) {
    for (auto &R : b->inflight) b->event_pool.push_back(R.ev);
    b->inflight.clear();
    b->stage_used = 0;
}

int next_buffer(const HostMap &hm) {
    for (int k = 0; k < 3; ++k)
        if (k != hm.kc && k != hm.kcur) return k;
    return (hm.kc + 1) % 3;
}

// Which maps run on the small-map path.  Per rotation, the multi-kernel passes cost a fixed
// launch chain plus a per-term streaming cost over their maps; the persistent kernel costs the
// per-term time of one workgroup on its largest map (its workgroups run side by side, in waves
// of `resident` groups).  Maps are taken onto the small path in increasing length while that
// lowers the modelled time.  TPGP_HIP_SMALL_LIMIT fixes the cut instead (a length bound).
void choose_paths(tpgp_hip_batch *b, const tpgp_hip_rotation *req, size_t n_rot) {
    const uint32_t M = b->M;
    std::vector<uint32_t> idx;
    idx.reserve(M);
    for (uint32_t m = 0; m < M; ++m) {
        b->small[m] = 0;
        bool active = false;
        for (size_t r = 0; r < n_rot && !active; ++r) active = req[r * M + m].active != 0;
        if (active) idx.push_back(m);
    }
    if (b->validate) return;
    if (b->small_fixed) {
        for (uint32_t m : idx) b->small[m] = b->maps[m].est_len <= b->small_limit;
        return;
    }
    std::sort(idx.begin(), idx.end(),
              [&](uint32_t x, uint32_t y) { return b->maps[x].est_len < b->maps[y].est_len; });
    double rest = 0;  // terms left on the multi-kernel path
    for (uint32_t m : idx) rest += (double)b->maps[m].est_len;
    // one small-map workgroup per CU (its LDS holds a whole map); a map that fits the LDS
    // capacity with room to grow runs there, larger ones on global memory
    const double resident = (double)b->ctx->cus;
    const double lds_cap = b->W == 1 ? LdsCap<1>::value : LdsCap<2>::value;
    auto small_us = [&](double len) {
        if (len <= 0.8 * lds_cap) return b->cost_lds_us + b->cost_lds_ns * len * 1e-3;
        return b->cost_small_ns * (b->W == 2 ? 2.0 : 1.0) * len * 1e-3;
    };
    auto cost = [&](size_t k, double large_terms) {  // first k maps small
        double t = 0;
        if (k < idx.size()) t += b->cost_launch_us + b->cost_multi_ns * large_terms * 1e-3;
        if (k > 0)
            t += small_us((double)b->maps[idx[k - 1]].est_len) * std::ceil((double)k / resident);
        return t;
    };
    size_t best = 0;
    double best_t = cost(0, rest), large = rest;
    for (size_t k = 1; k <= idx.size(); ++k) {
        large -= (double)b->maps[idx[k - 1]].est_len;
        const double t = cost(k, large);
        if (t < best_t) {
            best_t = t;
            best = k;
        }
    }
    for (size_t k = 0; k < best; ++k) b->small[idx[k]] = 1;
}

// Tile kernels run one tile per workgroup: the grid covers the estimated tile count with a
// margin (the kernels' grid-stride loops absorb an underestimate; surplus groups exit).
unsigned tile_grid(uint64_t tiles) {
    const uint64_t g = tiles + tiles / 16 + 64;
    return (unsigned)(g < (1u << 22) ? g : (1u << 22));
}
// The tuned value of a launch-shape parameter for a pass over `terms` terms of a batch with W
// key words (0: the default).
unsigned tuned(const tpgp_hip_batch *b, unsigned param, uint64_t terms) {
    unsigned bucket = terms ? 63u - (unsigned)__builtin_clzll(terms) : 0u;
    if (bucket >= TPGP_HIP_TUNE_BUCKETS) bucket = TPGP_HIP_TUNE_BUCKETS - 1;
    return b->ctx->tune[b->W - 1][param][bucket];
}
// Grid and rounds per thread of a partition pass over the maps with want[m].
unsigned partition_grid(const tpgp_hip_batch *b, const std::vector<int> &want, int *rounds) {
    uint64_t terms = 0;
    for (uint32_t m = 0; m < b->M; ++m)
        if (want[m] && !b->small[m]) terms += b->maps[m].est_len;
    const unsigned r = tuned(b, TPGP_HIP_TUNE_PARTITION_ROUNDS, terms);
    *rounds = r ? (int)r : kPartRounds;
    const uint64_t tile = (uint64_t)kBlock * *rounds;
    uint64_t t = 0;
    for (uint32_t m = 0; m < b->M; ++m)
        if (want[m] && !b->small[m]) t += b->maps[m].est_len / tile + 1;
    return tile_grid(t);
}
unsigned rotation_grid(tpgp_hip_batch *b, const tpgp_hip_rotation *req) {
    uint64_t t = 0;
    for (uint32_t m = 0; m < b->M; ++m)
        if (req[m].active && !b->small[m]) t += (3 * b->maps[m].est_len / 2) / kMergeTile + 16;
    return tile_grid(t);
}

hipEvent_t pooled_event(tpgp_hip_batch *b) {
    if (!b->event_pool.empty()) {
        hipEvent_t e = b->event_pool.back();
        b->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
// Bracket the next launches of one kind with events when profiling is on.
struct ProfileScope {
    tpgp_hip_batch *b;
    unsigned kind;
    hipEvent_t start = nullptr;
    ProfileScope(tpgp_hip_batch *bb, unsigned k) : b(bb), kind(k) {
        if (b->capturing || !(b->profiling >> kind & 1u)) return;
        start = pooled_event(b);
        if (start) (void)hipEventRecord(start, b->ctx->stream);
    }
    ~ProfileScope() {
        if (!start) return;
        hipEvent_t stop = pooled_event(b);
        if (!stop) {
            b->event_pool.push_back(start);
            return;
        }
        (void)hipEventRecord(stop, b->ctx->stream);
        b->pending.push_back({kind, start, stop});
    }
};
// After the stream has drained: accumulate elapsed times and recycle the events.
void collect_profile(tpgp_hip_batch *b) {
    for (auto &p : b->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            b->prof_ms[p.kind] += ms;
            b->prof_count[p.kind] += 1;
        }
        b->event_pool.push_back(p.start);
        b->event_pool.push_back(p.stop);
    }
    b->pending.clear();
}

// Enqueue a partition pass for the maps with want[m] (positions pa/pb given per map).
// Partition descriptors for the maps with want[m] (positions pa/pb moved to the top), and the
// host-side state they imply: the working buffer and the digit layout of each map.
unsigned plan_partition(tpgp_hip_batch *b, const std::vector<int> &want,
                        const std::vector<uint32_t> &pa, const std::vector<uint32_t> &pb,
                        PassMap *h, int *rounds) {
    for (uint32_t m = 0; m < b->M; ++m) {
        if (!want[m]) continue;
        HostMap &hm = b->maps[m];
        h[m].active = 1;
        h[m].in_buf = (uint32_t)hm.kcur;
        h[m].out_buf = (uint32_t)next_buffer(hm);
        h[m].pa = pa[m];
        h[m].pb = pb[m];
        hm.kcur = (int)h[m].out_buf;
        hm.gap = false;  // the partition writes the map contiguously
        hm.est_top = hm.est_len;
        // sigma: move sites at pa, pb to the top, keep the others' order
        size_t n = hm.n_spins;
        uint8_t sa = hm.sig[pa[m]], sb = hm.sig[pb[m]];
        std::vector<uint8_t> rest;
        rest.reserve(n);
        for (size_t p = 0; p < n; ++p)
            if (p != pa[m] && p != pb[m]) rest.push_back(hm.sig[p]);
        rest.push_back(sb);
        rest.push_back(sa);
        hm.sig = rest;
    }
    return partition_grid(b, want, rounds);
}

template <int W, int R>
void launch_partition_tiles(tpgp_hip_batch *b, const PassMap *d, unsigned pgrid) {
    hipStream_t st = b->ctx->stream;
    hipLaunchKernelGGL(k_part_plan, dim3(1), dim3(kBlock), 0, st, b->dev, d,
                       (uint32_t)(kBlock * R));
    {
        ProfileScope ph(b, TPGP_HIP_PROFILE_HISTOGRAM);
        hipLaunchKernelGGL((k_part_count<W, R>), dim3(pgrid), dim3(kBlock), 0, st, b->dev, d);
    }
    hipLaunchKernelGGL(k_part_scan, dim3(b->M * 16), dim3(kBlock), 0, st, b->dev, d);
    {
        ProfileScope ps(b, TPGP_HIP_PROFILE_SCATTER);
        hipLaunchKernelGGL((k_part_scatter<W, R>), dim3(pgrid), dim3(kBlock), 0, st, b->dev, d);
    }
}

void launch_partition(tpgp_hip_batch *b, const PassMap *d, unsigned pgrid, int rounds) {
    hipStream_t st = b->ctx->stream;
    ProfileScope whole(b, TPGP_HIP_PROFILE_PARTITION);
    if (b->W == 1) {
        if (rounds == 4) launch_partition_tiles<1, 4>(b, d, pgrid);
        else if (rounds == 12) launch_partition_tiles<1, 12>(b, d, pgrid);
        else launch_partition_tiles<1, kPartRounds>(b, d, pgrid);
    } else {
        if (rounds == 4) launch_partition_tiles<2, 4>(b, d, pgrid);
        else if (rounds == 12) launch_partition_tiles<2, 12>(b, d, pgrid);
        else launch_partition_tiles<2, kPartRounds>(b, d, pgrid);
    }
    if (b->validate) {
        uint32_t tag = (++b->pass_count << 1) | 0u;
        if (b->W == 1)
            hipLaunchKernelGGL(k_validate<1>, dim3(256), dim3(kBlock), 0, st, b->dev, d, tag);
        else
            hipLaunchKernelGGL(k_validate<2>, dim3(256), dim3(kBlock), 0, st, b->dev, d, tag);
    }
}

hipError_t enqueue_partition(tpgp_hip_batch *b, const std::vector<int> &want,
                             const std::vector<uint32_t> &pa, const std::vector<uint32_t> &pb) {
    PassMap *d;
    PassMap *h = stage(b, b->M, &d);
    int rounds = kPartRounds;
    unsigned pgrid = plan_partition(b, want, pa, pb, h, &rounds);
    hipError_t e = push_stage(b, h, d, b->M);
    if (e != hipSuccess) return e;
    launch_partition(b, d, pgrid, rounds);
    return hipGetLastError();
}

size_t position_of(const HostMap &hm, size_t site) {
    for (size_t p = 0; p < hm.n_spins; ++p)
        if (hm.sig[p] == site) return p;
    return SIZE_MAX;
}

// Reference-layout key from a stored key (digit p holds site sigma[p]).
void to_reference(const std::vector<uint8_t> &sigma, const uint64_t *stored, uint64_t *ref,
                  int W) {
    ref[0] = 0;
    if (W > 1) ref[1] = 0;
    for (size_t p = 0; p < sigma.size(); ++p) {
        unsigned d = (unsigned)(stored[p >> 5] >> (2 * (p & 31))) & 3u;
        size_t s = sigma[p];
        ref[s >> 5] |= (uint64_t)d << (2 * (s & 31));
    }
}

}  // namespace

namespace {
// Calls that synchronise or move device buffers cannot run inside a step being recorded.
bool recording_step(const tpgp_hip_batch *b, char *error, size_t cap) {
    if (!b->capturing) return false;
    msg(error, cap, "A step graph is being recorded: end it first.");
    return true;
}
}  // namespace

extern "C" tpgp_hip_batch *tpgp_hip_batch_create(tpgp_hip_context *ctx, size_t n_words,
                                                 size_t n_maps, const size_t *n_spins,
                                                 const size_t *capacities, char *error,
                                                 size_t cap) {
    msg(error, cap, "");
    if (!ctx || (n_words != 1 && n_words != 2) || !n_maps || n_maps > (1u << 20) || !n_spins ||
        !capacities) {
        msg(error, cap, "Invalid resident-batch arguments.");
        return nullptr;
    }
    tpgp_hip_batch *b = new (std::nothrow) tpgp_hip_batch();
    if (!b) {
        msg(error, cap, "Could not allocate the batch host state.");
        return nullptr;
    }
    b->ctx = ctx;
    b->W = (int)n_words;
    b->M = (uint32_t)n_maps;
    b->maps.resize(n_maps);
    uint64_t off = 0;
    for (size_t m = 0; m < n_maps; ++m) {
        size_t n = n_spins[m];
        if (n < 2 || n > 32 * n_words || (n_words == 2 && n <= 32) || capacities[m] == 0 ||
            capacities[m] > (1ull << 37)) {
            msg(error, cap, "Invalid spin count or capacity for map %zu.", m);
            delete b;
            return nullptr;
        }
        HostMap &hm = b->maps[m];
        hm.n_spins = n;
        hm.cap = capacities[m];
        hm.off = off;
        off += (capacities[m] + 63) & ~63ull;
        hm.sig.resize(n);
        for (size_t p = 0; p < n; ++p) hm.sig[p] = (uint8_t)p;
        hm.sig_c = hm.sig;
    }
    b->total_cap = off;
    b->max_rec = n_maps <= 16 ? 4096u : (uint32_t)std::max<size_t>(1024, 65536 / n_maps);
    b->max_rd = n_maps <= 16 ? 1024u : 64u;
    b->stage_cap = std::min<size_t>(std::max<size_t>(n_maps * 1024, 1u << 16) * sizeof(PassMap),
                                    (size_t)128 << 20);
    b->grid = (unsigned)ctx->cus * 8;
    const char *v = getenv("TPGP_HIP_VALIDATE");
    b->validate = v && *v && *v != '0';
    // small-map path: worth it while one workgroup per map beats ~10 launches per rotation;
    // with many maps the workgroups fill the GPU, so larger maps qualify
    const char *sl = getenv("TPGP_HIP_SMALL_LIMIT");
    b->small_fixed = sl && *sl;
    b->small_limit = b->small_fixed ? strtoull(sl, nullptr, 10) : 0;
    b->small.assign(n_maps, 0);
    const char *gr = getenv("TPGP_HIP_GRAPHS");
    b->graphs_on = !(gr && *gr == '0');
    const char *ap = getenv("TPGP_HIP_APPEND");
    b->append_on = !(ap && *ap == '0');
    const char *sf = getenv("TPGP_HIP_SCAN");
    b->scan_form = !sf ? 0 : !strcmp(sf, "maps") ? 1 : !strcmp(sf, "levels") ? 2
                 : !strcmp(sf, "direct") ? 3 : 0;
    std::lock_guard<std::recursive_mutex> lock(ctx->mutex);
    DeviceGuard guard(ctx->device);
    if (!guard.ok || allocate_device(b, error, cap) != TPGP_HIP_OK) {
        if (!guard.ok) msg(error, cap, "HIP device selection failed.");
        free_all(b);
        delete b;
        return nullptr;
    }
    return b;
}

extern "C" void tpgp_hip_batch_destroy(tpgp_hip_batch *b) {
    if (!b) return;
    if (b->capturing) {  // a step left recording: drop it (nothing of it ran)
        int captured = 0;
        (void)tpgp_hip_batch_graph_end(b, 0, &captured, nullptr, 0);
    }
    {
        std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
        DeviceGuard guard(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        free_all(b);
    }
    delete b;
}

extern "C" void *tpgp_hip_batch_stream(const tpgp_hip_batch *b) {
    return b ? (void *)b->ctx->stream : nullptr;
}

extern "C" int tpgp_hip_batch_info(const tpgp_hip_batch *b, size_t map, size_t *len,
                                   size_t *capacity) {
    if (!b || map >= b->M) return TPGP_HIP_EINVAL;
    if (len) *len = b->maps[map].committed_len;
    if (capacity) *capacity = b->maps[map].cap;
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_batch_upload(tpgp_hip_batch *b, size_t map, const uint64_t *keys,
                                     const double *values, size_t len, char *error, size_t cap) {
    msg(error, cap, "");
    if (!b || map >= b->M || (len && (!keys || !values))) {
        msg(error, cap, "Invalid upload arguments.");
        return TPGP_HIP_EINVAL;
    }
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    HostMap &hm = b->maps[map];
    if (len > hm.cap) {
        msg(error, cap, "Map %zu: %zu terms exceed the slot capacity %llu.", map, len,
            (unsigned long long)hm.cap);
        return TPGP_HIP_ECAPACITY;
    }
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    hipStream_t st = b->ctx->stream;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "HIP upload", error, cap);
    // discard pending work of this map, install the new state as committed (contiguous)
    hm.kcur = hm.kc;
    hm.gap = hm.gap_c = false;
    for (size_t p = 0; p < hm.n_spins; ++p) hm.sig[p] = (uint8_t)p;
    hm.sig_c = hm.sig;
    double fl = INFINITY;
    bool zero = false;  // a stored 0 (even the identity's) would be dropped by any rotation
    for (size_t t = 0; t < len; ++t) {
        bool ident = true;
        for (int w = 0; w < b->W; ++w) ident = ident && keys[t * b->W + w] == 0;
        if (!(values[t] != 0.0)) zero = true;
        if (!ident) fl = std::min(fl, std::fabs(values[t]));
    }
    hm.floor = hm.floor_c = (zero || !std::isfinite(fl)) ? 0.0 : fl;
    const int k = hm.kc;
    if (len) {
        e = hipMemcpyAsync(b->dev.arena[k].keys + hm.off * b->W, keys, len * b->W * 8,
                           hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->dev.arena[k].vals + hm.off, values, len * 8,
                               hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return drain_fail(st, e, "HIP upload transfer", error, cap);
    }
    hipLaunchKernelGGL(k_set_len, dim3(1), dim3(1), 0, st, b->dev.len, (uint64_t)map * 3 + k,
                       (uint64_t)len);
    if (b->W == 1)
        hipLaunchKernelGGL(k_regions<1>, dim3(1), dim3(64), 0, st, b->dev, (uint32_t)map,
                           (uint32_t)k);
    else
        hipLaunchKernelGGL(k_regions<2>, dim3(1), dim3(64), 0, st, b->dev, (uint32_t)map,
                           (uint32_t)k);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "HIP upload", error, cap);
    hm.committed_len = len;
    hm.est_len = len;
    hm.nrec = hm.nrd = 0;
    hm.app_recs.clear();
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_batch_assign(tpgp_hip_batch *b, size_t map, size_t n_spins,
                                     const uint64_t *keys, const double *values, size_t len,
                                     char *error, size_t cap) {
    msg(error, cap, "");
    if (!b || map >= b->M || n_spins < 2 || n_spins > 32u * (unsigned)b->W ||
        (b->W == 2 && n_spins <= 32)) {
        msg(error, cap, "Invalid assign arguments (map %zu, %zu spins).", map, n_spins);
        return TPGP_HIP_EINVAL;
    }
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    HostMap &hm = b->maps[map];
    if (hm.n_spins != n_spins) {
        std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
        DeviceGuard guard(b->ctx->device);
        hipError_t e = hipStreamSynchronize(b->ctx->stream);
        hm.n_spins = n_spins;
        hm.sig.assign(n_spins, 0);
        hm.sig_c.assign(n_spins, 0);
        const MapSlot slot{hm.off, hm.cap, (uint32_t)n_spins, 0};
        if (e == hipSuccess)
            e = hipMemcpy(b->d_slot + map, &slot, sizeof(slot), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "HIP assign", error, cap);
    }
    return tpgp_hip_batch_upload(b, map, keys, values, len, error, cap);
}

namespace {
// Rotation descriptors for the active requests, and the host-side state they imply.
unsigned plan_rotation(tpgp_hip_batch *b, const tpgp_hip_rotation *req, uint32_t rotation,
                       PassMap *h) {
    for (uint32_t m = 0; m < b->M; ++m) {
        const tpgp_hip_rotation &r = req[m];
        if (!r.active) continue;
        HostMap &hm = b->maps[m];
        size_t n = hm.n_spins;
        bool i_top = hm.sig[n - 1] == r.i;
        const bool kept = r.threshold < hm.floor;
        // Append when the unchanged classes can stay where they are: the working buffer is not
        // the committed one (whose table and terms a rollback returns to), and its free space
        // holds the planned bound on what the rotation writes (an underestimate is caught on
        // the device as an overflow; the step is rolled back and the bound raised at sync).
        const double grown = (double)hm.est_len * hm.growth;
        const uint64_t est = (uint64_t)(grown + grown / 8) + 1024;
        const uint64_t add = (uint64_t)(kAppendLevels[hm.app_level] * (double)est) + 1024;
        const bool append = b->append_on && kept && !b->small[m] && hm.kcur != hm.kc &&
                            hm.est_top + add <= hm.cap;
        h[m].active = 1;
        h[m].in_buf = (uint32_t)hm.kcur;
        h[m].out_buf = append ? (uint32_t)hm.kcur : (uint32_t)next_buffer(hm);
        h[m].qt = i_top ? r.qi : r.qj;
        h[m].qs = i_top ? r.qj : r.qi;
        h[m].flags = (r.validated ? 1u : 0u) | (r.log ? 2u : 0u) |
                     (kept ? kPassCommutingKept : 0u) | (append ? kPassAppend : 0u);
        hm.gap = append;
        hm.est_top = append ? hm.est_top + add : est;
        if (append) hm.app_recs.push_back(hm.nrec);
        // afterwards every kept non-identity coefficient exceeds this threshold
        hm.floor = std::nextafter(r.threshold, INFINITY);
        h[m].rec_slot = hm.nrec;
        h[m].rot_seq = rotation;
        h[m].c = r.c;
        h[m].s = r.s;
        h[m].rc = r.validated ? r.rc : 0.0;
        h[m].rs = r.validated ? r.rs : 0.0;
        h[m].threshold = r.threshold;
        hm.kcur = (int)h[m].out_buf;
        hm.nrec++;
        if (r.log) b->logged.push_back(LoggedLayout{m, rotation, hm.sig});
    }
    return rotation_grid(b, req);
}

// Rotations whose maps average at most this many tiles scan in one launch (k_rot_scan_maps).
constexpr uint64_t kSmallScanTiles = 4 * kChunkTiles;

void launch_rotation(tpgp_hip_batch *b, const PassMap *d, unsigned rgrid, bool validated) {
    hipStream_t st = b->ctx->stream;
    // this rotation's tile estimate, from its own grid (rgrid = tile_grid(tiles)); planning may
    // run ahead of launching, so the batch's last estimate can belong to a later rotation
    const uint64_t est_tiles = rgrid > 64 ? (uint64_t)(rgrid - 64) * 16 / 17 + 1 : 1;
    {
        ProfileScope whole(b, TPGP_HIP_PROFILE_ROTATION);
        const int plan_in_splits = b->M <= kPlanMaps;  // the split pass plans too
        if (!plan_in_splits) {
            ProfileScope pp(b, TPGP_HIP_PROFILE_PLAN);
            hipLaunchKernelGGL(k_rot_plan, dim3(1), dim3(kBlock), 0, st, b->dev, d);
        }
        {
            ProfileScope pc(b, TPGP_HIP_PROFILE_SPLITS);
            const unsigned sgrid = tile_grid(est_tiles / 32 + 1);  // 8 lanes per tile
            if (b->W == 1)
                hipLaunchKernelGGL(k_rot_splits<1>, dim3(sgrid), dim3(kBlock), 0, st, b->dev, d,
                                   plan_in_splits);
            else
                hipLaunchKernelGGL(k_rot_splits<2>, dim3(sgrid), dim3(kBlock), 0, st, b->dev, d,
                                   plan_in_splits);
        }
        {
            ProfileScope pc(b, TPGP_HIP_PROFILE_COUNT);
            // the count pass visits one tile per anticommuting pair and the commuting classes
            // truncation may change: about a third of all tiles (more loop in the grid)
            const unsigned tc = tuned(b, TPGP_HIP_TUNE_COUNT_GRID, est_tiles * kMergeTile);
            const unsigned cgrid = tc ? (unsigned)((uint64_t)rgrid * tc / 6) + 64
                                      : rgrid / 2 + 64;  // rgrid: about one workgroup per tile
            if (b->W == 1 && validated)
                hipLaunchKernelGGL((k_rot_count<1, true>), dim3(cgrid), dim3(kBlock), 0, st,
                                   b->dev, d);
            else if (b->W == 1)
                hipLaunchKernelGGL((k_rot_count<1, false>), dim3(cgrid), dim3(kBlock), 0, st,
                                   b->dev, d);
            else if (validated)
                hipLaunchKernelGGL((k_rot_count<2, true>), dim3(cgrid), dim3(kBlock), 0, st,
                                   b->dev, d);
            else
                hipLaunchKernelGGL((k_rot_count<2, false>), dim3(cgrid), dim3(kBlock), 0, st,
                                   b->dev, d);
        }
        int finish = 0;  // the write pass ends the map-level scan (direct chunk offsets)
        {
            ProfileScope psn(b, TPGP_HIP_PROFILE_SCAN);
            const uint64_t want = est_tiles / kChunkTiles + b->M + 1;
            const unsigned cgrid = (unsigned)(want < (1u << 20) ? want : (1u << 20));
            // mid-size rotations scan with direct chunk offsets (k_rot_scan1 with direct, the
            // map-level end inside the write pass): one launch less than two levels
            const int form =
                b->scan_form                                               ? b->scan_form
                : est_tiles <= (uint64_t)kSmallScanTiles * b->M            ? 1
                : est_tiles <= kDirectScanChunks * kChunkTiles - b->M * 16 ? 3
                                                                           : 2;
            if (form == 1) {
                hipLaunchKernelGGL(k_rot_scan_maps, dim3(b->M), dim3(kBlock), 0, st, b->dev, d);
            } else if (form == 3) {
                finish = 1;
                hipLaunchKernelGGL(k_rot_scan1, dim3(cgrid), dim3(kBlock), 0, st, b->dev, d, 1);
            } else {
                hipLaunchKernelGGL(k_rot_scan1, dim3(cgrid), dim3(kBlock), 0, st, b->dev, d, 0);
                hipLaunchKernelGGL(k_rot_scan2, dim3(b->M), dim3(kBlock), 0, st, b->dev, d);
            }
        }
        {
            ProfileScope pm(b, TPGP_HIP_PROFILE_MERGE);
            // every tile but the higher class's of each anticommuting pair
            const unsigned tw = tuned(b, TPGP_HIP_TUNE_WRITE_GRID, est_tiles * kMergeTile);
            const unsigned wgrid = tw ? (unsigned)((uint64_t)rgrid * tw / 6) + 64
                                      : rgrid / 3 * 2 + 64;
            if (b->W == 1)
                hipLaunchKernelGGL(k_rot_write<1>, dim3(wgrid), dim3(kBlock), 0, st, b->dev, d,
                                   finish);
            else
                hipLaunchKernelGGL(k_rot_write<2>, dim3(wgrid), dim3(kBlock), 0, st, b->dev, d,
                                   finish);
        }
    }
    if (b->validate) {
        uint32_t tag = (++b->pass_count << 1) | 1u;
        if (b->W == 1)
            hipLaunchKernelGGL(k_validate<1>, dim3(256), dim3(kBlock), 0, st, b->dev, d, tag);
        else
            hipLaunchKernelGGL(k_validate<2>, dim3(256), dim3(kBlock), 0, st, b->dev, d, tag);
    }
}

}  // namespace

extern "C" int tpgp_hip_batch_rotate_many(tpgp_hip_batch *b, const tpgp_hip_rotation *req,
                                          const uint32_t *rotations, size_t n_rot, char *error,
                                          size_t cap) {
    msg(error, cap, "");
    if (!b || (!req && n_rot) || (!rotations && n_rot)) {
        msg(error, cap, "Invalid rotation arguments.");
        return TPGP_HIP_EINVAL;
    }
    const uint32_t M = b->M;
    std::vector<uint32_t> count(M, 0);
    for (size_t r = 0; r < n_rot; ++r)
        for (uint32_t m = 0; m < M; ++m) {
            const tpgp_hip_rotation &q = req[r * M + m];
            if (!q.active) continue;
            const HostMap &hm = b->maps[m];
            if (q.i >= hm.n_spins || q.j >= hm.n_spins || q.i == q.j || q.qi > 3 || q.qj > 3 ||
                !std::isfinite(q.c) || !std::isfinite(q.s) || !std::isfinite(q.rc) ||
                !std::isfinite(q.rs) || !std::isfinite(q.threshold) || q.threshold < 0) {
                msg(error, cap, "Invalid rotation request for map %u.", m);
                return TPGP_HIP_EINVAL;
            }
            if (hm.nrec + ++count[m] > b->max_rec) {
                msg(error, cap, "Too many rotations since the last sync (map %u).", m);
                return TPGP_HIP_EINVAL;
            }
        }
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    // rotations go out in chunks whose descriptors fill at most a quarter of the staging ring,
    // so the host stages one chunk while the device runs the previous ones
    const size_t per_rot = 2 * (size_t)M * sizeof(PassMap);
    const size_t chunk = std::max<size_t>(1, b->stage_cap / 4 / per_rot);
    std::vector<int> want(M);
    std::vector<uint32_t> pa(M), pb(M);
    struct Launch {
        int kind;  // 0 rotation slot, 1 partition slot
        size_t at;
        unsigned grid;
        bool validated;  // rotation: some map computes radii
        int rounds;      // partition: elements per thread (tile = kBlock * rounds)
    };
    auto any_validated = [M](const tpgp_hip_rotation *rq) {
        for (uint32_t m = 0; m < M; ++m)
            if (rq[m].active && rq[m].validated) return true;
        return false;
    };
    std::vector<Launch> launches;
    // path of every map for this call (by its last known length)
    choose_paths(b, req, n_rot);
    std::vector<uint32_t> small_ids;
    for (uint32_t m = 0; m < M; ++m)
        if (b->small[m]) small_ids.push_back(m);
    for (size_t r0 = 0; r0 < n_rot; r0 += chunk) {
    const size_t r1 = std::min(n_rot, r0 + chunk);
    PassMap *d0;
    PassMap *h0 = stage(b, 2 * (r1 - r0) * M, &d0);
    launches.clear();
    size_t used = 0;
    bool any_small = false;
    for (size_t r = r0; r < r1; ++r) {
        const tpgp_hip_rotation *q = req + r * M;
        // bring each active map's pair (i, j) to the top two digits
        bool any = false;
        for (uint32_t m = 0; m < M; ++m) {
            want[m] = 0;
            if (!q[m].active) continue;
            HostMap &hm = b->maps[m];
            size_t n = hm.n_spins;
            uint8_t top = hm.sig[n - 1], sec = hm.sig[n - 2];
            if ((top == q[m].i && sec == q[m].j) || (top == q[m].j && sec == q[m].i)) continue;
            want[m] = 1;
            any = true;
            pa[m] = (uint32_t)position_of(hm, q[m].i);
            pb[m] = (uint32_t)position_of(hm, q[m].j);
        }
        if (any) {
            int rounds = kPartRounds;
            const unsigned g = plan_partition(b, want, pa, pb, h0 + used, &rounds);
            launches.push_back({1, used, g, false, rounds});
            used += M;
        }
        launches.push_back({0, used, plan_rotation(b, q, rotations[r], h0 + used),
                            any_validated(q), 0});
        used += M;
    }
    // maps on the small-map path leave the multi-kernel passes to the persistent kernel
    std::vector<uint8_t> pass_large(used / M, 0);
    for (const Launch &L : launches) {
        const size_t k = L.at / M;
        for (uint32_t m = 0; m < M; ++m) {
            PassMap &P = h0[k * M + m];
            if (!P.active) continue;
            if (b->small[m]) {
                P.active = 0;
                P.flags |= kPassSmall | (L.kind == 0 ? 0u : kPassPartition);
                any_small = true;
            } else {
                pass_large[k] = 1;
            }
        }
    }
    hipError_t e = push_stage(b, h0, d0, used);
    if (e != hipSuccess) return hip_fail(e, "HIP descriptor staging", error, cap);
    for (const Launch &L : launches) {
        if (!pass_large[L.at / M]) continue;
        if (L.kind == 0)
            launch_rotation(b, d0 + L.at, L.grid, L.validated);
        else
            launch_partition(b, d0 + L.at, L.grid, L.rounds);
    }
    if (any_small) {
        // the map list goes with this chunk's descriptors: a ring region is only guaranteed
        // until later uploads reuse it, so every launch reads a copy staged for it
        uint32_t *d_small = nullptr;
        uint32_t *h_small = stage(b, small_ids.size(), &d_small);
        std::memcpy(h_small, small_ids.data(), small_ids.size() * sizeof(uint32_t));
        if ((e = push_stage(b, h_small, d_small, small_ids.size())) != hipSuccess)
            return hip_fail(e, "HIP descriptor staging", error, cap);
        ProfileScope ps(b, TPGP_HIP_PROFILE_SMALL);
        const unsigned ns = (unsigned)small_ids.size();
        const uint32_t n_ops = (uint32_t)(used / M);
        // the LDS kernel first; the global one finishes what it left (small_resume)
        if (b->W == 1) {
            hipLaunchKernelGGL(k_small_lds<1>, dim3(ns), dim3(kLdsThreads), 0, b->ctx->stream,
                               b->dev, (const PassMap *)d0, n_ops, d_small, ns);
            hipLaunchKernelGGL(k_small<1>, dim3(ns), dim3(kBlock), 0, b->ctx->stream, b->dev,
                               (const PassMap *)d0, n_ops, d_small, ns);
        } else {
            hipLaunchKernelGGL(k_small_lds<2>, dim3(ns), dim3(kLdsThreads), 0, b->ctx->stream,
                               b->dev, (const PassMap *)d0, n_ops, d_small, ns);
            hipLaunchKernelGGL(k_small<2>, dim3(ns), dim3(kBlock), 0, b->ctx->stream, b->dev,
                               (const PassMap *)d0, n_ops, d_small, ns);
        }
    }
    if ((e = hipGetLastError()) != hipSuccess)
        return hip_fail(e, "HIP rotation launch", error, cap);
    }
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_batch_rotate(tpgp_hip_batch *b, const tpgp_hip_rotation *req,
                                     uint32_t rotation, char *error, size_t cap) {
    return tpgp_hip_batch_rotate_many(b, req, &rotation, 1, error, cap);
}

extern "C" int tpgp_hip_batch_readout(tpgp_hip_batch *b, unsigned pauli, const int *active,
                                      char *error, size_t cap) {
    msg(error, cap, "");
    if (!b || pauli < 1 || pauli > 3) {
        msg(error, cap, "Invalid readout arguments.");
        return TPGP_HIP_EINVAL;
    }
    for (uint32_t m = 0; m < b->M; ++m)
        if ((!active || active[m]) && b->maps[m].nrd >= b->max_rd) {
            msg(error, cap, "Too many readouts since the last sync (map %u).", m);
            return TPGP_HIP_EINVAL;
        }
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    ReadoutMap *d;
    ReadoutMap *h = stage(b, b->M, &d);
    for (uint32_t m = 0; m < b->M; ++m) {
        if (active && !active[m]) continue;
        HostMap &hm = b->maps[m];
        h[m].active = 1;
        h[m].in_buf = (uint32_t)hm.kcur;
        h[m].rec_slot = hm.nrd++;
        for (size_t p = 0; p < hm.n_spins; ++p) h[m].pos[hm.sig[p]] = (uint32_t)p;
    }
    hipError_t e = push_stage(b, h, d, b->M);
    if (e != hipSuccess) return hip_fail(e, "HIP readout staging", error, cap);
    if (b->W == 1)
        hipLaunchKernelGGL(k_readout<1>, dim3(b->M), dim3(64), 0, b->ctx->stream, b->dev,
                           (const ReadoutMap *)d, pauli);
    else
        hipLaunchKernelGGL(k_readout<2>, dim3(b->M), dim3(64), 0, b->ctx->stream, b->dev,
                           (const ReadoutMap *)d, pauli);
    if ((e = hipGetLastError()) != hipSuccess)
        return hip_fail(e, "HIP readout launch", error, cap);
    return TPGP_HIP_OK;
}

namespace {
// After a step: the planning bound on what one appending rotation writes follows the largest
// fraction a rotation of this map appended (up at once, with a margin; down one level at a time
// when well clear), and rises two levels when the step overflowed the slot.
void learn_append(HostMap &hm, const tpgp_hip_record *rec) {
    double most = 0.0;
    bool over = false;
    for (uint32_t r = 0; r < hm.nrec; ++r) over = over || rec[r].overflow;
    for (uint32_t r : hm.app_recs)  // wrote len_out - in_place after the extent
        if (r < hm.nrec && rec[r].len_in)
            most = std::max(most, (double)(rec[r].len_out - rec[r].in_place) /
                                      (double)rec[r].len_in);
    hm.app_recs.clear();
    if (most == 0.0 && !over) return;  // nothing appended: nothing learned
    uint32_t need = 0;
    while (need < kAppendTop && kAppendLevels[need] < most * 1.15 + 0.02) ++need;
    if (over) need = std::max(need, std::min(kAppendTop, hm.app_level + 2));
    if (need > hm.app_level)
        hm.app_level = need;
    else if (hm.app_level > 0 && kAppendLevels[hm.app_level - 1] >= most * 1.6 + 0.05)
        --hm.app_level;
}
}  // namespace

extern "C" int tpgp_hip_batch_sum_squares(tpgp_hip_batch *b, size_t map, double out[2],
                                          char *error, size_t cap) {
    msg(error, cap, "");
    if (!b || map >= b->M || !out) {
        msg(error, cap, "Invalid sum-of-squares arguments.");
        return TPGP_HIP_EINVAL;
    }
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    HostMap &hm = b->maps[map];
    if (hm.kcur != hm.kc || hm.nrec || hm.nrd) {
        msg(error, cap, "Map %zu has uncommitted work; sync and commit or roll back first.", map);
        return TPGP_HIP_EINVAL;
    }
    out[0] = out[1] = 0.0;
    const uint64_t len = hm.committed_len;
    if (!len) return TPGP_HIP_OK;
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    const double *vals = b->dev.arena[hm.kc].vals + hm.off;
    const uint64_t *row = region_row(b->dev.regions, (uint32_t)map, (uint32_t)hm.kc);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(kSumBlocks, (len + 255) / 256);
    hipLaunchKernelGGL(k_sumsq, dim3(grid), dim3(256), 0, b->ctx->stream, vals, row, len,
                       b->dev.sumsq);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "HIP sum-of-squares launch", error, cap);
    double part[2 * kSumBlocks];
    e = hipMemcpyAsync(part, b->dev.sumsq, 2 * grid * sizeof(double), hipMemcpyDeviceToHost,
                       b->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "HIP sum-of-squares transfer", error, cap);
    double hi = 0.0, lo = 0.0;
    for (uint32_t k = 0; k < grid; ++k) dd_add(hi, lo, part[2 * k], part[2 * k + 1]);
    out[0] = hi;
    out[1] = lo;
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_batch_sync(tpgp_hip_batch *b, tpgp_hip_record *records,
                                   size_t max_records, size_t *n_records, double *readouts,
                                   size_t max_readouts, size_t *n_readouts, uint64_t *lengths,
                                   tpgp_hip_log_entry *log, size_t log_capacity, size_t *n_log,
                                   char *error, size_t cap) {
    msg(error, cap, "");
    if (!b) return TPGP_HIP_EINVAL;
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    hipStream_t st = b->ctx->stream;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "HIP resident execution", error, cap);
    collect_profile(b);
    const uint32_t M = b->M;
    uint32_t nrec = 0, nrd = 0;
    for (uint32_t m = 0; m < M; ++m) {
        nrec = std::max(nrec, b->maps[m].nrec);
        nrd = std::max(nrd, b->maps[m].nrd);
    }
    if (n_records) *n_records = nrec;
    if (n_readouts) *n_readouts = nrd;
    if (records && nrec && max_records < nrec) {
        msg(error, cap, "Record buffer too small (%u needed).", nrec);
        return TPGP_HIP_EINVAL;
    }
    std::vector<uint64_t> peak(M, 0);  // largest output of each map's rotations (overflowed too)
    if (nrec) {
        std::vector<tpgp_hip_record> tmp((size_t)M * b->max_rec);
        e = hipMemcpy(tmp.data(), b->dev.rec, tmp.size() * sizeof(tpgp_hip_record),
                      hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "HIP record transfer", error, cap);
        for (uint32_t m = 0; m < M; ++m) {
            if (records)
                for (uint32_t r = 0; r < b->maps[m].nrec; ++r)
                    records[(size_t)m * max_records + r] = tmp[(size_t)m * b->max_rec + r];
            const tpgp_hip_record *rm = tmp.data() + (size_t)m * b->max_rec;
            for (uint32_t r = 0; r < b->maps[m].nrec; ++r)
                peak[m] = std::max<uint64_t>(peak[m], rm[r].len_out);
            if (b->maps[m].nrec && rm[0].len_in)
                b->maps[m].growth =
                    std::min(4.0, std::max(1.0, (double)peak[m] / (double)rm[0].len_in));
            learn_append(b->maps[m], tmp.data() + (size_t)m * b->max_rec);
        }
    }
    if (readouts && nrd) {
        if (max_readouts < nrd) {
            msg(error, cap, "Readout buffer too small (%u needed).", nrd);
            return TPGP_HIP_EINVAL;
        }
        std::vector<double> tmp((size_t)M * b->max_rd * 64);
        e = hipMemcpy(tmp.data(), b->dev.rd, tmp.size() * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "HIP readout transfer", error, cap);
        for (uint32_t m = 0; m < M; ++m)
            for (uint32_t q = 0; q < b->maps[m].nrd; ++q)
                std::memcpy(readouts + ((size_t)m * max_readouts + q) * 64,
                            tmp.data() + ((size_t)m * b->max_rd + q) * 64, 64 * sizeof(double));
    }
    {
        std::vector<uint64_t> len((size_t)M * 3);
        e = hipMemcpy(len.data(), b->dev.len, len.size() * sizeof(uint64_t),
                      hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "HIP length transfer", error, cap);
        for (uint32_t m = 0; m < M; ++m) {
            const uint64_t l = len[(size_t)m * 3 + b->maps[m].kcur];
            // the step's peak counts too: a map emptied by an overflow is rolled back and its
            // step runs again at the size it reached
            b->maps[m].est_len = std::max<uint64_t>({l, b->maps[m].est_len / 2, peak[m]});
            if (lengths) lengths[m] = l;
        }
    }
    uint32_t internal = 0;
    e = hipMemcpy(&internal, b->dev.counter + 1, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "HIP status transfer", error, cap);
    if (b->validate) {
        uint64_t dbg[16];
        e = hipMemcpy(dbg, b->dev.debug, sizeof(dbg), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "HIP validation transfer", error, cap);
        if (dbg[0]) {
            msg(error, cap,
                "Validation failed after pass %llu (%s): map %llu index %lld len %llu class %llu "
                "region [%llu,%llu) key %#llx prev %#llx",
                (unsigned long long)(dbg[0] >> 1), (dbg[0] & 1) ? "rotation" : "partition",
                (unsigned long long)dbg[1], (long long)dbg[2], (unsigned long long)dbg[3],
                (unsigned long long)dbg[4], (unsigned long long)dbg[5],
                (unsigned long long)dbg[6], (unsigned long long)dbg[7],
                (unsigned long long)dbg[8]);
            return TPGP_HIP_EDEVICE;
        }
    }
    if (internal) {
        msg(error, cap, "Internal merge-path inconsistency detected on the device.");
        return TPGP_HIP_EDEVICE;
    }
    uint64_t nlog = 0;
    e = hipMemcpy(&nlog, b->dev.log_count, sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "HIP log-count transfer", error, cap);
    int status = TPGP_HIP_OK;
    if (n_log) *n_log = 0;
    if (nlog) {
        if (nlog > b->dev.log_cap || (log && nlog > log_capacity)) {
            // the caller rolls back, grows the logs (tpgp_hip_batch_reserve_log) and retries
            msg(error, cap, "Ledger log overflow (%llu entries).", (unsigned long long)nlog);
            if (n_log) *n_log = nlog;
            status = TPGP_HIP_ECAPACITY;
        } else if (log) {
            std::vector<LogDev> tmp(nlog);
            e = hipMemcpy(tmp.data(), b->dev.log, nlog * sizeof(LogDev), hipMemcpyDeviceToHost);
            if (e != hipSuccess) return hip_fail(e, "HIP log transfer", error, cap);
            std::unordered_map<uint64_t, const std::vector<uint8_t> *> layout;
            for (const LoggedLayout &ll : b->logged)
                layout[((uint64_t)ll.map << 32) | ll.rotation] = &ll.sigma;
            size_t out = 0;
            for (uint64_t t = 0; t < nlog; ++t) {
                const LogDev &L = tmp[t];
                auto it = layout.find(((uint64_t)L.map << 32) | L.rotation);
                if (it == layout.end()) continue;
                const std::vector<uint8_t> *sig = it->second;
                tpgp_hip_log_entry &o = log[out++];
                uint64_t stored[2] = {L.key[0], L.key[1]};
                uint64_t sw[2] = {stored[0], b->W > 1 ? stored[1] : 0};
                to_reference(*sig, sw, o.key, b->W);
                if (b->W == 1) o.key[1] = 0;
                o.value = L.value;
                o.aux = L.aux;
                o.map = L.map;
                o.rotation = L.rotation;
                o.kind = L.kind;
                o.pad = 0;
            }
            if (n_log) *n_log = out;
        }
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, b->dev.log_count, 1ull, 0ull);
    }
    b->logged.clear();
    for (uint32_t m = 0; m < M; ++m) {
        b->maps[m].nrec = b->maps[m].nrd = 0;
        b->maps[m].app_recs.clear();
    }
    release_stage(b);
    return status;
}

extern "C" int tpgp_hip_batch_commit(tpgp_hip_batch *b, const int *mask, char *error,
                                     size_t cap) {
    msg(error, cap, "");
    if (!b) return TPGP_HIP_EINVAL;
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    hipError_t e = hipStreamSynchronize(b->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "HIP commit", error, cap);
    std::vector<uint64_t> len((size_t)b->M * 3);
    e = hipMemcpy(len.data(), b->dev.len, len.size() * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "HIP commit", error, cap);
    for (uint32_t m = 0; m < b->M; ++m) {
        HostMap &hm = b->maps[m];
        if (!mask || mask[m]) {
            hm.kc = hm.kcur;
            hm.sig_c = hm.sig;
            hm.floor_c = hm.floor;
            hm.gap_c = hm.gap;
            hm.committed_len = len[(size_t)m * 3 + hm.kc];
        } else {
            hm.kcur = hm.kc;
            hm.sig = hm.sig_c;
            hm.floor = hm.floor_c;
            hm.gap = hm.gap_c;
        }
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, b->ctx->stream, b->dev.overflow,
                       (uint64_t)b->M, 0u);
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_batch_rollback(tpgp_hip_batch *b, char *error, size_t cap) {
    msg(error, cap, "");
    if (!b) return TPGP_HIP_EINVAL;
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (auto &hm : b->maps) {
        hm.kcur = hm.kc;
        hm.sig = hm.sig_c;
        hm.floor = hm.floor_c;
        hm.gap = hm.gap_c;
        hm.nrec = hm.nrd = 0;
        hm.app_recs.clear();
    }
    b->logged.clear();
    release_stage(b);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, b->ctx->stream, b->dev.overflow,
                       (uint64_t)b->M, 0u);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, b->ctx->stream, b->dev.log_count, 1ull,
                       0ull);
    hipError_t e = hipStreamSynchronize(b->ctx->stream);
    return e == hipSuccess ? TPGP_HIP_OK : hip_fail(e, "HIP rollback", error, cap);
}

extern "C" int tpgp_hip_batch_download(tpgp_hip_batch *b, size_t map, uint64_t *keys,
                                       double *values, size_t capacity, size_t *len, char *error,
                                       size_t cap) {
    msg(error, cap, "");
    if (!b || map >= b->M || !len) {
        msg(error, cap, "Invalid download arguments.");
        return TPGP_HIP_EINVAL;
    }
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    HostMap &hm = b->maps[map];
    *len = hm.committed_len;
    if (!keys && !values && capacity == 0) return TPGP_HIP_OK;
    if (capacity < hm.committed_len || (hm.committed_len && (!keys || !values))) {
        msg(error, cap, "Download buffers hold %zu terms, %llu needed.", capacity,
            (unsigned long long)hm.committed_len);
        return TPGP_HIP_EINVAL;
    }
    // Reorder a scratch copy of the committed state into the reference layout with a
    // least-significant-pair-first sequence of stable partitions: pairs (1,0), (3,2), ...,
    // finishing with (n-1, n-2); each pass moves its pair above all others.
    if (hm.kcur != hm.kc || hm.nrec || hm.nrd) {
        msg(error, cap, "Map %zu has uncommitted work; sync and commit or roll back first.", map);
        return TPGP_HIP_EINVAL;
    }
    {
        std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
        DeviceGuard guard(b->ctx->device);
        hipError_t e = hipStreamSynchronize(b->ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "HIP download", error, cap);
        const size_t n = hm.n_spins;
        std::vector<int> want(b->M, 0);
        want[map] = 1;
        std::vector<uint32_t> pa(b->M, 0), pb(b->M, 0);
        std::vector<std::pair<size_t, size_t>> pairs;
        for (size_t lo = 0; lo + 1 < n; lo += 2) pairs.push_back({lo + 1, lo});
        if (n % 2) pairs.push_back({n - 1, n - 2});
        bool identity = true;
        for (size_t p = 0; p < n; ++p) identity = identity && hm.sig[p] == p;
        // a committed state with classes out of place is laid out contiguously by one
        // partition that keeps the digit layout (the top pair moved to the top again)
        if (identity && hm.gap_c) pairs.assign(1, {n - 1, n - 2});
        if (!identity || hm.gap_c) {
            for (auto &pr : pairs) {
                pa[map] = (uint32_t)position_of(hm, pr.first);
                pb[map] = (uint32_t)position_of(hm, pr.second);
                if ((e = enqueue_partition(b, want, pa, pb)) != hipSuccess)
                    return hip_fail(e, "HIP download reorder", error, cap);
            }
        }
        const int k = hm.kcur;
        if (hm.committed_len) {
            e = hipMemcpyAsync(keys, b->dev.arena[k].keys + hm.off * b->W,
                               hm.committed_len * b->W * 8, hipMemcpyDeviceToHost,
                               b->ctx->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(values, b->dev.arena[k].vals + hm.off, hm.committed_len * 8,
                                   hipMemcpyDeviceToHost, b->ctx->stream);
            if (e != hipSuccess)
                return drain_fail(b->ctx->stream, e, "HIP download transfer", error, cap);
        }
        e = hipStreamSynchronize(b->ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "HIP download", error, cap);
        // the reordered copy becomes the committed state (same map, reference layout)
        hm.kc = hm.kcur;
        hm.sig_c = hm.sig;
        hm.gap_c = hm.gap;
        release_stage(b);
    }
    return TPGP_HIP_OK;
}

namespace {
// Re-lay the arenas with slot m holding new_cap[m] terms, keeping every map's committed state
// (moved to buffer 0 with its length, regions and digit layout).  The committed states wait in
// one temporary device buffer (device-to-device copies); only if that allocation fails do they
// take the way through host memory.  A capacity below a map's committed length is rejected
// before anything changes.
int relayout(tpgp_hip_batch *b, const std::vector<uint64_t> &new_cap, char *error, size_t cap) {
    hipStream_t st = b->ctx->stream;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "HIP reserve", error, cap);
    const uint32_t M = b->M;
    const int W = b->W;
    std::vector<uint64_t> len((size_t)M * 3), reg((size_t)M * 3 * kRegStride);
    if ((e = hipMemcpy(len.data(), b->dev.len, len.size() * 8, hipMemcpyDeviceToHost)) !=
            hipSuccess ||
        (e = hipMemcpy(reg.data(), b->dev.regions, reg.size() * 8, hipMemcpyDeviceToHost)) !=
            hipSuccess)
        return hip_fail(e, "HIP reserve", error, cap);
    std::vector<uint64_t> n(M), at(M);
    uint64_t total = 0;
    for (uint32_t m = 0; m < M; ++m) {
        n[m] = len[(size_t)m * 3 + b->maps[m].kc];
        at[m] = total;
        total += n[m];
        if (new_cap[m] < n[m] || !new_cap[m]) {
            msg(error, cap, "Map %u: capacity %llu cannot hold its %llu committed terms.", m,
                (unsigned long long)new_cap[m], (unsigned long long)n[m]);
            return TPGP_HIP_EINVAL;
        }
    }
    uint64_t *tk = nullptr;
    double *tv = nullptr;
    bool on_device = hipMalloc((void **)&tk, (total * W + 1) * 8) == hipSuccess &&
                     hipMalloc((void **)&tv, (total + 1) * 8) == hipSuccess;
    std::vector<uint64_t> hk;
    std::vector<double> hv;
    if (!on_device) {
        (void)hipGetLastError();
        if (tk) (void)hipFree(tk);
        tk = nullptr;
        hk.resize(total * W);
        hv.resize(total);
    }
    for (uint32_t m = 0; m < M && e == hipSuccess; ++m) {
        const HostMap &hm = b->maps[m];
        if (!n[m]) continue;
        uint64_t *dk = on_device ? tk + at[m] * W : hk.data() + at[m] * W;
        double *dv = on_device ? tv + at[m] : hv.data() + at[m];
        const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        // the committed map in key order: its classes, runs of adjacent ones copied together
        const uint64_t *row = region_row(reg.data(), m, (uint32_t)hm.kc);
        for (int c = 0; c < 16 && e == hipSuccess;) {
            const uint64_t lo = row[c], src = row[kRegPhys + c];
            int d = c + 1;  // extend while the next class follows in the slot too
            while (d < 16 && row[kRegPhys + d] == src + (row[d] - lo)) ++d;
            const uint64_t cnt = row[d] - lo;
            c = d;
            if (!cnt) continue;
            e = hipMemcpyAsync(dk + lo * W, b->dev.arena[hm.kc].keys + (hm.off + src) * W,
                               cnt * W * 8, kind, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dv + lo, b->dev.arena[hm.kc].vals + hm.off + src, cnt * 8,
                                   kind, st);
        }
    }
    // wait even after a failed enqueue: copies already queued still write into hk / hv
    const hipError_t drained = hipStreamSynchronize(st);
    if (e == hipSuccess) e = drained;
    if (e != hipSuccess) {
        if (tk) (void)hipFree(tk);
        if (tv) (void)hipFree(tv);
        return hip_fail(e, "HIP reserve", error, cap);
    }
    free_all(b);
    std::vector<uint64_t> old_cap(M);
    auto lay_out = [&](bool grown) {
        uint64_t off = 0;
        for (uint32_t m = 0; m < M; ++m) {
            HostMap &hm = b->maps[m];
            hm.cap = grown ? new_cap[m] : old_cap[m];
            hm.off = off;
            off += (hm.cap + 63) & ~63ull;
        }
        b->total_cap = off;
    };
    for (uint32_t m = 0; m < M; ++m) old_cap[m] = b->maps[m].cap;
    lay_out(true);
    int s = allocate_device(b, error, cap);
    int result = TPGP_HIP_OK;
    if (s == TPGP_HIP_ENOMEM) {
        // the new layout does not fit: go back to the previous one (it fitted a moment ago,
        // and its memory has just been released), keep every committed state, report ENOMEM
        char again[256];
        free_all(b);
        lay_out(false);
        result = s;
        s = allocate_device(b, again, sizeof(again));
    }
    if (s != TPGP_HIP_OK) {
        if (tk) (void)hipFree(tk);
        if (tv) (void)hipFree(tv);
        b->poisoned = true;
        return s;
    }
    std::vector<uint64_t> nlen((size_t)M * 3, 0), nreg((size_t)M * 3 * kRegStride, 0);
    for (uint32_t m = 0; m < M && e == hipSuccess; ++m) {
        HostMap &hm = b->maps[m];
        nlen[(size_t)m * 3] = n[m];
        const uint64_t *row = region_row(reg.data(), m, (uint32_t)hm.kc);
        uint64_t *nrow = region_row(nreg.data(), m, 0);  // buffer 0, contiguous
        for (int c = 0; c < 17; ++c) nrow[c] = row[c];
        for (int c = 0; c < 16; ++c) nrow[kRegPhys + c] = row[c];
        nrow[kRegTop] = row[16];
        if (n[m]) {
            const uint64_t *sk = on_device ? tk + at[m] * W : hk.data() + at[m] * W;
            const double *sv = on_device ? tv + at[m] : hv.data() + at[m];
            const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
            e = hipMemcpyAsync(b->dev.arena[0].keys + hm.off * W, sk, n[m] * W * 8, kind, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(b->dev.arena[0].vals + hm.off, sv, n[m] * 8, kind, st);
        }
        hm.kc = hm.kcur = 0;
        hm.sig = hm.sig_c;
        hm.floor = hm.floor_c;
        hm.gap = hm.gap_c = false;
        hm.nrec = hm.nrd = 0;
        hm.app_recs.clear();
        hm.est_len = n[m];
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->dev.len, nlen.data(), nlen.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->dev.regions, nreg.data(), nreg.size() * 8, hipMemcpyHostToDevice,
                           st);
    // wait even after a failed enqueue: queued copies still read tk / tv and hk / hv
    const hipError_t done = hipStreamSynchronize(st);
    if (e == hipSuccess) e = done;
    if (tk) (void)hipFree(tk);
    if (tv) (void)hipFree(tv);
    b->logged.clear();
    if (e != hipSuccess) return hip_fail(e, "HIP reserve", error, cap);
    return result;  // ENOMEM: the slots kept their previous capacities (error says why)
}
}  // namespace

extern "C" int tpgp_hip_batch_reserve_many(tpgp_hip_batch *b, const size_t *capacities,
                                           char *error, size_t cap) {
    msg(error, cap, "");
    if (!b || !capacities) {
        msg(error, cap, "Invalid reserve arguments.");
        return TPGP_HIP_EINVAL;
    }
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    bool grow = false;
    std::vector<uint64_t> want(b->M);
    for (uint32_t m = 0; m < b->M; ++m) {
        if (capacities[m] > (1ull << 37)) {
            msg(error, cap, "Invalid reserve capacity for map %u.", m);
            return TPGP_HIP_EINVAL;
        }
        want[m] = std::max<uint64_t>(capacities[m], b->maps[m].cap);
        grow |= capacities[m] > b->maps[m].cap;
    }
    if (!grow) return TPGP_HIP_OK;
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    return relayout(b, want, error, cap);
}

extern "C" int tpgp_hip_batch_resize(tpgp_hip_batch *b, const size_t *capacities, char *error,
                                     size_t cap) {
    msg(error, cap, "");
    if (!b || !capacities) {
        msg(error, cap, "Invalid resize arguments.");
        return TPGP_HIP_EINVAL;
    }
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    bool change = false;
    std::vector<uint64_t> want(b->M);
    for (uint32_t m = 0; m < b->M; ++m) {
        if (capacities[m] > (1ull << 37)) {
            msg(error, cap, "Invalid resize capacity for map %u.", m);
            return TPGP_HIP_EINVAL;
        }
        want[m] = capacities[m] ? capacities[m] : b->maps[m].cap;
        change |= want[m] != b->maps[m].cap;
    }
    if (!change) return TPGP_HIP_OK;
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    return relayout(b, want, error, cap);
}

extern "C" int tpgp_hip_batch_reserve(tpgp_hip_batch *b, size_t map, size_t capacity,
                                      char *error, size_t cap) {
    msg(error, cap, "");
    if (!b || map >= b->M || capacity > (1ull << 37)) {
        msg(error, cap, "Invalid reserve arguments.");
        return TPGP_HIP_EINVAL;
    }
    if (capacity <= b->maps[map].cap) return TPGP_HIP_OK;
    std::vector<size_t> caps(b->M, 0);
    caps[map] = capacity;
    return tpgp_hip_batch_reserve_many(b, caps.data(), error, cap);
}

extern "C" int tpgp_hip_batch_reserve_log(tpgp_hip_batch *b, size_t entries, char *error,
                                          size_t cap) {
    msg(error, cap, "");
    if (!b || entries > (1ull << 34)) {
        msg(error, cap, "Invalid log reserve arguments.");
        return TPGP_HIP_EINVAL;
    }
    if (recording_step(b, error, cap)) return TPGP_HIP_EINVAL;
    if (entries <= b->dev.log_cap) return TPGP_HIP_OK;
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    hipError_t e = hipStreamSynchronize(b->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "HIP log reserve", error, cap);
    LogDev *fresh = nullptr;
    if ((e = dalloc(b, &fresh, entries)) != hipSuccess) {
        hip_fail(e, "HIP log reserve", error, cap);
        return e == hipErrorOutOfMemory ? TPGP_HIP_ENOMEM : TPGP_HIP_EDEVICE;
    }
    // the old log is released with the batch's other allocations
    auto it = std::find(b->allocations.begin(), b->allocations.end(), (void *)b->dev.log);
    if (it != b->allocations.end()) {
        (void)hipFree(*it);
        b->allocations.erase(it);
    }
    b->dev.log = fresh;
    b->dev.log_cap = entries;
    b->log_entries = entries;
    drop_graphs(b);  // recorded steps write to the old log
    return TPGP_HIP_OK;
}

namespace {
// The host state of every map that a step's planning reads and advances.
std::vector<uint8_t> host_state(const tpgp_hip_batch *b) {
    std::vector<uint8_t> out;
    for (const HostMap &hm : b->maps) {
        const uint32_t f[6] = {(uint32_t)hm.n_spins, (uint32_t)hm.kc, (uint32_t)hm.kcur, hm.nrec,
                               hm.nrd, (hm.gap ? 1u : 0u) | (hm.app_level << 1)};
        // the planned extent matters only while the working buffer is not the committed one
        // (appending starts after a partition or a copying rotation resets it)
        const uint64_t top = hm.kcur != hm.kc ? hm.est_top : 0;
        uint64_t fl;
        std::memcpy(&fl, &hm.floor, 8);
        const uint8_t *p = reinterpret_cast<const uint8_t *>(f);
        out.insert(out.end(), p, p + sizeof(f));
        p = reinterpret_cast<const uint8_t *>(&fl);
        out.insert(out.end(), p, p + 8);
        p = reinterpret_cast<const uint8_t *>(&top);
        out.insert(out.end(), p, p + 8);
        out.insert(out.end(), hm.sig.begin(), hm.sig.end());
    }
    return out;
}
void set_host_state(tpgp_hip_batch *b, const std::vector<uint8_t> &st) {
    size_t at = 0;
    for (HostMap &hm : b->maps) {
        uint32_t f[6];
        std::memcpy(f, st.data() + at, sizeof(f));
        at += sizeof(f);
        std::memcpy(&hm.floor, st.data() + at, 8);
        at += 8;
        std::memcpy(&hm.est_top, st.data() + at, 8);
        at += 8;
        hm.kcur = (int)f[2];
        hm.nrec = f[3];
        hm.nrd = f[4];
        hm.gap = (f[5] & 1u) != 0;
        hm.app_level = f[5] >> 1;
        std::memcpy(hm.sig.data(), st.data() + at, hm.n_spins);
        at += hm.n_spins;
    }
}
uint64_t fingerprint(uint64_t key, const std::vector<uint8_t> &st) {
    uint64_t h = 1469598103934665603ull ^ key;  // FNV-1a
    for (uint8_t c : st) h = (h ^ c) * 1099511628211ull;
    return h;
}
bool graphs_usable(const tpgp_hip_batch *b) {
    return b->graphs_on && !b->validate && !b->profiling && !b->poisoned;
}
constexpr size_t kMaxGraphs = 6;
}  // namespace

extern "C" int tpgp_hip_batch_graph_replay(tpgp_hip_batch *b, uint64_t key, int *replayed,
                                           int *capture, char *error, size_t cap) {
    msg(error, cap, "");
    if (!b || !replayed || b->capturing) {
        msg(error, cap, "Invalid graph replay arguments.");
        return TPGP_HIP_EINVAL;
    }
    *replayed = 0;
    if (capture) *capture = 0;
    if (!graphs_usable(b)) return TPGP_HIP_OK;
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    const std::vector<uint8_t> st = host_state(b);
    for (size_t k = 0; k < b->graphs.size(); ++k) {
        StepGraph &g = b->graphs[k];
        if (g.key != key || g.gen != b->gen || g.before != st) continue;
        // launch shapes were planned for the lengths at capture: record again once a map has
        // grown or shrunk well past them
        bool stale = false;
        for (uint32_t m = 0; m < b->M; ++m)
            stale |= 2 * b->maps[m].est_len > 3 * g.est[m] + 8192 ||
                     3 * b->maps[m].est_len + 8192 < 2 * g.est[m];
        if (stale) {
            drop_graph(g);
            b->graphs.erase(b->graphs.begin() + (long)k);
            if (capture) *capture = 1;
            return TPGP_HIP_OK;
        }
        hipError_t e = hipGraphLaunch(g.exec, b->ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "HIP graph launch", error, cap);
        set_host_state(b, g.after);
        b->logged.insert(b->logged.end(), g.logged.begin(), g.logged.end());
        g.last_use = ++b->use_clock;
        *replayed = 1;
        return TPGP_HIP_OK;
    }
    // record a state the step has started from before: it is likely to recur again
    const uint64_t fp = fingerprint(key, st);
    const bool recurs = std::find(b->seen.begin(), b->seen.end(), fp) != b->seen.end();
    if (!recurs) {
        b->seen.push_back(fp);
        if (b->seen.size() > 16) b->seen.erase(b->seen.begin());
    }
    if (capture) *capture = recurs;
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_batch_graph_begin(tpgp_hip_batch *b, uint64_t key, char *error,
                                          size_t cap) {
    msg(error, cap, "");
    if (!b || b->capturing || !graphs_usable(b)) {
        msg(error, cap, "Invalid graph capture arguments.");
        return TPGP_HIP_EINVAL;
    }
    b->ctx->mutex.lock();  // held until tpgp_hip_batch_graph_end: nothing else joins the graph
    DeviceGuard guard(b->ctx->device);
    if (!b->graph_desc_cap) b->graph_desc_cap = b->stage_cap;
    StepGraph &g = b->recording;
    g = StepGraph{};
    hipError_t e = hipMalloc((void **)&g.desc, b->graph_desc_cap);
    if (e == hipSuccess) e = hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        if (g.desc) (void)hipFree(g.desc);
        g.desc = nullptr;
        b->ctx->mutex.unlock();
        return hip_fail(e, "HIP graph capture", error, cap);
    }
    g.key = key;
    g.gen = b->gen;
    g.before = host_state(b);
    for (const HostMap &hm : b->maps) g.est.push_back(hm.est_len);
    b->cap_host.assign(b->graph_desc_cap, 0);
    b->cap_used = 0;
    b->spill.clear();
    b->capture_spilled = false;
    b->logged_mark = b->logged.size();
    b->capturing = true;
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_batch_graph_end(tpgp_hip_batch *b, int keep, int *captured,
                                        char *error, size_t cap) {
    msg(error, cap, "");
    if (!b || !captured || !b->capturing) {
        msg(error, cap, "Invalid graph end arguments.");
        return TPGP_HIP_EINVAL;
    }
    *captured = 0;
    DeviceGuard guard(b->ctx->device);
    StepGraph &g = b->recording;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(b->ctx->stream, &graph);
    b->capturing = false;
    bool ok = keep && e == hipSuccess && graph && !b->capture_spilled;
    if (ok) ok = (e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0)) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (ok)
        ok = (e = hipMemcpy(g.desc, b->cap_host.data(), b->cap_used, hipMemcpyHostToDevice)) ==
             hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        // nothing of the step ran: return the maps to the state it was planned from
        set_host_state(b, g.before);
        b->logged.resize(b->logged_mark);
        if (b->capture_spilled) b->graph_desc_cap *= 2;  // room for the next attempt
        drop_graph(g);
        b->spill.clear();
        b->ctx->mutex.unlock();
        return TPGP_HIP_OK;  // *captured = 0: enqueue the step again, without capture
    }
    g.after = host_state(b);
    g.logged.assign(b->logged.begin() + (long)b->logged_mark, b->logged.end());
    g.last_use = ++b->use_clock;
    if (b->graphs.size() >= kMaxGraphs) {  // the least recently used goes
        size_t lru = 0;
        for (size_t k = 1; k < b->graphs.size(); ++k)
            if (b->graphs[k].last_use < b->graphs[lru].last_use) lru = k;
        drop_graph(b->graphs[lru]);
        b->graphs.erase(b->graphs.begin() + (long)lru);
    }
    b->graphs.push_back(g);
    g = StepGraph{};  // the batch's list owns the graph now
    e = hipGraphLaunch(b->graphs.back().exec, b->ctx->stream);
    b->cap_host.clear();
    b->cap_host.shrink_to_fit();
    b->ctx->mutex.unlock();
    if (e != hipSuccess) return hip_fail(e, "HIP graph launch", error, cap);
    *captured = 1;
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_batch_profile(tpgp_hip_batch *b, unsigned mask) {
    if (!b) return TPGP_HIP_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(b->ctx->mutex);
    DeviceGuard guard(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    collect_profile(b);
    b->profiling = mask & TPGP_HIP_PROFILE_ALL;
    for (int k = 0; k < 10; ++k) {
        b->prof_ms[k] = 0;
        b->prof_count[k] = 0;
    }
    return TPGP_HIP_OK;
}

extern "C" int tpgp_hip_batch_profile_read(const tpgp_hip_batch *b, unsigned kind,
                                           double *milliseconds, uint64_t *launches) {
    if (!b || kind > 9) return TPGP_HIP_EINVAL;
    if (milliseconds) *milliseconds = b->prof_ms[kind];
    if (launches) *launches = b->prof_count[kind];
    return TPGP_HIP_OK;
}
