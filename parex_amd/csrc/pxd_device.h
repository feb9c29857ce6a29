// Device-side building blocks of the expander-decomposition engine for CDNA4
// (gfx950, 64-lane wavefronts): exact fixed-point accumulation, ordered-float
// keys and wave/block scans.  Included by pxd_engine.hip only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pxd {

constexpr int WAVE = 64;
constexpr uint32_t INVALID = 0xFFFFFFFFu;
constexpr double RW_STAY = 0.1;          // lib/core/definitions.h:48
constexpr int FIX_SHIFT = 72;            // walk sums: signed fixed point, scale 2^72

// error bits reported through the engine's status word
constexpr uint32_t ERR_FIX_RANGE = 1u;   // |x/dInt| >= 2^31 or row longer than 2^24 arcs
constexpr uint32_t ERR_ZERO_DEG = 4u;    // valid arc into a node with dInt == 0

// ------------------------------------------------------------------ keys

// Ordered image of a float (cuda/devRandomWalk.h:50-61, low half of packKey):
// unsigned order equals float order with -0 < +0.
__device__ __forceinline__ uint32_t ord_f32(float v) {
    uint32_t i = __float_as_uint(v);
    return i ^ ((i & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u);
}
__device__ __forceinline__ float unord_f32(uint32_t o) {
    uint32_t i = (o & 0x80000000u) ? (o ^ 0x80000000u) : ~o;
    return __uint_as_float(i);
}

// ------------------------------------------------------------------ fixed point

// Signed 128-bit accumulator as two 64-bit words (two's complement).
struct Fix128 {
    uint64_t lo;
    uint64_t hi;
};

__device__ __forceinline__ void fix_add128(Fix128& a, uint64_t blo, uint64_t bhi) {
    uint64_t lo = a.lo + blo;
    a.hi = a.hi + bhi + (lo < a.lo ? 1ull : 0ull);
    a.lo = lo;
}

// Walk terms.  The term of y is trunc(y * 2^72) (truncation toward zero).
// With |y| * 2^72 = M * 2^sh (53-bit significand M) it is M >> -sh for sh < 0
// (zero and subnormal y shift out entirely) and M << sh for sh >= 0, signed.
// It is computed once by the kernel that writes y and stored in one 64-bit
// word: q = (t << 6) | l with the term equal to t * 2^l, t signed, |t| < 2^53,
// 0 <= l <= 50.  A zero word adds nothing.  |y| >= 2^31 (sh > 50, also inf
// and nan) sets `ovf`: the engine then reports an error instead of a sum.
// Readers (fix_add_packed) only shift and add.
__device__ __forceinline__ uint64_t fix_pack(double y, uint32_t& ovf) {
    const uint64_t b = (uint64_t)__double_as_longlong(y);
    const int E = (int)((b >> 52) & 0x7FFu);
    const uint64_t M = (b & ((1ull << 52) - 1)) | (1ull << 52);
    const int sh = E - (1075 - FIX_SHIFT);
    ovf |= (uint32_t)(sh > 50);
    const uint32_t l = (uint32_t)min(max(sh, 0), 50);
    const uint32_t r = (uint32_t)min(max(-sh, 0), 63);
    const int64_t m = (int64_t)b >> 63;                            // 0 or -1
    const int64_t t = (int64_t)(((M >> r) ^ (uint64_t)m) - (uint64_t)m);
    return ((uint64_t)t << 6) | l;
}

// Adds a packed term (fix_pack) to the accumulator: t << l across the two
// words, exact.
__device__ __forceinline__ void fix_add_packed(Fix128& a, uint64_t q) {
    const uint32_t l = (uint32_t)q & 63u;
    const int64_t t = (int64_t)q >> 6;
    const uint64_t qlo = (uint64_t)t << l;
    const uint64_t qhi = (uint64_t)((t >> 1) >> (63 - l));         // t >> (64 - l), sign-filled
    fix_add128(a, qlo, qhi);
}

// Split accumulator for runs of packed terms: value = hi * 2^FIX_LIMB + lo
// with lo unsigned.  A term t * 2^l adds its low FIX_LIMB bits to lo and the
// rest (t >> (FIX_LIMB - l), l <= 50 < FIX_LIMB) to hi, so no carry moves per
// term; fixacc_norm carries lo's bits above FIX_LIMB into hi and must run at
// least every 2^(64 - FIX_LIMB) = 4096 terms.  Exact, like Fix128, with fewer
// instructions per term; fixacc_fix converts the result.
constexpr int FIX_LIMB = 52;
struct FixAcc {
    uint64_t lo;
    int64_t hi;
};
__device__ __forceinline__ void fixacc_add(FixAcc& a, uint64_t q) {
    const int64_t t = (int64_t)q >> 6;
    // shift counts mod 64 (the hardware's own masking): l = q mod 64, and
    // (FIX_LIMB - q) mod 64 = FIX_LIMB - l
    const uint32_t sl = (uint32_t)q & 63u, sh = ((uint32_t)FIX_LIMB - (uint32_t)q) & 63u;
    a.lo += ((uint64_t)t << sl) & ((1ull << FIX_LIMB) - 1ull);
    a.hi += t >> sh;
}
__device__ __forceinline__ void fixacc_norm(FixAcc& a) {
    a.hi += (int64_t)(a.lo >> FIX_LIMB);
    a.lo &= (1ull << FIX_LIMB) - 1ull;
}
__device__ __forceinline__ Fix128 fixacc_fix(FixAcc a) {
    Fix128 r = {(uint64_t)a.hi << FIX_LIMB, (uint64_t)(a.hi >> (64 - FIX_LIMB))};
    fix_add128(r, a.lo, 0ull);
    return r;
}

// Correctly rounded conversion of a value below 2^63 (exact halves, one rounding).
__device__ __forceinline__ double u63_to_f64(uint64_t v) {
    return ldexp((double)(uint32_t)(v >> 32), 32) + (double)(uint32_t)v;
}

// Accumulator -> double, one correct rounding, then exact scaling by 2^-72.
__device__ __forceinline__ double fix_to_f64(Fix128 a) {
    bool neg = (int64_t)a.hi < 0;
    uint64_t lo = a.lo, hi = a.hi;
    if (neg) {
        lo = ~lo + 1ull;
        hi = ~hi + (lo == 0ull ? 1ull : 0ull);
    }
    double d;
    if (hi == 0ull && (lo >> 63) == 0ull) {
        d = u63_to_f64(lo);
    } else {
        int nb = hi ? (128 - __clzll((long long)hi)) : (64 - __clzll((long long)lo));
        int k = nb - 63;                                   // 1..65
        uint64_t top, rest;
        if (k < 64) {
            top = (lo >> k) | (hi << (64 - k));
            rest = lo & ((1ull << k) - 1ull);
        } else {
            top = hi >> (k - 64);
            rest = lo | (k > 64 ? (hi & ((1ull << (k - 64)) - 1ull)) : 0ull);
        }
        if (rest) top |= 1ull;                            // sticky bit
        d = ldexp(u63_to_f64(top), k);
    }
    if (neg) d = -d;
    return ldexp(d, -FIX_SHIFT);
}

// int64 -> double with one correct rounding
__device__ __forceinline__ double i64_to_f64(int64_t v) {
    return ldexp((double)(int32_t)(v >> 32), 32) + (double)(uint32_t)(uint64_t)v;
}

// ------------------------------------------------------------------ wave helpers

template <int G>
__device__ __forceinline__ Fix128 group_sum_fix(Fix128 a) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        uint64_t olo = __shfl_xor(a.lo, off, G);
        uint64_t ohi = __shfl_xor(a.hi, off, G);
        fix_add128(a, olo, ohi);
    }
    return a;
}

// FixAcc over a group of G lanes, normalized inputs (lo < 2^FIX_LIMB: the
// lo sum stays below 2^(FIX_LIMB + 6) for G <= 64)
template <int G>
__device__ __forceinline__ FixAcc group_sum_fixacc(FixAcc a) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        a.lo += (uint64_t)__shfl_xor(a.lo, off, G);
        a.hi += (int64_t)__shfl_xor(a.hi, off, G);
    }
    return a;
}

template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

// ballot restricted to the calling lane's group of G lanes (bit i = lane i of
// the group), built from group-width exchanges so that groups of one wave may
// run different trip counts
template <int G>
__device__ __forceinline__ uint64_t group_ballot(bool pred) {
    uint64_t b = pred ? (1ull << (threadIdx.x & (G - 1))) : 0ull;
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) b |= __shfl_xor(b, off, G);
    return b;
}

template <int G>
__device__ __forceinline__ uint32_t group_or(uint32_t v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v |= __shfl_xor(v, off, G);
    return v;
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (WAVE - 1); }

// The value stays in its VGPRs as computed: the optimizer cannot fold the
// load that produced it into a later select (no instruction is emitted).
__device__ __forceinline__ void opaque_u64(uint64_t& v) { __asm__ volatile("" : "+v"(v)); }
__device__ __forceinline__ void opaque_u32(uint32_t& v) { __asm__ volatile("" : "+v"(v)); }
__device__ __forceinline__ void opaque(uint64_t& v) { opaque_u64(v); }
__device__ __forceinline__ void opaque(uint32_t& v) { opaque_u32(v); }

// U gathers g[idx[k]] through an LDS window sw[0, win) that holds g[wb, wb +
// win) when `use`.  Every slot reads its LDS word first (word 0 when the
// index misses the window); the results are pinned, then the valid misses
// load from global memory.  A miss's global load then writes a register no
// LDS read is still writing, so the misses of all U slots are issued back to
// back and waited for once (a per-slot choice between the two loads makes the
// compiler wait out each miss before the next slot's LDS read).  ZERO: an
// invalid slot gives 0, otherwise whatever its LDS read returned.
template <int U, bool ZERO, class T>
__device__ __forceinline__ void win_gather(T (&val)[U], const uint32_t (&idx)[U], const bool (&valid)[U],
                                           const T* sw, const T* __restrict__ g, uint32_t wb, uint32_t win,
                                           bool use) {
    bool miss[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const uint32_t off = idx[k] - wb;
        const bool inw = use && off < win;
        const T t = sw[inw ? off : 0u];
        val[k] = (ZERO && !valid[k]) ? T(0) : t;
        miss[k] = valid[k] && !inw;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) opaque(val[k]);
#pragma unroll
    for (int k = 0; k < U; ++k)
        if (miss[k]) val[k] = g[idx[k]];
}

// XCD-aware split of a grid-stride loop over n items.  Workgroups are
// dispatched to the 8 XCDs round-robin (block b runs on XCD b % 8, each XCD
// with its own L2), so XCD x takes the contiguous x-th eighth of the items and
// its blocks stride through that eighth: neighbouring items, whose gathers
// touch the same lines, share one L2.  The grid must be a multiple of 8; every
// item is visited exactly once.  `per` = items a block covers per stride,
// `local` = this thread's item within the block's share.
constexpr uint32_t NXCD = 8;
struct XcdSpan {
    uint32_t first, end, stride;
};
__device__ __forceinline__ XcdSpan xcd_span(uint32_t n, uint32_t per, uint32_t local) {
    const uint32_t x = blockIdx.x % NXCD, lb = blockIdx.x / NXCD, nb = gridDim.x / NXCD;
    const uint32_t chunk = n / NXCD + (n % NXCD != 0);
    const uint32_t b = min(n, x * chunk), e = min(n, b + chunk);
    return XcdSpan{b + lb * per + local, e, nb * per};
}

// The tiles of a tile loop over n items (tiles of `tile` consecutive items),
// split the same way: XCD x strides over the x-th eighth of the tiles.  Loops
// over positions use it: a cluster is a contiguous position range, so its
// gathers and scattered writes to node arrays (the cluster's nodes) stay in
// one XCD's L2 instead of touching the same lines from all eight.  The grid
// must be a multiple of 8.
__device__ __forceinline__ XcdSpan xcd_tiles(uint32_t n, uint32_t tile) {
    return xcd_span(n / tile + (n % tile != 0), 1, 0);
}

// blockIdx.x renumbered so that the blocks of each XCD (b % 8 == x) take one
// contiguous range of ids; a bijection of 0 .. gridDim.x - 1 for any grid.
// For kernels that give each block one tile by its id.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q = nb / NXCD, r = nb % NXCD, x = b % NXCD;
    return x * q + min(x, r) + b / NXCD;
}

// Device-coherent (agent scope) relaxed accesses for words that other threads
// of the same launch may write.  Each XCD has its own L2; plain accesses to
// such words are data races, these compile to sc1 accesses that every XCD
// sees alike.
__device__ __forceinline__ uint32_t ld_dev(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int G>
__device__ __forceinline__ uint32_t group_min(uint32_t v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor(v, off, G));
    return v;
}

// Append slot for every lane with pred set: one atomic per wave, slots in lane
// order.  All 64 lanes of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    const uint64_t b = __ballot(pred);
    uint32_t base = 0;
    if (b != 0ull && lane_id() == 0) base = atomicAdd(counter, (uint32_t)__popcll(b));
    base = __shfl(base, 0);
    return base + (uint32_t)__popcll(b & ((1ull << lane_id()) - 1ull));
}

// Inclusive scan over the wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        uint32_t t = __shfl_up(v, off);
        if (lane >= (uint32_t)off) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; `lds` holds >= BS/64 words.
template <int BS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    constexpr int NW = BS / WAVE;
    uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
    uint32_t inc = wave_incl_scan(v);
    if (lane == WAVE - 1) lds[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        uint32_t w = lane < (uint32_t)NW ? lds[lane] : 0u;
        w = wave_incl_scan(w);
        if (lane < (uint32_t)NW) lds[lane] = w;
    }
    __syncthreads();
    uint32_t before = wid ? lds[wid - 1] : 0u;
    total = lds[NW - 1];
    __syncthreads();
    return before + inc - v;
}

}  // namespace pxd
