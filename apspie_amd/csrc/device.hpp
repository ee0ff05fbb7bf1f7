// device.hpp -- wave64 building blocks shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace apspie {
namespace dev {

constexpr int WAVE = 64;

__device__ __forceinline__ uint64_t ld_acquire_agent(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t x, int lane) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint64_t y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_incl_scan32(uint32_t x, int lane) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
    }
    return x;
}

// Inclusive max-scan of non-negative ints across a full wave, on the VALU:
// Hillis-Steele steps inside each 16-lane row (DPP row_shr 1, 2, 4, 8; lanes
// shifted in from before the row read 0), then lane 15 of rows 0 and 2 into
// rows 1 and 3 (row_bcast:15), then lane 31 into rows 2 and 3
// (row_bcast:31).  Six DPP moves instead of six LDS-routed shuffles, each of
// which waits a full LDS round trip.  All 64 lanes must be active.
__device__ __forceinline__ int wave_incl_max_nonneg(int x) {
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true));  // row_shr:1
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true));  // row_shr:2
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true));  // row_shr:4
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true));  // row_shr:8
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return x;
}

// The same scan for sums of uint32 (0 is the identity of both).
__device__ __forceinline__ uint32_t wave_incl_sum_dpp(uint32_t x) {
    int v = int(x);
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
    return uint32_t(v);
}

// Lane 63's value, as a wave-uniform scalar.
__device__ __forceinline__ int wave_last(int x) { return __builtin_amdgcn_readlane(x, WAVE - 1); }

// Wave totals through the DPP scans (all 64 lanes active); the result is
// wave-uniform.
__device__ __forceinline__ uint32_t wave_total32(uint32_t x) { return uint32_t(wave_last(int(wave_incl_sum_dpp(x)))); }

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
    // lanes shifted in from outside a row, and rows outside the mask, read
    // the identity ~0u (bound_ctrl off keeps the `old` operand)
    x = min(x, uint32_t(__builtin_amdgcn_update_dpp(-1, int(x), 0x111, 0xf, 0xf, false)));
    x = min(x, uint32_t(__builtin_amdgcn_update_dpp(-1, int(x), 0x112, 0xf, 0xf, false)));
    x = min(x, uint32_t(__builtin_amdgcn_update_dpp(-1, int(x), 0x114, 0xf, 0xf, false)));
    x = min(x, uint32_t(__builtin_amdgcn_update_dpp(-1, int(x), 0x118, 0xf, 0xf, false)));
    x = min(x, uint32_t(__builtin_amdgcn_update_dpp(-1, int(x), 0x142, 0xa, 0xf, false)));
    x = min(x, uint32_t(__builtin_amdgcn_update_dpp(-1, int(x), 0x143, 0xc, 0xf, false)));
    return uint32_t(wave_last(int(x)));
}

// Orders this thread's returning atomics, whose results were folded into
// dep, before every later memory operation of the thread: the value of a
// returning atomic is back only once memory has performed it, and the empty
// asm both needs dep (so the wave waits for those values) and is a compiler
// barrier.  The "last workgroup finishes" pattern needs no more when the last
// workgroup reads only such counters; a device-scope fence would also do, but
// on gfx950 it writes back and invalidates the whole L2.
__device__ __forceinline__ void after_values(uint32_t dep) { asm volatile("" ::"v"(dep) : "memory"); }

// Sum across the wave; every lane receives the total.
// Makes a wave's earlier LDS writes visible to its lanes' later LDS reads:
// a wavefront-scope release/acquire fence pair around a code-motion barrier,
// so the compiler keeps a read of another lane's slot after the write.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t x) {
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, WAVE);
    return x;
}

__device__ __forceinline__ uint32_t wave_sum32(uint32_t x) {
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, WAVE);
    return x;
}

// Decoupled look-back over per-workgroup status words (Merrill & Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016), done by
// one full wave: each step inspects LB_DEPTH * 64 predecessors (lane l, slot
// j looks at bid - 1 - (j * 64 + l)) and votes with ballots, so the
// inclusive prefix front is found in few memory round trips.
// Status word: bits 62-63 flag (0 none, 1 aggregate, 2 inclusive prefix),
// bits 0-61 the value.  Values are sums that must stay below 2^62.
// Called by all 64 lanes of one wave; returns the exclusive prefix of `bid`.
constexpr int LB_DEPTH = 4;

__device__ __forceinline__ uint64_t lookback_exclusive(uint64_t *status, int bid, uint64_t agg,
                                                       int lane) {
    constexpr uint64_t FLAG_AGG = 1ull << 62, FLAG_INC = 2ull << 62, VMASK = (1ull << 62) - 1;
    if (bid == 0) {
        if (lane == 0) st_agent(&status[0], FLAG_INC | agg);
        return 0;
    }
    if (lane == 0) st_agent(&status[bid], FLAG_AGG | agg);
    uint64_t excl = 0;
    int base = bid - 1;
    for (;;) {
        uint64_t s[LB_DEPTH];
        for (;;) {
            bool ready = true;
#pragma unroll
            for (int j = 0; j < LB_DEPTH; ++j) {
                const int idx = base - (j * WAVE + lane);
                s[j] = idx >= 0 ? ld_acquire_agent(&status[idx]) : FLAG_INC;
                ready = ready && (s[j] >> 62) != 0;
            }
            if (__all(ready)) break;
            __builtin_amdgcn_s_sleep(1);
        }
        // nearest predecessor holding an inclusive prefix: slot j*, lane l*
        int jstar = -1, lstar = 0;
#pragma unroll
        for (int j = LB_DEPTH - 1; j >= 0; --j) {
            const uint64_t inc = __ballot((s[j] >> 62) == 2);
            if (inc) {
                jstar = j;
                lstar = __ffsll((unsigned long long)inc) - 1;
            }
        }
        uint64_t part = 0;
#pragma unroll
        for (int j = 0; j < LB_DEPTH; ++j) {
            const bool take = jstar < 0 || j < jstar || (j == jstar && lane <= lstar);
            part += take ? (s[j] & VMASK) : 0;
        }
        excl += wave_sum(part);
        if (jstar >= 0) break;
        base -= LB_DEPTH * WAVE;
    }
    if (lane == 0) st_agent(&status[bid], FLAG_INC | (excl + agg));
    return excl;
}

// Bit i set iff byte i of the 16-byte word is non-zero.
__device__ __forceinline__ uint32_t nonzero_bytes16(uint4 w) {
    uint32_t b = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        b |= ((w.x >> (8 * i)) & 0xffu) ? (1u << i) : 0u;
        b |= ((w.y >> (8 * i)) & 0xffu) ? (1u << (i + 4)) : 0u;
        b |= ((w.z >> (8 * i)) & 0xffu) ? (1u << (i + 8)) : 0u;
        b |= ((w.w >> (8 * i)) & 0xffu) ? (1u << (i + 12)) : 0u;
    }
    return b;
}

// Bit i set iff byte i of the 8-byte word is non-zero.
__device__ __forceinline__ uint32_t nonzero_bytes8(uint2 w) {
    uint32_t b = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        b |= ((w.x >> (8 * i)) & 0xffu) ? (1u << i) : 0u;
        b |= ((w.y >> (8 * i)) & 0xffu) ? (1u << (i + 4)) : 0u;
    }
    return b;
}

}  // namespace dev
}  // namespace apspie
