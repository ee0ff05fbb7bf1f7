// hint.hpp -- distance bound codes of the single-source SSSP rounds
// (traverse.hip, binned.hpp).
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace apspie {
namespace trav {

// Distance bound codes: a byte per vertex whose decoded value bounds the
// vertex's distance from above.
// The compaction writes the code of a vertex's distance whenever it dropped,
// and distances only fall during a round, so at every moment of a round the
// bound holds.  An edge whose
// candidate is not below its target's bound cannot lower the target's
// distance and needs no 4-byte distance probe (push) or candidate entry
// (binned); the byte array is a quarter of the distances' size, so its random
// probes hit the caches far more often.  Codes: 0 -> 0; c = 1..127 -> c;
// c = 128..254 -> the float whose bits are those of 127.0 plus (c - 127) <<
// 17 (128.0 in steps of 2 to 256.0, then steps of 4 to 504.0); 255 -> no
// bound.  Integer distances up to 128 are coded exactly.
constexpr uint32_t HINT_F127 = 0x42fe0000u;  // bits of 127.0f
__device__ __forceinline__ uint32_t hint_code(uint32_t bits) {
    if (bits == 0u) return 0u;
    if (bits <= HINT_F127) return uint32_t(ceilf(__uint_as_float(bits)));  // 1..127, at least the distance
    const uint32_t q = (bits - HINT_F127 + 0x1ffffu) >> 17;               // rounds up
    return q <= 127u ? 127u + q : 255u;
}
__device__ __forceinline__ uint32_t hint_bound(uint32_t c) {
    if (c <= 127u) return __float_as_uint(float(c));  // 0 -> bits 0
    return c == 255u ? ~0u : HINT_F127 + ((c - 127u) << 17);
}

}  // namespace trav
}  // namespace apspie
