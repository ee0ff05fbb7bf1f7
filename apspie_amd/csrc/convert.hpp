// convert.hpp -- host helpers of the device format conversions (convert.hip)
// shared with the other sparse kernels.
#pragma once

#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace apspie {
namespace cvt {

// Workgroups for a grid-stride loop over n items, `per` items per workgroup
// per step, capped at 32 per CU.
unsigned grid_for(apspie_ctx *ctx, int64_t n, unsigned per);

// off[r] = first position e with key[e] >= r, r in [0, m] (keys sorted, in
// [0, m); APSPIE_EINVAL otherwise).  Synchronises the stream.
void compress(apspie_ctx *ctx, const int32_t *key, int64_t n, int m, int32_t *off);

// out[0..n] = exclusive scan of n non-negative int32 (out[n] = total); the
// caller guarantees the total stays below 2^31.  Uses the S_STATUS slot.
void scan32(apspie_ctx *ctx, const int32_t *in, int32_t *out, int64_t n);

// The same scan into int64 (any total).
void scan_wide(apspie_ctx *ctx, const int32_t *in, int64_t *out, int64_t n);

// Stable least-significant-digit radix sort of the pairs (key[start][e],
// idx[start][e]) by the low `bits` bits of the key (non-negative keys), 8
// bits a pass, ping-ponging between the two buffers of each array.  Returns
// which buffer (0 or 1) holds the result.  ident: idx[start] is not read,
// the first pass takes each item's position (and at least one pass runs).
// Uses the S_AUX4, S_AUX5 and S_STATUS slots.
int radix_pairs(apspie_ctx *ctx, int32_t *key[2], int32_t *idx[2], int64_t n, int bits, int start = 0,
                bool ident = false);

// Stable LSD radix sort of (key[0][e], val[0][e]) by the low `bits` bits of
// the key (non-negative keys) inside each segment [seg[r], seg[r + 1]) (host
// offsets covering [0, n) in order), 8 bits a pass from bit first_shift up
// (0: a full sort; bits - 8: one pass that groups each segment by the top 8
// bits, stably), each key's FP64 value moving with it.  A pass's tiles never cross a segment, and the histogram of
// tile j of segment r sits at (segment r's first slot) + digit * (tiles of
// r) + (j - r's first tile), so one exclusive scan gives every (segment,
// digit, tile) its place.  Returns the buffer (0 or 1) holding the result.
// `tiles` receives the tile table uploaded from it: keep it until the stream
// synchronises.  Uses the S_AUX4, S_AUX5, S_AUX6 and S_STATUS slots.
int seg_radix_kv(apspie_ctx *ctx, int32_t *key[2], double *val[2], int64_t n, const std::vector<int32_t> &seg,
                 int bits, std::vector<int32_t> &tiles, int first_shift = 0);

// The CSC of a (rows ascending in each column) into caller memory; cv may be
// null.  Uses the S_AUX0..S_AUX5 and S_STATUS slots.
void csr2csc_into(apspie_ctx *ctx, const apspie_csr *a, int32_t *cp, int32_t *ri, float *cv);

}  // namespace cvt
}  // namespace apspie
