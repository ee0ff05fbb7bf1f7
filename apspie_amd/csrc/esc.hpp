// esc.hpp -- expand-sort-compress steps of the outer-product SpGEMM
// (outer.hip): products sorted stably by column and then by row
// (convert.hpp's radix pairs) are grouped into runs of equal (row, column).
// bits_for also sizes the row-wise SpGEMM's long-row sort (spgemm.hip).
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace apspie {
namespace esc {

// key[s] = the row of product perm[s], idx[s] = perm[s] (perm may alias idx:
// each thread reads its element before writing it)
static __global__ __launch_bounds__(256) void k_rows_by(int64_t n, const int32_t *perm,
                                                        const int32_t *__restrict__ prow, int32_t *key,
                                                        int32_t *idx) {
    for (int64_t s = int64_t(blockIdx.x) * 256 + threadIdx.x; s < n; s += int64_t(gridDim.x) * 256) {
        const int p = perm[s];
        key[s] = prow[p];
        idx[s] = p;
    }
}

// flag[s] = 1 where a new (row, column) starts in the sorted products
static __global__ __launch_bounds__(256) void k_runflag(int64_t n, const int32_t *__restrict__ rows,
                                                        const int32_t *__restrict__ perm,
                                                        const int32_t *__restrict__ pcol, int32_t *flag) {
    for (int64_t s = int64_t(blockIdx.x) * 256 + threadIdx.x; s < n; s += int64_t(gridDim.x) * 256)
        flag[s] = (s == 0 || rows[s] != rows[s - 1] || pcol[perm[s]] != pcol[perm[s - 1]]) ? 1 : 0;
}

// smallest b with 2^b >= x
inline int bits_for(int64_t x) {
    int b = 0;
    while ((int64_t(1) << b) < x) ++b;
    return b;
}

}  // namespace esc
}  // namespace apspie
