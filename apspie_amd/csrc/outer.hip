// outer.hip -- the reference's own SpGEMM scheme, kept beside the row-wise
// hash SpGEMM (spgemm.hip) for comparison: C = A * B^T as a sum of outer
// products of matching columns over 2D strips of rows (spgemm,
// src/spgemm.cuh:131-342, with the strip DCSC of csr_to_dcsc,
// src/matrix.cpp:286-468).
//
//   1. The CSC of A and of B on the device (convert.hip), rows ascending.
//   2. The DCSC of every strip of part_size rows: the CSC's runs of entries
//      that share a column and a strip ("segments"), in column order;
//      colseg[k] = the first segment of column k.
//   3. The column intersections of every strip pair (i, j): for each column
//      k, every (A segment, B segment) pair of that column is a triple
//      (i, j, k); a stable sort by the pair id i * nsb + j lists each
//      pair's matched columns in increasing order (what the reference's
//      SetOpPairs intersection gives pair by pair).
//   4. Products: triple t owns lenA * lenB consecutive slots (exclusive
//      scan); every tile of slots finds its triples from a tile ->
//      first-triple table and a binary search in LDS; a slot writes (row of
//      A, row of B, a * b in FP64).
//   5. Expand-sort-compress in place of the reference's per-pair global hash
//      tables (FP32 atomics, 131071 slots a pair, entries dropped on
//      overflow): stable radix sorts by column, then by row, keep the
//      products of each output entry in generation order (k ascending), so
//      the FP64 run sums are deterministic and equal a row-wise Gustavson's
//      in double.
#include <algorithm>
#include <climits>
#include <cstring>

#include "convert.hpp"
#include "device.hpp"
#include "esc.hpp"
#include "internal.hpp"

namespace apspie {
namespace outer {

using namespace dev;
using esc::bits_for;
using esc::k_rows_by;
using esc::k_runflag;

constexpr int TS = 1024;  // product slots per tile (256 threads x 4)

// flag[e] = 1 where a segment starts: the first entry of a column, or an
// entry whose strip differs from the previous entry's (one wave per column)
__global__ __launch_bounds__(256) void k_segflag(int ncol, const int32_t *__restrict__ cp,
                                                 const int32_t *__restrict__ ri, int P, int32_t *flag) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t nw = int64_t(gridDim.x) * (256 / WAVE);
    for (int64_t k = int64_t(blockIdx.x) * (256 / WAVE) + threadIdx.x / WAVE; k < ncol; k += nw) {
        const int a = cp[k], b = cp[k + 1];
        for (int e = a + lane; e < b; e += WAVE) flag[e] = (e == a || ri[e] / P != ri[e - 1] / P) ? 1 : 0;
    }
}

// segment pos[e] of every flagged entry e: its first entry and its column
__global__ __launch_bounds__(256) void k_segwrite(int ncol, const int32_t *__restrict__ cp,
                                                  const int32_t *__restrict__ flag, const int32_t *__restrict__ pos,
                                                  int32_t *seg_start, int32_t *seg_col) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t nw = int64_t(gridDim.x) * (256 / WAVE);
    for (int64_t k = int64_t(blockIdx.x) * (256 / WAVE) + threadIdx.x / WAVE; k < ncol; k += nw) {
        const int a = cp[k], b = cp[k + 1];
        for (int e = a + lane; e < b; e += WAVE)
            if (flag[e]) {
                seg_start[pos[e]] = e;
                seg_col[pos[e]] = int32_t(k);
            }
    }
}

__global__ void k_put(int32_t *p, int32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// tcnt[k] = (A segments of column k) * (B segments of column k); *total gets
// the exact 64-bit sum (the int32 counts are only scanned when it fits)
__global__ __launch_bounds__(256) void k_tcount(int ncol, const int32_t *__restrict__ csA,
                                                const int32_t *__restrict__ csB, int32_t *tcnt,
                                                unsigned long long *total) {
    uint64_t mine = 0;
    for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < ncol; k += int64_t(gridDim.x) * 256) {
        const int64_t c = int64_t(csA[k + 1] - csA[k]) * int64_t(csB[k + 1] - csB[k]);
        tcnt[k] = int32_t(std::min<int64_t>(c, INT_MAX));
        mine += uint64_t(c);
    }
    mine = wave_sum(mine);
    if ((threadIdx.x & (WAVE - 1)) == 0 && mine) atomicAdd(total, (unsigned long long)mine);
}

// one wave per column: the triple (sa, sb) of every A segment sa and B
// segment sb of column k, at tpos[k] + q, keyed by its strip pair
__global__ __launch_bounds__(256) void k_temit(int ncol, const int32_t *__restrict__ csA,
                                               const int32_t *__restrict__ csB, const int32_t *__restrict__ tpos,
                                               const int32_t *__restrict__ ssA, const int32_t *__restrict__ riA,
                                               const int32_t *__restrict__ ssB, const int32_t *__restrict__ riB,
                                               int P, int nsb, int32_t *tA, int32_t *tB, int32_t *tkey) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t nw = int64_t(gridDim.x) * (256 / WAVE);
    for (int64_t k = int64_t(blockIdx.x) * (256 / WAVE) + threadIdx.x / WAVE; k < ncol; k += nw) {
        const int a0 = csA[k], na = csA[k + 1] - a0, b0 = csB[k], nb = csB[k + 1] - b0;
        const int base = tpos[k];
        for (int q = lane; q < na * nb; q += WAVE) {  // na * nb < 2^31: the total was checked
            const int sa = a0 + q / nb, sb = b0 + q % nb;
            const int t = base + q;
            tA[t] = sa;
            tB[t] = sb;
            tkey[t] = (riA[ssA[sa]] / P) * nsb + riB[ssB[sb]] / P;
        }
    }
}

// the triples in strip-pair order, and their product counts
__global__ __launch_bounds__(256) void k_tgather(int64_t ntr, const int32_t *__restrict__ perm,
                                                 const int32_t *__restrict__ tA, const int32_t *__restrict__ tB,
                                                 const int32_t *__restrict__ ssA, const int32_t *__restrict__ ssB,
                                                 int32_t *oA, int32_t *oB, int32_t *pcnt,
                                                 unsigned long long *total) {
    uint64_t mine = 0;
    for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < ntr; t += int64_t(gridDim.x) * 256) {
        const int p = perm[t];
        const int sa = tA[p], sb = tB[p];
        oA[t] = sa;
        oB[t] = sb;
        const int64_t c = int64_t(ssA[sa + 1] - ssA[sa]) * int64_t(ssB[sb + 1] - ssB[sb]);
        pcnt[t] = int32_t(std::min<int64_t>(c, INT_MAX));
        mine += uint64_t(c);
    }
    mine = wave_sum(mine);
    if ((threadIdx.x & (WAVE - 1)) == 0 && mine) atomicAdd(total, (unsigned long long)mine);
}

// table[x] = the triple holding slot x * TS (every triple owns >= 1 slot)
__global__ __launch_bounds__(256) void k_ptiles(int64_t ntr, const int32_t *__restrict__ poff, int32_t *table) {
    for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < ntr; t += int64_t(gridDim.x) * 256) {
        const int64_t o = poff[t], e = poff[t + 1];
        for (int64_t x = (o + TS - 1) / TS; x * TS < e; ++x) table[x] = int32_t(t);
    }
}

// product slots, a tile of TS per step: the tile's triples are table[tile]
// .. table[tile + 1] (at most TS + 1 of them), their offsets go to LDS and
// every slot finds its triple by binary search there
__global__ __launch_bounds__(256) void k_products(int64_t ptot, int64_t ntr, int ntiles,
                                                  const int32_t *__restrict__ table,
                                                  const int32_t *__restrict__ poff, const int32_t *__restrict__ oA,
                                                  const int32_t *__restrict__ oB, const int32_t *__restrict__ ssA,
                                                  const int32_t *__restrict__ ssB, const int32_t *__restrict__ riA,
                                                  const float *__restrict__ cvA, const int32_t *__restrict__ riB,
                                                  const float *__restrict__ cvB, int32_t *prow, int32_t *pcol,
                                                  double *pval, int32_t *kcol) {
    __shared__ int32_t s_off[TS + 2];
    __shared__ int s_f0, s_n;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        if (threadIdx.x == 0) {
            const int f0 = table[tile];
            const int f1 = tile + 1 < ntiles ? table[tile + 1] : int(ntr - 1);
            s_f0 = f0;
            s_n = f1 - f0 + 1;
        }
        __syncthreads();
        const int f0 = s_f0, nt = s_n;
        for (int i = threadIdx.x; i <= nt; i += 256) s_off[i] = poff[f0 + i];  // f0 + nt <= ntr
        __syncthreads();
        const int64_t s0 = int64_t(tile) * TS;
        for (int j = threadIdx.x; j < TS; j += 256) {
            const int64_t s = s0 + j;
            if (s >= ptot) break;
            int lo = 0, hi = nt - 1;  // last triple whose first slot <= s
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= s) lo = mid;
                else hi = mid - 1;
            }
            const int t = f0 + lo;
            const int q = int(s - s_off[lo]);
            const int sa = oA[t], sb = oB[t];
            const int eb0 = ssB[sb], lb = ssB[sb + 1] - eb0;
            const int ia = q / lb;
            const int ea = ssA[sa] + ia, eb = eb0 + (q - ia * lb);
            const int r = riA[ea], c = riB[eb];
            prow[s] = r;
            pcol[s] = c;
            kcol[s] = c;
            pval[s] = double(cvA ? cvA[ea] : 1.0f) * double(cvB ? cvB[eb] : 1.0f);
        }
        __syncthreads();
    }
}

// one output entry per run: its products summed in FP64 in sorted order
__global__ __launch_bounds__(256) void k_runs(int64_t n, const int32_t *__restrict__ flag,
                                              const int32_t *__restrict__ pos, const int32_t *__restrict__ rows,
                                              const int32_t *__restrict__ perm, const int32_t *__restrict__ pcol,
                                              const double *__restrict__ pval, int32_t *crow, int32_t *ccol,
                                              float *cval) {
    for (int64_t s = int64_t(blockIdx.x) * 256 + threadIdx.x; s < n; s += int64_t(gridDim.x) * 256) {
        if (!flag[s]) continue;
        const int j = pos[s];
        double acc = pval[perm[s]];
        for (int64_t x = s + 1; x < n && !flag[x]; ++x) acc += pval[perm[x]];
        crow[j] = rows[s];
        ccol[j] = pcol[perm[s]];
        cval[j] = float(acc);
    }
}

struct Dcsc {
    int32_t *cp, *ri;
    float *cv;
    int32_t *seg_start;  // nseg + 1 (the last = nnz)
    int32_t *colseg;     // ncol + 1
    int64_t nseg;
};

}  // namespace outer
}  // namespace apspie

using namespace apspie;

extern "C" int apspie_spgemm_outer(apspie_ctx *ctx, const apspie_csr *a, const apspie_csr *b, int32_t part_size,
                                   apspie_csr_owned *c, apspie_stats *stats) {
    if (!ctx || !c) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    memset(c, 0, sizeof(*c));
    return guarded(ctx, [&] {
        using namespace outer;
        validate_csr(a, false);
        validate_csr(b, false);
        APSPIE_CHECK(a->n == b->n, APSPIE_EINVAL, "A and B must have the same number of columns (C = A * B^T)");
        APSPIE_CHECK(part_size >= 1, APSPIE_EINVAL, "part_size must be positive");
        const int m = a->m, p = b->m, ncol = a->n, P = part_size;
        const int64_t nsa = (std::max(m, 1) + int64_t(P) - 1) / P, nsb = (std::max(p, 1) + int64_t(P) - 1) / P;
        APSPIE_CHECK(nsa * nsb <= INT_MAX, APSPIE_EINVAL, "too many strip pairs for part_size");
        auto S = [&](int k, int64_t bytes) { return ctx->slot(S_OUTER0 + k, size_t(std::max<int64_t>(bytes, 16))); };
        auto grid = [&](int64_t n, unsigned per) { return cvt::grid_for(ctx, n, per); };
        int64_t ntr = 0, ptot = 0;
        APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        // C's arrays, once their size is known (nothing to release on the
        // error paths before this)
        auto alloc_c = [&](int32_t nnzc) {
            c->m = m;
            c->n = p;
            c->nnz = nnzc;
            if (hipMalloc(&c->row_ptr, (size_t(m) + 1) * 4) != hipSuccess ||
                hipMalloc(&c->col_ind, size_t(std::max(nnzc, 1)) * 4) != hipSuccess ||
                hipMalloc(&c->val, size_t(std::max(nnzc, 1)) * 4) != hipSuccess) {
                (void)hipGetLastError();
                apspie_csr_free(ctx, c);
                throw Error(APSPIE_ENOMEM, "output allocation failed");
            }
        };
        auto finish_empty = [&] {
            alloc_c(0);
            APSPIE_HIP(hipMemsetAsync(c->row_ptr, 0, (size_t(m) + 1) * 4, ctx->stream));
        };
        if (a->nnz == 0 || b->nnz == 0) {
            finish_empty();
        } else {
            // 1-2. CSC and strip segments of both operands
            Dcsc D[2];
            const apspie_csr *X[2] = {a, b};
            for (int w = 0; w < 2; ++w) {
                const int64_t nnz = X[w]->nnz;
                const int o = 10 * w;  // slots S_OUTER0 + o .. + o + 7
                Dcsc &d = D[w];
                d.cp = (int32_t *)S(o + 0, (int64_t(ncol) + 1) * 4);
                d.ri = (int32_t *)S(o + 1, nnz * 4);
                d.cv = X[w]->val ? (float *)S(o + 2, nnz * 4) : nullptr;
                cvt::csr2csc_into(ctx, X[w], d.cp, d.ri, d.cv);
                int32_t *flag = (int32_t *)S(o + 3, nnz * 4);
                int32_t *pos = (int32_t *)S(o + 4, (nnz + 1) * 4);
                k_segflag<<<grid(int64_t(ncol) * WAVE, 256), 256, 0, ctx->stream>>>(ncol, d.cp, d.ri, P, flag);
                cvt::scan32(ctx, flag, pos, nnz);
                int32_t nseg = 0;
                APSPIE_HIP(hipMemcpyAsync(&nseg, pos + nnz, 4, hipMemcpyDeviceToHost, ctx->stream));
                APSPIE_HIP(hipStreamSynchronize(ctx->stream));
                d.nseg = nseg;
                d.seg_start = (int32_t *)S(o + 5, (int64_t(nseg) + 1) * 4);
                int32_t *seg_col = (int32_t *)S(o + 6, int64_t(nseg) * 4);
                d.colseg = (int32_t *)S(o + 7, (int64_t(ncol) + 1) * 4);
                k_segwrite<<<grid(int64_t(ncol) * WAVE, 256), 256, 0, ctx->stream>>>(ncol, d.cp, flag, pos,
                                                                                      d.seg_start, seg_col);
                k_put<<<1, 64, 0, ctx->stream>>>(d.seg_start + nseg, int32_t(nnz));
                APSPIE_HIP(hipGetLastError());
                cvt::compress(ctx, seg_col, nseg, ncol, d.colseg);
            }
            // 3. strip-pair column intersections (triples), grouped by pair
            int32_t *tcnt = (int32_t *)S(20, int64_t(ncol) * 4);
            int32_t *tpos = (int32_t *)S(21, (int64_t(ncol) + 1) * 4);
            unsigned long long *tot = (unsigned long long *)S(22, 64);
            APSPIE_HIP(hipMemsetAsync(tot, 0, 16, ctx->stream));
            k_tcount<<<grid(ncol, 256), 256, 0, ctx->stream>>>(ncol, D[0].colseg, D[1].colseg, tcnt, tot);
            unsigned long long htot[2] = {0, 0};
            APSPIE_HIP(hipMemcpyAsync(htot, tot, 8, hipMemcpyDeviceToHost, ctx->stream));
            APSPIE_HIP(hipStreamSynchronize(ctx->stream));
            APSPIE_CHECK(htot[0] < (1ull << 31), APSPIE_ELIMIT, "2^31 or more strip-pair column matches");
            ntr = int64_t(htot[0]);
            if (ntr == 0) {
                finish_empty();
            } else {
                cvt::scan32(ctx, tcnt, tpos, ncol);
                int32_t *tA = (int32_t *)S(23, ntr * 4), *tB = (int32_t *)S(24, ntr * 4);
                int32_t *tkey[2] = {(int32_t *)S(25, ntr * 4), (int32_t *)S(26, ntr * 4)};
                int32_t *tidx[2] = {(int32_t *)S(27, ntr * 4), (int32_t *)S(28, ntr * 4)};
                k_temit<<<grid(int64_t(ncol) * WAVE, 256), 256, 0, ctx->stream>>>(
                    ncol, D[0].colseg, D[1].colseg, tpos, D[0].seg_start, D[0].ri, D[1].seg_start, D[1].ri, P,
                    int(nsb), tA, tB, tkey[0]);
                APSPIE_HIP(hipGetLastError());
                const int tc = cvt::radix_pairs(ctx, tkey, tidx, ntr, bits_for(nsa * nsb), 0, true);
                // the sorted triples go to the free pair of buffers; the
                // product counts over the sorted keys (no longer needed)
                int32_t *oA = tkey[tc ^ 1], *oB = tidx[tc ^ 1], *pcnt = tkey[tc];
                int32_t *poff = (int32_t *)S(29, (ntr + 1) * 4);
                APSPIE_HIP(hipMemsetAsync(tot + 1, 0, 8, ctx->stream));
                k_tgather<<<grid(ntr, 256), 256, 0, ctx->stream>>>(ntr, tidx[tc], tA, tB, D[0].seg_start,
                                                                    D[1].seg_start, oA, oB, pcnt, tot + 1);
                APSPIE_HIP(hipMemcpyAsync(htot + 1, tot + 1, 8, hipMemcpyDeviceToHost, ctx->stream));
                APSPIE_HIP(hipStreamSynchronize(ctx->stream));
                APSPIE_CHECK(htot[1] < (1ull << 31), APSPIE_ELIMIT, "2^31 or more products");
                ptot = int64_t(htot[1]);
                cvt::scan32(ctx, pcnt, poff, ntr);
                // 4. products
                const int ntiles = int((ptot + TS - 1) / TS);
                int32_t *table = (int32_t *)S(30, int64_t(ntiles) * 4);
                int32_t *prow = (int32_t *)S(31, ptot * 4), *pcol = (int32_t *)S(32, ptot * 4);
                double *pval = (double *)S(33, ptot * 8);
                int32_t *key[2] = {(int32_t *)S(34, (ptot + 1) * 4), (int32_t *)S(35, (ptot + 1) * 4)};
                int32_t *idx[2] = {(int32_t *)S(36, (ptot + 1) * 4), (int32_t *)S(37, (ptot + 1) * 4)};
                k_ptiles<<<grid(ntr, 256), 256, 0, ctx->stream>>>(ntr, poff, table);
                k_products<<<std::max(1u, std::min((unsigned)ntiles, unsigned(ctx->cu_count) * 8)), 256, 0,
                             ctx->stream>>>(ptot, ntr, ntiles, table, poff, oA, oB, D[0].seg_start, D[1].seg_start,
                                            D[0].ri, D[0].cv, D[1].ri, D[1].cv, prow, pcol, pval, key[0]);
                APSPIE_HIP(hipGetLastError());
                // 5. sort by (row, column): by column, then stably by row
                const int c1 = cvt::radix_pairs(ctx, key, idx, ptot, bits_for(p), 0, true);
                k_rows_by<<<grid(ptot, 256), 256, 0, ctx->stream>>>(ptot, idx[c1], prow, key[0], idx[0]);
                const int c2 = cvt::radix_pairs(ctx, key, idx, ptot, bits_for(m));
                const int32_t *rows = key[c2], *perm = idx[c2];
                int32_t *flag = key[c2 ^ 1], *pos = idx[c2 ^ 1];
                k_runflag<<<grid(ptot, 256), 256, 0, ctx->stream>>>(ptot, rows, perm, pcol, flag);
                cvt::scan32(ctx, flag, pos, ptot);
                int32_t nnzc = 0;
                APSPIE_HIP(hipMemcpyAsync(&nnzc, pos + ptot, 4, hipMemcpyDeviceToHost, ctx->stream));
                APSPIE_HIP(hipStreamSynchronize(ctx->stream));
                alloc_c(nnzc);
                int32_t *crow = prow;  // free once the rows were gathered
                k_runs<<<grid(ptot, 256), 256, 0, ctx->stream>>>(ptot, flag, pos, rows, perm, pcol, pval, crow,
                                                               c->col_ind, c->val);
                APSPIE_HIP(hipGetLastError());
                cvt::compress(ctx, crow, nnzc, m, c->row_ptr);
            }
        }
        APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        if (stats) {
            float ms = 0.f;
            APSPIE_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            memset(stats, 0, sizeof(*stats));
            stats->edges = ptot;         // products
            stats->vertices = c->nnz;    // nnz(C)
            stats->improved = ntr;       // strip-pair column matches
            stats->iterations = 1;
            stats->kernel_ms = ms;
            stats->bytes_alg = 8.0 * (double(m) + 1) + 16.0 * double(a->nnz) + 8.0 * double(ptot) +
                               8.0 * double(c->nnz);
        }
    });
}
