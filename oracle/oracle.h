/*
 * oracle.h -- CPU restatement of apspie's CPU reference paths.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * kernels in apspie_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it and
 * never falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the apspie checkout).  Plain C11, single-threaded, as the reference's CPU
 * paths are.
 */
#ifndef APSPIE_ORACLE_H
#define APSPIE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Host CSR owned by the oracle (malloc'd; release with orc_csr_free). */
typedef struct {
    int32_t m, n;
    int64_t nnz;
    int32_t *row_ptr; /* m+1 */
    int32_t *col_ind; /* nnz */
    float *val;       /* nnz, may be NULL */
} orc_csr;

/* MT19937 (Matsumoto & Nishimura 1998), restating src/mt19937ar.cpp:57-137
 * (init_genrand, genrand_int32). */
typedef struct {
    uint32_t s[624];
    int idx;
} orc_mt;
void orc_mt_seed(orc_mt *g, uint32_t seed);
uint32_t orc_mt_next(orc_mt *g);
/* Convenience for tests: n draws after seeding. */
void orc_mt_fill(uint32_t seed, int64_t n, uint32_t *out);

void orc_csr_free(orc_csr *a);

/* MatrixMarket loader: readEdge + readMtx + [makeSymmetric] + buildMatrix,
 * src/util.cuh:355-430, :522-643 and src/matrix.cpp:55-111.
 * Returns 0 on success, <0 on error. */
int orc_read_mtx(const char *path, int undirected, orc_csr *out);

/* RMAT(scale, edge_factor, a=.57, b=.19, c=.19, d=.05) with one mt19937
 * stream; self-loops and duplicates dropped; undirected adds reverse edges.
 * Sorted columns.  (Synthetic input definition; the reference has none.) */
int orc_rmat(int scale, int edge_factor, uint32_t seed, int undirected, orc_csr *out);
/* Integer FP32 weights 1 + (genrand_int32() >> 26) in CSR order. */
int orc_rmat_weights(orc_csr *a, uint32_t seed);
int32_t orc_argmax_degree(const orc_csr *a);

/* FIFO BFS, src/testBfs.cpp:6-72.  depth[v] = -1 if unreached.  Returns the
 * search depth (max depth + 1).  e_exp = sum of degrees of dequeued vertices,
 * v_reached = reached vertices (either may be NULL). */
int32_t orc_bfs(const orc_csr *a, int32_t src, int32_t stop, int32_t *depth,
                int64_t *e_exp, int64_t *v_reached);

/* Heap Dijkstra, src/testSssp.cpp:33-115.  dist[v] = -1 if unreachable. */
int32_t orc_sssp_dijkstra(const orc_csr *a, int32_t src, float *dist);

/* Synchronous (Jacobi) Bellman-Ford with a changed-vertex frontier: the
 * schedule the GPU path follows.  dist[v] = -1 if unreachable.  Counts:
 * iterations (rounds with a non-empty frontier), e_relax = sum over rounds of
 * the frontier's out-degrees, sum_f = sum of frontier sizes, r_imp = sum over
 * rounds of vertices whose distance dropped. */
int32_t orc_sssp_bf(const orc_csr *a, int32_t src, float *dist, int64_t *e_relax,
                    int64_t *sum_f, int64_t *r_imp);

/* Greedy MIS in descending (prio[v], -v) order -- SimpleReferenceMis,
 * test/mis/main.cu:26-88, generalised from index order to priority order.
 * out[v] = 1 in the set, 0 not. */
void orc_mis_greedy(const orc_csr *a, const uint32_t *prio, int32_t *out);
/* SimpleReferenceMis exactly: src first, then index order. */
void orc_mis_index(const orc_csr *a, int32_t src, int32_t *out);

/* One Boolean SpMSpV as spmspvCsr, src/spmspvMis.cuh:81-147: out[v] = 1 iff
 * v is in the row of some u with in[u] != 0. */
void orc_spmspv_or(const orc_csr *a, const int32_t *in, int32_t *out);

/* One min-plus SpMSpV step as spmspvSssp, src/spmspvSssp.cuh:92-291:
 * frontier = {u : -1 < x[u] < 1e38}; y[v] = min over frontier u of
 * x[u] + A(u,v); y = 1.70141e38 where nothing reaches. */
void orc_sssp_step(const orc_csr *a, const float *x, float *y);

/* Gustavson C = A*B with double accumulation, sorted columns, structural
 * zeros kept (semantics of src/spgemm.cuh:344-445 and the hash kernel,
 * src/spgemmKernel.cuh:179-208).  products = sum_i sum_k nnz(B(k,:)). */
int orc_spgemm(const orc_csr *a, const orc_csr *b, orc_csr *c, int64_t *products);

/* Strict lower triangle (row i keeps j < i), sorted columns -- buildLower,
 * test/tc/main.cu:39-55. */
int orc_strict_lower(const orc_csr *a, orc_csr *l);
/* Masked triangle count sum_{(i,k) in L} |L(i,:) & L(k,:)|, int64; the
 * count of src/tc.cuh:168-199 + test/tc/main.cu:238-249.  wedges = sum over
 * (i,k) in L of nnz(L(k,:)). */
int64_t orc_tc(const orc_csr *l, int64_t *wedges);
int64_t orc_tc_rows(const orc_csr *l, int32_t r0, int32_t r1, int64_t *wedges);
void orc_mxv(const orc_csr *a, int sr, const float *x, const uint8_t *listed, float *y, uint8_t *reached);

#ifdef __cplusplus
}
#endif
#endif
