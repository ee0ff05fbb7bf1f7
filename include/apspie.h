/*
 * apspie.h -- C ABI of the MI355X (gfx950) SpMSpV traversal and hash-SpGEMM
 * library.
 *
 * Every function returns 0 on success and a negative APSPIE_E* code on
 * failure; apspie_last_error() holds the message.  No function calls exit().
 * Device pointers are plain HIP device addresses (hipMalloc / apspie_malloc /
 * torch tensors' data_ptr()).  Work runs on the context's stream; the plain
 * (blocking) entry points synchronise that stream before returning.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the apspie checkout).  The reference has no FFI of its own: its hot path
 * is header-only CUDA C++ templates called from test/<algo>/main.cu.
 */
#ifndef APSPIE_H
#define APSPIE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define APSPIE_OK 0
#define APSPIE_EINVAL (-1)   /* bad argument (null pointer, id out of range) */
#define APSPIE_ENOMEM (-2)   /* device or host allocation failed */
#define APSPIE_EHIP (-3)     /* HIP runtime error */
#define APSPIE_EIO (-4)      /* file / parse error */
#define APSPIE_ELIMIT (-5)   /* size beyond a documented limit */
#define APSPIE_ENOCONV (-6)  /* iteration cap reached without convergence */
#define APSPIE_EBUSY (-7)    /* serving: admission refused / queue full */

/* Distance given to vertices SSSP does not reach: the reference's sentinel
 * 1.70141e38 (src/spmspvSssp.cuh:26, :214). */
#define APSPIE_SSSP_INF 1.70141e38f

typedef struct apspie_ctx apspie_ctx;

/* Device CSR (borrowed device pointers).  Replaces the raw
 * (m, edge, d_csrRowPtr, d_csrColInd, d_csrVal) argument groups of
 * src/spmspvBfs.cuh:96, src/sssp.cuh:28, src/spmspvMis.cuh:150 and the
 * d_matrix fields of src/matrix.hpp:6-37.  col_ind values are < n;
 * nnz < 2^31.  val may be NULL (all ones).
 *
 * The csc_* fields are optional (NULL when absent): the same matrix by
 * columns, csc_col_ptr[n+1] / csc_row_ind[nnz] / csc_val[nnz] -- the
 * d_cscColPtrA / d_cscRowIndA / d_cscValA arrays the reference's drivers build
 * beside the CSR with csr2csc (test/sssp/main.cu:110, src/util.cuh:474).  With
 * them BFS and SSSP run their dense rounds as pull steps over in-edges; the
 * results and statistics do not change.  For a symmetric matrix they may
 * alias the CSR arrays. */
typedef struct {
    int32_t m, n;
    int64_t nnz;
    const int32_t *row_ptr;
    const int32_t *col_ind;
    const float *val;
    const int32_t *csc_col_ptr;
    const int32_t *csc_row_ind;
    const float *csc_val;
} apspie_csr;

/* Host CSR allocated by the library (apspie_host_csr_free releases it). */
typedef struct {
    int32_t m, n;
    int64_t nnz;
    int32_t *row_ptr;
    int32_t *col_ind;
    float *val; /* NULL when the input is unweighted and no weights were asked for */
} apspie_host_csr;

/* Device CSR allocated by the library (apspie_csr_free releases it), with
 * the optional CSC of the same matrix (apspie_csr_attach_csc). */
typedef struct {
    int32_t m, n;
    int64_t nnz;
    int32_t *row_ptr;
    int32_t *col_ind;
    float *val;
    int32_t *csc_col_ptr;
    int32_t *csc_row_ind;
    float *csc_val;
} apspie_csr_owned;

/* Per-call statistics.  For BFS: edges = expanded edges (the reference's
 * "traversed edges", src/spmspvBfs.cuh:217), vertices = reached vertices.
 * For SSSP: edges = relaxed edges over all rounds, vertices = sum of frontier
 * sizes, improved = (vertex, round) pairs whose distance dropped.  bytes_alg
 * is the algorithmic HBM traffic of the traversal, kernel_ms the device time
 * between events recorded around it on the context stream. */
typedef struct {
    int64_t edges;
    int64_t vertices;
    int64_t improved;
    int32_t iterations;
    int32_t relax_rounds; /* SSSP with binned rounds on: the leading rounds whose relaxation ran in one
                             k_relax launch each (push or binned); the later ones ran the plain push
                             expansion.  0 otherwise */
    double bytes_alg;
    double kernel_ms;
} apspie_stats;

/* ---- context -------------------------------------------------------------
 * Replaces mgpu::CudaContext (passed to every reference entry point) and the
 * per-call d_scratch / allocScratch of src/scratch.hpp:5-31,
 * src/util.cuh:646-672: the context owns a stream and a grow-only scratch
 * pool.  One context per host thread. */
int apspie_ctx_create(int device, apspie_ctx **out);
int apspie_ctx_destroy(apspie_ctx *ctx);
const char *apspie_last_error(const apspie_ctx *ctx);
/* The context's stream as a hipStream_t cast to void*. */
void *apspie_ctx_stream(apspie_ctx *ctx);
int apspie_ctx_synchronize(apspie_ctx *ctx);
/* Compute-unit count of the context's device. */
int apspie_ctx_cu_count(const apspie_ctx *ctx);

/* Per-kernel timing, the counterpart of the reference's GpuTimer event pairs
 * (src/util.cuh:318-352) and per-stage printf timings.  While enabled, entry
 * points bracket every launch with HIP events on the context stream and
 * accumulate device time per kernel class. */
#define APSPIE_KCLASS_EXPAND 0   /* frontier edge expansion */
#define APSPIE_KCLASS_COMPACT 1  /* frontier compaction + scan */
#define APSPIE_KCLASS_OTHER 2    /* initialisation and everything else */
#define APSPIE_KCLASS_PULL 3     /* pull steps over in-edges (BFS bottom-up, SSSP pull) */
#define APSPIE_KCLASS_BIN 4      /* SSSP binned relaxation: candidates written by bucket */
#define APSPIE_KCLASS_APPLY 5    /* SSSP binned relaxation: buckets folded into the distances */
#define APSPIE_KCLASS_ROUND 6    /* BFS with a CSC: one launch per round, push or bottom-up */
#define APSPIE_KCLASS_COUNT 8
int apspie_ctx_set_profiling(apspie_ctx *ctx, int enable);
/* Totals since the last reset: device milliseconds and launch count. */
int apspie_ctx_profile_read(apspie_ctx *ctx, int kclass, double *ms, int64_t *launches);
int apspie_ctx_profile_reset(apspie_ctx *ctx);

/* Tunable knobs of the traversals (the reference has none: it always pushes
 * and fixes its launch shapes).  Direction thresholds apply to graphs that
 * carry their CSC, except binned SSSP rounds (m <= 2^23).  bfs_alpha <= 0
 * turns BFS bottom-up steps off, sssp_pull_ratio <= 0 SSSP pull steps,
 * sssp_bin_ratio <= 0 binned SSSP rounds.  Results never depend on them. */
typedef struct {
    int32_t bfs_alpha;        /* bottom-up once frontier edges * alpha > unexplored edges (default 14) */
    int32_t bfs_beta;         /* stay bottom-up while frontier vertices * beta >= m (default 24) */
    int32_t sssp_pull_ratio;  /* pull once frontier edges * ratio > nnz (default 0: off) */
    int32_t sssp_bin_ratio;   /* binned relaxation once frontier edges * ratio > nnz (default 0: off) */
    int32_t expand_wg_per_cu; /* expansion workgroups per compute unit, 1..32 (default 8) */
    int32_t round_chunk;      /* rounds launched before the first frontier check, 1..32 (default 4) */
    int32_t batch_chunk;      /* rounds per captured hipGraph of the batched forms, 1..16 (default 4) */
    int32_t sssp_batch_lanes; /* batched SSSP lanes per edge: 16 (k <= 16 only) or 64; 0 = by k (default) */
    int32_t sssp_delta;       /* SSSP near/far step: a changed vertex waits while its distance is at or
                                 past a threshold that rises by this much when the frontier below it
                                 runs dry; 0 = off (plain frontier Bellman-Ford), 0..2^24 */
    int32_t spgemm_long_rows; /* SpGEMM rows beyond the LDS bins (over 2048 products): 1 = expand-
                                 sort-compress (default; global hash tables when their products reach
                                 2^31), 0 = global hash tables */
    int32_t spgemm_passes;    /* SpGEMM rows of the LDS bins (hash tables up to 256 products, LDS
                                 sorts up to 2048): 1 = one pass into a staging area (default; two
                                 when their products reach 2^31), 2 = a counting pass, then the
                                 product pass */
    int32_t spgemm_esc_min;   /* SpGEMM with spgemm_long_rows = 1: rows with more products than this go
                                 to expand-sort-compress instead of the LDS bins; 256, 512, 1024
                                 (default) or 2048 */
    int32_t bfs_group_cap;    /* BFS with a CSC: entries one round may append to a queue group region
                                 (16 regions); a round whose new frontier does not fit, and whose
                                 successor runs top-down, is queued by a pass over the vertex range
                                 instead.  0 = the whole region (default), else 1..2^30 */
    int32_t sssp_hint;        /* SSSP push rounds: 1 = keep a byte per vertex whose code bounds the vertex's
                                 distance from above and probe a target's 4-byte distance only when an
                                 edge's candidate is below that bound (default), 0 = probe every target */
} apspie_tuning;
int apspie_ctx_set_tuning(apspie_ctx *ctx, const apspie_tuning *t);
int apspie_ctx_get_tuning(const apspie_ctx *ctx, apspie_tuning *t);

/* Autotuner: times candidate values of the knobs above on graph g -- BFS
 * and SSSP from `src` (kinds bit 0 and 1), batched BFS and SSSP from `k`
 * vertices with out-edges spread over the graph (bits 2 and 3) -- one knob at
 * a time, keeps the fastest in the context and returns it in *out.  Each
 * candidate is the median of three timed calls after a warm-up. */
#define APSPIE_TUNE_BFS 1u
#define APSPIE_TUNE_SSSP 2u
#define APSPIE_TUNE_BFS_BATCH 4u
#define APSPIE_TUNE_SSSP_BATCH 8u
int apspie_autotune(apspie_ctx *ctx, const apspie_csr *g, int32_t src, int32_t k, uint32_t kinds,
                    apspie_tuning *out);

/* Device memory helpers for callers without a framework allocator
 * (the reference's cudaMalloc / cudaMemcpy calls in test/<algo>/main.cu). */
int apspie_malloc(apspie_ctx *ctx, void **ptr, size_t bytes);
int apspie_free(apspie_ctx *ctx, void *ptr);
int apspie_memcpy_h2d(apspie_ctx *ctx, void *dst, const void *src, size_t bytes);
int apspie_memcpy_d2h(apspie_ctx *ctx, void *dst, const void *src, size_t bytes);

/* STREAM-like device copy of `bytes` (read + write counted), `iters` times;
 * *gbps = best rate.  Reports the measured HBM bandwidth beside the nominal
 * peak (the reference has no counterpart). */
int apspie_hbm_copy_bandwidth(apspie_ctx *ctx, size_t bytes, int iters, double *gbps);

/* ---- graph loading -------------------------------------------------------- */
#define APSPIE_LOAD_UNDIRECTED 1u   /* makeSymmetric (src/util.cuh:572-643) */
#define APSPIE_LOAD_SORT_COLUMNS 2u /* sort each row's columns (not in the reference) */
/* Read the %%MatrixMarket banner (the reference skips every % line): field
 * pattern -> no values, real / integer -> a value on every line; symmetry
 * symmetric -> off-diagonal entries mirrored as makeSymmetric does,
 * skew-symmetric -> mirrored with negated values; array format, complex and
 * hermitian are refused (APSPIE_EIO).  Entries keep readMtx's reading
 * (line "a b" is column a-1 of row b-1). */
#define APSPIE_LOAD_MM_BANNER 4u

/* MatrixMarket reader: readEdge + readMtx + [makeSymmetric] + buildMatrix,
 * src/util.cuh:355-430, :522-643, src/matrix.cpp:55-111.  Line "a b [w]"
 * puts column a-1 into row b-1, as readMtx does. */
int apspie_load_mtx(const char *path, uint32_t flags, apspie_host_csr *out);
void apspie_host_csr_free(apspie_host_csr *h);

/* makeSymmetric (src/util.cuh:572-643) on host COO arrays that hold room
 * for 2 * edge entries: the reversed copy of every entry that is not a
 * self-loop is appended, the entries are sorted by (row, col) and exact
 * repeats dropped (self-loops stay); values move with their entries (the
 * reference's compaction moves the indices only).  coo_val may be NULL.
 * *out_edge = the new entry count. */
int apspie_host_coo_make_symmetric(int64_t edge, int32_t *coo_col, int32_t *coo_row, float *coo_val,
                                   int64_t *out_edge);
/* buildMatrix's CSR construction (src/matrix.cpp:55-111; also the host-array
 * form test/bfs/main.cu:83 calls): stable counting sort of host COO entries
 * by row into caller-allocated row_ptr[m+1], col_ind[nnz], val[nnz] (val
 * NULL: no values; coo_val NULL: ones).  APSPIE_EINVAL for a row outside
 * [0, m). */
int apspie_host_coo_to_csr(int32_t m, int64_t nnz, const int32_t *coo_row, const int32_t *coo_col,
                           const float *coo_val, int32_t *row_ptr, int32_t *col_ind, float *val);

/* Binary CSR cache (no counterpart in the reference, which re-parses its .mtx
 * text on every run): header ("APSPCSR1", version, flags, m, n, nnz), the
 * arrays, a 64-bit check word.  Loading verifies the check word and the
 * structure (offsets, column range); APSPIE_EIO otherwise. */
int apspie_csr_save(const char *path, const apspie_host_csr *h);
int apspie_csr_load(const char *path, apspie_host_csr *out);

#define APSPIE_RMAT_UNDIRECTED 1u
#define APSPIE_RMAT_WEIGHTED 2u
/* Seeded RMAT(scale, edge_factor, a=.57, b=.19, c=.19, d=.05) with one
 * MT19937 stream (seed), self-loops and duplicates removed, sorted columns;
 * weights 1 + (genrand_int32() >> 26) from a second stream (weight_seed).
 * Replaces the reference's downloaded datasets (dataset/large/). */
int apspie_rmat(int scale, int edge_factor, uint32_t seed, uint32_t flags, uint32_t weight_seed,
                apspie_host_csr *out);

/* Seeded m x n sparse matrix with power-law row lengths (mean mean_len,
 * at most max_len), uniform distinct sorted columns, FP32 values in (-1, 1):
 * the SpGEMM input (the reference reads .mtx files, test/spgemm/main.cu). */
int apspie_powerlaw_csr(int32_t m, int32_t n, double mean_len, int32_t max_len, uint32_t seed,
                        apspie_host_csr *out);

/* Strict lower triangle of a host CSR (row i keeps j < i), sorted columns:
 * the L of triangle counting.  Replaces buildLower (test/tc/main.cu:39-55). */
int apspie_strict_lower(const apspie_host_csr *a, apspie_host_csr *l);

/* Transpose of a host CSR (at = A^T by rows, i.e. A's CSC): rows of `at` list
 * their entries in increasing source-row order, values moved with them.
 * Replaces csr2csc (src/util.cuh:474-490, cusparseScsr2csc). */
int apspie_host_csr_transpose(const apspie_host_csr *a, apspie_host_csr *at);

/* Host CSR -> device (cudaMemcpy block of test/bfs/main.cu:104-120). */
int apspie_csr_upload(apspie_ctx *ctx, const apspie_host_csr *h, apspie_csr_owned *out);
/* Give an uploaded CSR its CSC: `csc` is the host transpose
 * (apspie_host_csr_transpose), uploaded into c->csc_*; NULL declares the
 * matrix symmetric, values included, and aliases the CSR arrays.  The d_cscColPtrA /
 * d_cscRowIndA / d_cscValA set of test/sssp/main.cu:69-110. */
int apspie_csr_attach_csc(apspie_ctx *ctx, apspie_csr_owned *c, const apspie_host_csr *csc);
/* apspie_csr_attach_csc with the CSC built on the device (apspie_csr2csc). */
int apspie_csr_attach_csc_device(apspie_ctx *ctx, apspie_csr_owned *c);
/* Row offsets d_row_ptr[m+1] of row-sorted COO row indices; APSPIE_EINVAL if
 * the rows are not sorted or out of [0, m).  Replaces coo2csr
 * (src/util.cuh:456-471, cusparseXcoo2csr). */
int apspie_coo_to_csr(apspie_ctx *ctx, const int32_t *d_rows, int64_t nnz, int32_t m, int32_t *d_row_ptr);
/* The CSC of a device CSR into caller-allocated d_col_ptr[n+1],
 * d_row_ind[nnz] and d_val[nnz] (NULL: structure only, as
 * CUSPARSE_ACTION_SYMBOLIC): rows ascending inside each column, values moved
 * with them (the order apspie_host_csr_transpose gives).  Replaces csr2csc
 * (src/util.cuh:473-511, cusparseScsr2csc). */
int apspie_csr2csc(apspie_ctx *ctx, const apspie_csr *a, int32_t *d_col_ptr, int32_t *d_row_ind, float *d_val);
int apspie_csr_free(apspie_ctx *ctx, apspie_csr_owned *c);
/* View an owned CSR as a borrowed one. */
apspie_csr apspie_csr_view(const apspie_csr_owned *c);

/* ---- plain path ----------------------------------------------------------- */

/* Level-synchronous BFS over Boolean OR-AND SpMSpV from `src`; d_depth[m]
 * receives the level of each vertex, -1 if unreached.  Replaces
 * spmspvBfs (src/spmspvBfs.cuh:96) and bfs (src/bfs.cuh:33), which take the
 * level count from the CPU BFS; this runs until the frontier is empty.
 * When g carries its CSC, levels whose frontier holds many edges run
 * bottom-up (each unvisited vertex looks for a parent among its in-edges). */
int apspie_bfs(apspie_ctx *ctx, const apspie_csr *g, int32_t src, int32_t *d_depth,
               apspie_stats *stats);

/* SSSP by synchronous min-plus Bellman-Ford over a changed-vertex frontier,
 * to convergence; d_dist[m] receives distances, APSPIE_SSSP_INF when
 * unreachable.  Weights must be finite and >= 0.  Replaces sssp
 * (src/sssp.cuh:28).  When g carries its CSC, rounds whose frontier holds
 * many edges run as pull steps (each vertex takes the min over its in-edges
 * from the frontier, no atomics); the rounds and their results are the same. */
int apspie_sssp(apspie_ctx *ctx, const apspie_csr *g, int32_t src, float *d_dist,
                apspie_stats *stats);

/* apspie_sssp without any allocation in the call.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  The traversal's scratch
 * is a fixed 48 MiB device array of the library (one per GPU, looked up when
 * a context is created and held until the process ends) instead of the
 * context's pool, so the graph may need at most 48 MiB of it: about 20 B per
 * vertex plus 32 B per 1024 edges (m up to about 2.3 million); larger graphs
 * get APSPIE_ELIMIT.  Push rounds only (the pull and binned knobs are
 * ignored; results and counters are apspie_sssp's).  Per-kernel timing is
 * off for the call.  Calls on one GPU share its array: the library
 * serialises them, and any other use of that memory while a call runs would
 * race. */
int apspie_sssp_noalloc(apspie_ctx *ctx, const apspie_csr *g, int32_t src, float *d_dist, apspie_stats *stats);

/* Round log of the last frontier call (apspie_bfs, apspie_sssp, ...) on this context: for
 * round i (i < 64), the frontier it produced (vertices with out-degree > 0
 * and their out-edges) and whether the next round expanded it by push (0)
 * or pull (1).  *n = rounds written.  The per-level frontier sizes the
 * reference prints in its debug builds (src/spmspvBfs.cuh:209-210). */
typedef struct {
    int32_t frontier;
    int32_t pull;
    int64_t edges;
} apspie_round;
int apspie_ctx_rounds(apspie_ctx *ctx, apspie_round *out, int max, int *n);

/* One min-plus SpMSpV step: frontier = {u : -1 < x[u] < 1e38},
 * y[v] = min_u x[u] + A(u,v), APSPIE_SSSP_INF where nothing reaches.
 * Replaces spmspvSssp (src/spmspvSssp.cuh:92). */
int apspie_sssp_step(apspie_ctx *ctx, const apspie_csr *g, const float *d_x, float *d_y,
                     apspie_stats *stats);

/* Maximal independent set of a symmetric graph with fixed priorities: the
 * result is the greedy MIS that visits vertices by descending key
 * (d_prio[v], -v) -- Luby's rule "join iff your key beats every undecided
 * neighbour" iterated to its fixed point.  d_mis[m] receives 1 (in the set)
 * or 0.  Replaces spmspvMis (src/spmspvMis.cuh:150), whose FP32 cuRAND
 * priorities and `delta` threshold schedule give way to uint32 priorities
 * (the schedule does not change the fixed point). */
int apspie_mis(apspie_ctx *ctx, const apspie_csr *g, const uint32_t *d_prio, int32_t *d_mis,
               apspie_stats *stats);

/* One Boolean SpMSpV: d_out[v] = 1 iff v is in the row of some u with
 * d_in[u] != 0, else 0.  Replaces spmspvCsr (src/spmspvMis.cuh:81), the
 * reference's MIS validity check. */
int apspie_spmspv_or(apspie_ctx *ctx, const apspie_csr *g, const int32_t *d_in, int32_t *d_out);

/* C = A * B (CSR, FP32 values, FP64 accumulation), sorted columns, structural
 * zeros kept.  C's arrays are allocated by the library (release with
 * apspie_csr_free).  stats: edges = products, vertices = nnz(C).  Replaces
 * spgemm(d_matrix *C, d_matrix *A, d_matrix *B, partSize, partNum, ctx)
 * (src/spgemm.cuh:132), whose C fields are callee-allocated the same way
 * (src/spgemmKernel.cuh:669-681), and its cuSPARSE twin
 * spgemm<T>(C, A, B) (src/spgemm.cuh:345). */
int apspie_spgemm(apspie_ctx *ctx, const apspie_csr *a, const apspie_csr *b, apspie_csr_owned *c,
                  apspie_stats *stats);

/* Row offsets from per-row counts: d_row_ptr[0..n] = exclusive scan of the n
 * non-negative d_counts, *total (may be NULL) = their sum, exact in 64 bits.
 * The Scan with a host-returned total that turns row lengths into C's row
 * pointers (src/spgemmKernel.cuh:667).  APSPIE_ELIMIT when the total reaches
 * 2^31 - 4096, the most a CSR here holds (d_row_ptr then holds wrapped
 * offsets); apspie_spgemm refuses such a C the same way before allocating it. */
int apspie_rowptr_from_counts(apspie_ctx *ctx, const int32_t *d_counts, int32_t n, int32_t *d_row_ptr,
                              int64_t *total);

/* C = A * B^T, C(r, c) = sum_k A(r, k) * B(c, k), by the reference's own
 * scheme: row strips of part_size rows of A and of B in DCSC, the column
 * intersection of every strip pair, and the outer products of the matched
 * columns (spgemm, src/spgemm.cuh:131-342; csr_to_dcsc,
 * src/matrix.cpp:286-468).  Kept for comparison with apspie_spgemm; the
 * products are sorted and summed in FP64 in increasing k (deterministic, the
 * same sums as a row-wise product in double) where the reference hashes them
 * with FP32 atomics.  A and B must have the same number of columns; C (m(A) x
 * m(B), sorted columns, structural zeros kept) is allocated by the library.
 * part_size >= 1 and ceil(m(A)/part_size) * ceil(m(B)/part_size) < 2^31.
 * stats: edges = products, vertices = nnz(C), improved = strip-pair column
 * matches.  APSPIE_ELIMIT when the products or matches reach 2^31. */
int apspie_spgemm_outer(apspie_ctx *ctx, const apspie_csr *a, const apspie_csr *b, int32_t part_size,
                        apspie_csr_owned *c, apspie_stats *stats);

/* Triangle count of an undirected graph from its strict lower triangle L
 * (row i holds the sorted columns j < i): *count = sum over (i,k) in L of
 * |L(i,:) & L(k,:)|, exact int64.  stats: edges = wedges of that sum (sum
 * over (i,k) of nnz(L(k,:))), vertices = nnz(L), improved = set probes made
 * (each edge intersects the shorter of its two rows with the longer, so
 * sum over (i,k) of min(nnz(L(i,:)), nnz(L(k,:)))).  Replaces the TC pipeline of test/tc/main.cu:148-261
 * (cuSPARSE L*L^T, ewiseMultTc src/tc.cuh:168-199, float mgpu::Reduce). */
int apspie_tc(apspie_ctx *ctx, const apspie_csr *l, int64_t *count, apspie_stats *stats);

/* ---- general-semiring SpMSpV ----------------------------------------------
 * The reference's mXv (src/mXv.cuh:110-389, dense-in and sparse-in forms):
 * y[c] = (+)_{u present, (u,c) in A} x[u] (*) A(u,c), y = the semiring's
 * identity where nothing reaches.  PLUS_TIMES sums in FP64 and rounds once
 * (identity 0); MIN_PLUS is apspie_sssp_step's min-plus (identity and
 * presence rule the reference's: 1.70141e38, -1 < x < 1e38); MAX_TIMES has
 * identity -inf; OR_AND works on 0 / 1 (identity 0).  Dense form: x[m], u
 * present when x[u] differs from the identity; y[n]. */
#define APSPIE_SR_PLUS_TIMES 0
#define APSPIE_SR_MIN_PLUS 1
#define APSPIE_SR_MAX_TIMES 2
#define APSPIE_SR_OR_AND 3
/* presence of a dense x entry: it differs from the semiring's identity, or
 * (the reference's mXv / bfs frontier, bitify in src/mXv.cuh:50-55) x > 0.5 */
#define APSPIE_MXV_NOT_IDENTITY 0
#define APSPIE_MXV_ABOVE_HALF 1
int apspie_mxv(apspie_ctx *ctx, const apspie_csr *a, int semiring, int presence, const float *d_x, float *d_y,
               apspie_stats *stats);
/* Sparse form: the nx listed indices of x are present (any value); the
 * reached columns come back in ascending order (d_yind, d_yval, room for n)
 * with *ny their count. */
int apspie_mxv_sparse(apspie_ctx *ctx, const apspie_csr *a, int semiring, const int32_t *d_xind,
                      const float *d_xval, int32_t nx, int32_t *d_yind, float *d_yval, int32_t *ny,
                      apspie_stats *stats);

/* ---- batched path ----------------------------------------------------------
 * No counterpart in the reference (it runs one source per call); each result
 * row equals the plain entry point's output for that source. */

/* BFS from k <= 256 sources at once (device array d_srcs[k]); d_depth[k*m]
 * holds source s's depths in row s (-1 unreached). */
int apspie_bfs_batch(apspie_ctx *ctx, const apspie_csr *g, const int32_t *d_srcs, int32_t k, int32_t *d_depth,
                     apspie_stats *stats);

/* SSSP from k <= 64 sources at once; d_dist[k*m] holds source s's distances
 * in row s (APSPIE_SSSP_INF unreachable). */
int apspie_sssp_batch(apspie_ctx *ctx, const apspie_csr *g, const int32_t *d_srcs, int32_t k, float *d_dist,
                      apspie_stats *stats);

/* ---- serving ---------------------------------------------------------------
 * No counterpart in the reference.  A pool of device-resident graphs
 * (borrowed: they must outlive the server); BFS / SSSP requests are queued and
 * served in batches by `lanes` worker threads, each with its own stream, so
 * batches overlap on the device.  Requests that arrive meanwhile join a
 * running batch of their kind and graph that has a free lane, at its next
 * chunk of rounds (continuous batching: distances are the same whatever round
 * a source starts in, and a joined BFS source's depths count from the round
 * that picks it), or else the next batch a lane forms.  Admission is by device-memory footprint against
 * budget_bytes.  A request whose source has no out-edges needs no traversal
 * and is answered on the host without a lane or device memory.  All entry
 * points are thread-safe. */
typedef struct apspie_server apspie_server;
typedef struct {
    int lanes;            /* concurrent batches (default 4) */
    int max_bfs_batch;    /* <= 256 (default 256) */
    int max_sssp_batch;   /* <= 64 (default 64) */
    int plain_bfs_upto;   /* BFS batches of at most this many requests run as plain single-source
                             launches, which are faster there (default 3; < 0: never) */
    int plain_sssp_upto;  /* the same for SSSP (default 1; < 0: never) */
    size_t budget_bytes;  /* device memory for in-flight requests (default 32 GiB) */
} apspie_server_opts;
typedef struct {
    int64_t submitted, completed, rejected, batches;
    double mean_batch;
    double p50_ms, p99_ms, mean_ms, max_ms;
    double qps;      /* completed / (last completion - first arrival) */
    int64_t trivial; /* answered on the host: the source has no out-edges */
    int64_t joined;  /* requests that joined a batch already running (continuous batching) */
} apspie_server_report;

#define APSPIE_REQ_BFS 0
#define APSPIE_REQ_SSSP 1
/* Seconds on the server's steady clock (for arrival times). */
double apspie_now(void);
int apspie_server_create(int device, const apspie_csr *graphs, int ngraphs, const apspie_server_opts *opts,
                         apspie_server **out);
/* Queue one request.  host_out: m int32 depths (BFS) or float distances
 * (SSSP), or NULL to discard.  arrival < 0 means now.  APSPIE_EBUSY when the
 * request's footprint exceeds the whole budget. */
int apspie_server_submit(apspie_server *s, int kind, int graph, int32_t src, void *host_out, double arrival,
                         uint64_t *ticket);
/* Block until the ticket completes; returns its status. */
int apspie_server_wait(apspie_server *s, uint64_t ticket, double *latency_ms);
int apspie_server_drain(apspie_server *s);
int apspie_server_stats(apspie_server *s, apspie_server_report *st);
int apspie_server_reset_stats(apspie_server *s);
/* Sets the traversal knobs of every lane's context (apspie_ctx_set_tuning;
 * a lane running a batch takes them when the batch ends).  Results never
 * depend on them. */
int apspie_server_set_tuning(apspie_server *s, const apspie_tuning *t);
const char *apspie_server_last_error(apspie_server *s);
int apspie_server_destroy(apspie_server *s);

#ifdef __cplusplus
}
#endif
#endif
