// apspie_ref.hpp -- the reference's C++ host API (names, argument order and
// meaning of ctcyang/apspie's entry points) over the gfx950 C ABI of
// include/apspie.h.  Header-only; link libapspie.so.
//
// A caller of the reference switches by replacing its includes
// (src/spmspvBfs.cuh, src/bfs.cuh, src/sssp.cuh, src/spmspvSssp.cuh,
// src/spmspvMis.cuh, src/spgemm.cuh, src/matrix.hpp, src/util.cuh) with this
// header; calls keep their form, e.g. test/bfs/main.cu:140
//     bfs(source, edge, m, h_csrRowPtrA, d_csrRowPtrA, d_csrColIndA, d_bfsResult, depth, *context);
// Device pointers are HIP device pointers.  Differences, all documented at the
// functions: loop bounds taken from the CPU (`depth`) are ignored (traversals
// run to convergence), errors throw apspie_ref::Error instead of exit(), and
// MIS priorities are compared as keys (float bits of non-negative priorities
// order like the floats, so cuRAND-style FP32 priorities keep their meaning).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/apspie.h"

namespace apspie_ref {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, apspie_ctx *ctx, const char *what) {
    if (rc != APSPIE_OK)
        throw Error(rc, std::string(what) + ": " + (ctx ? apspie_last_error(ctx) : "error"));
}

}  // namespace apspie_ref

// ---------------------------------------------------------------------------
// context: mgpu::CudaContext / mgpu::CreateCudaDevice (test/bfs/main.cu:27)
namespace mgpu {

class CudaContext {
   public:
    explicit CudaContext(int device) { apspie_ref::check(apspie_ctx_create(device, &ctx_), nullptr, "ctx"); }
    ~CudaContext() { apspie_ctx_destroy(ctx_); }
    CudaContext(const CudaContext &) = delete;
    CudaContext &operator=(const CudaContext &) = delete;
    apspie_ctx *get() const { return ctx_; }

   private:
    apspie_ctx *ctx_ = nullptr;
};

typedef std::shared_ptr<CudaContext> ContextPtr;

// The most recently created context, used by the calls that take none in the
// reference (buildMatrix, matrix_delete, copy_matrix_device).
inline CudaContext *&default_context() {
    static thread_local CudaContext *c = nullptr;
    return c;
}

inline ContextPtr CreateCudaDevice(int device) {
    ContextPtr p = std::make_shared<CudaContext>(device);
    default_context() = p.get();
    return p;
}

}  // namespace mgpu

namespace apspie_ref {

inline apspie_csr csr_of(int m, int edge, const int *row_ptr, const int *col_ind, const float *val) {
    apspie_csr g{};
    g.m = m;
    g.n = m;
    g.nnz = edge;
    g.row_ptr = row_ptr;
    g.col_ind = col_ind;
    g.val = val;
    return g;
}

}  // namespace apspie_ref

// ---------------------------------------------------------------------------
// plain path

// spmspvBfs, src/spmspvBfs.cuh:96.  `depth` (the CPU BFS's level count) is
// ignored; h_csrRowPtr is unused as in the reference.
inline void spmspvBfs(const int vertex, const int edge, const int m, const int * /*h_csrRowPtr*/,
                      const int *d_csrRowPtr, const int *d_csrColInd, int *d_bfsResult, const int /*depth*/,
                      mgpu::CudaContext &context) {
    const apspie_csr g = apspie_ref::csr_of(m, edge, d_csrRowPtr, d_csrColInd, nullptr);
    apspie_ref::check(apspie_bfs(context.get(), &g, vertex, d_bfsResult, nullptr), context.get(), "spmspvBfs");
}

// bfs, src/bfs.cuh:33 (values are all ones, so only the structure matters),
// and the form test/bfs/main.cu:140 calls it with.
template <typename T>
inline void bfs(const int vertex, const int edge, const int m, const T * /*d_cscValA*/, const int *d_cscColPtrA,
                const int *d_cscRowIndA, int *d_bfsResult, const int depth, mgpu::CudaContext &context) {
    spmspvBfs(vertex, edge, m, nullptr, d_cscColPtrA, d_cscRowIndA, d_bfsResult, depth, context);
}
inline void bfs(const int vertex, const int edge, const int m, const int *h_csrRowPtr, const int *d_csrRowPtr,
                const int *d_csrColInd, int *d_bfsResult, const int depth, mgpu::CudaContext &context) {
    spmspvBfs(vertex, edge, m, h_csrRowPtr, d_csrRowPtr, d_csrColInd, d_bfsResult, depth, context);
}

// sssp, src/sssp.cuh:28: distances to convergence with the real weights
// (the reference switches to unit weights after the first step and stops at
// the CPU BFS depth); unreachable = 1.70141e38.
template <typename Value>
inline void sssp(const int vertex, const int edge, const int m, const Value *d_csrValA, const int *d_csrRowPtrA,
                 const int *d_csrColIndA, Value *d_ssspResult, const int /*depth*/, mgpu::CudaContext &context) {
    static_assert(sizeof(Value) == 4, "FP32 distances");
    const apspie_csr g = apspie_ref::csr_of(m, edge, d_csrRowPtrA, d_csrColIndA, (const float *)d_csrValA);
    apspie_ref::check(apspie_sssp(context.get(), &g, vertex, (float *)d_ssspResult, nullptr), context.get(), "sssp");
}

// spmspvSssp, src/spmspvSssp.cuh:92: one min-plus step y = A^T x over the
// frontier {u : -1 < x[u] < 1e38}.
template <typename Value>
inline void spmspvSssp(const Value *d_x, const int edge, const int m, const Value *d_csrVal, const int *d_csrRowPtr,
                       const int *d_csrColInd, Value *d_y, mgpu::CudaContext &context) {
    static_assert(sizeof(Value) == 4, "FP32 vectors");
    const apspie_csr g = apspie_ref::csr_of(m, edge, d_csrRowPtr, d_csrColInd, (const float *)d_csrVal);
    apspie_ref::check(apspie_sssp_step(context.get(), &g, (const float *)d_x, (float *)d_y, nullptr), context.get(),
                      "spmspvSssp");
}

// spmspvMis, src/spmspvMis.cuh:150.  The result is the greedy MIS in
// descending (priority, -id) order, the fixed point the reference's Luby
// rounds aim at; `delta` (threshold schedule) does not change it.  FP32
// priorities must be >= 0 (their bit patterns are used as keys).
template <typename T>
inline void spmspvMis(const int edge, const int m, const int * /*h_csrRowPtr*/, const int *d_csrRowPtr,
                      const int *d_csrColInd, const T *d_randVec, int *d_misResult, const float /*delta*/,
                      mgpu::CudaContext &context) {
    static_assert(sizeof(T) == 4, "32-bit priorities");
    const apspie_csr g = apspie_ref::csr_of(m, edge, d_csrRowPtr, d_csrColInd, nullptr);
    apspie_ref::check(apspie_mis(context.get(), &g, (const uint32_t *)d_randVec, d_misResult, nullptr), context.get(),
                      "spmspvMis");
}

// spmspvCsr, src/spmspvMis.cuh:81: out[v] = 1 iff some u with in[u] != 0 has
// v in its row.
inline void spmspvCsr(const int *d_inputVector, const int edge, const int m, const int *d_csrRowPtr,
                      const int *d_csrColInd, int *d_spmspvResult, mgpu::CudaContext &context) {
    const apspie_csr g = apspie_ref::csr_of(m, edge, d_csrRowPtr, d_csrColInd, nullptr);
    apspie_ref::check(apspie_spmspv_or(context.get(), &g, d_inputVector, d_spmspvResult), context.get(),
                      "spmspvCsr");
}

// ---------------------------------------------------------------------------
// matrices: d_matrix (src/matrix.hpp:6-37).  The reference stores CSR in the
// fields it names csc (row pointer in d_cscColPtr, column ids in d_cscRowInd).
typedef struct matrix {
    int *d_cscColPtr = nullptr;
    int *d_cscRowInd = nullptr;
    float *d_cscVal = nullptr;
    int *h_cscColPtr = nullptr;
    int *h_cscRowInd = nullptr;
    float *h_cscVal = nullptr;
    int m = 0, n = 0, nnz = 0;
    apspie_csr_owned owned{};  // device arrays owned through the library
} d_matrix;

inline void matrix_new(d_matrix *A, int m, int n) {
    *A = d_matrix();
    A->m = m;
    A->n = n;
    A->h_cscColPtr = (int *)calloc(size_t(m) + 1, sizeof(int));
}

inline void matrix_delete(d_matrix *A, mgpu::CudaContext &context = *mgpu::default_context()) {
    if (A->owned.row_ptr) apspie_csr_free(context.get(), &A->owned);
    free(A->h_cscColPtr);
    free(A->h_cscRowInd);
    free(A->h_cscVal);
    *A = d_matrix();
}

// buildMatrix, src/matrix.cpp:55-111: stable counting sort of COO by the
// first index array into CSR, then copy to the device.
template <typename typeVal>
inline void buildMatrix(d_matrix *A, int numEdge, const int *h_cooRowInd, const int *h_cooColInd,
                        const typeVal *h_cooVal, mgpu::CudaContext &context = *mgpu::default_context()) {
    A->nnz = numEdge;
    std::vector<int> cur(size_t(A->m) + 1, 0);
    for (int i = 0; i < numEdge; ++i) cur[h_cooRowInd[i] + 1]++;
    for (int i = 0; i < A->m; ++i) cur[i + 1] += cur[i];
    free(A->h_cscRowInd);
    free(A->h_cscVal);
    A->h_cscRowInd = (int *)malloc(size_t(std::max(numEdge, 1)) * sizeof(int));
    A->h_cscVal = (float *)malloc(size_t(std::max(numEdge, 1)) * sizeof(float));
    memcpy(A->h_cscColPtr, cur.data(), (size_t(A->m) + 1) * sizeof(int));
    for (int i = 0; i < numEdge; ++i) {
        const int d = cur[h_cooRowInd[i]]++;
        A->h_cscRowInd[d] = h_cooColInd[i];
        A->h_cscVal[d] = float(h_cooVal[i]);
    }
    apspie_host_csr h{};
    h.m = A->m;
    h.n = A->n;
    h.nnz = numEdge;
    h.row_ptr = A->h_cscColPtr;
    h.col_ind = A->h_cscRowInd;
    h.val = A->h_cscVal;
    if (A->owned.row_ptr) apspie_csr_free(context.get(), &A->owned);
    apspie_ref::check(apspie_csr_upload(context.get(), &h, &A->owned), context.get(), "buildMatrix");
    A->d_cscColPtr = A->owned.row_ptr;
    A->d_cscRowInd = A->owned.col_ind;
    A->d_cscVal = A->owned.val;
}

// copy_matrix_device, src/matrix.cpp:167: device arrays -> host arrays.
inline void copy_matrix_device(d_matrix *A, mgpu::CudaContext &context = *mgpu::default_context()) {
    free(A->h_cscRowInd);
    free(A->h_cscVal);
    A->h_cscRowInd = (int *)malloc(size_t(std::max(A->nnz, 1)) * sizeof(int));
    A->h_cscVal = (float *)malloc(size_t(std::max(A->nnz, 1)) * sizeof(float));
    apspie_ctx *c = context.get();
    apspie_ref::check(apspie_memcpy_d2h(c, A->h_cscColPtr, A->d_cscColPtr, (size_t(A->m) + 1) * 4), c, "copy");
    apspie_ref::check(apspie_memcpy_d2h(c, A->h_cscRowInd, A->d_cscRowInd, size_t(A->nnz) * 4), c, "copy");
    apspie_ref::check(apspie_memcpy_d2h(c, A->h_cscVal, A->d_cscVal, size_t(A->nnz) * 4), c, "copy");
}

// spgemm, src/spgemm.cuh:132: C(r, c) = sum_k A(r, k) * B(c, k), i.e.
// C = A * B^T with B given by rows (test/spgemm/main.cu passes D = A^T, so
// C = A * A).  B^T is formed on the host from B's host arrays; partSize and
// partNum (the reference's 2D partitioning) are not needed.  C's device
// arrays are allocated here, as the reference does.
inline void spgemm(d_matrix *C, d_matrix *A, d_matrix *B, const int /*partSize*/, const int /*partNum*/,
                   mgpu::CudaContext &context) {
    apspie_ctx *c = context.get();
    // B^T on the host
    std::vector<int> tp(size_t(B->n) + 1, 0), ti(size_t(B->nnz));
    std::vector<float> tv(size_t(B->nnz));
    for (int p = 0; p < B->nnz; ++p) tp[B->h_cscRowInd[p] + 1]++;
    for (int j = 0; j < B->n; ++j) tp[j + 1] += tp[j];
    std::vector<int> cur(tp.begin(), tp.end() - 1);
    for (int r = 0; r < B->m; ++r)
        for (int p = B->h_cscColPtr[r]; p < B->h_cscColPtr[r + 1]; ++p) {
            const int d = cur[B->h_cscRowInd[p]]++;
            ti[d] = r;
            tv[d] = B->h_cscVal ? B->h_cscVal[p] : 1.0f;
        }
    apspie_host_csr hbt{B->n, B->m, B->nnz, tp.data(), ti.data(), tv.data()};
    apspie_csr_owned bt{};
    apspie_ref::check(apspie_csr_upload(c, &hbt, &bt), c, "spgemm B^T");
    const apspie_csr a = apspie_csr_view(&A->owned), b = apspie_csr_view(&bt);
    if (C->owned.row_ptr) apspie_csr_free(c, &C->owned);
    const int rc = apspie_spgemm(c, &a, &b, &C->owned, nullptr);
    apspie_csr_free(c, &bt);
    apspie_ref::check(rc, c, "spgemm");
    C->m = int(C->owned.m);
    C->n = int(C->owned.n);
    C->nnz = int(C->owned.nnz);
    C->d_cscColPtr = C->owned.row_ptr;
    C->d_cscRowInd = C->owned.col_ind;
    C->d_cscVal = C->owned.val;
    if (!C->h_cscColPtr) C->h_cscColPtr = (int *)calloc(size_t(C->m) + 1, sizeof(int));
}

// The reference's spgemm algorithm itself (src/spgemm.cuh:131-342): the same
// C = A * B^T by outer products of matching columns over row strips of
// partSize rows, for comparison with spgemm above.  partNum follows from
// partSize and the row counts.
inline void spgemmOuter(d_matrix *C, d_matrix *A, d_matrix *B, const int partSize, const int /*partNum*/,
                        mgpu::CudaContext &context) {
    apspie_ctx *c = context.get();
    const apspie_csr a = apspie_csr_view(&A->owned), b = apspie_csr_view(&B->owned);
    if (C->owned.row_ptr) apspie_csr_free(c, &C->owned);
    apspie_ref::check(apspie_spgemm_outer(c, &a, &b, partSize, &C->owned, nullptr), c, "spgemmOuter");
    C->m = int(C->owned.m);
    C->n = int(C->owned.n);
    C->nnz = int(C->owned.nnz);
    C->d_cscColPtr = C->owned.row_ptr;
    C->d_cscRowInd = C->owned.col_ind;
    C->d_cscVal = C->owned.val;
    if (!C->h_cscColPtr) C->h_cscColPtr = (int *)calloc(size_t(C->m) + 1, sizeof(int));
}

// Triangle counting, the pipeline of test/tc/main.cu:148-261 on the strict
// lower triangle L (CSR, sorted columns): exact int64 count.
inline long long tc(const int m, const int edgeL, const int *d_rowPtrL, const int *d_colIndL,
                    mgpu::CudaContext &context) {
    const apspie_csr l = apspie_ref::csr_of(m, edgeL, d_rowPtrL, d_colIndL, nullptr);
    int64_t count = 0;
    apspie_ref::check(apspie_tc(context.get(), &l, &count, nullptr), context.get(), "tc");
    return (long long)count;
}

// mXv, src/mXv.cuh:111 (dense form): y = sum over u with x[u] > 0.5 (bitify,
// src/mXv.cuh:50-55) of x[u] * M(u, :), M given by the "csc" arrays read as
// rows, as the reference gathers them; y = 0 where nothing reaches.  Returns
// the expanded edges (the reference's `total`).  The scratch argument of the
// reference is not needed.
template <typename T>
inline int mXv(const T *d_randVec, const int edge, const int m, const T *d_cscVal, const int *d_cscColPtr,
               const int *d_cscRowInd, T *d_mmResult, void * /*d_scratch*/, mgpu::CudaContext &context) {
    static_assert(sizeof(T) == sizeof(float), "FP32 vectors");
    const apspie_csr g = apspie_ref::csr_of(m, edge, d_cscColPtr, d_cscRowInd, (const float *)d_cscVal);
    apspie_stats st{};
    apspie_ref::check(apspie_mxv(context.get(), &g, APSPIE_SR_PLUS_TIMES, APSPIE_MXV_ABOVE_HALF,
                                 (const float *)d_randVec, (float *)d_mmResult, &st),
                      context.get(), "mXv");
    return int(st.edges);
}

// mXm, src/tc.cuh:314 (column-by-column SpGEMM over mXv): C(:, i) = A B(:, i)
// for every column i, all three matrices in CSC (m x m).  In CSC terms that
// is C^T = B^T A^T by rows, one row-wise SpGEMM of B's and A's arrays read as
// rows.  C's arrays are allocated here and handed back through the
// references (the reference allocates them into by-value pointers); returns
// nnz(C).  h_cscColPtrB is not needed.
template <typename T>
inline int mXm(const int edge, const int m, const T *d_cscValA, const int *d_cscColPtrA, const int *d_cscRowIndA,
               const T *d_cscValB, const int * /*h_cscColPtrB*/, const int *d_cscColPtrB, const int *d_cscRowIndB,
               T *&d_cscValC, int *&d_cscColPtrC, int *&d_cscRowIndC, mgpu::CudaContext &context,
               int edgeB = -1) {
    static_assert(sizeof(T) == sizeof(float), "FP32 values");
    apspie_ctx *c = context.get();
    int nnzb = edgeB;
    if (nnzb < 0)
        apspie_ref::check(apspie_memcpy_d2h(c, &nnzb, d_cscColPtrB + m, sizeof(int)), c, "mXm nnz(B)");
    const apspie_csr a = apspie_ref::csr_of(m, edge, d_cscColPtrA, d_cscRowIndA, (const float *)d_cscValA);
    const apspie_csr b = apspie_ref::csr_of(m, nnzb, d_cscColPtrB, d_cscRowIndB, (const float *)d_cscValB);
    apspie_csr_owned out{};
    apspie_ref::check(apspie_spgemm(c, &b, &a, &out, nullptr), c, "mXm");
    d_cscColPtrC = out.row_ptr;
    d_cscRowIndC = out.col_ind;
    d_cscValC = (T *)out.val;
    return int(out.nnz);
}

// ---------------------------------------------------------------------------
// loader: readMtx / makeSymmetric (src/util.cuh:370-643) as one call returning
// host CSR arrays with the reference's conventions.
namespace apspie_ref {

struct HostGraph {
    apspie_host_csr h{};
    HostGraph() = default;
    HostGraph(const HostGraph &) = delete;
    HostGraph &operator=(const HostGraph &) = delete;
    ~HostGraph() { apspie_host_csr_free(&h); }
};

inline void readMtxGraph(const char *path, bool undirected, HostGraph *out) {
    apspie_ref::check(apspie_load_mtx(path, undirected ? APSPIE_LOAD_UNDIRECTED : 0u, &out->h), nullptr, path);
}

}  // namespace apspie_ref
