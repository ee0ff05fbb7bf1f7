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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
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
// loader: the reference's host calls (src/util.cuh:355-511, src/matrix.cpp:55-111)
// with their signatures, so a driver written like test/bfs/main.cu:40-120 or
// test/sssp/main.cu:25-110 compiles against this header.  Errors throw
// apspie_ref::Error where the reference prints and carries on.

// readEdge, src/util.cuh:355: skips the lines that start with '%' and reads
// "m n edge".  The reference reads stdin whatever file it is given (its
// drivers freopen the .mtx onto stdin); this reads inputFile.
inline void readEdge(int &m, int &n, int &edge, FILE *inputFile) {
    if (!inputFile) throw apspie_ref::Error(APSPIE_EINVAL, "readEdge: null file");
    int c = fgetc(inputFile), old = 0;
    while (c != EOF) {
        if ((old == '\n' || old == 0) && c != '%') {
            ungetc(c, inputFile);
            break;
        }
        old = c;
        c = fgetc(inputFile);
    }
    if (fscanf(inputFile, "%d %d %d", &m, &n, &edge) != 3)
        throw apspie_ref::Error(APSPIE_EIO, "readEdge: no \"m n edge\" line");
}

// readMtx, src/util.cuh:370: `edge` lines "a b [w]" after readEdge; a - 1
// goes to h_cooColInd, b - 1 to h_cooRowInd.  The file is weighted iff the
// character right after the first line's second integer is a space; else
// every value is 1.  Returns whether it is weighted.  Reads stdin, as the
// reference does, unless another stream is given.
template <typename typeVal>
inline bool readMtx(int edge, int *h_cooColInd, int *h_cooRowInd, typeVal *h_cooVal, FILE *in = stdin) {
    bool weighted = true;
    int c = 0;
    for (int j = 0; j < edge; ++j) {
        if (fscanf(in, "%d", &h_cooColInd[j]) != 1 || fscanf(in, "%d", &h_cooRowInd[j]) != 1)
            throw apspie_ref::Error(APSPIE_EIO, "readMtx: not enough rows in mtx file");
        if (j == 0) c = fgetc(in);
        if (c != ' ') {
            h_cooVal[j] = typeVal(1);
            if (j == 0) weighted = false;
        } else {
            float w = 0.f;
            if (fscanf(in, "%f", &w) != 1) throw apspie_ref::Error(APSPIE_EIO, "readMtx: bad value");
            h_cooVal[j] = typeVal(w);
        }
        h_cooRowInd[j]--;
        h_cooColInd[j]--;
    }
    return weighted;
}

// makeSymmetric, src/util.cuh:572: the arrays hold room for 2 * edge
// entries; appends the reversed non-loop entries, sorts by (row, col), drops
// exact repeats; returns the new entry count.  Values move with their entries.
template <typename typeVal>
inline int makeSymmetric(int edge, int *h_cooColInd, int *h_cooRowInd, typeVal *h_cooVal) {
    int64_t out = 0;
    if (std::is_same<typeVal, float>::value) {
        apspie_ref::check(apspie_host_coo_make_symmetric(edge, h_cooColInd, h_cooRowInd, (float *)h_cooVal, &out),
                          nullptr, "makeSymmetric");
    } else {
        std::vector<float> v(size_t(2) * size_t(edge > 0 ? edge : 0));
        for (int i = 0; i < edge; ++i) v[i] = float(h_cooVal[i]);
        apspie_ref::check(apspie_host_coo_make_symmetric(edge, h_cooColInd, h_cooRowInd, v.data(), &out), nullptr,
                          "makeSymmetric");
        for (int64_t i = 0; i < out; ++i) h_cooVal[i] = typeVal(v[i]);
    }
    return int(out);
}

// The host-array buildMatrix test/bfs/main.cu:83 calls: stable counting sort
// of the COO entries by h_cooRowInd into caller-allocated CSR arrays.
template <typename typeVal>
inline void buildMatrix(int *h_csrRowPtr, int *h_csrColInd, typeVal *h_csrVal, const int m, const int edge,
                        const int *h_cooRowInd, const int *h_cooColInd, const typeVal *h_cooVal) {
    std::vector<float> cv(size_t(edge > 0 ? edge : 0)), csr(cv.size());
    for (int i = 0; i < edge; ++i) cv[i] = float(h_cooVal[i]);
    apspie_ref::check(apspie_host_coo_to_csr(m, edge, h_cooRowInd, h_cooColInd, cv.data(), h_csrRowPtr, h_csrColInd,
                                             csr.data()),
                      nullptr, "buildMatrix");
    for (int i = 0; i < edge; ++i) h_csrVal[i] = typeVal(csr[i]);
}

// coo2csr, src/util.cuh:456 (cusparseXcoo2csr): row offsets d_csrRowPtrA[m+1]
// of the row-sorted COO row indices, on the most recently created context.
inline void coo2csr(const int *d_cooRowIndA, const int edge, const int m, int *d_csrRowPtrA,
                    mgpu::CudaContext &context = *mgpu::default_context()) {
    apspie_ref::check(apspie_coo_to_csr(context.get(), d_cooRowIndA, edge, m, d_csrRowPtrA), context.get(), "coo2csr");
}

// csr2csc, src/util.cuh:473 (cusparseScsr2csc) for an m x m matrix: the CSC
// arrays d_cscColPtrA[m+1], d_cscRowIndA[edge] and d_cscValA[edge] (rows
// ascending in each column).  The reference asks cuSPARSE for the structure
// only (CUSPARSE_ACTION_SYMBOLIC) and leaves d_cscValA unwritten; the values
// are moved here too.
template <typename typeVal>
inline void csr2csc(const int m, const int edge, const typeVal *d_csrValA, const int *d_csrRowPtrA,
                    const int *d_csrColIndA, typeVal *d_cscValA, int *d_cscRowIndA, int *d_cscColPtrA,
                    mgpu::CudaContext &context = *mgpu::default_context()) {
    static_assert(sizeof(typeVal) == 4, "FP32 values");
    const apspie_csr g = apspie_ref::csr_of(m, edge, d_csrRowPtrA, d_csrColIndA, (const float *)d_csrValA);
    apspie_ref::check(apspie_csr2csc(context.get(), &g, d_cscColPtrA, d_cscRowIndA,
                                     d_csrValA ? (float *)d_cscValA : nullptr),
                      context.get(), "csr2csc");
}

// d_scratch / allocScratch (src/scratch.hpp:5-31, src/util.cuh:646): the
// reference's per-call scratch arrays.  The context owns a pooled scratch
// here, so the struct is empty and only keeps callers' signatures.
struct d_scratch {};
inline void allocScratch(d_scratch **d, const int /*edge*/, const int /*m*/) { *d = new d_scratch(); }

// mXv, src/mXv.cuh:250 (sparse form): the nnz listed entries
// (d_randVecInd, d_randVecVal) times M, M given by the "csc" arrays read as
// rows: the reached indices come back ascending in d_resultInd with their
// plus-times sums in d_resultVal (room for m), and nnz becomes their count.
// As in the reference, nnz == 0 returns at once.
template <typename T>
inline void mXv(const int *d_randVecInd, const T *d_randVecVal, const int edge, const int m, int &nnz,
                const T *d_cscVal, const int *d_cscColPtr, const int *d_cscRowInd, int *d_resultInd, T *d_resultVal,
                d_scratch * /*d*/, mgpu::CudaContext &context) {
    static_assert(sizeof(T) == sizeof(float), "FP32 vectors");
    if (nnz == 0) return;
    const apspie_csr g = apspie_ref::csr_of(m, edge, d_cscColPtr, d_cscRowInd, (const float *)d_cscVal);
    int32_t ny = 0;
    apspie_ref::check(apspie_mxv_sparse(context.get(), &g, APSPIE_SR_PLUS_TIMES, d_randVecInd,
                                        (const float *)d_randVecVal, nnz, d_resultInd, (float *)d_resultVal, &ny,
                                        nullptr),
                      context.get(), "mXv");
    nnz = ny;
}

// verify, src/util.cuh:261 (with CompareResults :186-258): prints CORRECT or
// the first mismatches and returns their count.  Exact comparison (the
// reference allows FP32 results 10 % off and equates -1 with > 1e38; the
// results here are bit-exact, with 1.70141e38 for unreachable vertices).
template <typename value>
inline int verify(const int m, const value *h_result, const value *h_resultCPU) {
    int errors = 0;
    for (int i = 0; i < m; ++i)
        if (!(h_result[i] == h_resultCPU[i])) {
            if (errors < 10) printf("mismatch at %d: %g vs %g\n", i, double(h_result[i]), double(h_resultCPU[i]));
            errors++;
        }
    printf(errors ? "INCORRECT: %d errors\n" : "CORRECT\n", errors);
    return errors;
}

// The whole loader as one call: readMtx + [makeSymmetric] + buildMatrix
// returning host CSR arrays (apspie_load_mtx).
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
