// Drives the reference-named C++ host API (apspie_amd/host/apspie_ref.hpp)
// the way the reference's test drivers do (test/bfs, sssp, mis, spgemm, tc
// main.cu, plus mXv and mXm) and checks every result against the CPU oracle.  Prints
// "REF_API OK" and exits 0 when all checks pass.
#include <hip/hip_runtime.h>

#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../apspie_amd/host/apspie_ref.hpp"
#include "../../oracle/oracle.h"

static int failures = 0;
#define EXPECT(cond, what)                                   \
    do {                                                     \
        if (!(cond)) {                                       \
            printf("FAIL %s (%s:%d)\n", what, __FILE__, __LINE__); \
            failures++;                                      \
        }                                                    \
    } while (0)

template <class T>
static T *dev_copy(const T *h, size_t n) {
    T *d = nullptr;
    if (hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    if (n) (void)hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice);
    return d;
}

template <class T>
static std::vector<T> host_copy(const T *d, size_t n) {
    std::vector<T> h(n);
    if (n) (void)hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost);
    return h;
}

static orc_csr as_orc(const apspie_host_csr &h) {
    orc_csr o;
    o.m = h.m;
    o.n = h.n;
    o.nnz = h.nnz;
    o.row_ptr = h.row_ptr;
    o.col_ind = h.col_ind;
    o.val = h.val;
    return o;
}


// test/bfs/main.cu:40-150 step by step: readEdge, readMtx, makeSymmetric,
// the host-array buildMatrix, device copies, bfs, verify against the CPU BFS.
// Two slips of that driver are not repeated: it passes makeSymmetric the
// already doubled edge count (writing past its 2 * edge arrays) and reads the
// values into h_csrValA while symmetrizing h_cooValA.
static void bfs_driver(const char *path, bool undirected, mgpu::CudaContext &context) {
    int m = 0, n = 0, edge = 0;
    FILE *f = fopen(path, "r");
    EXPECT(f != nullptr, "open mtx");
    if (!f) return;
    readEdge(m, n, edge, f);
    if (undirected) edge = 2 * edge;
    typedef float typeVal;
    std::vector<typeVal> h_csrValA(edge), h_cooValA(edge);
    std::vector<int> h_csrColIndA(edge), h_csrRowPtrA(m + 1), h_cooColIndA(edge), h_cooRowIndA(edge);
    std::vector<int> h_bfsResult(m), h_bfsResultCPU(m);
    if (undirected) {
        readMtx<typeVal>(edge / 2, h_cooColIndA.data(), h_cooRowIndA.data(), h_cooValA.data(), f);
        edge = makeSymmetric(edge / 2, h_cooColIndA.data(), h_cooRowIndA.data(), h_cooValA.data());
    } else {
        readMtx<typeVal>(edge, h_cooColIndA.data(), h_cooRowIndA.data(), h_cooValA.data(), f);
    }
    fclose(f);
    buildMatrix<typeVal>(h_csrRowPtrA.data(), h_csrColIndA.data(), h_csrValA.data(), m, edge, h_cooRowIndA.data(),
                         h_cooColIndA.data(), h_cooValA.data());
    // the loader in one call gives the same CSR
    apspie_ref::HostGraph hg;
    apspie_ref::readMtxGraph(path, undirected, &hg);
    bool same = hg.h.nnz == edge && hg.h.m == m;
    for (int i = 0; same && i <= m; ++i) same &= hg.h.row_ptr[i] == h_csrRowPtrA[i];
    for (int p = 0; same && p < edge; ++p) same &= hg.h.col_ind[p] == h_csrColIndA[p];
    EXPECT(same, "readEdge + readMtx + makeSymmetric + buildMatrix == apspie_load_mtx");
    int *d_bfsResult = nullptr, *d_csrRowPtrA = dev_copy(h_csrRowPtrA.data(), size_t(m) + 1),
        *d_csrColIndA = dev_copy(h_csrColIndA.data(), size_t(edge));
    (void)hipMalloc(&d_bfsResult, size_t(m) * sizeof(int));
    orc_csr o;
    o.m = m;
    o.n = m;
    o.nnz = edge;
    o.row_ptr = h_csrRowPtrA.data();
    o.col_ind = h_csrColIndA.data();
    o.val = nullptr;
    for (int source = 0; source < m; source += (m > 64 ? 7 : 1)) {
        int depth = 1000;
        depth = orc_bfs(&o, source, depth, h_bfsResultCPU.data(), nullptr, nullptr);  // bfsCPU
        bfs(source, edge, m, h_csrRowPtrA.data(), d_csrRowPtrA, d_csrColIndA, d_bfsResult, depth, context);
        (void)hipMemcpy(h_bfsResult.data(), d_bfsResult, size_t(m) * sizeof(int), hipMemcpyDeviceToHost);
        EXPECT(verify(m, h_bfsResult.data(), h_bfsResultCPU.data()) == 0, "bfs driver");
    }
    for (void *p : {(void *)d_bfsResult, (void *)d_csrRowPtrA, (void *)d_csrColIndA}) (void)hipFree(p);
}

// test/sssp/main.cu:25-130: readMtx straight into the CSR column and value
// arrays (the file lists its rows in order), coo2csr on the device, csr2csc,
// sssp, verify against the CPU SSSP.  The file is written here from a loaded
// graph with integer weights, rows in order.
static void sssp_driver(const char *src_path, mgpu::CudaContext &context) {
    apspie_ref::HostGraph hg;
    apspie_ref::readMtxGraph(src_path, true, &hg);
    char path[] = "/tmp/apspie_sssp_driver_XXXXXX";
    const int fd = mkstemp(path);
    EXPECT(fd >= 0, "temp file");
    if (fd < 0) return;
    FILE *w = fdopen(fd, "w");
    fprintf(w, "%%%%MatrixMarket matrix coordinate real general\n%d %d %lld\n", hg.h.m, hg.h.m, (long long)hg.h.nnz);
    for (int r = 0; r < hg.h.m; ++r)
        for (int p = hg.h.row_ptr[r]; p < hg.h.row_ptr[r + 1]; ++p)
            fprintf(w, "%d %d %d\n", hg.h.col_ind[p] + 1, r + 1, 1 + (p * 37) % 64);
    fclose(w);
    int m = 0, n = 0, edge = 0;
    FILE *f = fopen(path, "r");
    readEdge(m, n, edge, f);
    typedef float Value;
    std::vector<Value> h_csrValA(edge), h_ssspResult(m), h_ssspResultCPU(m);
    std::vector<int> h_csrRowPtrA(m + 1), h_csrColIndA(edge), h_cooRowIndA(edge);
    readMtx<Value>(edge, h_csrColIndA.data(), h_cooRowIndA.data(), h_csrValA.data(), f);
    fclose(f);
    unlink(path);
    Value *d_csrValA = dev_copy(h_csrValA.data(), size_t(edge)), *d_cscValA = nullptr, *d_ssspResult = nullptr;
    int *d_csrColIndA = dev_copy(h_csrColIndA.data(), size_t(edge)),
        *d_cooRowIndA = dev_copy(h_cooRowIndA.data(), size_t(edge)), *d_csrRowPtrA = nullptr, *d_cscRowIndA = nullptr,
        *d_cscColPtrA = nullptr;
    (void)hipMalloc(&d_ssspResult, size_t(m) * sizeof(float));
    (void)hipMalloc(&d_csrRowPtrA, (size_t(m) + 1) * sizeof(int));
    (void)hipMalloc(&d_cscValA, size_t(edge) * sizeof(Value));
    (void)hipMalloc(&d_cscRowIndA, size_t(edge) * sizeof(int));
    (void)hipMalloc(&d_cscColPtrA, (size_t(m) + 1) * sizeof(int));
    coo2csr(d_cooRowIndA, edge, m, d_csrRowPtrA);
    (void)hipMemcpy(h_csrRowPtrA.data(), d_csrRowPtrA, (size_t(m) + 1) * sizeof(int), hipMemcpyDeviceToHost);
    bool rp = true;
    for (int i = 0; i <= m; ++i) rp &= h_csrRowPtrA[i] == hg.h.row_ptr[i];
    EXPECT(rp, "coo2csr");
    csr2csc<Value>(m, edge, d_csrValA, d_csrRowPtrA, d_csrColIndA, d_cscValA, d_cscRowIndA, d_cscColPtrA);
    apspie_host_csr hc{m, m, edge, h_csrRowPtrA.data(), h_csrColIndA.data(), h_csrValA.data()}, ht{};
    apspie_ref::check(apspie_host_csr_transpose(&hc, &ht), nullptr, "transpose");
    EXPECT(host_copy(d_cscColPtrA, size_t(m) + 1) == std::vector<int>(ht.row_ptr, ht.row_ptr + m + 1) &&
               host_copy(d_cscRowIndA, size_t(edge)) == std::vector<int>(ht.col_ind, ht.col_ind + edge) &&
               host_copy(d_cscValA, size_t(edge)) == std::vector<float>(ht.val, ht.val + edge),
           "csr2csc");
    apspie_host_csr_free(&ht);
    orc_csr o;
    o.m = m;
    o.n = m;
    o.nnz = edge;
    o.row_ptr = h_csrRowPtrA.data();
    o.col_ind = h_csrColIndA.data();
    o.val = h_csrValA.data();
    for (int source = 0; source < m; source += (m > 64 ? 11 : 1)) {
        orc_sssp_dijkstra(&o, source, h_ssspResultCPU.data());  // ssspCPU
        for (float &x : h_ssspResultCPU)
            if (x == -1.0f) x = APSPIE_SSSP_INF;
        sssp<Value>(source, edge, m, d_csrValA, d_csrRowPtrA, d_csrColIndA, d_ssspResult, 1000, context);
        (void)hipMemcpy(h_ssspResult.data(), d_ssspResult, size_t(m) * sizeof(float), hipMemcpyDeviceToHost);
        EXPECT(verify<float>(m, h_ssspResult.data(), h_ssspResultCPU.data()) == 0, "sssp driver");
    }
    // mXv, sparse form (src/mXv.cuh:250): listed entries every third vertex
    std::vector<int> xi;
    std::vector<float> xv, xd(m, 0.0f);
    std::vector<uint8_t> listed(m, 0);
    for (int v = 0; v < m; v += 3) {
        xi.push_back(v);
        xv.push_back(0.5f + float(v % 7));
        xd[v] = xv.back();
        listed[v] = 1;
    }
    int nnz = int(xi.size());
    int *d_xi = dev_copy(xi.data(), xi.size()), *d_yi = nullptr;
    float *d_xv = dev_copy(xv.data(), xv.size()), *d_yv = nullptr;
    (void)hipMalloc(&d_yi, size_t(m) * sizeof(int));
    (void)hipMalloc(&d_yv, size_t(m) * sizeof(float));
    d_scratch *d = nullptr;
    allocScratch(&d, edge, m);
    mXv<float>(d_xi, d_xv, edge, m, nnz, d_csrValA, d_csrRowPtrA, d_csrColIndA, d_yi, d_yv, d, context);
    delete d;
    std::vector<float> yo(m);
    std::vector<uint8_t> reached(m);
    orc_mxv(&o, 0, xd.data(), listed.data(), yo.data(), reached.data());
    std::vector<int> wi;
    std::vector<float> wv;
    for (int v = 0; v < m; ++v)
        if (reached[v]) {
            wi.push_back(v);
            wv.push_back(yo[v]);
        }
    EXPECT(nnz == int(wi.size()) && host_copy(d_yi, size_t(nnz)) == wi && host_copy(d_yv, size_t(nnz)) == wv,
           "mXv sparse");
    for (void *p : {(void *)d_csrValA, (void *)d_cscValA, (void *)d_ssspResult, (void *)d_csrColIndA,
                    (void *)d_cooRowIndA, (void *)d_csrRowPtrA, (void *)d_cscRowIndA, (void *)d_cscColPtrA,
                    (void *)d_xi, (void *)d_yi, (void *)d_xv, (void *)d_yv})
        (void)hipFree(p);
}

int main(int argc, char **argv) {
    const char *mtx = argc > 1 ? argv[1] : "tests/golden/mtx/chesapeake.mtx";
    mgpu::ContextPtr context = mgpu::CreateCudaDevice(0);
    try {
        bfs_driver(mtx, false, *context);
        bfs_driver(mtx, true, *context);
        sssp_driver(mtx, *context);
        // BFS (test/bfs/main.cu): undirected chesapeake, every source
        apspie_ref::HostGraph hg;
        apspie_ref::readMtxGraph(mtx, true, &hg);
        const apspie_host_csr &h = hg.h;
        const int m = h.m, edge = int(h.nnz);
        int *d_rp = dev_copy(h.row_ptr, size_t(m) + 1), *d_ci = dev_copy(h.col_ind, size_t(edge));
        int *d_res = nullptr;
        (void)hipMalloc(&d_res, size_t(m) * 4);
        orc_csr o = as_orc(h);
        std::vector<int> want(m);
        for (int src = 0; src < m; ++src) {
            bfs(src, edge, m, h.row_ptr, d_rp, d_ci, d_res, 1000, *context);
            orc_bfs(&o, src, 1 << 30, want.data(), nullptr, nullptr);
            EXPECT(host_copy(d_res, m) == want, "bfs");
        }
        // SSSP (test/sssp/main.cu) with integer weights
        std::vector<float> w(edge);
        for (int p = 0; p < edge; ++p) w[p] = float(1 + (p * 7919) % 64);
        float *d_w = dev_copy(w.data(), size_t(edge)), *d_dist = nullptr;
        (void)hipMalloc(&d_dist, size_t(m) * 4);
        o.val = w.data();
        std::vector<float> wd(m);
        for (int src = 0; src < m; src += 5) {
            sssp<float>(src, edge, m, d_w, d_rp, d_ci, d_dist, 1000, *context);
            orc_sssp_dijkstra(&o, src, wd.data());
            std::vector<float> got = host_copy(d_dist, m);
            bool ok = true;
            for (int v = 0; v < m; ++v) ok &= wd[v] == -1.0f ? got[v] == APSPIE_SSSP_INF : got[v] == wd[v];
            EXPECT(ok, "sssp");
        }
        // MIS (test/mis/main.cu) with FP32 priorities, checked by spmspvCsr +
        // the reference's XOR rule (src/util.cuh:277)
        std::vector<float> prio(m);
        for (int v = 0; v < m; ++v) prio[v] = float((v * 2654435761u) % 1000) / 1000.0f + 0.001f;
        float *d_prio = dev_copy(prio.data(), size_t(m));
        int *d_mis = nullptr, *d_has = nullptr;
        (void)hipMalloc(&d_mis, size_t(m) * 4);
        (void)hipMalloc(&d_has, size_t(m) * 4);
        spmspvMis<float>(edge, m, h.row_ptr, d_rp, d_ci, d_prio, d_mis, 0.01f, *context);
        spmspvCsr(d_mis, edge, m, d_rp, d_ci, d_has, *context);
        std::vector<int> mis = host_copy(d_mis, m), has = host_copy(d_has, m);
        std::vector<uint32_t> keys(m);
        for (int v = 0; v < m; ++v) memcpy(&keys[v], &prio[v], 4);
        std::vector<int> wm(m);
        o.val = nullptr;
        orc_mis_greedy(&o, keys.data(), wm.data());
        EXPECT(mis == wm, "mis");
        bool xr = true;
        for (int v = 0; v < m; ++v) xr &= mis[v] + has[v] == 1;
        EXPECT(xr, "mis xor check");
        // SpGEMM (test/spgemm/main.cu): E = A * D^T with D = A^T built from
        // swapped COO arrays, i.e. E = A * A
        std::vector<int> cr, cc;
        std::vector<float> cv;
        for (int r = 0; r < m; ++r)
            for (int p = h.row_ptr[r]; p < h.row_ptr[r + 1]; ++p) {
                cr.push_back(r);
                cc.push_back(h.col_ind[p]);
                cv.push_back(0.25f + float(p % 5));
            }
        d_matrix A, D, E;
        matrix_new(&A, m, m);
        matrix_new(&D, m, m);
        matrix_new(&E, m, m);
        buildMatrix<float>(&A, edge, cr.data(), cc.data(), cv.data());
        buildMatrix<float>(&D, edge, cc.data(), cr.data(), cv.data());
        spgemm(&E, &A, &D, 0, 0, *context);
        copy_matrix_device(&E);
        orc_csr oa;
        oa.m = m;
        oa.n = m;
        oa.nnz = edge;
        oa.row_ptr = A.h_cscColPtr;
        oa.col_ind = A.h_cscRowInd;
        oa.val = A.h_cscVal;
        orc_csr oc;
        int64_t prods = 0;
        orc_spgemm(&oa, &oa, &oc, &prods);
        bool sg = E.nnz == oc.nnz;
        for (int i = 0; sg && i <= m; ++i) sg &= E.h_cscColPtr[i] == oc.row_ptr[i];
        for (int p = 0; sg && p < E.nnz; ++p)
            sg &= E.h_cscRowInd[p] == oc.col_ind[p] &&
                  std::fabs(E.h_cscVal[p] - oc.val[p]) <= 1e-5f * std::fabs(oc.val[p]);
        EXPECT(sg, "spgemm");
        // the reference's own algorithm, strips of 3 rows: the same product
        d_matrix E2;
        matrix_new(&E2, m, m);
        spgemmOuter(&E2, &A, &D, 3, (m + 2) / 3, *context);
        copy_matrix_device(&E2);
        bool so = E2.nnz == oc.nnz;
        for (int i = 0; so && i <= m; ++i) so &= E2.h_cscColPtr[i] == oc.row_ptr[i];
        for (int p = 0; so && p < E2.nnz; ++p)
            so &= E2.h_cscRowInd[p] == oc.col_ind[p] && E2.h_cscVal[p] == oc.val[p];
        EXPECT(so, "spgemmOuter");
        orc_csr_free(&oc);
        // TC (test/tc/main.cu) on the strict lower triangle
        apspie_host_csr l{};
        apspie_ref::check(apspie_strict_lower(&h, &l), nullptr, "lower");
        int *d_lrp = dev_copy(l.row_ptr, size_t(m) + 1), *d_lci = dev_copy(l.col_ind, size_t(l.nnz));
        orc_csr ol = as_orc(l);
        const long long t = tc(m, int(l.nnz), d_lrp, d_lci, *context);
        EXPECT(t == orc_tc(&ol, nullptr), "tc");
        apspie_host_csr_free(&l);
        // mXv (src/mXv.cuh:111): entries above 0.5 are the frontier
        std::vector<float> hx(m), xh(m), yo(m);
        long long want_tot = 0;
        for (int v = 0; v < m; ++v) {
            hx[v] = v % 3 == 0 ? 1.0f + float(v % 4) : (v % 3 == 1 ? 0.25f : 0.0f);
            xh[v] = hx[v] > 0.5f ? hx[v] : 0.0f;
            if (hx[v] > 0.5f) want_tot += h.row_ptr[v + 1] - h.row_ptr[v];
        }
        float *d_x = dev_copy(hx.data(), size_t(m)), *d_av = dev_copy(cv.data(), size_t(edge)), *d_y = nullptr;
        (void)hipMalloc(&d_y, size_t(m) * 4);
        const int tot = mXv(d_x, edge, m, d_av, d_rp, d_ci, d_y, nullptr, *context);
        orc_csr ov;
        ov.m = m;
        ov.n = m;
        ov.nnz = edge;
        ov.row_ptr = h.row_ptr;
        ov.col_ind = h.col_ind;
        ov.val = cv.data();
        orc_mxv(&ov, 0, xh.data(), nullptr, yo.data(), nullptr);
        EXPECT(host_copy(d_y, size_t(m)) == yo && tot == want_tot, "mXv");
        // mXm (src/tc.cuh:314): A's arrays read as CSC twice; the result's
        // arrays read as rows are A * A
        float *d_cval = nullptr;
        int *d_ccp = nullptr, *d_cri = nullptr;
        const int nnzc = mXm(edge, m, d_av, d_rp, d_ci, d_av, h.row_ptr, d_rp, d_ci, d_cval, d_ccp, d_cri, *context);
        orc_csr om;
        int64_t pm = 0;
        orc_spgemm(&ov, &ov, &om, &pm);
        const std::vector<int> gcp = host_copy(d_ccp, size_t(m) + 1), gri = host_copy(d_cri, size_t(nnzc));
        const std::vector<float> gcv = host_copy(d_cval, size_t(nnzc));
        bool mm = nnzc == om.nnz;
        for (int i = 0; mm && i <= m; ++i) mm &= gcp[i] == om.row_ptr[i];
        for (int q = 0; mm && q < nnzc; ++q)
            mm &= gri[q] == om.col_ind[q] && std::fabs(gcv[q] - om.val[q]) <= 1e-5f * std::fabs(om.val[q]);
        EXPECT(mm, "mXm");
        orc_csr_free(&om);
        for (void *q : {(void *)d_x, (void *)d_av, (void *)d_y, (void *)d_cval, (void *)d_ccp, (void *)d_cri})
            (void)hipFree(q);
        matrix_delete(&A);
        matrix_delete(&D);
        matrix_delete(&E);
        for (void *p : {(void *)d_rp, (void *)d_ci, (void *)d_res, (void *)d_w, (void *)d_dist, (void *)d_prio,
                        (void *)d_mis, (void *)d_has, (void *)d_lrp, (void *)d_lci})
            (void)hipFree(p);
    } catch (const std::exception &e) {
        printf("FAIL exception: %s\n", e.what());
        return 2;
    }
    if (failures) {
        printf("%d failures\n", failures);
        return 1;
    }
    printf("REF_API OK\n");
    return 0;
}
