/*
 * oracle.c -- CPU restatement of apspie's CPU reference paths.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Each function names the reference file:line it restates.  Paths are
 * relative to the apspie checkout.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* MT19937: src/mt19937ar.cpp:57-75 (init_genrand), :102-137 (genrand_int32) */
/* Written from the published recurrence x_{k+n} = x_{k+m} ^ ((x_k^u | x_{k+1}^l) A)
 * with n=624, m=397, r=31, a=0x9908b0df and the standard tempering. */

void orc_mt_seed(orc_mt *g, uint32_t seed) {
    g->s[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->s[i] = 1812433253u * (g->s[i - 1] ^ (g->s[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

static void mt_twist(orc_mt *g) {
    for (int i = 0; i < 624; ++i) {
        uint32_t x = (g->s[i] & 0x80000000u) | (g->s[(i + 1) % 624] & 0x7fffffffu);
        uint32_t xa = x >> 1;
        if (x & 1u) xa ^= 0x9908b0dfu;
        g->s[i] = g->s[(i + 397) % 624] ^ xa;
    }
    g->idx = 0;
}

uint32_t orc_mt_next(orc_mt *g) {
    if (g->idx >= 624) mt_twist(g);
    uint32_t y = g->s[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

void orc_mt_fill(uint32_t seed, int64_t n, uint32_t *out) {
    orc_mt g;
    orc_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = orc_mt_next(&g);
}

void orc_csr_free(orc_csr *a) {
    if (!a) return;
    free(a->row_ptr);
    free(a->col_ind);
    free(a->val);
    a->row_ptr = NULL;
    a->col_ind = NULL;
    a->val = NULL;
}

/* ------------------------------------------------------------------------ */
/* Loader */

typedef struct {
    int32_t row, col;
    float val;
    int64_t pos;
} coo_ent;

static int coo_cmp(const void *x, const void *y) {
    const coo_ent *a = (const coo_ent *)x, *b = (const coo_ent *)y;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->col != b->col) return a->col < b->col ? -1 : 1;
    /* position tie-break makes qsort stable */
    return a->pos < b->pos ? -1 : (a->pos > b->pos);
}

/* buildMatrix, src/matrix.cpp:55-111: stable counting sort by row. */
static int build_csr(int32_t m, int32_t n, int64_t nnz, const int32_t *row, const int32_t *col,
                     const float *val, orc_csr *out) {
    out->m = m;
    out->n = n;
    out->nnz = nnz;
    out->row_ptr = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
    out->col_ind = (int32_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
    out->val = (float *)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    if (!out->row_ptr || !out->col_ind || !out->val) return -2;
    for (int64_t i = 0; i < nnz; ++i) {
        if (row[i] < 0 || row[i] >= m || col[i] < 0 || col[i] >= n) return -3;
        out->row_ptr[row[i] + 1]++;
    }
    for (int32_t i = 0; i < m; ++i) out->row_ptr[i + 1] += out->row_ptr[i];
    int32_t *cur = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    memcpy(cur, out->row_ptr, ((size_t)m + 1) * sizeof(int32_t));
    for (int64_t i = 0; i < nnz; ++i) {
        int32_t d = cur[row[i]]++;
        out->col_ind[d] = col[i];
        out->val[d] = val ? val[i] : 1.0f;
    }
    free(cur);
    return 0;
}

/* readEdge src/util.cuh:355-367, readMtx :370-430, makeSymmetric :572-643,
 * then buildMatrix.  Line "a b [w]" stores a-1 in the column and b-1 in the
 * row (readMtx's first scanf fills h_cooColInd).  The graph is weighted iff
 * the character after the first line's second integer is a space. */
int orc_read_mtx(const char *path, int undirected, orc_csr *out) {
    memset(out, 0, sizeof(*out));
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    int c = fgetc(f), old = 0;
    while (c != EOF) {
        if ((old == '\n' || old == 0) && c != '%') {
            ungetc(c, f);
            break;
        }
        old = c;
        c = fgetc(f);
    }
    int m, n, nnz;
    if (fscanf(f, "%d %d %d", &m, &n, &nnz) != 3 || m < 0 || n < 0 || nnz < 0) {
        fclose(f);
        return -4;
    }
    int64_t cap = undirected ? 2 * (int64_t)nnz : nnz;
    if (cap == 0) cap = 1;
    int32_t *row = (int32_t *)malloc(cap * sizeof(int32_t));
    int32_t *col = (int32_t *)malloc(cap * sizeof(int32_t));
    float *val = (float *)malloc(cap * sizeof(float));
    int ch = 0;
    for (int j = 0; j < nnz; ++j) {
        if (fscanf(f, "%d", &col[j]) != 1 || fscanf(f, "%d", &row[j]) != 1) {
            free(row); free(col); free(val); fclose(f);
            return -5; /* "not enough rows in mtx file" */
        }
        if (j == 0) ch = fgetc(f);
        if (ch != ' ') {
            val[j] = 1.0f;
        } else if (fscanf(f, "%f", &val[j]) != 1) {
            free(row); free(col); free(val); fclose(f);
            return -5;
        }
        row[j]--;
        col[j]--;
    }
    fclose(f);
    int64_t e = nnz;
    if (undirected) {
        /* append reversed non-loop entries (util.cuh:576-586) */
        int64_t shift = 0;
        for (int64_t i = 0; i < nnz; ++i) {
            if (col[i] != row[i]) {
                row[nnz + i - shift] = col[i];
                col[nnz + i - shift] = row[i];
                val[nnz + i - shift] = val[i];
            } else {
                shift++;
            }
        }
        e = 2 * (int64_t)nnz - shift;
        /* sort by (row, col) (util.cuh:592-593), drop exact repeats
         * (util.cuh:602-641).  Self-loops stay (their removal is commented
         * out in the reference).  Values stay with their entries. */
        coo_ent *t = (coo_ent *)malloc((e ? e : 1) * sizeof(coo_ent));
        for (int64_t i = 0; i < e; ++i) {
            t[i].row = row[i]; t[i].col = col[i]; t[i].val = val[i]; t[i].pos = i;
        }
        qsort(t, e, sizeof(coo_ent), coo_cmp);
        int64_t k = 0;
        for (int64_t i = 0; i < e; ++i) {
            if (k > 0 && t[i].row == row[k - 1] && t[i].col == col[k - 1]) continue;
            row[k] = t[i].row; col[k] = t[i].col; val[k] = t[i].val; ++k;
        }
        free(t);
        e = k;
    }
    int rc = build_csr(m, n, e, row, col, val, out);
    free(row); free(col); free(val);
    if (rc) orc_csr_free(out);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* RMAT generator (synthetic input; the reference downloads its graphs). */

static int cmp_u64(const void *x, const void *y) {
    uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
    return a < b ? -1 : (a > b);
}

int orc_rmat(int scale, int edge_factor, uint32_t seed, int undirected, orc_csr *out) {
    memset(out, 0, sizeof(*out));
    if (scale < 1 || scale > 30 || edge_factor < 1) return -1;
    int64_t m = (int64_t)1 << scale;
    int64_t ne = m * edge_factor;
    uint64_t *keys = (uint64_t *)malloc((undirected ? 2 : 1) * ne * sizeof(uint64_t));
    if (!keys) return -2;
    orc_mt g;
    orc_mt_seed(&g, seed);
    int64_t k = 0;
    for (int64_t e = 0; e < ne; ++e) {
        uint32_t s = 0, d = 0;
        for (int l = 0; l < scale; ++l) {
            uint32_t bit = 1u << (scale - 1 - l);
            double u = (double)orc_mt_next(&g) * (1.0 / 4294967296.0);
            if (u < 0.57) {
            } else if (u < 0.76) {
                d |= bit;
            } else if (u < 0.95) {
                s |= bit;
            } else {
                s |= bit;
                d |= bit;
            }
        }
        if (s == d) continue;
        keys[k++] = ((uint64_t)s << 32) | d;
        if (undirected) keys[k++] = ((uint64_t)d << 32) | s;
    }
    qsort(keys, k, sizeof(uint64_t), cmp_u64);
    int64_t u = 0;
    for (int64_t i = 0; i < k; ++i)
        if (u == 0 || keys[i] != keys[u - 1]) keys[u++] = keys[i];
    out->m = (int32_t)m;
    out->n = (int32_t)m;
    out->nnz = u;
    out->row_ptr = (int32_t *)calloc(m + 1, sizeof(int32_t));
    out->col_ind = (int32_t *)malloc((u ? u : 1) * sizeof(int32_t));
    out->val = NULL;
    for (int64_t i = 0; i < u; ++i) {
        out->row_ptr[(keys[i] >> 32) + 1]++;
        out->col_ind[i] = (int32_t)(keys[i] & 0xffffffffu);
    }
    for (int64_t i = 0; i < m; ++i) out->row_ptr[i + 1] += out->row_ptr[i];
    free(keys);
    return 0;
}

int orc_rmat_weights(orc_csr *a, uint32_t seed) {
    free(a->val);
    a->val = (float *)malloc((a->nnz ? a->nnz : 1) * sizeof(float));
    if (!a->val) return -2;
    orc_mt g;
    orc_mt_seed(&g, seed);
    for (int64_t p = 0; p < a->nnz; ++p) a->val[p] = (float)(1u + (orc_mt_next(&g) >> 26));
    return 0;
}

int32_t orc_argmax_degree(const orc_csr *a) {
    int32_t best = 0, bd = -1;
    for (int32_t v = 0; v < a->m; ++v) {
        int32_t d = a->row_ptr[v + 1] - a->row_ptr[v];
        if (d > bd) { bd = d; best = v; }
    }
    return best;
}

/* ------------------------------------------------------------------------ */
/* BFS: SimpleReferenceBfs, src/testBfs.cpp:6-72 (FIFO queue, -1 unreached,
 * returns max depth + 1, stops expanding beyond `stop`). */
int32_t orc_bfs(const orc_csr *a, int32_t src, int32_t stop, int32_t *depth,
                int64_t *e_exp, int64_t *v_reached) {
    int32_t m = a->m;
    for (int32_t i = 0; i < m; ++i) depth[i] = -1;
    int32_t *q = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    int64_t head = 0, tail = 0, ee = 0;
    depth[src] = 0;
    q[tail++] = src;
    int32_t search_depth = 0;
    while (head < tail) {
        int32_t u = q[head++];
        int32_t nd = depth[u] + 1;
        if (nd > stop) break;
        int32_t b = a->row_ptr[u], e = a->row_ptr[u + 1];
        ee += e - b;
        for (int32_t p = b; p < e; ++p) {
            int32_t v = a->col_ind[p];
            if (depth[v] == -1) {
                depth[v] = nd;
                if (search_depth < nd) search_depth = nd;
                q[tail++] = v;
            }
        }
    }
    free(q);
    if (e_exp) *e_exp = ee;
    if (v_reached) *v_reached = tail;
    return search_depth + 1;
}

/* ------------------------------------------------------------------------ */
/* Binary min-heap of (dist bits << 32 | vertex): non-negative floats order
 * like their bit patterns. */
typedef struct {
    uint64_t *h;
    int64_t n, cap;
} heap64;

static void hpush(heap64 *hp, uint64_t x) {
    if (hp->n == hp->cap) {
        hp->cap = hp->cap ? 2 * hp->cap : 1024;
        hp->h = (uint64_t *)realloc(hp->h, hp->cap * sizeof(uint64_t));
    }
    int64_t i = hp->n++;
    while (i > 0) {
        int64_t p = (i - 1) >> 1;
        if (hp->h[p] <= x) break;
        hp->h[i] = hp->h[p];
        i = p;
    }
    hp->h[i] = x;
}

static uint64_t hpop(heap64 *hp) {
    uint64_t top = hp->h[0], x = hp->h[--hp->n];
    int64_t i = 0;
    for (;;) {
        int64_t c = 2 * i + 1;
        if (c >= hp->n) break;
        if (c + 1 < hp->n && hp->h[c + 1] < hp->h[c]) c++;
        if (hp->h[c] >= x) break;
        hp->h[i] = hp->h[c];
        i = c;
    }
    if (hp->n) hp->h[i] = x;
    return top;
}

static uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* SimpleReferenceSssp, src/testSssp.cpp:33-115: priority-queue Dijkstra that
 * relaxes with the vertex's current distance and pushes every improvement.
 * Its initial pushes of (i, h_csrValA[i]) for i != src (:49-56) never pass the
 * guard at :88 while i is unreached and only repeat relaxations once it is, so
 * they are left out; popped entries older than the vertex's current distance
 * are skipped, which relaxes the same edges with the same distances. */
int32_t orc_sssp_dijkstra(const orc_csr *a, int32_t src, float *dist) {
    int32_t m = a->m;
    for (int32_t i = 0; i < m; ++i) dist[i] = -1.0f;
    dist[src] = 0.0f;
    heap64 hp = {0, 0, 0};
    hpush(&hp, ((uint64_t)f2u(0.0f) << 32) | (uint32_t)src);
    while (hp.n) {
        uint64_t t = hpop(&hp);
        int32_t v = (int32_t)(t & 0xffffffffu);
        float d = u2f((uint32_t)(t >> 32));
        if (d > dist[v]) continue;
        for (int32_t p = a->row_ptr[v]; p < a->row_ptr[v + 1]; ++p) {
            int32_t w = a->col_ind[p];
            float alt = dist[v] + (a->val ? a->val[p] : 1.0f);
            if (dist[w] == -1.0f || alt < dist[w]) {
                dist[w] = alt;
                hpush(&hp, ((uint64_t)f2u(alt) << 32) | (uint32_t)w);
            }
        }
    }
    free(hp.h);
    return 0;
}

int32_t orc_sssp_bf(const orc_csr *a, int32_t src, float *dist, int64_t *e_relax,
                    int64_t *sum_f, int64_t *r_imp) {
    int32_t m = a->m;
    const float INF = 3.0e38f;
    float *cur = dist;
    float *nxt = (float *)malloc(((size_t)m + 1) * sizeof(float));
    int32_t *f = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    int32_t *g = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    for (int32_t i = 0; i < m; ++i) cur[i] = nxt[i] = INF;
    cur[src] = nxt[src] = 0.0f;
    int64_t nf = 1, er = 0, sf = 0, ri = 0;
    f[0] = src;
    int32_t iters = 0;
    while (nf > 0) {
        iters++;
        sf += nf;
        for (int64_t i = 0; i < nf; ++i) {
            int32_t u = f[i];
            er += a->row_ptr[u + 1] - a->row_ptr[u];
            for (int32_t p = a->row_ptr[u]; p < a->row_ptr[u + 1]; ++p) {
                int32_t v = a->col_ind[p];
                float nd = cur[u] + (a->val ? a->val[p] : 1.0f);
                if (nd < nxt[v]) nxt[v] = nd;
            }
        }
        /* next frontier = vertices whose distance dropped, in id order */
        int64_t ng = 0;
        for (int32_t v = 0; v < m; ++v)
            if (nxt[v] < cur[v]) {
                g[ng++] = v;
                cur[v] = nxt[v];
            }
        ri += ng;
        int32_t *t = f; f = g; g = t;
        nf = ng;
    }
    for (int32_t i = 0; i < m; ++i)
        if (cur[i] == INF) cur[i] = -1.0f;
    free(nxt); free(f); free(g);
    if (e_relax) *e_relax = er;
    if (sum_f) *sum_f = sf;
    if (r_imp) *r_imp = ri;
    return iters;
}

/* ------------------------------------------------------------------------ */
/* MIS */

static void mis_visit(const orc_csr *a, int32_t v, int32_t *state) {
    if (state[v] != -1) return;
    state[v] = 1;
    for (int32_t p = a->row_ptr[v]; p < a->row_ptr[v + 1]; ++p) {
        int32_t u = a->col_ind[p];
        if (state[u] == -1) state[u] = 0;
    }
}

/* Greedy MIS over vertices in descending key (prio[v], -v): the loop of
 * test/mis/main.cu:55-79 with index order replaced by key order.  This is the
 * fixed point of Luby's rule "join iff your key beats every undecided
 * neighbour" (lexicographically-first MIS). */
static int cmp_u64_desc(const void *x, const void *y) { return -cmp_u64(x, y); }

void orc_mis_greedy(const orc_csr *a, const uint32_t *prio, int32_t *out) {
    int32_t m = a->m;
    uint64_t *order = (uint64_t *)malloc(((size_t)m + 1) * sizeof(uint64_t));
    for (int32_t v = 0; v < m; ++v)
        order[v] = ((uint64_t)prio[v] << 32) | (uint32_t)(0xffffffffu - (uint32_t)v);
    qsort(order, m, sizeof(uint64_t), cmp_u64_desc);
    for (int32_t v = 0; v < m; ++v) out[v] = -1;
    for (int32_t i = 0; i < m; ++i) {
        int32_t v = (int32_t)(0xffffffffu - (uint32_t)(order[i] & 0xffffffffu));
        mis_visit(a, v, out);
    }
    free(order);
}

/* SimpleReferenceMis, test/mis/main.cu:26-88. */
void orc_mis_index(const orc_csr *a, int32_t src, int32_t *out) {
    for (int32_t v = 0; v < a->m; ++v) out[v] = -1;
    mis_visit(a, src, out);
    for (int32_t v = 0; v < a->m; ++v) mis_visit(a, v, out);
}

void orc_spmspv_or(const orc_csr *a, const int32_t *in, int32_t *out) {
    for (int32_t v = 0; v < a->m; ++v) out[v] = 0;
    for (int32_t u = 0; u < a->m; ++u) {
        if (!in[u]) continue;
        for (int32_t p = a->row_ptr[u]; p < a->row_ptr[u + 1]; ++p) out[a->col_ind[p]] = 1;
    }
}

void orc_sssp_step(const orc_csr *a, const float *x, float *y) {
    const float SENT = 1.70141e38f; /* src/spmspvSssp.cuh:26,214 */
    for (int32_t v = 0; v < a->m; ++v) y[v] = SENT;
    for (int32_t u = 0; u < a->m; ++u) {
        if (!(x[u] > -1.0f && x[u] < 1e38f)) continue; /* bitify, :31-36 */
        for (int32_t p = a->row_ptr[u]; p < a->row_ptr[u + 1]; ++p) {
            float nd = x[u] + (a->val ? a->val[p] : 1.0f);
            int32_t v = a->col_ind[p];
            if (nd < y[v]) y[v] = nd;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* SpGEMM: row-wise Gustavson with a dense double accumulator. */
static int cmp_i32(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return a < b ? -1 : (a > b);
}

int orc_spgemm(const orc_csr *a, const orc_csr *b, orc_csr *c, int64_t *products) {
    memset(c, 0, sizeof(*c));
    if (a->n != b->m) return -1;
    int32_t m = a->m, n = b->n;
    double *acc = (double *)calloc((size_t)n + 1, sizeof(double));
    int32_t *mark = (int32_t *)malloc(((size_t)n + 1) * sizeof(int32_t));
    int32_t *list = (int32_t *)malloc(((size_t)n + 1) * sizeof(int32_t));
    for (int32_t j = 0; j < n; ++j) mark[j] = -1;
    int64_t cap = a->nnz + 16, nnz = 0, prod = 0;
    c->m = m;
    c->n = n;
    c->row_ptr = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    c->col_ind = (int32_t *)malloc(cap * sizeof(int32_t));
    c->val = (float *)malloc(cap * sizeof(float));
    c->row_ptr[0] = 0;
    for (int32_t i = 0; i < m; ++i) {
        int32_t cnt = 0;
        for (int32_t p = a->row_ptr[i]; p < a->row_ptr[i + 1]; ++p) {
            int32_t k = a->col_ind[p];
            if (k < 0 || k >= b->m) continue;  /* a column of A outside B's rows: no products */
            double av = a->val ? a->val[p] : 1.0;
            prod += b->row_ptr[k + 1] - b->row_ptr[k];
            for (int32_t q = b->row_ptr[k]; q < b->row_ptr[k + 1]; ++q) {
                int32_t j = b->col_ind[q];
                if (j < 0 || j >= n) continue; /* a column outside B: dropped, as the product drops it */
                double bv = b->val ? b->val[q] : 1.0;
                if (mark[j] != i) {
                    mark[j] = i;
                    acc[j] = 0.0;
                    list[cnt++] = j;
                }
                acc[j] += av * bv;
            }
        }
        qsort(list, cnt, sizeof(int32_t), cmp_i32);
        if (nnz + cnt > cap) {
            while (nnz + cnt > cap) cap *= 2;
            c->col_ind = (int32_t *)realloc(c->col_ind, cap * sizeof(int32_t));
            c->val = (float *)realloc(c->val, cap * sizeof(float));
        }
        for (int32_t t = 0; t < cnt; ++t) {
            c->col_ind[nnz] = list[t];
            c->val[nnz] = (float)acc[list[t]];
            nnz++;
        }
        if (nnz > 0x7fffffff) return -2;
        c->row_ptr[i + 1] = (int32_t)nnz;
    }
    c->nnz = nnz;
    free(acc); free(mark); free(list);
    if (products) *products = prod;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Triangle counting */

int orc_strict_lower(const orc_csr *a, orc_csr *l) {
    memset(l, 0, sizeof(*l));
    int32_t m = a->m;
    int64_t cnt = 0;
    for (int32_t i = 0; i < m; ++i)
        for (int32_t p = a->row_ptr[i]; p < a->row_ptr[i + 1]; ++p)
            if (a->col_ind[p] < i) cnt++;
    l->m = m;
    l->n = a->n;
    l->nnz = cnt;
    l->row_ptr = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    l->col_ind = (int32_t *)malloc((cnt ? cnt : 1) * sizeof(int32_t));
    l->val = NULL;
    int64_t k = 0;
    l->row_ptr[0] = 0;
    for (int32_t i = 0; i < m; ++i) {
        int64_t b = k;
        for (int32_t p = a->row_ptr[i]; p < a->row_ptr[i + 1]; ++p)
            if (a->col_ind[p] < i) l->col_ind[k++] = a->col_ind[p];
        qsort(l->col_ind + b, k - b, sizeof(int32_t), cmp_i32);
        l->row_ptr[i + 1] = (int32_t)k;
    }
    return 0;
}

int64_t orc_tc(const orc_csr *l, int64_t *wedges) { return orc_tc_rows(l, 0, l->m, wedges); }

/* the part of the count whose mask row i lies in [r0, r1) (a bounded sample
   of the same work; the parts over a split of the rows add up to orc_tc) */
int64_t orc_tc_rows(const orc_csr *l, int32_t r0, int32_t r1, int64_t *wedges) {
    int64_t tri = 0, w = 0;
    for (int32_t i = r0; i < r1; ++i) {
        int32_t ib = l->row_ptr[i], ie = l->row_ptr[i + 1];
        for (int32_t p = ib; p < ie; ++p) {
            int32_t k = l->col_ind[p];
            int32_t kb = l->row_ptr[k], ke = l->row_ptr[k + 1];
            w += ke - kb;
            int32_t x = ib, y = kb;
            while (x < ie && y < ke) {
                int32_t cx = l->col_ind[x], cy = l->col_ind[y];
                if (cx == cy) { tri++; x++; y++; }
                else if (cx < cy) x++;
                else y++;
            }
        }
    }
    if (wedges) *wedges = w;
    return tri;
}

/* General-semiring SpMSpV, restating the reference's mXv (src/mXv.cuh:110-389)
   with the semirings the library offers: y[c] = (+) over present u with
   (u, c) in A of x[u] (*) A(u, c).  sr: 0 plus-times (double accumulation,
   identity 0), 1 min-plus (identity 1.70141e38, present -1 < x < 1e38), 2
   max-times (identity -inf), 3 or-and (0 / 1).  listed (nullable) marks the
   present entries of a sparse input; reached (nullable) receives the columns
   any present entry reaches. */
void orc_mxv(const orc_csr *a, int sr, const float *x, const uint8_t *listed, float *y, uint8_t *reached) {
    const float inf_sentinel = 1.70141e38f;
    double *acc = (double *)calloc((size_t)(a->n > 0 ? a->n : 1), sizeof(double));
    for (int32_t c = 0; c < a->n; ++c) {
        y[c] = sr == 1 ? inf_sentinel : (sr == 2 ? -INFINITY : 0.0f);
        if (reached) reached[c] = 0;
    }
    for (int32_t u = 0; u < a->m; ++u) {
        const float xu = x[u];
        int present;
        if (listed) present = listed[u] != 0;
        else if (sr == 1) present = xu > -1.0f && xu < 1e38f;
        else if (sr == 2) present = xu != -INFINITY;
        else present = xu != 0.0f;
        if (!present) continue;
        for (int32_t e = a->row_ptr[u]; e < a->row_ptr[u + 1]; ++e) {
            const int32_t c = a->col_ind[e];
            const float w = a->val ? a->val[e] : 1.0f;
            if (reached) reached[c] = 1;
            if (sr == 0) acc[c] += (double)xu * (double)w;
            else if (sr == 1) { const float t = xu + w; if (t < y[c]) y[c] = t; }
            else if (sr == 2) { const float t = xu * w; if (t > y[c]) y[c] = t; }
            else if (xu != 0.0f && w != 0.0f) y[c] = 1.0f;
        }
    }
    if (sr == 0)
        for (int32_t c = 0; c < a->n; ++c) y[c] = (float)acc[c];
    free(acc);
}

