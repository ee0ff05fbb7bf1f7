// graphio.cpp -- host-side graph construction: the MatrixMarket loader with
// the reference's semantics and the seeded RMAT generator.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "internal.hpp"

using namespace apspie;

namespace {

int hw_threads() {
    unsigned n = std::thread::hardware_concurrency();
    if (const char *e = getenv("OMP_NUM_THREADS")) n = std::min<unsigned>(n ? n : 1, (unsigned)atoi(e));
    return (int)std::max(1u, std::min(n ? n : 1u, 16u));
}

// Run f(lo, hi) over [0, n) split into contiguous ranges, one per thread.
template <class F>
void parallel_ranges(int64_t n, F &&f) {
    int nt = hw_threads();
    if (n < 65536) nt = 1;
    if (nt == 1) {
        f(int64_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    for (auto &x : th) x.join();
}

void host_csr_alloc(apspie_host_csr *h, int32_t m, int32_t n, int64_t nnz, bool with_val) {
    memset(h, 0, sizeof(*h));
    h->m = m;
    h->n = n;
    h->nnz = nnz;
    h->row_ptr = (int32_t *)malloc((size_t(m) + 1) * 4);
    h->col_ind = (int32_t *)malloc(size_t(std::max<int64_t>(nnz, 1)) * 4);
    h->val = with_val ? (float *)malloc(size_t(std::max<int64_t>(nnz, 1)) * 4) : nullptr;
    if (!h->row_ptr || !h->col_ind || (with_val && !h->val)) {
        apspie_host_csr_free(h);
        throw Error(APSPIE_ENOMEM, "host CSR allocation failed");
    }
}

// makeSymmetric (src/util.cuh:572-643) on COO arrays of capacity 2 * nnz:
// append the reversed copy of every entry that is not a self-loop (its value
// negated when skew), sort by (row, col), drop exact repeats (the first
// kept, values moving with their entries).  Returns the new count.
int64_t coo_symmetrize(int64_t nnz, int32_t *col, int32_t *row, float *val, bool skew) {
    int64_t e = nnz;
    for (int64_t i = 0; i < nnz; ++i)
        if (col[i] != row[i]) {
            row[e] = col[i];
            col[e] = row[i];
            if (val) val[e] = skew ? -val[i] : val[i];
            ++e;
        }
    std::vector<int64_t> idx(e);
    for (int64_t i = 0; i < e; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) {
        return row[x] != row[y] ? row[x] < row[y] : col[x] < col[y];
    });
    std::vector<int32_t> r2, c2;
    std::vector<float> v2;
    r2.reserve(e);
    c2.reserve(e);
    if (val) v2.reserve(e);
    for (int64_t i = 0; i < e; ++i) {
        const int64_t t = idx[i];
        if (!r2.empty() && r2.back() == row[t] && c2.back() == col[t]) continue;
        r2.push_back(row[t]);
        c2.push_back(col[t]);
        if (val) v2.push_back(val[t]);
    }
    e = (int64_t)r2.size();
    std::copy(r2.begin(), r2.end(), row);
    std::copy(c2.begin(), c2.end(), col);
    if (val) std::copy(v2.begin(), v2.end(), val);
    return e;
}

// buildMatrix (src/matrix.cpp:55-111): stable counting sort of COO entries
// by row into row_ptr[m+1] / col_ind[nnz] / val[nnz] (val and coo_val may be
// null).  False when a row lies outside [0, m).
bool coo_counting_sort(int32_t m, int64_t nnz, const int32_t *row, const int32_t *col, const float *coo_val,
                       int32_t *row_ptr, int32_t *col_ind, float *val) {
    std::vector<int64_t> cur(size_t(m) + 1, 0);
    for (int64_t i = 0; i < nnz; ++i) {
        if (row[i] < 0 || row[i] >= m) return false;
        cur[row[i] + 1]++;
    }
    for (int32_t i = 0; i < m; ++i) cur[i + 1] += cur[i];
    for (int32_t i = 0; i <= m; ++i) row_ptr[i] = int32_t(cur[i]);
    for (int64_t i = 0; i < nnz; ++i) {
        const int64_t d = cur[row[i]]++;
        col_ind[d] = col[i];
        if (val) val[d] = coo_val ? coo_val[i] : 1.0f;
    }
    return true;
}

// Sort each row's columns (and values) in place, rows split across threads.
void sort_rows(apspie_host_csr *h) {
    parallel_ranges(h->m, [&](int64_t lo, int64_t hi) {
        std::vector<std::pair<int32_t, float>> tmp;
        for (int64_t r = lo; r < hi; ++r) {
            int32_t b = h->row_ptr[r], e = h->row_ptr[r + 1];
            if (!h->val) {
                std::sort(h->col_ind + b, h->col_ind + e);
                continue;
            }
            tmp.clear();
            for (int32_t p = b; p < e; ++p) tmp.emplace_back(h->col_ind[p], h->val[p]);
            std::stable_sort(tmp.begin(), tmp.end(),
                             [](const auto &x, const auto &y) { return x.first < y.first; });
            for (int32_t p = b; p < e; ++p) {
                h->col_ind[p] = tmp[p - b].first;
                h->val[p] = tmp[p - b].second;
            }
        }
    });
}

// MT19937 (Matsumoto & Nishimura 1998) producing 624 tempered words per
// refill, so the twist and the tempering run as vectorisable loops; the
// output stream equals std::mt19937 and the reference's genrand_int32
// (src/mt19937ar.cpp:102-137) for the same init_genrand seed.
class BlockMT {
    uint32_t s_[624], out_[624];
    int idx_ = 624;
    static uint32_t mix(uint32_t a, uint32_t b, uint32_t c) {
        const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
        return c ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
    }
    void refill() {
        for (int i = 0; i < 227; ++i) s_[i] = mix(s_[i], s_[i + 1], s_[i + 397]);
        for (int i = 227; i < 623; ++i) s_[i] = mix(s_[i], s_[i + 1], s_[i - 227]);
        s_[623] = mix(s_[623], s_[0], s_[396]);
        for (int i = 0; i < 624; ++i) {
            uint32_t y = s_[i];
            y ^= y >> 11;
            y ^= (y << 7) & 0x9d2c5680u;
            y ^= (y << 15) & 0xefc60000u;
            y ^= y >> 18;
            out_[i] = y;
        }
        idx_ = 0;
    }

   public:
    explicit BlockMT(uint32_t seed) {
        s_[0] = seed;
        for (int i = 1; i < 624; ++i) s_[i] = 1812433253u * (s_[i - 1] ^ (s_[i - 1] >> 30)) + uint32_t(i);
    }
    inline uint32_t next() {
        if (idx_ >= 624) refill();
        return out_[idx_++];
    }
};

// ---------------------------------------------------------------------------
// MatrixMarket text scanner with scanf-like whitespace handling.
struct Scanner {
    std::vector<char> buf;
    size_t pos = 0;
    bool load(const char *path) {
        FILE *f = fopen(path, "rb");
        if (!f) return false;
        fseek(f, 0, SEEK_END);
        long n = ftell(f);
        fseek(f, 0, SEEK_SET);
        buf.resize(size_t(n) + 1);
        size_t got = fread(buf.data(), 1, size_t(n), f);
        fclose(f);
        buf.resize(got);
        buf.push_back('\0');
        return true;
    }
    int peek() const { return pos + 1 < buf.size() ? (unsigned char)buf[pos] : EOF; }
    int get() { return pos + 1 < buf.size() ? (unsigned char)buf[pos++] : EOF; }
    bool read_int(int &out) {
        const char *s = buf.data() + pos;
        char *end = nullptr;
        long v = strtol(s, &end, 10);
        if (end == s) return false;
        pos += size_t(end - s);
        out = (int)v;
        return true;
    }
    bool read_float(float &out) {
        const char *s = buf.data() + pos;
        char *end = nullptr;
        float v = strtof(s, &end);
        if (end == s) return false;
        pos += size_t(end - s);
        out = v;
        return true;
    }
};

}  // namespace

extern "C" {

void apspie_host_csr_free(apspie_host_csr *h) {
    if (!h) return;
    free(h->row_ptr);
    free(h->col_ind);
    free(h->val);
    memset(h, 0, sizeof(*h));
}

// readEdge (src/util.cuh:355-367) skips leading '%' lines and reads "m n nnz";
// readMtx (:370-430) reads nnz lines "a b [w]" into (col=a-1, row=b-1), weighted
// iff the character right after the first line's second integer is a space;
// makeSymmetric (:572-643) appends reversed non-loop entries, sorts by
// (row, col) and drops exact repeats (self-loops stay); buildMatrix
// (src/matrix.cpp:55-111) counting-sorts by row, stable.
int apspie_load_mtx(const char *path, uint32_t flags, apspie_host_csr *out) {
    if (!path || !out) return APSPIE_EINVAL;
    memset(out, 0, sizeof(*out));
    return guarded(nullptr, [&] {
        Scanner sc;
        APSPIE_CHECK(sc.load(path), APSPIE_EIO, std::string("cannot open ") + path);
        // %%MatrixMarket banner: field (real / integer / pattern) and symmetry
        // (general / symmetric / skew-symmetric), when asked for
        int banner_weighted = -1;  // -1: the reference's space-after rule
        bool mirror = false, skew = false;
        if (flags & APSPIE_LOAD_MM_BANNER) {
            std::string line;
            for (size_t i = 0; i + 1 < sc.buf.size() && sc.buf[i] != '\n'; ++i) line += char(std::tolower((unsigned char)sc.buf[i]));
            std::vector<std::string> tok;
            for (size_t i = 0; i < line.size();) {
                while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
                size_t j = i;
                while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
                if (j > i) tok.push_back(line.substr(i, j - i));
                i = j;
            }
            APSPIE_CHECK(tok.size() == 5 && tok[0] == "%%matrixmarket" && tok[1] == "matrix", APSPIE_EIO,
                         "no %%MatrixMarket matrix banner");
            APSPIE_CHECK(tok[2] == "coordinate", APSPIE_EIO, "only coordinate MatrixMarket files are read");
            if (tok[3] == "pattern") banner_weighted = 0;
            else if (tok[3] == "real" || tok[3] == "integer" || tok[3] == "double") banner_weighted = 1;
            else throw Error(APSPIE_EIO, "unsupported MatrixMarket field " + tok[3]);
            if (tok[4] == "symmetric") mirror = true;
            else if (tok[4] == "skew-symmetric") mirror = skew = true;
            else APSPIE_CHECK(tok[4] == "general", APSPIE_EIO, "unsupported MatrixMarket symmetry " + tok[4]);
        }
        int c = sc.get(), old = 0;
        while (c != EOF) {
            if ((old == '\n' || old == 0) && c != '%') {
                sc.pos--;
                break;
            }
            old = c;
            c = sc.get();
        }
        int m = 0, n = 0, nnz = 0;
        APSPIE_CHECK(sc.read_int(m) && sc.read_int(n) && sc.read_int(nnz), APSPIE_EIO, "bad mtx header");
        APSPIE_CHECK(m >= 0 && n >= 0 && nnz >= 0, APSPIE_EIO, "negative mtx size");
        const bool undirected = (flags & APSPIE_LOAD_UNDIRECTED) || mirror;
        const int64_t cap = undirected ? 2 * int64_t(nnz) : nnz;
        std::vector<int32_t> row(cap), col(cap);
        std::vector<float> val(cap);
        int ch = 0;
        bool weighted = false;
        for (int j = 0; j < nnz; ++j) {
            int a = 0, b = 0;
            APSPIE_CHECK(sc.read_int(a) && sc.read_int(b), APSPIE_EIO, "not enough rows in mtx file");
            if (j == 0) {
                ch = sc.get();
                weighted = banner_weighted >= 0 ? banner_weighted == 1 : ch == ' ';
                if (banner_weighted >= 0) sc.pos--;  // the character may start the value
            }
            float w = 1.0f;
            if (weighted) APSPIE_CHECK(sc.read_float(w), APSPIE_EIO, "bad weight");
            col[j] = a - 1;
            row[j] = b - 1;
            val[j] = w;
        }
        const int64_t e = undirected ? coo_symmetrize(nnz, col.data(), row.data(), val.data(), skew) : nnz;
        host_csr_alloc(out, m, n, e, weighted);
        for (int64_t i = 0; i < e; ++i)
            if (col[i] < 0 || col[i] >= n) {
                apspie_host_csr_free(out);
                throw Error(APSPIE_EIO, "mtx index out of range");
            }
        if (!coo_counting_sort(m, e, row.data(), col.data(), weighted ? val.data() : nullptr, out->row_ptr,
                               out->col_ind, out->val)) {
            apspie_host_csr_free(out);
            throw Error(APSPIE_EIO, "mtx index out of range");
        }
        if (flags & APSPIE_LOAD_SORT_COLUMNS) sort_rows(out);
    });
}

int apspie_rmat(int scale, int edge_factor, uint32_t seed, uint32_t flags, uint32_t weight_seed,
                apspie_host_csr *out) {
    if (!out) return APSPIE_EINVAL;
    memset(out, 0, sizeof(*out));
    return guarded(nullptr, [&] {
        APSPIE_CHECK(scale >= 1 && scale <= 30 && edge_factor >= 1, APSPIE_EINVAL, "bad rmat shape");
        const int64_t m = int64_t(1) << scale;
        const int64_t ne = m * edge_factor;
        const bool und = flags & APSPIE_RMAT_UNDIRECTED;
        APSPIE_CHECK((und ? 2 : 1) * ne < (int64_t(1) << 31) - 4096, APSPIE_ELIMIT, "rmat too large for int32 offsets");
        // u = r / 2^32 < p  <=>  r < ceil(p * 2^32): exact integer thresholds
        const uint64_t ta = (uint64_t)std::ceil(0.57 * 4294967296.0);
        const uint64_t tb = (uint64_t)std::ceil(0.76 * 4294967296.0);
        const uint64_t tc = (uint64_t)std::ceil(0.95 * 4294967296.0);
        std::vector<uint32_t> es(ne), ed(ne);
        {
            BlockMT g(seed);
            for (int64_t e = 0; e < ne; ++e) {
                uint32_t s = 0, d = 0;
                for (int l = scale - 1; l >= 0; --l) {
                    // quadrant q: 0 = a (0,0), 1 = b (0,1), 2 = c (1,0), 3 = d (1,1)
                    const uint64_t r = g.next();
                    const uint32_t q = uint32_t(r >= ta) + uint32_t(r >= tb) + uint32_t(r >= tc);
                    s |= (q >> 1) << l;
                    d |= (q & 1u) << l;
                }
                es[e] = s;
                ed[e] = d;
            }
        }
        // bucket by source row (both directions when undirected); positions
        // inside a row are claimed atomically and the rows sorted afterwards,
        // so the result does not depend on thread timing
        std::vector<int64_t> cnt(m + 1, 0);
        parallel_ranges(ne, [&](int64_t lo, int64_t hi) {
            for (int64_t e = lo; e < hi; ++e) {
                if (es[e] == ed[e]) continue;
                __atomic_fetch_add(&cnt[es[e] + 1], 1, __ATOMIC_RELAXED);
                if (und) __atomic_fetch_add(&cnt[ed[e] + 1], 1, __ATOMIC_RELAXED);
            }
        });
        for (int64_t i = 0; i < m; ++i) cnt[i + 1] += cnt[i];
        std::vector<int32_t> bucket(std::max<int64_t>(cnt[m], 1));
        {
            std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
            parallel_ranges(ne, [&](int64_t lo, int64_t hi) {
                for (int64_t e = lo; e < hi; ++e) {
                    const uint32_t s = es[e], d = ed[e];
                    if (s == d) continue;
                    bucket[__atomic_fetch_add(&cur[s], 1, __ATOMIC_RELAXED)] = (int32_t)d;
                    if (und) bucket[__atomic_fetch_add(&cur[d], 1, __ATOMIC_RELAXED)] = (int32_t)s;
                }
            });
        }
        std::vector<uint32_t>().swap(es);
        std::vector<uint32_t>().swap(ed);
        // sort + unique each row
        std::vector<int32_t> len(m);
        parallel_ranges(m, [&](int64_t lo, int64_t hi) {
            for (int64_t r = lo; r < hi; ++r) {
                int32_t *b = bucket.data() + cnt[r], *e = bucket.data() + cnt[r + 1];
                std::sort(b, e);
                len[r] = int32_t(std::unique(b, e) - b);
            }
        });
        int64_t nnz = 0;
        for (int64_t r = 0; r < m; ++r) nnz += len[r];
        const bool weighted = flags & APSPIE_RMAT_WEIGHTED;
        host_csr_alloc(out, (int32_t)m, (int32_t)m, nnz, weighted);
        out->row_ptr[0] = 0;
        for (int64_t r = 0; r < m; ++r) out->row_ptr[r + 1] = out->row_ptr[r] + len[r];
        parallel_ranges(m, [&](int64_t lo, int64_t hi) {
            for (int64_t r = lo; r < hi; ++r)
                memcpy(out->col_ind + out->row_ptr[r], bucket.data() + cnt[r], size_t(len[r]) * 4);
        });
        if (weighted) {
            BlockMT g(weight_seed);
            for (int64_t p = 0; p < nnz; ++p) out->val[p] = float(1u + (g.next() >> 26));
        }
    });
}

// buildLower (test/tc/main.cu:39-55): keep the entries strictly below the
// diagonal (row i keeps columns j < i), columns sorted.
// Each worker owns a range of columns and walks the rows in order, so the
// entries of every output row come out in increasing source-row order
// (the stable counting sort csr2csc performs).
// ---- binary CSR cache ------------------------------------------------------
// File: "APSPCSR1", u32 version (1), u32 flags (bit 0: values), i64 m, n, nnz,
// int32 row_ptr[m+1], int32 col_ind[nnz], [float val[nnz]], then u64 check =
// the word hash below over every byte before it.  Little-endian.

namespace {

constexpr char CSR_MAGIC[8] = {'A', 'P', 'S', 'P', 'C', 'S', 'R', '1'};

struct CsrFileHeader {
    char magic[8];
    uint32_t version, flags;
    int64_t m, n, nnz;
};

// 64-bit hash of a byte stream fed in pieces: 8-byte words mixed by
// multiply-xorshift (a tail shorter than a word is zero-padded)
struct WordHash {
    uint64_t h = 0x9e3779b97f4a7c15ull, carry = 0;
    int nc = 0;
    void word(uint64_t w) {
        h ^= w;
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 33;
    }
    void bytes(const void *p, size_t n) {
        const unsigned char *b = (const unsigned char *)p;
        size_t i = 0;
        while (nc && i < n) {
            carry |= uint64_t(b[i++]) << (8 * nc);
            if (++nc == 8) {
                word(carry);
                carry = 0;
                nc = 0;
            }
        }
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            memcpy(&w, b + i, 8);
            word(w);
        }
        for (; i < n; ++i) {
            carry |= uint64_t(b[i]) << (8 * nc);
            ++nc;
        }
    }
    uint64_t done() {
        if (nc) word(carry);
        word(0x5bd1e995ull);
        return h;
    }
};

}  // namespace

int apspie_csr_save(const char *path, const apspie_host_csr *h) {
    if (!path || !h || h->m < 0 || h->n < 0 || h->nnz < 0 || !h->row_ptr || (h->nnz && !h->col_ind))
        return APSPIE_EINVAL;
    return guarded(nullptr, [&] {
        FILE *f = fopen(path, "wb");
        APSPIE_CHECK(f, APSPIE_EIO, std::string("cannot create ") + path);
        CsrFileHeader hd;
        memcpy(hd.magic, CSR_MAGIC, 8);
        hd.version = 1;
        hd.flags = h->val ? 1u : 0u;
        hd.m = h->m;
        hd.n = h->n;
        hd.nnz = h->nnz;
        WordHash wh;
        bool ok = true;
        auto put = [&](const void *p, size_t n) {
            wh.bytes(p, n);
            ok = ok && fwrite(p, 1, n, f) == n;
        };
        put(&hd, sizeof(hd));
        put(h->row_ptr, (size_t(h->m) + 1) * 4);
        put(h->col_ind, size_t(h->nnz) * 4);
        if (h->val) put(h->val, size_t(h->nnz) * 4);
        const uint64_t chk = wh.done();
        ok = ok && fwrite(&chk, 1, 8, f) == 8;
        ok = (fclose(f) == 0) && ok;
        APSPIE_CHECK(ok, APSPIE_EIO, std::string("write failed: ") + path);
    });
}

int apspie_csr_load(const char *path, apspie_host_csr *out) {
    if (!path || !out) return APSPIE_EINVAL;
    memset(out, 0, sizeof(*out));
    return guarded(nullptr, [&] {
        FILE *f = fopen(path, "rb");
        APSPIE_CHECK(f, APSPIE_EIO, std::string("cannot open ") + path);
        CsrFileHeader hd;
        WordHash wh;
        auto get = [&](void *p, size_t n) {
            const bool ok = fread(p, 1, n, f) == n;
            if (ok) wh.bytes(p, n);
            return ok;
        };
        bool ok = get(&hd, sizeof(hd)) && memcmp(hd.magic, CSR_MAGIC, 8) == 0 && hd.version == 1 &&
                  hd.m >= 0 && hd.m < (int64_t(1) << 31) && hd.n >= 0 && hd.n < (int64_t(1) << 31) && hd.nnz >= 0 &&
                  hd.nnz < (int64_t(1) << 31) && (hd.flags & ~1u) == 0;
        if (!ok) {
            fclose(f);
            throw Error(APSPIE_EIO, std::string("not a CSR cache file: ") + path);
        }
        host_csr_alloc(out, int32_t(hd.m), int32_t(hd.n), hd.nnz, hd.flags & 1u);
        ok = get(out->row_ptr, (size_t(hd.m) + 1) * 4) && get(out->col_ind, size_t(hd.nnz) * 4) &&
             (!(hd.flags & 1u) || get(out->val, size_t(hd.nnz) * 4));
        uint64_t chk = 0;
        ok = ok && fread(&chk, 1, 8, f) == 8 && chk == wh.done();
        fclose(f);
        // structure: offsets from 0 to nnz, non-decreasing, columns in range
        for (int64_t i = 0; ok && i < hd.m; ++i) ok = out->row_ptr[i] <= out->row_ptr[i + 1];
        ok = ok && out->row_ptr[0] == 0 && out->row_ptr[hd.m] == hd.nnz;
        for (int64_t e = 0; ok && e < hd.nnz; ++e) ok = out->col_ind[e] >= 0 && out->col_ind[e] < hd.n;
        if (!ok) {
            apspie_host_csr_free(out);
            throw Error(APSPIE_EIO, std::string("truncated or corrupt CSR cache file: ") + path);
        }
    });
}

int apspie_host_coo_make_symmetric(int64_t edge, int32_t *coo_col, int32_t *coo_row, float *coo_val,
                                   int64_t *out_edge) {
    if (!out_edge || edge < 0 || (edge > 0 && (!coo_col || !coo_row))) return APSPIE_EINVAL;
    return guarded(nullptr, [&] { *out_edge = edge ? coo_symmetrize(edge, coo_col, coo_row, coo_val, false) : 0; });
}

int apspie_host_coo_to_csr(int32_t m, int64_t nnz, const int32_t *coo_row, const int32_t *coo_col,
                           const float *coo_val, int32_t *row_ptr, int32_t *col_ind, float *val) {
    if (m < 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!coo_row || !coo_col || !col_ind))) return APSPIE_EINVAL;
    return guarded(nullptr, [&] {
        APSPIE_CHECK(coo_counting_sort(m, nnz, coo_row, coo_col, coo_val, row_ptr, col_ind, val), APSPIE_EINVAL,
                     "COO row index outside [0, m)");
    });
}

int apspie_host_csr_transpose(const apspie_host_csr *a, apspie_host_csr *at) {
    if (!a || !at) return APSPIE_EINVAL;
    memset(at, 0, sizeof(*at));
    return guarded(nullptr, [&] {
        APSPIE_CHECK(a->m >= 0 && a->n >= 0 && a->nnz >= 0 && a->row_ptr, APSPIE_EINVAL, "bad input csr");
        APSPIE_CHECK(a->nnz == 0 || a->col_ind, APSPIE_EINVAL, "bad input csr");
        host_csr_alloc(at, a->n, a->m, a->nnz, a->val != nullptr);
        std::vector<int64_t> cnt(size_t(a->n) + 1, 0);
        parallel_ranges(a->n, [&](int64_t lo, int64_t hi) {
            for (int64_t p = 0; p < a->nnz; ++p) {
                const int64_t c = a->col_ind[p];
                if (c >= lo && c < hi) cnt[c + 1]++;
            }
        });
        for (int64_t c = 0; c < a->n; ++c) cnt[c + 1] += cnt[c];
        APSPIE_CHECK(cnt[a->n] == a->nnz, APSPIE_EINVAL, "column index out of range");
        for (int64_t c = 0; c <= a->n; ++c) at->row_ptr[c] = int32_t(cnt[c]);
        parallel_ranges(a->n, [&](int64_t lo, int64_t hi) {
            for (int32_t r = 0; r < a->m; ++r)
                for (int32_t p = a->row_ptr[r]; p < a->row_ptr[r + 1]; ++p) {
                    const int64_t c = a->col_ind[p];
                    if (c < lo || c >= hi) continue;
                    const int64_t d = cnt[c]++;
                    at->col_ind[d] = r;
                    if (a->val) at->val[d] = a->val[p];
                }
        });
    });
}

int apspie_strict_lower(const apspie_host_csr *a, apspie_host_csr *l) {
    if (!a || !l) return APSPIE_EINVAL;
    memset(l, 0, sizeof(*l));
    return guarded(nullptr, [&] {
        APSPIE_CHECK(a->m >= 0 && a->row_ptr, APSPIE_EINVAL, "bad input csr");
        std::vector<int32_t> rp(size_t(a->m) + 1, 0);
        for (int32_t i = 0; i < a->m; ++i) {
            int32_t c = 0;
            for (int32_t p = a->row_ptr[i]; p < a->row_ptr[i + 1]; ++p) c += a->col_ind[p] < i;
            rp[i + 1] = rp[i] + c;
        }
        host_csr_alloc(l, a->m, a->n, rp[a->m], false);
        memcpy(l->row_ptr, rp.data(), (size_t(a->m) + 1) * 4);
        parallel_ranges(a->m, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                int32_t k = rp[i];
                for (int32_t p = a->row_ptr[i]; p < a->row_ptr[i + 1]; ++p)
                    if (a->col_ind[p] < i) l->col_ind[k++] = a->col_ind[p];
                std::sort(l->col_ind + rp[i], l->col_ind + k);
            }
        });
    });
}

// Row lengths from a discrete power law P(len = k) ~ k^-alpha on
// [1, max_len], alpha solved by bisection so the mean is mean_len; columns
// uniform without replacement within a row (multiply-shift of one draw,
// redrawn on repeats), sorted; values U(-1, 1) in CSR order.  One MT19937
// stream, per row: length draw, column draws, then value draws.
int apspie_powerlaw_csr(int32_t m, int32_t n, double mean_len, int32_t max_len, uint32_t seed,
                        apspie_host_csr *out) {
    if (!out) return APSPIE_EINVAL;
    memset(out, 0, sizeof(*out));
    return guarded(nullptr, [&] {
        APSPIE_CHECK(m >= 0 && n >= 1 && max_len >= 1 && max_len <= n, APSPIE_EINVAL, "bad power-law shape");
        APSPIE_CHECK(mean_len >= 1.0 && mean_len <= max_len, APSPIE_EINVAL, "bad mean row length");
        auto mean_of = [&](double a) {
            double z = 0, s = 0;
            for (int k = 1; k <= max_len; ++k) {
                const double w = std::pow(double(k), -a);
                z += w;
                s += w * k;
            }
            return s / z;
        };
        double lo = -4.0, hi = 8.0;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (mean_of(mid) > mean_len) lo = mid;
            else hi = mid;
        }
        const double alpha = 0.5 * (lo + hi);
        std::vector<double> cdf(size_t(max_len) + 1, 0.0);
        for (int k = 1; k <= max_len; ++k) cdf[k] = cdf[k - 1] + std::pow(double(k), -alpha);
        for (int k = 1; k <= max_len; ++k) cdf[k] /= cdf[max_len];
        BlockMT g(seed);
        std::vector<int32_t> rp(size_t(m) + 1, 0), cols;
        std::vector<float> vals;
        std::vector<int32_t> mark(size_t(n), -1), row;
        cols.reserve(size_t(double(m) * mean_len * 1.1) + 16);
        vals.reserve(cols.capacity());
        for (int32_t i = 0; i < m; ++i) {
            const double u = (double(g.next()) + 0.5) / 4294967296.0;
            const int len = int(std::lower_bound(cdf.begin() + 1, cdf.end(), u) - cdf.begin());
            row.clear();
            while ((int)row.size() < len) {
                const int32_t c = int32_t((uint64_t(g.next()) * uint64_t(n)) >> 32);
                if (mark[c] == i) continue;
                mark[c] = i;
                row.push_back(c);
            }
            std::sort(row.begin(), row.end());
            for (int32_t c : row) {
                cols.push_back(c);
                vals.push_back(float((double(g.next()) + 0.5) / 4294967296.0 * 2.0 - 1.0));
            }
            APSPIE_CHECK(cols.size() < (size_t(1) << 31) - 4096, APSPIE_ELIMIT, "power-law matrix too large");
            rp[i + 1] = int32_t(cols.size());
        }
        host_csr_alloc(out, m, n, int64_t(cols.size()), true);
        memcpy(out->row_ptr, rp.data(), (size_t(m) + 1) * 4);
        memcpy(out->col_ind, cols.data(), cols.size() * 4);
        memcpy(out->val, vals.data(), vals.size() * 4);
    });
}

}  // extern "C"
