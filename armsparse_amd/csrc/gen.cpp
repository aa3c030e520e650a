// Seeded CSR-native generators and host format helpers (pure host C++).
// The reference's own generators (prepareSparse + toCSR, SDMM.cpp:329-370;
// SpMM.cpp:48-79; SpMV.cpp:39-50) are time-seeded and dense-first; these are
// counter-based and CSR-native so graph-scale inputs are reproducible.
#include "armsp_gen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <cstdlib>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

// Host parallelism with std::thread (no OpenMP runtime in the library, so it
// never brings a second one into a process that already has torch's).
int num_threads() {
    static const int n = [] {
        const char* e = std::getenv("OMP_NUM_THREADS");
        int v = e ? std::atoi(e) : 0;
        if (v <= 0) v = (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(v, 64));
    }();
    return n;
}

// body(tid, nthreads) on nt threads.
template <class F>
void run_threads(int nt, F body) {
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(body, t, nt);
    body(0, nt);
    for (auto& x : th) x.join();
}

// body(i) for i in [0, n), static contiguous chunks.
template <class F>
void parallel_for(int64_t n, F body) {
    const int nt = n < 65536 ? 1 : num_threads();
    run_threads(nt, [&](int t, int T) {
        const int64_t lo = n * t / T, hi = n * (t + 1) / T;
        for (int64_t i = lo; i < hi; ++i) body(i);
    });
}

// body(i) for i in [0, n), handed out `grain` at a time.
template <class F>
void parallel_dynamic(int64_t n, int64_t grain, F body) {
    std::atomic<int64_t> next{0};
    const int nt = n <= grain ? 1 : num_threads();
    run_threads(nt, [&](int, int) {
        for (int64_t lo; (lo = next.fetch_add(grain)) < n;)
            for (int64_t i = lo; i < std::min(n, lo + grain); ++i) body(i);
    });
}

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline uint64_t stream_key(uint64_t seed, uint64_t stream) { return mix64(seed + kGolden * (stream + 1)); }
inline uint64_t draw(uint64_t key, uint64_t i) { return mix64(key + kGolden * (i + 1)); }

// (r>>40)·2^-24 in [0,1); lo + (hi-lo)·u is exact for (lo,hi) = (-1,1).
inline float to_uniform(uint64_t r, float lo, float hi) {
    float u = (float)(uint32_t)(r >> 40) * 0x1p-24f;
    return lo + (hi - lo) * u;
}

inline float nonzero_value(uint64_t key, uint64_t i) {
    float v = to_uniform(draw(key, i), -1.0f, 1.0f);
    return v == 0.0f ? 0x1p-24f : v;
}

// The first `want` distinct values of draw(e), e = 0, 1, 2, ... (draw returns
// ~0 for a rejected draw), sorted ascending.  Draws are generated in parallel,
// partitioned into NB buckets by bucket_of(key) (monotone in the key) with
// each bucket keeping draw order, sorted and deduplicated per bucket keeping
// the earliest draw of every key, and cut at the want-th smallest first-draw
// index — the same set a sequential scan would collect, for any thread count.
template <class Draw, class Bucket>
std::vector<uint64_t> first_distinct(int64_t want, Draw draw, Bucket bucket_of, int NB) {
    if (want <= 0) return {};
    int64_t T = want + want / 4 + 1024;
    while (true) {
        std::vector<uint64_t> keys(T);
        parallel_for(T, [&](int64_t e) { keys[e] = draw((uint64_t)e); });
        const int nt = T < 65536 ? 1 : num_threads();
        std::vector<int64_t> cnt((size_t)nt * NB, 0);
        run_threads(nt, [&](int tid, int T_) {
            const int64_t lo = T * tid / T_, hi = T * (tid + 1) / T_;
            int64_t* c = &cnt[(size_t)tid * NB];
            for (int64_t e = lo; e < hi; ++e)
                if (keys[e] != ~0ULL) c[bucket_of(keys[e])]++;
        });
        std::vector<int64_t> boff(NB + 1, 0), off((size_t)nt * NB);
        int64_t acc = 0;
        for (int b = 0; b < NB; ++b) {
            boff[b] = acc;
            for (int t = 0; t < nt; ++t) {
                off[(size_t)t * NB + b] = acc;
                acc += cnt[(size_t)t * NB + b];
            }
        }
        boff[NB] = acc;
        std::vector<std::pair<uint64_t, int64_t>> pairs(acc);
        run_threads(nt, [&](int tid, int T_) {
            const int64_t lo = T * tid / T_, hi = T * (tid + 1) / T_;
            int64_t* o = &off[(size_t)tid * NB];
            for (int64_t e = lo; e < hi; ++e)
                if (keys[e] != ~0ULL) pairs[o[bucket_of(keys[e])]++] = {keys[e], e};
        });
        std::vector<uint64_t>().swap(keys);
        std::vector<int64_t> bcount(NB, 0);
        parallel_dynamic(NB, 4, [&](int64_t b) {
            auto first = pairs.begin() + boff[b], last = pairs.begin() + boff[b + 1];
            std::sort(first, last);  // by key, then by draw index
            auto w = first;
            for (auto it = first; it != last; ++it)
                if (w == first || (w - 1)->first != it->first) *w++ = *it;
            bcount[b] = w - first;
        });
        int64_t D = 0;
        for (int b = 0; b < NB; ++b) D += bcount[b];
        if (D < want) {
            T = T + T / 2;
            continue;
        }
        std::vector<int64_t> firsts;
        firsts.reserve(D);
        for (int b = 0; b < NB; ++b)
            for (int64_t i = 0; i < bcount[b]; ++i) firsts.push_back(pairs[boff[b] + i].second);
        std::nth_element(firsts.begin(), firsts.begin() + (want - 1), firsts.end());
        const int64_t tstar = firsts[want - 1];
        std::vector<uint64_t> out;
        out.reserve(want);
        for (int b = 0; b < NB; ++b)
            for (int64_t i = 0; i < bcount[b]; ++i)
                if (pairs[boff[b] + i].second <= tstar) out.push_back(pairs[boff[b] + i].first);
        return out;
    }
}

// Sorts keys (row<<32 | col) ascending: a parallel counting pass into buckets
// of rows (row >> rshift), then every bucket sorted on its own.
void sort_keys(std::vector<uint64_t>& keys, int rshift, int NB) {
    const int64_t n = (int64_t)keys.size();
    const int nt = n < 65536 ? 1 : num_threads();
    if (nt == 1) {
        std::sort(keys.begin(), keys.end());
        return;
    }
    auto bucket_of = [&](uint64_t k) { return (int)((k >> 32) >> rshift); };
    std::vector<int64_t> cnt((size_t)nt * NB, 0);
    run_threads(nt, [&](int tid, int T) {
        int64_t* c = &cnt[(size_t)tid * NB];
        for (int64_t i = n * tid / T; i < n * (tid + 1) / T; ++i) c[bucket_of(keys[i])]++;
    });
    std::vector<int64_t> boff(NB + 1, 0), off((size_t)nt * NB);
    int64_t acc = 0;
    for (int b = 0; b < NB; ++b) {
        boff[b] = acc;
        for (int t = 0; t < nt; ++t) {
            off[(size_t)t * NB + b] = acc;
            acc += cnt[(size_t)t * NB + b];
        }
    }
    boff[NB] = acc;
    std::vector<uint64_t> out(n);
    run_threads(nt, [&](int tid, int T) {
        int64_t* o = &off[(size_t)tid * NB];
        for (int64_t i = n * tid / T; i < n * (tid + 1) / T; ++i) out[o[bucket_of(keys[i])]++] = keys[i];
    });
    parallel_dynamic(NB, 4, [&](int64_t b) { std::sort(out.begin() + boff[b], out.begin() + boff[b + 1]); });
    keys.swap(out);
}

// Sorted unique keys (row<<32 | col) → canonical CSR.
void keys_to_csr(std::vector<uint64_t>& keys, int M, uint64_t vkey, int* rowptr, int* colind, float* val) {
    if (!std::is_sorted(keys.begin(), keys.end())) {
        int rshift = 0;
        while ((M >> rshift) > 4096) ++rshift;
        sort_keys(keys, rshift, (M >> rshift) + 1);
    }
    std::fill(rowptr, rowptr + M + 1, 0);
    for (uint64_t k : keys) rowptr[(k >> 32) + 1]++;
    for (int r = 0; r < M; ++r) rowptr[r + 1] += rowptr[r];
    parallel_for((int64_t)keys.size(), [&](int64_t i) {
        colind[i] = (int)(keys[i] & 0xffffffffULL);
        if (val) val[i] = nonzero_value(vkey, (uint64_t)i);
    });
}

}  // namespace

extern "C" {

uint64_t armsp_rng_u64(uint64_t seed, uint64_t stream, uint64_t i) { return draw(stream_key(seed, stream), i); }

float armsp_rng_uniform(uint64_t seed, uint64_t stream, uint64_t i, float lo, float hi) {
    return to_uniform(armsp_rng_u64(seed, stream, i), lo, hi);
}

int armsp_host_fill_uniform(int64_t n, uint64_t seed, uint64_t stream, float lo, float hi, float* out) {
    if (n < 0 || (n > 0 && !out)) return 1;
    const uint64_t key = stream_key(seed, stream);
    parallel_for(n, [&](int64_t i) { out[i] = to_uniform(draw(key, (uint64_t)i), lo, hi); });
    return 0;
}

int armsp_host_fill_normal(int64_t n, uint64_t seed, uint64_t stream, float* out) {
    if (n < 0 || (n > 0 && !out)) return 1;
    const uint64_t key = stream_key(seed, stream);
    parallel_for(n, [&](int64_t i) {
        double u1 = ((double)(draw(key, 2 * (uint64_t)i) >> 11) + 1.0) * 0x1p-53;  // (0,1]
        double u2 = (double)(draw(key, 2 * (uint64_t)i + 1) >> 11) * 0x1p-53;      // [0,1)
        out[i] = (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2));
    });
    return 0;
}

int armsp_gen_uniform_csr(int M, int K, int64_t nnz, uint64_t seed, int* rowptr, int* colind, float* val) {
    if (M <= 0 || K <= 0 || nnz < 0 || !rowptr || (nnz > 0 && !colind)) return 1;
    const uint64_t cells = (uint64_t)M * (uint64_t)K;
    if ((uint64_t)nnz > cells) return 1;
    const uint64_t pkey = stream_key(seed, ARMSP_STREAM_POS);
    const uint64_t Ku = (uint64_t)K;
    int rshift = 0;
    while ((M >> rshift) > 4096) ++rshift;
    std::vector<uint64_t> keys = first_distinct(
        nnz,
        [&](uint64_t i) {
            uint64_t p = draw(pkey, i) % cells;
            return ((p / Ku) << 32) | (p % Ku);
        },
        [&](uint64_t k) { return (int)((k >> 32) >> rshift); }, (M >> rshift) + 1);
    keys_to_csr(keys, M, stream_key(seed, ARMSP_STREAM_VAL), rowptr, colind, val);
    return 0;
}

int armsp_gen_rmat_csr(int scale, int64_t nnz, double a, double b, double c, uint64_t seed, int* rowptr,
                       int* colind, float* val) {
    if (scale < 1 || scale > 30 || nnz < 0 || !rowptr || (nnz > 0 && !colind)) return 1;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return 1;
    const int M = 1 << scale;
    if ((uint64_t)nnz > (uint64_t)M * (uint64_t)M / 2) return 1;
    // Quadrant thresholds on 16-bit draws.
    const uint32_t ta = (uint32_t)std::lround(a * 65536.0);
    const uint32_t tb = (uint32_t)std::lround((a + b) * 65536.0);
    const uint32_t tc = (uint32_t)std::lround((a + b + c) * 65536.0);
    const uint64_t rkey = stream_key(seed, ARMSP_STREAM_RMAT);
    const int rshift = scale > 12 ? scale - 12 : 0;
    std::vector<uint64_t> keys = first_distinct(
        nnz,
        [&](uint64_t e) {
            uint64_t row = 0, col = 0, q = 0;
            for (int t = 0; t < scale; ++t) {
                if ((t & 3) == 0) q = draw(rkey, e * 8 + (t >> 2));
                uint32_t u = (uint32_t)((q >> (16 * (t & 3))) & 0xffff);
                uint64_t rb = (u >= tb) ? 1 : 0;                         // quadrants c, d
                uint64_t cb = (u >= ta && u < tb) || (u >= tc) ? 1 : 0;  // quadrants b, d
                row = (row << 1) | rb;
                col = (col << 1) | cb;
            }
            return (row << 32) | col;
        },
        [&](uint64_t k) { return (int)((k >> 32) >> rshift); }, (M >> rshift) + 1);
    keys_to_csr(keys, M, stream_key(seed, ARMSP_STREAM_VAL), rowptr, colind, val);
    return 0;
}

// Distinct undirected pairs (lo<<32 | hi, lo < hi) of a Chung–Lu-style power-law graph.
static std::vector<uint64_t> powerlaw_pairs(int n, int64_t pairs, double gamma, uint64_t seed) {
    std::vector<double> cdf(n);
    double acc = 0.0, expo = -1.0 / (gamma - 1.0);
    for (int i = 0; i < n; ++i) {
        acc += std::pow((double)(i + 1), expo);
        cdf[i] = acc;
    }
    for (int i = 0; i < n; ++i) cdf[i] /= acc;
    auto pick = [&](uint64_t r) {
        double u = (double)(r >> 11) * 0x1p-53;
        int v = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        return v < n ? v : n - 1;
    };
    const uint64_t gkey = stream_key(seed, ARMSP_STREAM_GRAPH);
    int rshift = 0;
    while ((n >> rshift) > 4096) ++rshift;
    return first_distinct(
        pairs,
        [&](uint64_t i) -> uint64_t {
            int u = pick(draw(gkey, 2 * i)), v = pick(draw(gkey, 2 * i + 1));
            if (u == v) return (uint64_t)~0ULL;
            uint64_t lo = (uint64_t)std::min(u, v), hi = (uint64_t)std::max(u, v);
            return (lo << 32) | hi;
        },
        [&](uint64_t k) { return (int)((k >> 32) >> rshift); }, (n >> rshift) + 1);
}

// Both directions of every pair, plus the diagonal when `self_loops`, as canonical CSR.
static void symmetric_csr(const std::vector<uint64_t>& und, int n, bool self_loops, int* rowptr, int* colind) {
    const int64_t P = (int64_t)und.size();
    std::vector<uint64_t> keys((size_t)(2 * P + (self_loops ? n : 0)));
    parallel_for(P, [&](int64_t i) {
        const uint64_t k = und[i];
        keys[2 * i] = k;
        keys[2 * i + 1] = (k << 32) | (k >> 32);
    });
    if (self_loops)
        parallel_for(n, [&](int64_t i) { keys[2 * P + i] = ((uint64_t)i << 32) | (uint64_t)i; });
    keys_to_csr(keys, n, 0, rowptr, colind, nullptr);
}

int armsp_gen_powerlaw_graph(int n, int64_t pairs, double gamma, uint64_t seed, int* rowptr, int* colind) {
    if (n < 2 || pairs < 0 || gamma <= 1.0 || !rowptr || (pairs > 0 && !colind)) return 1;
    if ((uint64_t)pairs > (uint64_t)n * (uint64_t)(n - 1) / 4) return 1;
    symmetric_csr(powerlaw_pairs(n, pairs, gamma, seed), n, false, rowptr, colind);
    return 0;
}

int armsp_gen_gcn_graph(int n, int64_t pairs, double gamma, uint64_t seed, int* rowptr, int* colind) {
    if (n < 2 || pairs < 0 || gamma <= 1.0 || !rowptr || !colind) return 1;
    if ((uint64_t)pairs > (uint64_t)n * (uint64_t)(n - 1) / 4) return 1;
    if (2 * pairs + n > (int64_t)INT32_MAX) return 1;
    symmetric_csr(powerlaw_pairs(n, pairs, gamma, seed), n, true, rowptr, colind);
    return 0;
}

int64_t armsp_gen_corpus_csr(int kind, int M, double avg_row, double param, uint64_t seed, int64_t cap, int* rowptr,
                             int* colind, float* val) {
    if (M <= 0 || !rowptr || !colind || avg_row < 0) return -1;
    const uint64_t key = stream_key(seed, 100 + (uint64_t)kind);
    const uint64_t vkey = stream_key(seed, ARMSP_STREAM_VAL);
    std::vector<std::vector<int>> rows(M);
    if (kind == 1) {
        int scale = 0;
        while ((1 << scale) < M) ++scale;
        if ((1 << scale) != M) return -1;
        double a = param, b = 0.4 * (1.0 - a), c = b;
        const uint32_t ta = (uint32_t)std::lround(a * 65536.0);
        const uint32_t tb = (uint32_t)std::lround((a + b) * 65536.0);
        const uint32_t tc = (uint32_t)std::lround((a + b + c) * 65536.0);
        const int64_t E = (int64_t)std::llround(avg_row * M);
        std::vector<uint64_t> keys(E);
        parallel_for(E, [&](int64_t e) {
            uint64_t row = 0, col = 0, q = 0;
            for (int t = 0; t < scale; ++t) {
                if ((t & 3) == 0) q = draw(key, (uint64_t)e * 8 + (t >> 2));
                uint32_t u = (uint32_t)((q >> (16 * (t & 3))) & 0xffff);
                uint64_t rb = (u >= tb) ? 1 : 0;
                uint64_t cb = (u >= ta && u < tb) || (u >= tc) ? 1 : 0;
                row = (row << 1) | rb;
                col = (col << 1) | cb;
            }
            keys[e] = (row << 32) | col;
        });
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        if ((int64_t)keys.size() > cap) return -1;
        keys_to_csr(keys, M, vkey, rowptr, colind, val);
        return (int64_t)keys.size();
    }
    parallel_dynamic(M, 256, [&](int64_t ri) {
        const int r = (int)ri;
        std::vector<int>& out = rows[r];
        const uint64_t base = (uint64_t)r * 4096;
        if (kind == 3) {
            const int nb = std::max(1, (int)param), nblk = (M + 31) / 32, br = r / 32;
            std::vector<int> bc{br};
            for (int t = 1; t < nb; ++t) bc.push_back((int)(draw(key, (uint64_t)br * 4096 + t) % (uint64_t)nblk));
            std::sort(bc.begin(), bc.end());
            bc.erase(std::unique(bc.begin(), bc.end()), bc.end());
            for (int blk : bc)
                for (int t = 0; t < 32 && blk * 32 + t < M; ++t) out.push_back(blk * 32 + t);
            return;
        }
        double fl = std::floor(avg_row);
        int d = (int)fl + ((double)(draw(key, base) >> 11) * 0x1p-53 < avg_row - fl ? 1 : 0);
        int lo = 0, hi = M;
        if (kind == 2) {
            int bw = std::max(1, (int)param);
            lo = std::max(0, r - bw);
            hi = std::min(M, r + bw + 1);
        }
        const int width = hi - lo;
        if (d >= width) {
            for (int cidx = lo; cidx < hi; ++cidx) out.push_back(cidx);
            return;
        }
        for (int t = 0; t < d; ++t) out.push_back(lo + (int)(draw(key, base + 1 + t) % (uint64_t)width));
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    });
    int64_t total = 0;
    rowptr[0] = 0;
    for (int r = 0; r < M; ++r) {
        total += (int64_t)rows[r].size();
        if (total > cap || total > 0x7fffffff) return -1;
        rowptr[r + 1] = (int)total;
    }
    parallel_dynamic(M, 256, [&](int64_t r) {
        int64_t p = rowptr[r];
        for (int cidx : rows[r]) {
            colind[p] = cidx;
            if (val) val[p] = nonzero_value(vkey, (uint64_t)p);
            ++p;
        }
    });
    return total;
}

int64_t armsp_csr_to_vcsr(int M, int K, const int* rowptr, const int* colind, const float* val, float threshold,
                          int* out_rowptr, int* out_colind, float* out_val) {
    if (M < 0 || K < 0 || !rowptr || !out_rowptr) return -1;
    auto trig = [&](float v) { return threshold > 0.0f ? std::fabs(v) > threshold : v != 0.0f; };
    int64_t size = 0;
    out_rowptr[0] = 0;
    for (int r = 0; r < M; ++r) {
        int p = rowptr[r], e = rowptr[r + 1];
        while (p < e) {
            int blk = colind[p] >> 2;
            bool keep = false;
            int q = p;
            while (q < e && (colind[q] >> 2) == blk) {
                keep = keep || trig(val ? val[q] : 1.0f);
                ++q;
            }
            if (keep) {
                if (out_colind) {
                    for (int t = 0; t < 4; ++t) {
                        out_colind[size + t] = blk * 4 + t;
                        if (out_val) out_val[size + t] = 0.0f;
                    }
                    if (out_val)
                        for (int s = p; s < q; ++s) out_val[size + (colind[s] & 3)] = val ? val[s] : 1.0f;
                }
                size += 4;
            }
            p = q;
        }
        out_rowptr[r + 1] = (int)size;
    }
    return size;
}

int armsp_csr_to_datareuse(int M, int K, const int* rowptr, const int* colind, const float* val, int* out_rowptr,
                           int* out_colind, float* out_val, uint8_t* out_flag) {
    if (M < 0 || K < 0 || !rowptr || !out_rowptr) return 1;
    const int64_t nnz = rowptr[M];
    if (nnz > 0 && (!colind || !out_colind || !out_flag)) return 1;
    if (armsp_csr_check(M, K, nnz, rowptr, colind)) return 2;
    // entries of row r+1 already placed behind their partner in row r
    std::vector<uint8_t> taken;
    int64_t size = 0;
    for (int r = 0; r < M; ++r) {
        out_rowptr[r] = (int)size;
        const bool even = (r & 1) == 0;
        if (even) {
            const int q0 = r + 1 < M ? rowptr[r + 1] : 0, q1 = r + 1 < M ? rowptr[r + 2] : 0;
            taken.assign((size_t)(q1 - q0), 0);
            int q = q0;
            for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
                const int c = colind[p];
                out_colind[size] = c;
                if (out_val) out_val[size] = val ? val[p] : 1.0f;
                out_flag[size++] = 0;
                while (q < q1 && colind[q] < c) ++q;  // both rows ascend: one merge pass
                if (q < q1 && colind[q] == c) {
                    out_colind[size] = c;
                    if (out_val) out_val[size] = val ? val[q] : 1.0f;
                    out_flag[size++] = 1;
                    taken[(size_t)(q - q0)] = 1;
                }
            }
        } else {
            const int q0 = rowptr[r];
            for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
                if (taken[(size_t)(p - q0)]) continue;
                out_colind[size] = colind[p];
                if (out_val) out_val[size] = val ? val[p] : 1.0f;
                out_flag[size++] = 0;
            }
        }
    }
    out_rowptr[M] = (int)size;
    return 0;
}

int armsp_csr_check(int M, int K, int64_t nnz, const int* rowptr, const int* colind) {
    if (!rowptr) return 1;
    if (rowptr[0] != 0) return 1;
    for (int r = 0; r < M; ++r)
        if (rowptr[r + 1] < rowptr[r]) return 2;
    if (rowptr[M] != nnz) return 3;
    for (int r = 0; r < M; ++r)
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            if (colind[p] < 0 || colind[p] >= K) return 4;
            if (p > rowptr[r] && colind[p] <= colind[p - 1]) return 5;
        }
    return 0;
}

}  // extern "C"
