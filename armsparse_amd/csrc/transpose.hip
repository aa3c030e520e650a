// Stable CSR → CSC transpose and COO → CSR builder for gfx950.
//
// Reference: toCSC's column-major scan (SpMM.cpp:80-95) and scipy's tocsc in
// the GCN script (gcn_custom.py:48-55) produce, for every column, its entries
// with ascending rows; csr2cscKernel (gcn/pytorch-custom/spmm_kernel.cu:381-423)
// would call cuSPARSE for it.  Here: an LSD radix sort of (column, position)
// pairs, 8 bits per pass.  The input is already in row order, and every pass
// is stable (per-block digit histograms, a device-wide exclusive scan, and a
// scatter whose in-block ranks come from wave ballots in item order), so rows
// stay ascending inside each column.  The last pass gathers row ids and
// values; colptr is read off the sorted column keys.
//
// COO → CSR (the canonical CSR of toCSR, SDMM.cpp:355-370 / SpMM.cpp:64-79,
// built from unordered triplets on the device): the same stable sort by
// column, then by row, so entries end in (row, column) order with duplicates
// in input order; duplicates are summed in that order, entries whose sum is
// zero are dropped on request (toCSR keeps only `!= 0`), and rowptr is read off
// the sorted rows.
//
// CSR → vector-aware CSR (toVectorAwareCSR, SpMV.cpp:15-38) on the device: a
// decision per block-opening entry, an exclusive scan of the decisions, and a
// scatter of the stored blocks.
#include <climits>

#include "common.hpp"

namespace armsp {
namespace {

constexpr int TPB = 256, IPT = 16, TILE = TPB * IPT;  // items per block: 4 waves × 16 rounds × 64
constexpr int NB = 256;                               // 8-bit digits
constexpr int SCAN_CHUNK = 2048;                       // 256 threads × 8

__global__ __launch_bounds__(256) void k_expand_rows(int M, const int* __restrict__ rowptr, int* __restrict__ rows) {
    const int lane = threadIdx.x & 63;
    const int r = uniform((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    if (r >= M) return;
    const int b = rowptr[r], e = rowptr[r + 1];
    for (int p = b + lane; p < e; p += 64) rows[p] = r;
}

__global__ __launch_bounds__(256) void k_radix_hist(int n, int shift, const int* __restrict__ keys,
                                                    int* __restrict__ hist, int nblocks) {
    __shared__ int cnt[NB];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * TILE;
    for (int j = 0; j < IPT; ++j) {
        const int64_t e = base + (int64_t)j * TPB + threadIdx.x;
        if (e < n) atomicAdd(&cnt[(keys[e] >> shift) & (NB - 1)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// idx_in == nullptr: item e carries its own position.  final_pass: instead of
// idx_out, write rowind[dst] = rows[idx], valT[dst] = val[idx].
__global__ __launch_bounds__(256) void k_radix_scatter(int n, int shift, const int* __restrict__ keys_in,
                                                       const int* __restrict__ idx_in, const int* __restrict__ offs,
                                                       int nblocks, int* __restrict__ keys_out,
                                                       int* __restrict__ idx_out, const int* __restrict__ rows,
                                                       const float* __restrict__ val, int* __restrict__ rowind,
                                                       float* __restrict__ valT, int final_pass) {
    __shared__ int wcnt[4][NB];
    __shared__ int wpre[4][NB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int i = tid; i < 4 * NB; i += TPB) wcnt[i / NB][i % NB] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * TILE + (int64_t)w * (IPT * 64);
    const uint64_t lt = (1ull << lane) - 1ull;
    int key[IPT], idx[IPT], dig[IPT], pos[IPT];
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int64_t e = base + r * 64 + lane;
        const bool valid = e < n;
        key[r] = valid ? keys_in[e] : 0;
        idx[r] = valid ? (idx_in ? idx_in[e] : (int)e) : -1;
        const int d = (key[r] >> shift) & (NB - 1);
        dig[r] = d;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1;
            const uint64_t bal = __ballot(on);
            m &= on ? bal : ~bal;
        }
        int before = 0;
        if (valid) before = wcnt[w][d];
        pos[r] = before + __popcll(m & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid && (m >> lane) == 1ull) wcnt[w][d] = before + __popcll(m);  // highest peer updates
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    {
        int run = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            wpre[q][tid] = run;
            run += wcnt[q][tid];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        if (idx[r] < 0) continue;
        const int d = dig[r];
        const int dst = offs[(int64_t)d * nblocks + blockIdx.x] + wpre[w][d] + pos[r];
        keys_out[dst] = key[r];
        if (final_pass) {
            rowind[dst] = rows[idx[r]];
            if (valT) valT[dst] = val[idx[r]];
        } else {
            idx_out[dst] = idx[r];
        }
    }
}

// In-place exclusive scan of 2048-element chunks; sums[b] = chunk total.
__global__ __launch_bounds__(256) void k_scan_chunks(int64_t n, int* __restrict__ data, int* __restrict__ sums) {
    __shared__ int wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + (int64_t)tid * 8;
    int v[8], s = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        v[i] = base + i < n ? data[base + i] : 0;
        s += v[i];
    }
    int inc = s;  // inclusive wave scan of per-thread sums
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(inc, d);
        if (lane >= d) inc += o;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int wofs = 0;
    for (int q = 0; q < w; ++q) wofs += wsum[q];
    int run = wofs + inc - s;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (base + i < n) data[base + i] = run;
        run += v[i];
    }
    if (tid == TPB - 1) sums[blockIdx.x] = run;
}

__global__ void k_scan_add(int64_t n, int* __restrict__ data, const int* __restrict__ sums) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) data[i] += sums[i / SCAN_CHUNK];
}

// colptr[c] = first sorted position whose column is ≥ c.
__global__ void k_colptr(int nnz, int K, const int* __restrict__ keys, int* __restrict__ colptr) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p > nnz) return;
    const int k = p < nnz ? keys[p] : K;
    const int kp = p > 0 ? keys[p - 1] : -1;
    for (int c = kp + 1; c <= k; ++c) colptr[c] = (int)p;
}

// rowptr[r] = first position whose row is ≥ r, for the *n_dev sorted rows.
__global__ void k_rowptr_dev(const int* __restrict__ n_dev, int M, const int* __restrict__ keys,
                             int* __restrict__ rowptr) {
    const int n = *n_dev;
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p > n) return;
    const int k = p < n ? keys[p] : M;
    const int kp = p > 0 ? keys[p - 1] : -1;
    for (int c = kp + 1; c <= k; ++c) rowptr[c] = (int)p;
}

__global__ void k_coo_check(int nnz, int M, int K, const int* __restrict__ rows, const int* __restrict__ cols,
                            int* __restrict__ bad) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int r = rows[p], c = cols[p];
    if (r < 0 || r >= M || c < 0 || c >= K) atomicOr(bad, 1);
}

__global__ void k_gather_keys(int n, const int* __restrict__ idx, const int* __restrict__ src,
                              int* __restrict__ dst) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[idx[p]];
}

// keep[p] = 1 where p heads a run of equal (row, column) and the run is kept;
// sums[p] = the run's values added in order (runs of duplicates are short).
__global__ void k_coo_runs(int n, const int* __restrict__ srow, const int* __restrict__ scol,
                           const float* __restrict__ sval, float* __restrict__ sums, int* __restrict__ keep,
                           int drop_zeros) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int r = srow[p], c = scol[p];
    const bool head = p == 0 || srow[p - 1] != r || scol[p - 1] != c;
    int k = 0;
    if (head) {
        float s = sval ? sval[p] : 1.0f;
        for (int64_t q = p + 1; q < n && srow[q] == r && scol[q] == c; ++q)
            if (sval) s += sval[q];
        if (sval) sums[p] = s;
        k = (!drop_zeros || !sval || s != 0.0f) ? 1 : 0;
    }
    keep[p] = k;
}

__global__ void k_coo_compact(int n, const int* __restrict__ srow, const int* __restrict__ scol,
                              const float* __restrict__ sums, const int* __restrict__ keep,
                              const int* __restrict__ pos, int* __restrict__ urow, int* __restrict__ colind,
                              float* __restrict__ val, int* __restrict__ n_out) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (keep[p]) {
        const int d = pos[p];
        urow[d] = srow[p];
        colind[d] = scol[p];
        if (val) val[d] = sums[p];
    }
    if (p == n - 1) *n_out = pos[p] + keep[p];
}

int64_t scan_ws_ints(int64_t n) {
    int64_t total = 0;
    while (n > 1) {
        n = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
        total += n;
    }
    return total + 1;
}

void exclusive_scan(int64_t n, int* data, int* ws, hipStream_t s) {
    const int64_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    hipLaunchKernelGGL(k_scan_chunks, dim3((unsigned)nb), dim3(TPB), 0, s, n, data, ws);
    if (nb > 1) {
        exclusive_scan(nb, ws, ws + nb, s);
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, data, ws);
    }
}

// One stable LSD sort of keys (values in [0, range)) carrying positions:
// idx_in == nullptr sorts positions 0..n-1.  Non-final: sorted keys and
// positions land in *kout_final / *iout_final (ping-pong buffers A/B).  Final
// (gather_src != nullptr): the last pass writes gathered[dst] = gather_src[pos]
// and gvals[dst] = vals[pos] instead of positions.
void radix_sort(int n, int range, const int* keys_in, const int* idx_in, int* keysA, int* keysB, int* idxA,
                int* idxB, int* hist, int* scanws, const int* gather_src, const float* vals, int* gathered,
                float* gvals, const int** kout_final, const int** iout_final, hipStream_t s) {
    const int nblocks = (int)(((int64_t)n + TILE - 1) / TILE);
    const int64_t nhist = (int64_t)NB * nblocks;
    int passes = 1;
    while (range > 1 && ((range - 1) >> (8 * passes)) > 0) ++passes;
    const int* kin = keys_in;
    const int* iin = idx_in;
    int* kout = keysA;
    int* iout = idxA;
    for (int pass = 0; pass < passes; ++pass) {
        const bool last = gather_src && pass == passes - 1;
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(TPB), 0, s, n, 8 * pass, kin, hist, nblocks);
        exclusive_scan(nhist, hist, scanws, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(TPB), 0, s, n, 8 * pass, kin, iin, hist, nblocks,
                           kout, iout, gather_src, vals, gathered, gvals, last ? 1 : 0);
        kin = kout;
        iin = iout;
        kout = (kout == keysA) ? keysB : keysA;
        iout = (iout == idxA) ? idxB : idxA;
    }
    *kout_final = kin;
    *iout_final = iin;
}

// ------------------------------------------------ column blocks of a CSR matrix
// A CSR matrix split into nb blocks of columns [x·W, (x+1)·W), each a CSR matrix
// of its own over all M rows with the row's entries in their order.  Block x of
// row r is a run of the row's entries (columns increase along a row), found by a
// binary search per block boundary: one wave per row, lane x searching boundary
// x.  The (block, row) counts in block-major order are scanned once, which places
// every run; a second pass copies the runs with the row's wave.

// entries of row [a, e) with column < c (a binary search)
__device__ __forceinline__ int cols_below(const int* __restrict__ colind, int a, int e, int c) {
    int lo = a, hi = e;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (colind[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// cnt[x·M + r] = entries of row r in block x; cnt[nb·M] = 0 (the scan total)
__global__ __launch_bounds__(256) void k_split_count(int M, int nb, int W, const int* __restrict__ rowptr,
                                                     const int* __restrict__ colind, int* __restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int r = uniform((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    if (r == M && lane == 0) cnt[(int64_t)nb * M] = 0;
    if (r >= M) return;
    const int a = rowptr[r], e = rowptr[r + 1];
    const int sx = lane <= nb ? (lane == 0 ? a : (lane == nb ? e : cols_below(colind, a, e, lane * W))) : e;
    const int nx = __shfl(sx, lane + 1 <= nb ? lane + 1 : nb);
    if (lane < nb) cnt[(int64_t)lane * M + r] = nx - sx;
}

// out_rowptr[x·(M+1) + r] = pos[x·M + r] − pos[x·M], r = 0..M
__global__ void k_split_rowptr(int M, int nb, const int* __restrict__ pos, int* __restrict__ out_rowptr) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)nb * (M + 1)) return;
    const int x = (int)(i / (M + 1)), r = (int)(i - (int64_t)x * (M + 1));
    const int64_t b = (int64_t)x * M;
    out_rowptr[i] = pos[b + r] - pos[b];
}

// base[x] = pos[x·M], x = 0..nb (the blocks' first entries; base[nb] = nnz)
__global__ void k_split_base(int M, int nb, const int* __restrict__ pos, int* __restrict__ base) {
    const int x = threadIdx.x;
    if (x <= nb) base[x] = pos[(int64_t)x * M];
}

__global__ __launch_bounds__(256) void k_split_scatter(int M, int nb, int W, const int* __restrict__ rowptr,
                                                       const int* __restrict__ colind, const float* __restrict__ val,
                                                       const int* __restrict__ pos, int* __restrict__ out_colind,
                                                       float* __restrict__ out_val) {
    const int lane = threadIdx.x & 63;
    const int r = uniform((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    if (r >= M) return;
    const int a = rowptr[r], e = rowptr[r + 1];
    const int sx = lane <= nb ? (lane == 0 ? a : (lane == nb ? e : cols_below(colind, a, e, lane * W))) : e;
    for (int x = 0; x < nb; ++x) {
        const int s0 = __shfl(sx, x), s1 = __shfl(sx, x + 1);
        const int64_t d = pos[(int64_t)x * M + r];
        for (int p = s0 + lane; p < s1; p += 64) {
            out_colind[d + (p - s0)] = colind[p];
            if (out_val) out_val[d + (p - s0)] = val ? val[p] : 1.0f;
        }
    }
}

// err bit 2: a column out of [0, K) or not above the previous one in its row
// (start[p] = 1 where a non-empty row begins, from k_vcsr_rows)
__global__ void k_cols_increasing(int nnz, int K, const int* __restrict__ colind, const uint8_t* __restrict__ start,
                                  int* __restrict__ err) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int c = colind[p];
    const bool first = p == 0 || start[p] != 0;
    if (c < 0 || c >= K || (!first && colind[p - 1] >= c)) atomicOr(err, 2);
}

// ------------------------------------------------ CSR → vector-aware CSR
// toVectorAwareCSR (SpMV.cpp:15-38) on the device, from CSR instead of a dense
// matrix.  A row's aligned block of columns [4b, 4b+4) is stored iff one of the
// row's values in it triggers (v != 0, or |v| > threshold when threshold > 0:
// numaAware/numaAwareSpMV.cpp:48), as four slots holding the row's values (0
// where absent) with column indices 4b..4b+3.  The entry that opens a block (the
// row's first entry in it) decides it alone, reading the at most three entries
// after it (columns strictly increase within a row, so a block holds at most
// four).  An exclusive scan over those decisions gives every stored block its
// place, and the blocked rowptr is read off the scan at the row starts.

// start[p] = 1 where a non-empty row begins; err bit 0: malformed rowptr.
__global__ void k_vcsr_rows(int M, int nnz, const int* __restrict__ rowptr, uint8_t* __restrict__ start,
                            int* __restrict__ err) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= M) return;
    const int a = rowptr[r], b = rowptr[r + 1];
    if ((r == 0 && a != 0) || (r == M - 1 && b != nnz) || a < 0 || b < a || b > nnz) {
        atomicOr(err, 1);
        return;
    }
    if (a < b) start[a] = 1;
}

__global__ void k_vcsr_clear(int64_t n, uint8_t* __restrict__ start, int* __restrict__ err) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) start[i] = 0;
    if (i == 0) *err = 0;
}

__device__ __forceinline__ bool vcsr_trig(float v, float threshold) {
    return threshold > 0.0f ? fabsf(v) > threshold : v != 0.0f;
}

// keep[p] = 1 iff entry p opens a stored block, keep[nnz] = 0 (the scan then
// leaves the block count there).  err bit 1: a column out of [0, K) or a row
// whose columns do not strictly increase.
__global__ void k_vcsr_mark(int nnz, int K, const int* __restrict__ colind, const float* __restrict__ val,
                            float threshold, const uint8_t* __restrict__ start, int* __restrict__ keep,
                            int* __restrict__ err) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p > nnz) return;
    if (p == nnz) {
        keep[p] = 0;
        return;
    }
    const int c = colind[p];
    const bool first = p == 0 || start[p] != 0;
    const int cp = first ? -1 : colind[p - 1];
    if (c < 0 || c >= K || (!first && cp >= c)) atomicOr(err, 2);
    const bool opens = first || (cp >> 2) != (c >> 2);
    bool on = opens && vcsr_trig(val ? val[p] : 1.0f, threshold);
    for (int d = 1; opens && !on && d <= 3; ++d) {
        const int64_t q = p + d;
        if (q >= nnz || start[q] || (colind[q] >> 2) != (c >> 2)) break;
        on = vcsr_trig(val ? val[q] : 1.0f, threshold);
    }
    keep[p] = on ? 1 : 0;
}

// out_rowptr[r] = 4 · (stored blocks before row r), r = 0..M (with no rows, the one
// entry is 0 and rowptr is not read).
__global__ void k_vcsr_rowptr(int M, const int* __restrict__ rowptr, const int* __restrict__ pos,
                              int* __restrict__ out_rowptr) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r > M) return;
    out_rowptr[r] = M == 0 ? 0 : 4 * pos[rowptr[r]];
}

// The opening entry of every stored block writes its four slots.
__global__ void k_vcsr_fill(int nnz, const int* __restrict__ colind, const float* __restrict__ val,
                            const uint8_t* __restrict__ start, const int* __restrict__ pos,
                            int* __restrict__ out_colind, float* __restrict__ out_val) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int b = pos[p];
    if (pos[p + 1] == b) return;
    const int c = colind[p];
    const int c0 = (c >> 2) * 4;
    float slot[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    slot[c & 3] = val ? val[p] : 1.0f;
    for (int d = 1; d <= 3; ++d) {
        const int64_t q = p + d;
        if (q >= nnz || start[q] || (colind[q] >> 2) != (c >> 2)) break;
        slot[colind[q] & 3] = val ? val[q] : 1.0f;
    }
    const int64_t o = 4 * (int64_t)b;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        out_colind[o + t] = c0 + t;
        out_val[o + t] = slot[t];
    }
}

}  // namespace
}  // namespace armsp

extern "C" int armsp_csr_transpose(int M, int K, int nnz, const int* rowptr, const int* colind, const float* val,
                                   int* colptr, int* rowind, float* valT, void* stream) {
    using namespace armsp;
    if (M < 0 || K < 0 || nnz < 0) return ARMSP_ERROR_INVALID_VALUE;
    if (!colptr || (nnz > 0 && (!rowptr || !colind || !rowind))) return ARMSP_ERROR_INVALID_VALUE;
    if ((val == nullptr) != (valT == nullptr)) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    TimedLaunch t("transpose", s);
    if (nnz == 0 || M == 0) {
        hipLaunchKernelGGL(k_colptr, dim3(1), dim3(64), 0, s, 0, K, (const int*)nullptr, colptr);
        ARMSP_LAUNCHED();
        t.stop();
        return ARMSP_SUCCESS;
    }
    const int nblocks = (int)(((int64_t)nnz + TILE - 1) / TILE);
    const int64_t nhist = (int64_t)NB * nblocks;
    if (nhist > INT_MAX) return ARMSP_ERROR_NOT_SUPPORTED;
    const size_t n4 = ((size_t)nnz * 4 + 255) & ~(size_t)255;
    const size_t h4 = ((size_t)nhist * 4 + 255) & ~(size_t)255;
    const size_t s4 = ((size_t)scan_ws_ints(nhist) * 4 + 255) & ~(size_t)255;
    void* ws = nullptr;
    int st = ws_alloc(&ws, 5 * n4 + h4 + s4, s);
    if (st) return st;
    char* p = static_cast<char*>(ws);
    int* rows = reinterpret_cast<int*>(p);
    int* keysA = reinterpret_cast<int*>(p + n4);
    int* keysB = reinterpret_cast<int*>(p + 2 * n4);
    int* idxA = reinterpret_cast<int*>(p + 3 * n4);
    int* idxB = reinterpret_cast<int*>(p + 4 * n4);
    int* hist = reinterpret_cast<int*>(p + 5 * n4);
    int* scanws = reinterpret_cast<int*>(p + 5 * n4 + h4);

    hipLaunchKernelGGL(k_expand_rows, dim3((unsigned)((M + 3) / 4)), dim3(256), 0, s, M, rowptr, rows);
    const int* sorted_keys = nullptr;
    const int* sorted_idx = nullptr;
    radix_sort(nnz, K, colind, nullptr, keysA, keysB, idxA, idxB, hist, scanws, rows, val, rowind, valT,
               &sorted_keys, &sorted_idx, s);
    hipLaunchKernelGGL(k_colptr, dim3((unsigned)(((int64_t)nnz + 1 + 255) / 256)), dim3(256), 0, s, nnz, K,
                       sorted_keys, colptr);
    hipError_t e = hipGetLastError();
    ws_free(ws, s);
    if (e != hipSuccess) return hip_fail(e);
    t.stop();
    return ARMSP_SUCCESS;
}

extern "C" int armsp_coo_to_csr(int M, int K, int nnz, const int* rows, const int* cols, const float* vals,
                                int drop_zeros, int* rowptr, int* colind, float* val, int* nnz_out, void* stream) {
    using namespace armsp;
    if (M < 0 || K < 0 || nnz < 0) return ARMSP_ERROR_INVALID_VALUE;
    if (!rowptr || !nnz_out || (nnz > 0 && (!rows || !cols || !colind))) return ARMSP_ERROR_INVALID_VALUE;
    if (nnz > 0 && (vals == nullptr) != (val == nullptr)) return ARMSP_ERROR_INVALID_VALUE;
    if (nnz > 0 && (M == 0 || K == 0)) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    TimedLaunch t("coo_to_csr", s);
    const int nblocks = (int)(((int64_t)nnz + TILE - 1) / TILE);
    const int64_t nhist = (int64_t)NB * (nblocks > 0 ? nblocks : 1);
    if (nhist > INT_MAX) return ARMSP_ERROR_NOT_SUPPORTED;
    const int64_t nscan = nhist > nnz ? nhist : (int64_t)nnz;
    const size_t n4 = ((size_t)(nnz > 0 ? nnz : 1) * 4 + 255) & ~(size_t)255;
    const size_t h4 = ((size_t)nhist * 4 + 255) & ~(size_t)255;
    const size_t s4 = ((size_t)scan_ws_ints(nscan) * 4 + 255) & ~(size_t)255;
    void* ws = nullptr;
    int st = ws_alloc(&ws, 11 * n4 + h4 + s4 + 256, s);
    if (st) return st;
    char* p = static_cast<char*>(ws);
    int* keysA = reinterpret_cast<int*>(p);
    int* keysB = reinterpret_cast<int*>(p + n4);
    int* idxA = reinterpret_cast<int*>(p + 2 * n4);
    int* idxB = reinterpret_cast<int*>(p + 3 * n4);
    int* rperm = reinterpret_cast<int*>(p + 4 * n4);
    int* scol = reinterpret_cast<int*>(p + 5 * n4);
    float* sval = reinterpret_cast<float*>(p + 6 * n4);
    float* sums = reinterpret_cast<float*>(p + 7 * n4);
    int* keep = reinterpret_cast<int*>(p + 8 * n4);
    int* pos = reinterpret_cast<int*>(p + 9 * n4);
    int* urow = reinterpret_cast<int*>(p + 10 * n4);
    int* hist = reinterpret_cast<int*>(p + 11 * n4);
    int* scanws = reinterpret_cast<int*>(p + 11 * n4 + h4);
    int* bad = reinterpret_cast<int*>(p + 11 * n4 + h4 + s4);
    auto fail = [&](int code) {
        ws_free(ws, s);
        return code;
    };
    if (nnz == 0) {
        hipError_t e = hipMemsetAsync(nnz_out, 0, sizeof(int), s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_rowptr_dev, dim3(1), dim3(64), 0, s, (const int*)nnz_out, M, (const int*)nullptr,
                               rowptr);
            e = hipGetLastError();
        }
        if (e != hipSuccess) return fail(hip_fail(e));
        ws_free(ws, s);
        t.stop();
        return ARMSP_SUCCESS;
    }
    // every index in range, or an error before anything is written through one
    int bad_h = 0;
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(int), s);
    const unsigned g1 = (unsigned)(((int64_t)nnz + 255) / 256);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_coo_check, dim3(g1), dim3(256), 0, s, nnz, M, K, rows, cols, bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&bad_h, bad, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(hip_fail(e));
    if (bad_h) return fail(ARMSP_ERROR_INVALID_VALUE);

    // by column, then (stably) by row: (row, column) order, duplicates in input order
    const int* ck = nullptr;
    const int* ci = nullptr;
    radix_sort(nnz, K, cols, nullptr, keysA, keysB, idxA, idxB, hist, scanws, nullptr, nullptr, nullptr, nullptr,
               &ck, &ci, s);
    hipLaunchKernelGGL(k_gather_keys, dim3(g1), dim3(256), 0, s, nnz, ci, rows, rperm);
    // the column pass left its positions in idxA or idxB; sort rows into the other pair
    int* k2a = (ci == idxA) ? keysA : keysB;
    int* k2b = (ci == idxA) ? keysB : keysA;
    int* i2a = (ci == idxA) ? idxB : idxA;
    int* i2b = (ci == idxA) ? idxA : idxB;
    (void)i2b;
    const int* srow = nullptr;
    const int* si = nullptr;
    radix_sort(nnz, M, rperm, ci, k2a, k2b, i2a, const_cast<int*>(ci), hist, scanws, cols, vals, scol,
               vals ? sval : nullptr, &srow, &si, s);
    hipLaunchKernelGGL(k_coo_runs, dim3(g1), dim3(256), 0, s, nnz, srow, scol, vals ? sval : nullptr, sums, keep,
                       drop_zeros);
    e = hipMemcpyAsync(pos, keep, (size_t)nnz * sizeof(int), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return fail(hip_fail(e));
    exclusive_scan(nnz, pos, scanws, s);
    hipLaunchKernelGGL(k_coo_compact, dim3(g1), dim3(256), 0, s, nnz, srow, scol, sums, keep, pos, urow, colind,
                       val, nnz_out);
    hipLaunchKernelGGL(k_rowptr_dev, dim3((unsigned)(((int64_t)nnz + 1 + 255) / 256)), dim3(256), 0, s,
                       (const int*)nnz_out, M, urow, rowptr);
    e = hipGetLastError();
    ws_free(ws, s);
    if (e != hipSuccess) return hip_fail(e);
    t.stop();
    return ARMSP_SUCCESS;
}

extern "C" int armsp_csr_to_vcsr_device(int M, int K, int nnz, const int* rowptr, const int* colind,
                                        const float* val, float threshold, int* out_rowptr, int* out_colind,
                                        float* out_val, int64_t capacity, int64_t* padded, void* stream) {
    using namespace armsp;
    if (M < 0 || K < 0 || nnz < 0 || !padded || !out_rowptr || !rowptr) return ARMSP_ERROR_INVALID_VALUE;
    if (nnz > 0 && !colind) return ARMSP_ERROR_INVALID_VALUE;
    if ((out_colind == nullptr) != (out_val == nullptr)) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 && nnz > 0) return ARMSP_ERROR_INVALID_VALUE;
    if (nnz > INT_MAX - 2) return ARMSP_ERROR_NOT_SUPPORTED;
    hipStream_t s = as_stream(stream);
    TimedLaunch t("csr_to_vcsr", s);
    const int64_t n1 = (int64_t)nnz + 1;  // scanned entries: nnz decisions and the total
    const size_t p4 = ((size_t)(n1 + 1) * 4 + 255) & ~(size_t)255;
    const size_t s4 = ((size_t)scan_ws_ints(n1) * 4 + 255) & ~(size_t)255;
    const size_t b1 = ((size_t)(nnz > 0 ? nnz : 1) + 255) & ~(size_t)255;
    void* ws = nullptr;
    int st = ws_alloc(&ws, p4 + s4 + b1, s);
    if (st) return st;
    char* w = static_cast<char*>(ws);
    int* pos = reinterpret_cast<int*>(w);
    int* err = pos + n1;  // one int past the scanned range
    int* scanws = reinterpret_cast<int*>(w + p4);
    uint8_t* start = reinterpret_cast<uint8_t*>(w + p4 + s4);
    auto fail = [&](int code) {
        ws_free(ws, s);
        return code;
    };
    const unsigned gs = (unsigned)(((int64_t)(nnz > 0 ? nnz : 1) + 255) / 256);
    hipLaunchKernelGGL(k_vcsr_clear, dim3(gs), dim3(256), 0, s, (int64_t)nnz, start, err);
    if (M > 0)
        hipLaunchKernelGGL(k_vcsr_rows, dim3((unsigned)(((int64_t)M + 255) / 256)), dim3(256), 0, s, M, nnz, rowptr,
                           start, err);
    hipLaunchKernelGGL(k_vcsr_mark, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, s, nnz, K, colind, val,
                       threshold, start, pos, err);
    exclusive_scan(n1, pos, scanws, s);
    int h[2] = {0, 0};  // the block count and the error bits
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, pos + nnz, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(hip_fail(e));
    if (h[1]) return fail(ARMSP_ERROR_INVALID_VALUE);
    if (4 * (int64_t)h[0] > INT_MAX) return fail(ARMSP_ERROR_NOT_SUPPORTED);
    *padded = 4 * (int64_t)h[0];
    hipLaunchKernelGGL(k_vcsr_rowptr, dim3((unsigned)(((int64_t)M + 1 + 255) / 256)), dim3(256), 0, s, M, rowptr,
                       (const int*)pos, out_rowptr);
    if (out_colind) {
        if (capacity < *padded) return fail(ARMSP_ERROR_INVALID_VALUE);
        if (nnz > 0)
            hipLaunchKernelGGL(k_vcsr_fill, dim3(gs), dim3(256), 0, s, nnz, colind, val, (const uint8_t*)start,
                               (const int*)pos, out_colind, out_val);
    }
    e = hipGetLastError();
    ws_free(ws, s);
    if (e != hipSuccess) return hip_fail(e);
    t.stop();
    return ARMSP_SUCCESS;
}

extern "C" int armsp_csr_column_split(int M, int K, int nnz, const int* rowptr, const int* colind, const float* val,
                                      int nblocks, int* out_rowptr, int* out_colind, float* out_val,
                                      int64_t* block_start, void* stream) {
    using namespace armsp;
    if (M < 0 || K < 0 || nnz < 0 || nblocks < 1 || nblocks > 63) return ARMSP_ERROR_INVALID_VALUE;
    if (!rowptr || !out_rowptr || !block_start || (nnz > 0 && (!colind || !out_colind))) return ARMSP_ERROR_INVALID_VALUE;
    if (val && !out_val) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 && nnz > 0) return ARMSP_ERROR_INVALID_VALUE;
    if ((int64_t)nblocks * M + 1 > INT_MAX) return ARMSP_ERROR_NOT_SUPPORTED;
    hipStream_t s = as_stream(stream);
    TimedLaunch t("csr_column_split", s);
    const int W = K > 0 ? (K + nblocks - 1) / nblocks : 1;
    const int64_t ncnt = (int64_t)nblocks * M + 1;  // (block, row) counts and the total
    const size_t c4 = ((size_t)ncnt * 4 + 255) & ~(size_t)255;
    const size_t s4 = ((size_t)scan_ws_ints(ncnt) * 4 + 255) & ~(size_t)255;
    const size_t b1 = ((size_t)(nnz > 0 ? nnz : 1) + 255) & ~(size_t)255;
    void* ws = nullptr;
    int st = ws_alloc(&ws, c4 + s4 + b1 + 512, s);
    if (st) return st;
    char* w = static_cast<char*>(ws);
    int* pos = reinterpret_cast<int*>(w);
    int* scanws = reinterpret_cast<int*>(w + c4);
    uint8_t* start = reinterpret_cast<uint8_t*>(w + c4 + s4);
    int* err = reinterpret_cast<int*>(w + c4 + s4 + b1);
    int* base = err + 1;  // nblocks + 1 ints
    auto fail = [&](int code) {
        ws_free(ws, s);
        return code;
    };
    // canonical CSR: rowptr from 0 to nnz, non-decreasing; columns in range and increasing along a row
    const unsigned gs = (unsigned)(((int64_t)(nnz > 0 ? nnz : 1) + 255) / 256);
    hipLaunchKernelGGL(k_vcsr_clear, dim3(gs), dim3(256), 0, s, (int64_t)nnz, start, err);
    if (M > 0)
        hipLaunchKernelGGL(k_vcsr_rows, dim3((unsigned)(((int64_t)M + 255) / 256)), dim3(256), 0, s, M, nnz, rowptr,
                           start, err);
    if (nnz > 0) hipLaunchKernelGGL(k_cols_increasing, dim3(gs), dim3(256), 0, s, nnz, K, colind, start, err);
    int h = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(hip_fail(e));
    if (h) return fail(ARMSP_ERROR_INVALID_VALUE);
    const unsigned gr = (unsigned)(((int64_t)M + 1 + 3) / 4);  // a wave per row, and one for the total
    hipLaunchKernelGGL(k_split_count, dim3(gr), dim3(256), 0, s, M, nblocks, W, rowptr, colind, pos);
    exclusive_scan(ncnt, pos, scanws, s);
    hipLaunchKernelGGL(k_split_rowptr, dim3((unsigned)(((int64_t)nblocks * (M + 1) + 255) / 256)), dim3(256), 0, s, M,
                       nblocks, (const int*)pos, out_rowptr);
    if (nnz > 0)
        hipLaunchKernelGGL(k_split_scatter, dim3((unsigned)(((int64_t)M + 3) / 4)), dim3(256), 0, s, M, nblocks, W,
                           rowptr, colind, val, (const int*)pos, out_colind, out_val);
    hipLaunchKernelGGL(k_split_base, dim3(1), dim3(64), 0, s, M, nblocks, (const int*)pos, base);
    int hb[64];
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hb, base, sizeof(int) * (size_t)(nblocks + 1), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ws_free(ws, s);
    if (e != hipSuccess) return hip_fail(e);
    for (int x = 0; x <= nblocks; ++x) block_start[x] = hb[x];
    t.stop();
    return ARMSP_SUCCESS;
}
