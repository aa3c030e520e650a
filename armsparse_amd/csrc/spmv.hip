// CSR SpMV for gfx950.
//
// Reference: sparse_dense_vector (SpMV.cpp:155-179) runs an OpenMP loop over
// rows of a zero-padded 1×4-blocked CSR; numaAware/numaAwareSpMV.cpp:102-125
// splits the rows into four NUMA quarters.  Here:
//  ROW   — G lanes per row, lanes stride the row, butterfly reduction.
//  MERGE — merge-path over (row ends, nonzeros): each wave owns 64·IPT items,
//          stages row ends and products val·x in LDS, every lane walks IPT
//          consecutive items, a wave-level segmented scan joins rows cut
//          between lanes, and a carry per wave is added by a fixup kernel.
//  VCSR  — the reference's blocked format with its exact summation order:
//          every row's four partial sums s0..s3 are each added in block order
//          (by four lanes of a short row, by a wave for a medium row, by a
//          workgroup for a long one) and the row total is (s0+s2)+(s1+s3), as
//          vadd/vpadd compute it.
#include <climits>
#include <cstdlib>

#include "common.hpp"

namespace armsp {
namespace {

template <int G>
__global__ __launch_bounds__(256) void k_spmv_row(int M, const int* __restrict__ rowptr, const int* __restrict__ colind,
                                                  const float* __restrict__ val, const float* __restrict__ x,
                                                  float* __restrict__ y, int flags) {
    const int lane = threadIdx.x & (G - 1);
    const int row = (int)((blockIdx.x * 256u + threadIdx.x) / G);
    if (row >= M) return;
    const int beg = rowptr[row], end = rowptr[row + 1];
    float s = 0.0f;
    for (int p = beg + lane; p < end; p += G) s = fmaf(val ? val[p] : 1.0f, x[colind[p]], s);
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) s += __shfl_xor(s, m, G);
    if (lane == 0) y[row] = (flags & ARMSP_ACCUMULATE) ? y[row] + s : s;
}

template <int IPT>
__global__ __launch_bounds__(256) void k_spmv_merge(int M, int nnz, const int* __restrict__ rowptr,
                                                    const int* __restrict__ colind, const float* __restrict__ val,
                                                    const float* __restrict__ x, float* __restrict__ y, int flags,
                                                    const int* __restrict__ part, int* __restrict__ carry_row,
                                                    float* __restrict__ carry_val) {
    constexpr int TILE = 64 * IPT;
    __shared__ int s_re[4][TILE + 1];
    __shared__ float s_prod[4][TILE];
    __shared__ float s_out[4][TILE];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int wave = uniform((int)(blockIdx.x * 4u + wid));
    const int64_t total = (int64_t)M + nnz;
    const int64_t d0l = (int64_t)wave * TILE;
    if (d0l >= total) return;
    const int d0 = (int)d0l;
    const int d1 = (int)(d0l + TILE < total ? d0l + TILE : total);
    const int i0 = uniform(partition_at(part, wave, d0, M, nnz));  // merge_partition
    const int i1 = uniform(partition_at(part, wave + 1, d1, M, nnz));
    const int j0 = d0 - i0, j1 = d1 - i1;
    const int nrows = i1 - i0, nz = j1 - j0;
    int* re = s_re[wid];
    float* prod = s_prod[wid];
    float* out = s_out[wid];
    for (int k = lane; k < nrows; k += 64) re[k] = rowptr[i0 + k + 1];
    {
        // nonzero k = lane + 64·q: all IPT index loads, then all IPT gathers of x, in flight together
        int cs[IPT];
        float vs[IPT], xs[IPT];
#pragma unroll
        for (int q = 0; q < IPT; ++q) {
            const int k = lane + 64 * q;
            cs[q] = 0;
            vs[q] = 0.0f;
            if (k < nz) {
                cs[q] = __builtin_nontemporal_load(colind + j0 + k);
                vs[q] = val ? __builtin_nontemporal_load(val + j0 + k) : 1.0f;
            }
        }
#pragma unroll
        for (int q = 0; q < IPT; ++q) xs[q] = (lane + 64 * q < nz) ? x[cs[q]] : 0.0f;
#pragma unroll
        for (int q = 0; q < IPT; ++q) {
            const int k = lane + 64 * q;
            if (k < nz) prod[k] = vs[q] * xs[q];
        }
    }
    wave_sync_lds();

    // this lane's start on the tile's merge path
    const int items = nrows + nz;
    const int ld = lane * IPT < items ? lane * IPT : items;
    int lo = ld - nz > 0 ? ld - nz : 0, hi = ld < nrows ? ld : nrows;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (re[mid] - j0 <= ld - mid - 1)
            lo = mid + 1;
        else
            hi = mid;
    }
    int ri = lo, rj = ld - lo;
    const int ri_start = ri;
    float sum = 0.0f;
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
        if (ri + rj < items) {
            if (ri < nrows && re[ri] - j0 <= rj) {
                out[ri] = sum;
                sum = 0.0f;
                ++ri;
            } else {
                sum += prod[rj];
                ++rj;
            }
        }
    }
    // segmented inclusive scan of (row, partial) across lanes; rows are non-decreasing in lane order
    int key = ri;
    float S = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const float os = __shfl_up(S, d);
        const int ok = __shfl_up(key, d);
        if (lane >= d && ok == key) S += os;
    }
    const float prevS = __shfl_up(S, 1);
    if (lane > 0 && ri > ri_start) out[ri_start] += prevS;
    wave_sync_lds();
    const bool acc = (flags & ARMSP_ACCUMULATE) != 0;
    for (int k = lane; k < nrows; k += 64) y[i0 + k] = acc ? y[i0 + k] + out[k] : out[k];

    const float tail = read_lane(S, 63);
    if (lane == 0) {
        bool has = false;
        if (i1 < M) {
            const int rs = rowptr[i1];
            has = j1 > (rs > j0 ? rs : j0);
        }
        carry_row[wave] = has ? i1 : -1;
        carry_val[wave] = tail;
    }
}

// Adds each run of carries (consecutive waves cut inside the same row) into
// y.  A wave takes 64 consecutive carries; a segmented suffix sum (rows are
// non-decreasing along the carries, so equal tags mark one run) leaves each
// run's in-wave total on its head lane.  A run that goes on past the wave is
// finished by the wave that holds its head, 64 carries a load, and a run that
// began in an earlier wave is left to that wave.  Fixed order, no atomics.
__global__ __launch_bounds__(256) void k_spmv_fixup(int nwaves, const int* __restrict__ carry_row,
                                                    const float* __restrict__ carry_val, float* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int base = uniform((int)(blockIdx.x * 4u + (threadIdx.x >> 6))) * 64;
    if (base >= nwaves) return;
    const int w = base + lane;
    const int r = w < nwaves ? carry_row[w] : -2;
    const int rprev = (w == 0 || w > nwaves) ? -2 : carry_row[w - 1];
    float s = (w < nwaves && r >= 0) ? carry_val[w] : 0.0f;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const float t = __shfl_down(s, o);
        const int rt = __shfl_down(r, o);
        if (lane + o < 64 && rt == r) s += t;
    }
    // the run of lane 63 may go on into the next waves; the wave holding its head sums the rest
    const int r63 = read_lane(r, 63);
    const bool owner = __ballot(r == r63 && rprev != r63) != 0ull;
    float ext = 0.0f;
    if (r63 >= 0 && owner && base + 64 < nwaves && carry_row[base + 64] == r63) {
        float acc = 0.0f;
        for (int b = base + 64;; b += 64) {
            const int cr = (b + lane < nwaves) ? carry_row[b + lane] : -2;
            const uint64_t m = __ballot(cr == r63);
            const int cnt = (~m == 0ull) ? 64 : __builtin_ctzll(~m);
            if (lane < cnt) acc += carry_val[b + lane];
            if (cnt < 64) break;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        ext = acc;
    }
    if (r >= 0 && rprev != r) y[r] += (r == r63 ? s + ext : s);
}

// a·b rounded to fp32 by itself: the empty asm makes the product opaque, so the
// compiler cannot contract it with the following add into a fused multiply-add
// (not volatile: the scheduler may still move it, and the loads around it)
__device__ __forceinline__ float unfused_product(float a, float b) {
    float p = a * b;
    asm("" : "+v"(p));
    return p;
}

// Four lanes per row (rows of at most long_blocks blocks when long_blocks > 0),
// 64 rows per 256-thread block `blk`.
__device__ __forceinline__ void vcsr_short_rows(unsigned blk, int M, const int* __restrict__ rowptr,
                                                const int* __restrict__ colind, const float* __restrict__ val,
                                                const float* __restrict__ x, float* __restrict__ y, int fused,
                                                int long_blocks) {
    const int l = threadIdx.x & 3;
    const int row = (int)((blk * 256u + threadIdx.x) >> 2);
    if (row >= M) return;
    const int beg = rowptr[row], end = rowptr[row + 1];
    if (long_blocks > 0 && end - beg > 4 * long_blocks) return;  // a listed long row
    float s = 0.0f;
    // the lane's sum runs block after block in the reference's order; blocks come
    // in batches of 16 whose loads are issued while the previous batch is added (a
    // hub row of an R-MAT graph is tens of thousands of blocks walked by one
    // 4-lane group, so its latency must not be paid per block)
    constexpr int BB = 16;
    auto load = [&](int base, float (&v)[BB], float (&xv)[BB]) {
#pragma unroll
        for (int u = 0; u < BB; ++u) {
            v[u] = val[base + 4 * u + l];
            xv[u] = x[colind[base + 4 * u] + l];
        }
    };
    auto add = [&](const float (&v)[BB], const float (&xv)[BB]) {
        if (fused) {
#pragma unroll
            for (int u = 0; u < BB; ++u) s = fmaf(v[u], xv[u], s);
        } else {  // vmlaq_f32 is a multiply then an add: the product is rounded on its own
#pragma unroll
            for (int u = 0; u < BB; ++u) s += unfused_product(v[u], xv[u]);
        }
    };
    int idx = beg;
    if (idx + 4 * BB <= end) {
        float v0[BB], x0[BB], v1[BB], x1[BB];
        load(idx, v0, x0);
        while (true) {
            idx += 4 * BB;
            if (idx + 4 * BB > end) {
                add(v0, x0);
                break;
            }
            load(idx, v1, x1);
            add(v0, x0);
            idx += 4 * BB;
            if (idx + 4 * BB > end) {
                add(v1, x1);
                break;
            }
            load(idx, v0, x0);
            add(v1, x1);
        }
    }
    // the rest, fewer than BB blocks (most rows of a power-law graph are this short):
    // every load in flight at once, then the adds in block order (blocks past the
    // row are skipped, not added as zeros, which would turn a -0 sum into +0)
    const int nrem = (end - idx) >> 2;
    if (nrem > 0) {
        float v[BB], xv[BB];
#pragma unroll
        for (int u = 0; u < BB; ++u) {
            v[u] = 0.0f;
            xv[u] = 0.0f;
            if (u < nrem) {
                v[u] = val[idx + 4 * u + l];
                xv[u] = x[colind[idx + 4 * u] + l];
            }
        }
        if (fused) {
#pragma unroll
            for (int u = 0; u < BB; ++u)
                if (u < nrem) s = fmaf(v[u], xv[u], s);
        } else {
#pragma unroll
            for (int u = 0; u < BB; ++u)
                if (u < nrem) s += unfused_product(v[u], xv[u]);
        }
    }
    const float t = __fadd_rn(s, __shfl_xor(s, 2, 4));   // lanes 0/2: s0+s2, lanes 1/3: s1+s3
    const float tot = __fadd_rn(t, __shfl_xor(t, 1, 4));  // (s0+s2)+(s1+s3)
    if (l == 0) y[row] = __fadd_rn(y[row], tot);
}

__global__ __launch_bounds__(256) void k_spmv_vcsr(int M, const int* __restrict__ rowptr,
                                                   const int* __restrict__ colind, const float* __restrict__ val,
                                                   const float* __restrict__ x, float* __restrict__ y, int fused) {
    vcsr_short_rows(blockIdx.x, M, rowptr, colind, val, x, y, fused, 0);
}

// One launch, three walkers: rows of at most VCSR_MED_MIN blocks get four lanes
// each, rows of up to `long_blocks` blocks a wave each (vcsr_medium_rows), and
// longer rows a workgroup each: k_vcsr_find_long lists the medium and the long
// rows, and workgroups of 256 lanes take listed long rows from a shared counter
// while the other workgroups of the same launch walk the rest (so the walkers
// overlap, and the hub rows, listed and started first, are the critical path).  A
// workgroup walks its row in batches of 256 blocks, one block per lane: each
// lane keeps the loads of its next D batches in flight (values and x, 16 bytes
// each; a block's column index one ring further ahead, so no x load waits for
// its index), stages the block in LDS transposed — one array per position l of
// the 4-block — and wave 0's lanes 0-3 keep the reference's four partial sums
// s_l, adding each batch in block order four blocks per LDS read while the
// next batch is staged into the other LDS buffer.  With the unfused reading
// the lanes form the rounded products a·b, so lanes 0-3 only add; with the
// fused one they stage values and x and lanes 0-3 run the fmaf chain.  Either
// way each s_l sees the reference's sequence of operations: bit-exact.
constexpr int VCSR_LONG_DEFAULT = 1024;    // 512, 1024, 2048, 4096: 381, 351, 390, 385 us (1M R-MAT)
constexpr int VCSR_MED_MIN = 16;          // rows of more blocks than this (and at most long_blocks) take a wave each
constexpr int VCSR_MED_AHEAD = 4;         // steps of 16 blocks in flight per medium-row wave
constexpr int VCSR_BATCH = 256;           // blocks per batch, one per lane
constexpr int VCSR_PAD = VCSR_BATCH + 4;  // row pitch: lanes 0-3 read from distinct LDS banks
constexpr int VCSR_AHEAD = 4;             // batches in flight per lane

// Workgroups of medium-row waves per CU (1M R-MAT, 64.5M padded entries: 1, 2, 4, 8,
// 16, 32 per CU give 987, 596, 383, 355, 351, 357 us).
constexpr int VCSR_MED_WG_PER_CU = 16;

// Rows longer than this many blocks take the workgroup-per-row path
// (ARMSP_VCSR_LONG_BLOCKS overrides it, read once per process).
int vcsr_long_blocks() {
    static const int v = [] {
        const char* e = getenv("ARMSP_VCSR_LONG_BLOCKS");
        const int n = e ? atoi(e) : 0;
        return n > 0 ? n : VCSR_LONG_DEFAULT;
    }();
    return v;
}

// Lists the long rows (more than long_blocks blocks) in list[0, count[0]) and the
// medium rows (more than VCSR_MED_MIN blocks, at most long_blocks) in
// list_med[0, count[2]); the order within a list does not matter, each row is
// summed by one walker.  A workgroup of 1024 rows ranks its listed rows by wave
// ballots and a scan of the 16 wave counts, and takes its slots with one atomic
// per list (an atomic per row, or per wave, serialises on the counter: 1.5 ms and
// 117 us for the 1M R-MAT's 10^5 medium rows).
constexpr int VCSR_FIND_ROWS = 1024;  // rows per workgroup of k_vcsr_find_long

__global__ __launch_bounds__(VCSR_FIND_ROWS) void k_vcsr_find_long(int M, const int* __restrict__ rowptr,
                                                                   int long_blocks, int* __restrict__ list,
                                                                   int* __restrict__ list_med,
                                                                   int* __restrict__ count) {
    constexpr int NW = VCSR_FIND_ROWS / 64;
    __shared__ int s_off[2][NW];  // per wave: its rows' first slot within the workgroup
    __shared__ int s_base[2];     // the workgroup's first slot in each list
    const int r = blockIdx.x * VCSR_FIND_ROWS + threadIdx.x;
    const int n = r < M ? rowptr[r + 1] - rowptr[r] : 0;
    const bool is_long = r < M && n > 4 * long_blocks;
    const bool is_med = r < M && !is_long && n > 4 * VCSR_MED_MIN;
    const uint64_t bl = __ballot(is_long), bm = __ballot(is_med);
    const int lane = (int)__lane_id(), wid = (int)(threadIdx.x >> 6);
    if (lane == 0) {
        s_off[0][wid] = __popcll(bl);
        s_off[1][wid] = __popcll(bm);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int t = (int)threadIdx.x;
        int acc = 0;
        for (int w = 0; w < NW; ++w) {
            const int c = s_off[t][w];
            s_off[t][w] = acc;
            acc += c;
        }
        s_base[t] = acc > 0 ? atomicAdd(&count[t == 0 ? 0 : 2], acc) : 0;
    }
    __syncthreads();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (is_long) list[s_base[0] + s_off[0][wid] + __popcll(bl & below)] = r;
    if (is_med) list_med[s_base[1] + s_off[1][wid] + __popcll(bm & below)] = r;
}

// count[0] = listed long rows, count[1] = the long rows' work counter, count[2] =
// listed medium rows
__global__ void k_zero_count(int* count) {
    count[0] = 0;
    count[1] = 0;
    count[2] = 0;
}

// Medium rows, one wave each: wave gw of nmw takes list_med[gw], list_med[gw + nmw],
// ... (a fixed assignment, no shared counter).  A step covers 16 blocks, lane
// 4q + l holding position l of block q; each lane keeps its next VCSR_MED_AHEAD
// steps' value and x loads in flight (the column index one ring further ahead).
// The lanes form the rounded products (the unfused reading) or keep value and x
// (fused), and lane l's partial sum s_l takes block 0..15's position-l term by a
// shuffle from lane 4q + l, in block order: the reference's sequence of
// operations per s_l, as in the other walkers.
template <bool FUSED>
__device__ void vcsr_medium_rows(int gw, int nmw, const int* __restrict__ list_med, const int* __restrict__ count,
                                 const int* __restrict__ rowptr, const int* __restrict__ colind,
                                 const float* __restrict__ val, const float* __restrict__ x, float* __restrict__ y) {
    constexpr int D = VCSR_MED_AHEAD;
    const int lane = threadIdx.x & 63;
    const int q = lane >> 2, l = lane & 3;
    const int nmed = __builtin_amdgcn_readfirstlane(count[2]);
    for (int i = gw; i < nmed; i += nmw) {
        const int r = __builtin_amdgcn_readfirstlane(list_med[i]);
        const int beg = __builtin_amdgcn_readfirstlane(rowptr[r]);
        const int nb = (__builtin_amdgcn_readfirstlane(rowptr[r + 1]) - beg) >> 2;
        const int nsteps = (nb + 15) >> 4;
        auto col_of = [&](int k) {
            const int b = k * 16 + q;
            return b < nb ? colind[beg + 4 * b] : 0;
        };
        auto fetch = [&](int k, int c, float& v, float& xv) {
            const int b = k * 16 + q;
            v = 0.0f;
            xv = 0.0f;
            if (b < nb) {
                v = val[beg + 4 * b + l];
                xv = x[c + l];
            }
        };
        int rc[D];
        float rv[D], rx[D];
#pragma unroll
        for (int u = 0; u < D; ++u) rc[u] = col_of(u);
#pragma unroll
        for (int u = 0; u < D; ++u) {
            fetch(u, rc[u], rv[u], rx[u]);
            rc[u] = col_of(u + D);
        }
        float s = 0.0f;
        for (int k0 = 0; k0 < nsteps; k0 += D) {
#pragma unroll
            for (int u = 0; u < D; ++u) {
                const int k = k0 + u;
                if (k >= nsteps) break;  // uniform: nsteps is the wave's
                const int cnt = min(16, nb - k * 16);
                const float pv = rv[u], px = rx[u];
                float p = 0.0f;
                if constexpr (!FUSED) p = unfused_product(pv, px);  // vmlaq_f32 rounds the product on its own
                fetch(k + D, rc[u], rv[u], rx[u]);  // in flight for D steps
                rc[u] = col_of(k + 2 * D);
#pragma unroll
                for (int b = 0; b < 16; ++b) {
                    if (b < cnt) {  // uniform
                        if constexpr (FUSED) {
                            const float a = __shfl(pv, 4 * b + l);
                            const float c = __shfl(px, 4 * b + l);
                            s = fmaf(a, c, s);
                        } else {
                            s = __fadd_rn(s, __shfl(p, 4 * b + l));
                        }
                    }
                }
            }
        }
        // every lane group now holds s0..s3 of the row
        const float t = __fadd_rn(s, __shfl_xor(s, 2, 4));   // lanes 0/2: s0+s2, lanes 1/3: s1+s3
        const float tot = __fadd_rn(t, __shfl_xor(t, 1, 4));  // (s0+s2)+(s1+s3)
        if (lane == 0) y[r] = __fadd_rn(y[r], tot);
    }
}

template <bool FUSED>
__device__ void vcsr_long_rows(const int* __restrict__ list, int* __restrict__ count,
                               const int* __restrict__ rowptr, const int* __restrict__ colind,
                               const float* __restrict__ val, const float* __restrict__ x, float* __restrict__ y,
                               float (*s_a)[4][VCSR_PAD], float (*s_b)[4][VCSR_PAD], int* s_row) {
    constexpr int D = VCSR_AHEAD;
    const int t = threadIdx.x;
    const int nlong = count[0];
    while (true) {
        if (t == 0) *s_row = atomicAdd(&count[1], 1);
        __syncthreads();
        const int i = *s_row;
        __syncthreads();  // every lane has read the row before lane 0 takes the next
        if (i >= nlong) break;  // the same for the whole workgroup once the list is drained
        const int r = list[i];
        const int beg = rowptr[r];
        const int nb = (rowptr[r + 1] - beg) >> 2;
        const int nbat = (nb + VCSR_BATCH - 1) / VCSR_BATCH;
        // lane t owns block k·256 + t of batch k
        auto col_of = [&](int k) {
            const int b = k * VCSR_BATCH + t;
            return b < nb ? colind[beg + 4 * b] : -1;
        };
        float4 rv[D], rx[D];
        int rc[D];
        auto fetch = [&](int k, int c, float4& v, float4& xv) {
            const int b = k * VCSR_BATCH + t;
            v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            xv = v;
            if (b < nb) {
                v = *reinterpret_cast<const float4*>(val + beg + 4 * b);
                xv = *reinterpret_cast<const float4*>(x + c);
            }
        };
#pragma unroll
        for (int u = 0; u < D; ++u) rc[u] = col_of(u);
#pragma unroll
        for (int u = 0; u < D; ++u) {
            fetch(u, rc[u], rv[u], rx[u]);
            rc[u] = col_of(u + D);
        }
        float s = 0.0f;
        for (int k0 = 0; k0 < nbat; k0 += D) {
#pragma unroll
            for (int u = 0; u < D; ++u) {
                const int k = k0 + u;
                if (k >= nbat) break;  // uniform: nbat is the workgroup's
                const int buf = k & 1;
                if constexpr (FUSED) {
                    s_a[buf][0][t] = rv[u].x;
                    s_a[buf][1][t] = rv[u].y;
                    s_a[buf][2][t] = rv[u].z;
                    s_a[buf][3][t] = rv[u].w;
                    s_b[buf][0][t] = rx[u].x;
                    s_b[buf][1][t] = rx[u].y;
                    s_b[buf][2][t] = rx[u].z;
                    s_b[buf][3][t] = rx[u].w;
                } else {  // vmlaq_f32 rounds the product on its own
                    s_a[buf][0][t] = unfused_product(rv[u].x, rx[u].x);
                    s_a[buf][1][t] = unfused_product(rv[u].y, rx[u].y);
                    s_a[buf][2][t] = unfused_product(rv[u].z, rx[u].z);
                    s_a[buf][3][t] = unfused_product(rv[u].w, rx[u].w);
                }
                fetch(k + D, rc[u], rv[u], rx[u]);  // in flight for D batches
                rc[u] = col_of(k + 2 * D);
                // batch k staged; wave 0 has also finished adding batch k-1, whose
                // buffer batch k+1 is about to overwrite
                __syncthreads();
                if (t < 4) {
                    const int cnt = min(VCSR_BATCH, nb - k * VCSR_BATCH);
                    const float* pa = &s_a[buf][t][0];
                    const float* pb = &s_b[buf][FUSED ? t : 0][0];
                    int b = 0;
                    for (; b + 4 <= cnt; b += 4) {
                        const float4 a4 = *reinterpret_cast<const float4*>(pa + b);
                        if constexpr (FUSED) {
                            const float4 x4 = *reinterpret_cast<const float4*>(pb + b);
                            s = fmaf(a4.x, x4.x, s);
                            s = fmaf(a4.y, x4.y, s);
                            s = fmaf(a4.z, x4.z, s);
                            s = fmaf(a4.w, x4.w, s);
                        } else {
                            s = __fadd_rn(s, a4.x);
                            s = __fadd_rn(s, a4.y);
                            s = __fadd_rn(s, a4.z);
                            s = __fadd_rn(s, a4.w);
                        }
                    }
                    for (; b < cnt; ++b) {
                        if constexpr (FUSED)
                            s = fmaf(pa[b], pb[b], s);
                        else
                            s = __fadd_rn(s, pa[b]);
                    }
                }
            }
        }
        if (t < 64) {  // wave 0: lanes 0-3 hold s0..s3
            const float tt = __fadd_rn(s, __shfl_xor(s, 2, 4));  // lanes 0/2: s0+s2, lanes 1/3: s1+s3
            const float tot = __fadd_rn(tt, __shfl_xor(tt, 1, 4));  // (s0+s2)+(s1+s3)
            if (t == 0) y[r] = __fadd_rn(y[r], tot);
        }
        __syncthreads();  // the LDS buffers are free for the next row
    }
}

// One launch: workgroups [0, n_long_wg) walk the listed long rows, the next
// n_med_wg the listed medium rows (a wave each), the rest the short rows (64 per
// workgroup).
template <bool FUSED>
__global__ __launch_bounds__(256) void k_spmv_vcsr_split(int M, const int* __restrict__ rowptr,
                                                         const int* __restrict__ colind,
                                                         const float* __restrict__ val, const float* __restrict__ x,
                                                         float* __restrict__ y, const int* __restrict__ list,
                                                         const int* __restrict__ list_med, int* __restrict__ count,
                                                         int n_long_wg, int n_med_wg) {
    __shared__ __align__(16) float s_a[2][4][VCSR_PAD];
    __shared__ __align__(16) float s_b[FUSED ? 2 : 1][4][FUSED ? VCSR_PAD : 4];
    __shared__ int s_row;
    const int b = (int)blockIdx.x;
    if (b < n_long_wg) {
        vcsr_long_rows<FUSED>(list, count, rowptr, colind, val, x, y, s_a,
                              reinterpret_cast<float (*)[4][VCSR_PAD]>(&s_b[0][0][0]), &s_row);
        return;
    }
    if (b < n_long_wg + n_med_wg) {
        const int gw = (b - n_long_wg) * 4 + (int)(threadIdx.x >> 6);
        vcsr_medium_rows<FUSED>(gw, n_med_wg * 4, list_med, count, rowptr, colind, val, x, y);
        return;
    }
    // the short walker takes what neither list holds: rows of at most VCSR_MED_MIN
    // blocks (the long-row limit is never below it; a constant here lets the
    // compiler drop the walker's batch loop for longer rows: 66 VGPRs, not 111)
    vcsr_short_rows((unsigned)(b - n_long_wg - n_med_wg), M, rowptr, colind, val, x, y, FUSED ? 1 : 0, VCSR_MED_MIN);
}

template <int IPT>
int launch_merge(int M, int nnz, const int* rp, const int* ci, const float* va, const float* x, float* y, int flags,
                 hipStream_t s) {
    constexpr int TILE = 64 * IPT;
    const int64_t total = (int64_t)M + nnz;
    const int64_t nwaves = (total + TILE - 1) / TILE;
    if (nwaves > INT_MAX / 4) return ARMSP_ERROR_NOT_SUPPORTED;
    void* ws = nullptr;
    const size_t rb = ((size_t)nwaves * 4 + 255) & ~(size_t)255;
    int st = ws_alloc(&ws, 2 * rb + (size_t)(nwaves + 1) * 4, s);
    if (st) return st;
    int* crow = static_cast<int*>(ws);
    float* cval = reinterpret_cast<float*>(static_cast<char*>(ws) + rb);
    int* part = reinterpret_cast<int*>(static_cast<char*>(ws) + 2 * rb);
    {
        TimedLaunch t("merge_partition", s);
        st = merge_partition(M, nnz, TILE, nwaves, rp, part, s);
        if (st) {
            ws_free(ws, s);
            return st;
        }
        t.stop();
    }
    {
        TimedLaunch t("spmv", s);
        hipLaunchKernelGGL((k_spmv_merge<IPT>), dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, s, M, nnz, rp, ci,
                           va, x, y, flags, part, crow, cval);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            ws_free(ws, s);
            return hip_fail(e);
        }
        t.stop();
    }
    {
        TimedLaunch t("spmv_fixup", s);
        hipLaunchKernelGGL(k_spmv_fixup, dim3((unsigned)((nwaves + 255) / 256)), dim3(256), 0, s, (int)nwaves, crow,
                           cval, y);
        hipError_t e = hipGetLastError();
        ws_free(ws, s);
        if (e != hipSuccess) return hip_fail(e);
        t.stop();
    }
    return ARMSP_SUCCESS;
}

}  // namespace
}  // namespace armsp

extern "C" {

int armsp_spmv_f32(int M, int K, int nnz, const int* rowptr, const int* colind, const float* val, const float* x,
                   float* y, int flags, const armsp_config* cfg, void* stream) {
    using namespace armsp;
    if (M < 0 || K < 0 || nnz < 0) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0) return ARMSP_SUCCESS;
    if (!rowptr || !y || (nnz > 0 && (!colind || !x))) return ARMSP_ERROR_INVALID_VALUE;
    if (flags & ~ARMSP_ACCUMULATE) return ARMSP_ERROR_INVALID_VALUE;
    if ((int64_t)M + nnz > INT_MAX) return ARMSP_ERROR_NOT_SUPPORTED;
    hipStream_t s = as_stream(stream);
    armsp_config c = cfg ? *cfg : armsp_config{0, 0, 0, 0, 0};
    int sched = c.schedule == ARMSP_SCHED_AUTO ? ARMSP_SCHED_MERGE : c.schedule;
    if (sched == ARMSP_SCHED_NNZ) sched = ARMSP_SCHED_MERGE;
    if (sched == ARMSP_SCHED_ROW) {
        int G = c.group;
        if (G <= 0) {
            const int avg = nnz / M;
            G = 1;
            while (G < 64 && G < avg) G <<= 1;
        }
        TimedLaunch t("spmv", s);
        const dim3 grid((unsigned)(((int64_t)M * G + 255) / 256));
        switch (G) {
#define ARMSP_SPMV_ROW(GG)                                                                                          \
    case GG:                                                                                                        \
        hipLaunchKernelGGL((k_spmv_row<GG>), grid, dim3(256), 0, s, M, rowptr, colind, val, x, y, flags); \
        break;
            ARMSP_SPMV_ROW(1)
            ARMSP_SPMV_ROW(2)
            ARMSP_SPMV_ROW(4)
            ARMSP_SPMV_ROW(8)
            ARMSP_SPMV_ROW(16)
            ARMSP_SPMV_ROW(32)
            ARMSP_SPMV_ROW(64)
#undef ARMSP_SPMV_ROW
            default: return ARMSP_ERROR_INVALID_VALUE;
        }
        ARMSP_LAUNCHED();
        t.stop();
        return ARMSP_SUCCESS;
    }
    if (sched != ARMSP_SCHED_MERGE) return ARMSP_ERROR_INVALID_VALUE;
    // AUTO: 16 items per lane when that still leaves 16 waves per CU (more gathers in
    // flight per lane: 1M R-MAT, 2^24 nnz, 91.6 -> 87.5 us), else 8
    const int64_t items = (int64_t)M + nnz;
    const int ipt = c.chunk > 0 ? c.chunk / 64 : (items / (64 * 16) >= 16 * (int64_t)device_cus() ? 16 : 8);
    switch (ipt) {
        case 4: return launch_merge<4>(M, nnz, rowptr, colind, val, x, y, flags, s);
        case 8: return launch_merge<8>(M, nnz, rowptr, colind, val, x, y, flags, s);
        case 16: return launch_merge<16>(M, nnz, rowptr, colind, val, x, y, flags, s);
        default: return ARMSP_ERROR_INVALID_VALUE;
    }
}

int armsp_spmv_vcsr_f32(int M, int nnz_padded, const int* rowptr, const int* colind, const float* val, const float* x,
                        float* y, int flags, void* stream) {
    using namespace armsp;
    if (M < 0 || nnz_padded < 0 || (nnz_padded & 3)) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0) return ARMSP_SUCCESS;
    if (!rowptr || !y || (nnz_padded > 0 && (!colind || !val || !x))) return ARMSP_ERROR_INVALID_VALUE;
    if (flags & ~(ARMSP_ACCUMULATE | ARMSP_FUSED_MLA | ARMSP_VCSR_BALANCED)) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    if (flags & ARMSP_VCSR_BALANCED) {
        // the padded blocks are a CSR matrix of their own: merge-path SpMV, y += A·x
        if (flags & ARMSP_FUSED_MLA) return ARMSP_ERROR_INVALID_VALUE;
        if (nnz_padded == 0) return ARMSP_SUCCESS;
        return launch_merge<8>(M, nnz_padded, rowptr, colind, val, x, y, ARMSP_ACCUMULATE, s);
    }
    const int fused = (flags & ARMSP_FUSED_MLA) ? 1 : 0;
    // Rows of more than VCSR_MED_MIN blocks take a wave each, rows of more than
    // long_blocks a workgroup each; the workgroup walker's 16-byte loads need
    // 16-byte-aligned values and x (VCSR rows hold whole 4-blocks and block columns
    // are multiples of 4, so rowptr entries and the gathered x offsets are multiples
    // of 4 as well), so with unaligned operands every long row goes to the waves.
    const bool aligned = ((uintptr_t)val % 16) == 0 && ((uintptr_t)x % 16) == 0;
    // (a long-row limit below VCSR_MED_MIN acts as VCSR_MED_MIN: rows up to it are the short walker's)
    const int lb = vcsr_long_blocks();
    const int long_blocks = aligned ? (lb < VCSR_MED_MIN ? VCSR_MED_MIN : lb) : INT_MAX / 8;
    const bool split_long = nnz_padded / 4 > VCSR_MED_MIN;
    void* ws = nullptr;
    int* list = nullptr;
    int* list_med = nullptr;
    int* count = nullptr;
    if (split_long) {
        const int st = ws_alloc(&ws, (2 * (size_t)M + 64) * sizeof(int), s);
        if (st) return st;
        count = static_cast<int*>(ws);
        list = count + 64;
        list_med = list + M;
        hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, s, count);
        hipLaunchKernelGGL(k_vcsr_find_long, dim3((unsigned)((M + VCSR_FIND_ROWS - 1) / VCSR_FIND_ROWS)),
                           dim3(VCSR_FIND_ROWS), 0, s, M, rowptr,
                           long_blocks, list, list_med, count);
    }
    TimedLaunch t("spmv_vcsr", s);
    const unsigned short_wg = (unsigned)(((int64_t)M * 4 + 255) / 256);
    if (split_long) {  // long-row workgroups first (the hub rows are the critical path), then medium-row waves
        const int n_long_wg = 2 * device_cus();
        const int n_med_wg = VCSR_MED_WG_PER_CU * device_cus();
        const dim3 grid((unsigned)(n_long_wg + n_med_wg) + short_wg);
        if (fused)
            hipLaunchKernelGGL(k_spmv_vcsr_split<true>, grid, dim3(256), 0, s, M, rowptr, colind, val, x, y, list,
                               list_med, count, n_long_wg, n_med_wg);
        else
            hipLaunchKernelGGL(k_spmv_vcsr_split<false>, grid, dim3(256), 0, s, M, rowptr, colind, val, x, y, list,
                               list_med, count, n_long_wg, n_med_wg);
    } else {
        hipLaunchKernelGGL(k_spmv_vcsr, dim3(short_wg), dim3(256), 0, s, M, rowptr, colind, val, x, y, fused);
    }
    const hipError_t e = hipGetLastError();
    if (ws) ws_free(ws, s);
    if (e != hipSuccess) return hip_fail(e);
    t.stop();
    return ARMSP_SUCCESS;
}

}  // extern "C"
