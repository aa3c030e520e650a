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
//          four lanes per row keep the four partial sums s0..s3 and the row
//          total is (s0+s2)+(s1+s3), as vadd/vpadd compute it.
#include <climits>

#include "common.hpp"

namespace armsp {
namespace {

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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

__global__ __launch_bounds__(256) void k_spmv_vcsr(int M, const int* __restrict__ rowptr,
                                                   const int* __restrict__ colind, const float* __restrict__ val,
                                                   const float* __restrict__ x, float* __restrict__ y, int fused,
                                                   int long_blocks) {
    const int l = threadIdx.x & 3;
    const int row = (int)((blockIdx.x * 256u + threadIdx.x) >> 2);
    if (row >= M) return;
    const int beg = rowptr[row], end = rowptr[row + 1];
    if (long_blocks > 0 && end - beg > 4 * long_blocks) return;  // k_spmv_vcsr_long's row
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
    if (fused) {
        for (; idx < end; idx += 4) s = fmaf(val[idx + l], x[colind[idx] + l], s);
    } else {
        for (; idx < end; idx += 4) s += unfused_product(val[idx + l], x[colind[idx] + l]);
    }
    const float t = __fadd_rn(s, __shfl_xor(s, 2, 4));   // lanes 0/2: s0+s2, lanes 1/3: s1+s3
    const float tot = __fadd_rn(t, __shfl_xor(t, 1, 4));  // (s0+s2)+(s1+s3)
    if (l == 0) y[row] = __fadd_rn(y[row], tot);
}

// Rows of more than VCSR_LONG blocks: k_vcsr_find_long lists them, and a wave
// per listed row walks it in batches of 256 blocks.  All 64 lanes load a batch
// (16-byte loads of values and x) into LDS while lanes 0-3 add the previous
// batch in block order — each lane l keeps the reference's partial sum s_l, so
// the result stays bit-exact — and the global loads of the next batch are in
// flight during those adds.
constexpr int VCSR_LONG = 512;
constexpr int VCSR_BATCH = 256;

__global__ void k_vcsr_find_long(int M, const int* __restrict__ rowptr, int* __restrict__ list,
                                 int* __restrict__ count) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    if (rowptr[r + 1] - rowptr[r] > 4 * VCSR_LONG) list[atomicAdd(count, 1)] = r;
}

__global__ void k_zero_count(int* count) { *count = 0; }

__global__ __launch_bounds__(256) void k_spmv_vcsr_long(const int* __restrict__ list, const int* __restrict__ count,
                                                        const int* __restrict__ rowptr,
                                                        const int* __restrict__ colind,
                                                        const float* __restrict__ val, const float* __restrict__ x,
                                                        float* __restrict__ y, int fused) {
    __shared__ float4 s_v[4][2][VCSR_BATCH];
    __shared__ float4 s_x[4][2][VCSR_BATCH];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nlong = *count;
    const int nwaves = (int)gridDim.x * 4;
    for (int i = uniform((int)(blockIdx.x * 4u + wid)); i < nlong; i += nwaves) {
        const int r = uniform(list[i]);
        const int beg = uniform(rowptr[r]);
        const int nb = uniform((rowptr[r + 1] - beg) >> 2);
        const int nbat = (nb + VCSR_BATCH - 1) / VCSR_BATCH;
        float4 rv[VCSR_BATCH / 64], rx[VCSR_BATCH / 64];
        auto fetch = [&](int k) {
#pragma unroll
            for (int q = 0; q < VCSR_BATCH / 64; ++q) {
                const int b = k * VCSR_BATCH + lane + 64 * q;
                rv[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                rx[q] = rv[q];
                if (b < nb) {
                    rv[q] = *reinterpret_cast<const float4*>(val + beg + 4 * b);
                    rx[q] = *reinterpret_cast<const float4*>(x + colind[beg + 4 * b]);
                }
            }
        };
        float s = 0.0f;
        fetch(0);
        for (int k = 0; k < nbat; ++k) {
            const int buf = k & 1;
#pragma unroll
            for (int q = 0; q < VCSR_BATCH / 64; ++q) {
                s_v[wid][buf][lane + 64 * q] = rv[q];
                s_x[wid][buf][lane + 64 * q] = rx[q];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (k + 1 < nbat) fetch(k + 1);  // in flight during the adds below
            const int cnt = min(VCSR_BATCH, nb - k * VCSR_BATCH);
            if (lane < 4) {
                const float* pv = reinterpret_cast<const float*>(&s_v[wid][buf][0]) + lane;
                const float* px = reinterpret_cast<const float*>(&s_x[wid][buf][0]) + lane;
                // 16 blocks' operands read from LDS ahead of their in-order adds
                int b = 0;
                for (; b + 16 <= cnt; b += 16) {
                    float a[16], c[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u) {
                        a[u] = pv[4 * (b + u)];
                        c[u] = px[4 * (b + u)];
                    }
                    if (fused) {
#pragma unroll
                        for (int u = 0; u < 16; ++u) s = fmaf(a[u], c[u], s);
                    } else {
#pragma unroll
                        for (int u = 0; u < 16; ++u) s += unfused_product(a[u], c[u]);
                    }
                }
                if (fused) {
                    for (; b < cnt; ++b) s = fmaf(pv[4 * b], px[4 * b], s);
                } else {
                    for (; b < cnt; ++b) s += unfused_product(pv[4 * b], px[4 * b]);
                }
            }
        }
        const float t = __fadd_rn(s, __shfl_xor(s, 2, 4));   // lanes 0/2: s0+s2, lanes 1/3: s1+s3
        const float tot = __fadd_rn(t, __shfl_xor(t, 1, 4));  // (s0+s2)+(s1+s3)
        if (lane == 0) y[r] = __fadd_rn(y[r], tot);
        // the next row's first batch reuses this row's LDS: every lane is past the adds
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
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
    const int ipt = c.chunk > 0 ? c.chunk / 64 : 8;
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
    // rows of more than VCSR_LONG blocks (possible only above that many blocks in
    // all) go to the wave-per-row kernel
    const bool split_long = nnz_padded / 4 > VCSR_LONG;
    void* ws = nullptr;
    int* list = nullptr;
    int* count = nullptr;
    if (split_long) {
        const int st = ws_alloc(&ws, ((size_t)M + 64) * sizeof(int), s);
        if (st) return st;
        count = static_cast<int*>(ws);
        list = count + 64;
        hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, s, count);
        hipLaunchKernelGGL(k_vcsr_find_long, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, M, rowptr, list,
                           count);
    }
    TimedLaunch t("spmv_vcsr", s);
    hipLaunchKernelGGL(k_spmv_vcsr, dim3((unsigned)(((int64_t)M * 4 + 255) / 256)), dim3(256), 0, s, M, rowptr,
                       colind, val, x, y, fused, split_long ? VCSR_LONG : 0);
    if (split_long)
        hipLaunchKernelGGL(k_spmv_vcsr_long, dim3((unsigned)(2 * device_cus())), dim3(256), 0, s, list, count, rowptr,
                           colind, val, x, y, fused);
    const hipError_t e = hipGetLastError();
    if (ws) ws_free(ws, s);
    if (e != hipSuccess) return hip_fail(e);
    t.stop();
    return ARMSP_SUCCESS;
}

}  // extern "C"
