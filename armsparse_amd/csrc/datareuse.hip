// SpMM on the 2-row DataReuse format (armsp_csr_to_datareuse).
//
// Reference: DataReuseSpmm (suitesparse_sdmm.cpp:25-43, datasets/suitesparse.cpp
// :23-50) saxpys the B row of every entry of row m into C row m + flag, rows
// taken in pairs by one thread (schedule(static, 2)); toDataReuseFormat puts
// an odd row's entry right behind the even row's entry of the same column, so
// the pair shares that B row while it is in cache.
//
// Here a group of G lanes owns a row pair (2p, 2p+1) and V columns: both C
// rows live in registers, the even row's entries go to the row their flag
// names — a flagged entry reuses the B row its partner just loaded (no second
// gather) — then the odd row's own entries.  That is the reference's order of
// additions into each C row, with the same fused multiply-adds, so the result
// is bit-identical to the sequential restatement.  Flags in odd rows, and in
// the last row of an odd M, are ignored (the format never sets them).
#include "common.hpp"

namespace armsp {
namespace {

template <int V>
__device__ __forceinline__ void ld(const float* __restrict__ p, float (&o)[V]) {
    if constexpr (V == 4) {
        const float4 t = *reinterpret_cast<const float4*>(p);
        o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = p[v];
    }
}

template <int V>
__device__ __forceinline__ void st(float* __restrict__ p, const float (&o)[V]) {
    if constexpr (V == 4) {
        *reinterpret_cast<float4*>(p) = make_float4(o[0], o[1], o[2], o[3]);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) p[v] = o[v];
    }
}

template <int V, int G>
__global__ __launch_bounds__(256) void k_spmm_datareuse(int M, int N, const int* __restrict__ rowptr,
                                                        const int* __restrict__ colind,
                                                        const float* __restrict__ val,
                                                        const uint8_t* __restrict__ flag, const float* __restrict__ B,
                                                        int ldb, float* __restrict__ C, int ldc, int accumulate) {
    const int lane = threadIdx.x & (G - 1);
    const int pair = (int)((blockIdx.x * 256u + threadIdx.x) / G);
    const int m = 2 * pair;
    if (m >= M) return;
    const int col = ((int)blockIdx.y * G + lane) * V;
    if (col >= N) return;
    const bool has_odd = m + 1 < M;
    float a0[V], a1[V];
    if (accumulate) {
        ld<V>(C + (size_t)m * ldc + col, a0);
        if (has_odd) ld<V>(C + (size_t)(m + 1) * ldc + col, a1);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) a0[v] = 0.0f;
    }
    if (!accumulate || !has_odd) {
#pragma unroll
        for (int v = 0; v < V; ++v) a1[v] = 0.0f;
    }
    float b[V];
#pragma unroll
    for (int v = 0; v < V; ++v) b[v] = 0.0f;
    int prev_c = -1;
    for (int p = rowptr[m]; p < rowptr[m + 1]; ++p) {
        const int c = colind[p];
        const float a = val ? val[p] : 1.0f;
        const bool to_odd = flag[p] != 0 && has_odd;
        if (!(to_odd && c == prev_c)) ld<V>(B + (size_t)c * ldb + col, b);  // a flagged partner reuses b
        prev_c = c;
        if (to_odd) {
#pragma unroll
            for (int v = 0; v < V; ++v) a1[v] = fmaf(a, b[v], a1[v]);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) a0[v] = fmaf(a, b[v], a0[v]);
        }
    }
    st<V>(C + (size_t)m * ldc + col, a0);
    if (!has_odd) return;
    for (int p = rowptr[m + 1]; p < rowptr[m + 2]; ++p) {
        const float a = val ? val[p] : 1.0f;
        ld<V>(B + (size_t)colind[p] * ldb + col, b);
#pragma unroll
        for (int v = 0; v < V; ++v) a1[v] = fmaf(a, b[v], a1[v]);
    }
    st<V>(C + (size_t)(m + 1) * ldc + col, a1);
}

template <int V>
int launch_dr(int G, dim3 grid, hipStream_t s, int M, int N, const int* rp, const int* ci, const float* va,
              const uint8_t* fl, const float* B, int ldb, float* C, int ldc, int acc) {
    switch (G) {
#define ARMSP_DR_CASE(GG)                                                                                          \
    case GG:                                                                                                       \
        hipLaunchKernelGGL((k_spmm_datareuse<V, GG>), grid, dim3(256), 0, s, M, N, rp, ci, va, fl, B, ldb, C, ldc, \
                           acc);                                                                                   \
        break;
        ARMSP_DR_CASE(4)
        ARMSP_DR_CASE(8)
        ARMSP_DR_CASE(16)
        ARMSP_DR_CASE(32)
        ARMSP_DR_CASE(64)
#undef ARMSP_DR_CASE
        default: return ARMSP_ERROR_INVALID_VALUE;
    }
    return ARMSP_SUCCESS;
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace
}  // namespace armsp

extern "C" int armsp_spmm_datareuse_f32(int M, int K, int N, int nnz, const int* rowptr, const int* colind,
                                        const float* val, const uint8_t* flag, const float* B, int ldb, float* C,
                                        int ldc, int flags, void* stream) {
    using namespace armsp;
    if (M < 0 || K < 0 || N < 0 || nnz < 0) return ARMSP_ERROR_INVALID_VALUE;
    if (flags & ~ARMSP_ACCUMULATE) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0) return ARMSP_SUCCESS;
    if (!rowptr || !C || ldc < N) return ARMSP_ERROR_INVALID_VALUE;
    if (nnz > 0 && (!colind || !flag || !B || ldb < N || K == 0)) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    const int V = (N % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 && aligned16(B) && aligned16(C)) ? 4 : 1;
    int G = 4;
    while (G < 64 && G * V < N) G <<= 1;
    const int64_t pairs = ((int64_t)M + 1) / 2;
    const dim3 grid((unsigned)((pairs * G + 255) / 256), (unsigned)((N + G * V - 1) / (G * V)));
    if (grid.y > 65535) return ARMSP_ERROR_NOT_SUPPORTED;
    const int acc = (flags & ARMSP_ACCUMULATE) ? 1 : 0;
    TimedLaunch t("spmm_datareuse", s);
    const int st = V == 4 ? launch_dr<4>(G, grid, s, M, N, rowptr, colind, val, flag, B, ldb, C, ldc, acc)
                          : launch_dr<1>(G, grid, s, M, N, rowptr, colind, val, flag, B, ldb, C, ldc, acc);
    if (st) return st;
    ARMSP_LAUNCHED();
    t.stop();
    return ARMSP_SUCCESS;
}
