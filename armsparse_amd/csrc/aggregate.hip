// Max aggregation over CSR neighbourhoods (the GraphSAGE / DGL max reducer).
//
// Reference: gcn/dgl-custom/binary_reduce_max.cu — topoSimpleSPMMMaxKernel
// (:151-166) and the tiled variants (:26-149) compute, for every row r and
// column j, C[r][j] = max over the row's nonzeros of B[colind][j] (copy of the
// source features, no edge values), starting from max_init() = -10000 (:22-24)
// and folding with `acc > x ? acc : x` (:18-20) in CSR order.  The same fold
// here, with an optional edge value multiplied in (x = val·B[c][j]), the
// initial value a parameter (the reference's -10000 by default in the Python
// layer), and the winning nonzero position recorded per output for the
// backward pass: dB[colind[p]][j] += val[p]·dC[r][j] for p = argmax[r][j].
//
// Layout: a group of G lanes per row, V columns per lane; each lane issues U
// gathers of its row together and folds them in CSR order, so the result and
// the argmax (a later equal value wins, as `acc > x ? acc : x` does) are the
// sequential fold's.
#include "common.hpp"

namespace armsp {
namespace {

template <int V>
__device__ __forceinline__ void load_v(const float* __restrict__ p, float (&o)[V]) {
    if constexpr (V == 4) {
        const float4 t = *reinterpret_cast<const float4*>(p);
        o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = p[v];
    }
}

template <int V>
__device__ __forceinline__ void store_v(float* __restrict__ p, const float (&o)[V]) {
    if constexpr (V == 4) {
        *reinterpret_cast<float4*>(p) = make_float4(o[0], o[1], o[2], o[3]);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) p[v] = o[v];
    }
}

template <int V, int G>
__global__ __launch_bounds__(256) void k_spmm_max(int M, int N, const int* __restrict__ rowptr,
                                                  const int* __restrict__ colind, const float* __restrict__ val,
                                                  const float* __restrict__ B, int ldb, float init,
                                                  float* __restrict__ C, int ldc, int* __restrict__ argmax) {
    constexpr int U = 4;
    const int lane = threadIdx.x & (G - 1);
    const int row = (int)((blockIdx.x * 256u + threadIdx.x) / G);
    if (row >= M) return;
    const int col = ((int)blockIdx.y * G + lane) * V;
    if (col >= N) return;
    const int beg = rowptr[row], end = rowptr[row + 1];
    float acc[V];
    int arg[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        acc[v] = init;
        arg[v] = -1;
    }
    auto fold = [&](int p, float a, const float (&b)[V]) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const float x = a * b[v];
            const bool keep = acc[v] > x;  // the reference's `acc > x ? acc : x`
            acc[v] = keep ? acc[v] : x;
            arg[v] = keep ? arg[v] : p;
        }
    };
    int p = beg;
    for (; p + U <= end; p += U) {
        float b[U][V], a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = colind[p + u];
            a[u] = val ? val[p + u] : 1.0f;
            load_v<V>(B + (size_t)c * ldb + col, b[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) fold(p + u, a[u], b[u]);
    }
    for (; p < end; ++p) {
        float b[V];
        const float a = val ? val[p] : 1.0f;
        load_v<V>(B + (size_t)colind[p] * ldb + col, b);
        fold(p, a, b);
    }
    store_v<V>(C + (size_t)row * ldc + col, acc);
    if (argmax) {
#pragma unroll
        for (int v = 0; v < V; ++v) argmax[(size_t)row * N + col + v] = arg[v];
    }
}

// dB[colind[p]][j] += val[p]·dC[r][j] where p = argmax[r][j] (p = -1: the
// initial value won, no gradient; positions and columns out of range are
// skipped).  Many outputs can share a source row, so the adds are atomic
// (their order, and so the last bits, vary run to run).
__global__ __launch_bounds__(256) void k_spmm_max_backward(int M, int K, int N, int nnz,
                                                           const int* __restrict__ colind,
                                                           const float* __restrict__ val,
                                                           const float* __restrict__ dC, int ldc,
                                                           const int* __restrict__ argmax, float* __restrict__ dB,
                                                           int ldb) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)M * N) return;
    const int r = (int)(i / N), j = (int)(i % N);
    const int p = argmax[i];
    if (p < 0 || p >= nnz) return;
    const int c = colind[p];
    if (c < 0 || c >= K) return;
    const float g = dC[(size_t)r * ldc + j] * (val ? val[p] : 1.0f);
    atomicAdd(dB + (size_t)c * ldb + j, g);
}

template <int V>
int launch_max(int G, dim3 grid, hipStream_t s, int M, int N, const int* rp, const int* ci, const float* va,
               const float* B, int ldb, float init, float* C, int ldc, int* arg) {
    switch (G) {
#define ARMSP_MAX_CASE(GG)                                                                                     \
    case GG:                                                                                                   \
        hipLaunchKernelGGL((k_spmm_max<V, GG>), grid, dim3(256), 0, s, M, N, rp, ci, va, B, ldb, init, C, ldc, \
                           arg);                                                                               \
        break;
        ARMSP_MAX_CASE(4)
        ARMSP_MAX_CASE(8)
        ARMSP_MAX_CASE(16)
        ARMSP_MAX_CASE(32)
        ARMSP_MAX_CASE(64)
#undef ARMSP_MAX_CASE
        default: return ARMSP_ERROR_INVALID_VALUE;
    }
    return ARMSP_SUCCESS;
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace
}  // namespace armsp

extern "C" int armsp_spmm_max_f32(int M, int K, int N, int nnz, const int* rowptr, const int* colind,
                                  const float* val, const float* B, int ldb, float init, float* C, int ldc,
                                  int* argmax, void* stream) {
    using namespace armsp;
    if (M < 0 || K < 0 || N < 0 || nnz < 0) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0) return ARMSP_SUCCESS;
    if (!rowptr || !C || ldc < N) return ARMSP_ERROR_INVALID_VALUE;
    if (nnz > 0 && (!colind || !B || ldb < N || K == 0)) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    const int V = (N % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 && aligned16(B) && aligned16(C)) ? 4 : 1;
    int G = 4;
    while (G < 64 && G * V < N) G <<= 1;
    const dim3 grid((unsigned)(((int64_t)M * G + 255) / 256), (unsigned)((N + G * V - 1) / (G * V)));
    if (grid.y > 65535) return ARMSP_ERROR_NOT_SUPPORTED;
    TimedLaunch t("spmm_max", s);
    const int st = V == 4 ? launch_max<4>(G, grid, s, M, N, rowptr, colind, val, B, ldb, init, C, ldc, argmax)
                          : launch_max<1>(G, grid, s, M, N, rowptr, colind, val, B, ldb, init, C, ldc, argmax);
    if (st) return st;
    ARMSP_LAUNCHED();
    t.stop();
    return ARMSP_SUCCESS;
}

extern "C" int armsp_spmm_max_backward_f32(int M, int K, int N, int nnz, const int* colind, const float* val,
                                           const float* dC, int ldc, const int* argmax, float* dB, int ldb,
                                           void* stream) {
    using namespace armsp;
    if (M < 0 || K < 0 || N < 0 || nnz < 0) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0 || nnz == 0) return ARMSP_SUCCESS;
    if (!colind || !dC || !argmax || !dB || ldc < N || ldb < N || K == 0) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    const int64_t n = (int64_t)M * N;
    TimedLaunch t("spmm_max", s);
    hipLaunchKernelGGL(k_spmm_max_backward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, M, K, N, nnz, colind,
                       val, dC, ldc, argmax, dB, ldb);
    ARMSP_LAUNCHED();
    t.stop();
    return ARMSP_SUCCESS;
}
