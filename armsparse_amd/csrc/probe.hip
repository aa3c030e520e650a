// Measurement probe of the memory system: the B-row gathers of the SpMM
// kernels without the arithmetic, the staging of CSR streams or the output.
// Each wave takes `per_wave` row indices, stages them in LDS, and gathers
// its column slice of every indexed row of B: `group` lanes cover a row slice
// with `vec` floats each (P = 64 / group rows per load instruction), four load
// instructions form a batch and two batches are in flight, as in the SpMM
// kernels.  With S = N / (group·vec) slices, workgroup b serves slice (b mod 8)
// mod S, so each XCD's L2 sees one slice of every row (the SpMM slice kernel's
// mapping).  The loaded values are summed into `out` so nothing is elided.
// Comparing its rate on a matrix's column indices with its rate on relabelled
// or uniform indices separates what the gather pattern costs from what the
// SpMM kernel around it adds.  No reference counterpart.
#include "common.hpp"

namespace armsp {
namespace {

constexpr int kProbeU = 4;
constexpr int kProbeMaxPerWave = 1024;

template <int V>
__device__ __forceinline__ void probe_load(__amdgpu_buffer_rsrc_t rs, uint32_t off, float (&o)[V]) {
    if constexpr (V == 4) {
        const auto t = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
        o[0] = __uint_as_float(t[0]);
        o[1] = __uint_as_float(t[1]);
        o[2] = __uint_as_float(t[2]);
        o[3] = __uint_as_float(t[3]);
    } else {
        const auto t = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
        o[0] = __uint_as_float(t[0]);
        o[1] = __uint_as_float(t[1]);
    }
}

template <int V, int G>
__global__ __launch_bounds__(256) void k_probe_gather(int n, int per_wave, const int* __restrict__ idx,
                                                      const float* __restrict__ B, int ldb, int S,
                                                      float* __restrict__ out) {
    constexpr int P = 64 / G;
    constexpr int SW = G * V;
    constexpr int U = kProbeU;
    __shared__ uint32_t s_off[4][kProbeMaxPerWave];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int p = lane / G, l = lane % G;
    const int x = (int)(blockIdx.x & 7u);
    const int sl = x % S;
    const int g = (int)(blockIdx.x >> 3) * (8 / S) + x / S;
    const int wave = g * 4 + wid;
    const int64_t beg = (int64_t)wave * per_wave;
    if (beg >= n) return;
    const int cnt = (int)min((int64_t)per_wave, (int64_t)n - beg);
    const uint32_t rowbytes = (uint32_t)ldb * 4u;
    for (int k = lane; k < cnt; k += 64) s_off[wid][k] = (uint32_t)idx[beg + k] * rowbytes;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(B + sl * SW), (short)0, (int)0xffffffff, 0x00020000);
    const uint32_t voff = (uint32_t)(l * V * 4);
    float acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;
    const int kend = (cnt / (P * U)) * (P * U);
    float bA[U][V], bB[U][V];
    auto issue = [&](int kb, float (&bb)[U][V]) {
#pragma unroll
        for (int u = 0; u < U; ++u) probe_load<V>(rs, voff + s_off[wid][kb + P * u + p], bb[u]);
    };
    auto use = [&](const float (&bb)[U][V]) {
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] += bb[u][v];
    };
    int k = 0;
    if (k < kend) issue(k, bA);
    while (k < kend) {
        if (k + P * U < kend) issue(k + P * U, bB);
        use(bA);
        k += P * U;
        if (k >= kend) break;
        if (k + P * U < kend) issue(k + P * U, bA);
        use(bB);
        k += P * U;
    }
    float s = 0.0f;
#pragma unroll
    for (int v = 0; v < V; ++v) s += acc[v];
    out[((size_t)wave * S + sl) * 64 + lane] = s;
}

}  // namespace
}  // namespace armsp

extern "C" int armsp_probe_gather_f32(int n, const int* idx, const float* B, int K, int ldb, int N, int group,
                                      int vec, float* out, void* stream) {
    using namespace armsp;
    if (n < 0 || K <= 0 || N <= 0 || ldb < N || !idx || !B || !out) return ARMSP_ERROR_INVALID_VALUE;
    if ((uint64_t)K * (uint64_t)ldb * 4u > 0xffffffffull) return ARMSP_ERROR_NOT_SUPPORTED;
    const int sw = group * vec;
    if ((vec != 2 && vec != 4) || (group != 8 && group != 16 && group != 32 && group != 64) || N % sw)
        return ARMSP_ERROR_INVALID_VALUE;
    const int S = N / sw;
    if (8 % S) return ARMSP_ERROR_NOT_SUPPORTED;
    if (n == 0) return ARMSP_SUCCESS;
    const int per_wave = 512;
    const int64_t nw = ((int64_t)n + per_wave - 1) / per_wave;
    const int64_t G = (nw + 3) / 4;
    const int64_t nblk = ((G + 8 / S - 1) / (8 / S)) * 8;
    if (nblk > INT32_MAX) return ARMSP_ERROR_NOT_SUPPORTED;
    hipStream_t s = as_stream(stream);
    TimedLaunch t("probe", s);
#define ARMSP_PROBE(VV, GG)                                                                                         \
    if (vec == VV && group == GG)                                                                                   \
        hipLaunchKernelGGL((k_probe_gather<VV, GG>), dim3((unsigned)nblk), dim3(256), 0, s, n, per_wave, idx, B, ldb, \
                           S, out);
    ARMSP_PROBE(2, 8)
    ARMSP_PROBE(2, 16)
    ARMSP_PROBE(2, 32)
    ARMSP_PROBE(2, 64)
    ARMSP_PROBE(4, 8)
    ARMSP_PROBE(4, 16)
    ARMSP_PROBE(4, 32)
    ARMSP_PROBE(4, 64)
#undef ARMSP_PROBE
    ARMSP_LAUNCHED();
    t.stop();
    return ARMSP_SUCCESS;
}
