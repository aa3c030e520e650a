// Dense helpers: counter-based fills (same RNG as the host generators),
// fp32 → bf16 conversion, per-row fp8-e4m3 quantisation.
#include <cmath>

#include "common.hpp"

namespace armsp {
namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t host_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ void k_fill_uniform(int64_t n, uint64_t key, float lo, float hi, float* __restrict__ out) {
#pragma clang fp contract(off)  // lo + (hi-lo)·u unfused, as on the host
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t r = mix64(key + kGolden * (uint64_t)(i + 1));
        const float u = (float)(uint32_t)(r >> 40) * 0x1p-24f;
        out[i] = lo + (hi - lo) * u;
    }
}

__global__ void k_fill_normal(int64_t n, uint64_t key, float* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double u1 = ((double)(mix64(key + kGolden * (uint64_t)(2 * i + 1)) >> 11) + 1.0) * 0x1p-53;
        const double u2 = (double)(mix64(key + kGolden * (uint64_t)(2 * i + 2)) >> 11) * 0x1p-53;
        out[i] = (float)(sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
    }
}

__global__ void k_to_bf16(int64_t n, const float* __restrict__ x, uint16_t* __restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = f32_to_bf16_rne(x[i]);
}

// rows of N floats at pitch ldx → rows of N bf16 at pitch ldy
__global__ void k_rows_to_bf16_2d(int64_t total, int N, const float* __restrict__ x, int64_t ldx,
                                  uint16_t* __restrict__ y, int64_t ldy) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t m = i / N, n = i - m * N;
        y[m * ldy + n] = f32_to_bf16_rne(x[m * ldx + n]);
    }
}

// fp32 → OCP e4m3fn, round to nearest even, saturating at ±448.
__device__ __forceinline__ uint32_t f32_to_e4m3(float x) {
    const uint32_t bits = __float_as_uint(x);
    const uint32_t sign = (bits >> 24) & 0x80u;
    const float a = fabsf(x);
    if (a != a) return sign | 0x7fu;
    if (a >= 448.0f) return sign | 0x7eu;
    if (a < 0x1p-6f) return sign | (uint32_t)rintf(a * 512.0f);  // subnormals (8 rounds up to the least normal)
    const uint32_t ab = __float_as_uint(a);
    int e = (int)(ab >> 23) - 127;
    uint32_t mant = ab & 0x7fffffu;
    uint32_t m3 = mant >> 20, rem = mant & 0xfffffu;
    if (rem > 0x80000u || (rem == 0x80000u && (m3 & 1u))) ++m3;
    if (m3 == 8u) {
        m3 = 0;
        ++e;
    }
    uint32_t code = ((uint32_t)(e + 7) << 3) | m3;
    if (code > 0x7eu) code = 0x7eu;
    return sign | code;
}

__global__ __launch_bounds__(256) void k_quant_fp8(int K, int N, const float* __restrict__ B, int ldb,
                                                   uint8_t* __restrict__ Q, float* __restrict__ scale) {
    const int lane = threadIdx.x & 63;
    const int k = (int)(blockIdx.x * 4u + (threadIdx.x >> 6));
    if (k >= K) return;
    const float* row = B + (size_t)k * ldb;
    float mx = 0.0f;
    for (int j = lane; j < N; j += 64) mx = fmaxf(mx, fabsf(row[j]));
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) mx = fmaxf(mx, __shfl_xor(mx, m));
    const float sc = mx > 0.0f ? mx / 448.0f : 1.0f;
    if (lane == 0) scale[k] = sc;
    uint8_t* qrow = Q + (size_t)k * ldb;
    for (int j = lane; j < N; j += 64) qrow[j] = (uint8_t)f32_to_e4m3(row[j] / sc);
}

// out[k][j] = decode(Q[k][j])·scale[k]; four codes per thread
__global__ __launch_bounds__(256) void k_dequant_fp8(int K, int N, const uint8_t* __restrict__ Q, int ldq,
                                                     const float* __restrict__ scale, float* __restrict__ out,
                                                     int ldo) {
    const int64_t quads = (int64_t)K * ((N + 3) / 4);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < quads; t += (int64_t)gridDim.x * blockDim.x) {
        const int qpr = (N + 3) / 4;
        const int k = (int)(t / qpr), j = (int)(t - (int64_t)k * qpr) * 4;
        const uint8_t* q = Q + (size_t)k * ldq + j;
        uint32_t w = 0;
        for (int i = 0; i < 4 && j + i < N; ++i) w |= (uint32_t)q[i] << (8 * i);
        float f[4];
        e4m3x4_to_f32(w, f);
        const float sc = scale ? scale[k] : 1.0f;
        for (int i = 0; i < 4 && j + i < N; ++i) out[(size_t)k * ldo + j + i] = f[i] * sc;
    }
}

inline dim3 grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

}  // namespace
}  // namespace armsp

extern "C" {

int armsp_fill_uniform_f32(int64_t n, uint64_t seed, uint64_t stream_id, float lo, float hi, float* out,
                           void* stream) {
    using namespace armsp;
    if (n < 0 || (n > 0 && !out)) return ARMSP_ERROR_INVALID_VALUE;
    if (n == 0) return ARMSP_SUCCESS;
    const uint64_t key = host_mix64(seed + kGolden * (stream_id + 1));
    hipLaunchKernelGGL(k_fill_uniform, grid_for(n), dim3(256), 0, as_stream(stream), n, key, lo, hi, out);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_fill_normal_f32(int64_t n, uint64_t seed, uint64_t stream_id, float* out, void* stream) {
    using namespace armsp;
    if (n < 0 || (n > 0 && !out)) return ARMSP_ERROR_INVALID_VALUE;
    if (n == 0) return ARMSP_SUCCESS;
    const uint64_t key = host_mix64(seed + kGolden * (stream_id + 1));
    hipLaunchKernelGGL(k_fill_normal, grid_for(n), dim3(256), 0, as_stream(stream), n, key, out);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_convert_f32_to_bf16(int64_t n, const float* x, uint16_t* y, void* stream) {
    using namespace armsp;
    if (n < 0 || (n > 0 && (!x || !y))) return ARMSP_ERROR_INVALID_VALUE;
    if (n == 0) return ARMSP_SUCCESS;
    hipLaunchKernelGGL(k_to_bf16, grid_for(n), dim3(256), 0, as_stream(stream), n, x, y);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_convert_rows_f32_to_bf16(int M, int N, const float* x, int64_t ldx, uint16_t* y, int64_t ldy,
                                   void* stream) {
    using namespace armsp;
    if (M < 0 || N < 0 || ldx < N || ldy < N) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0) return ARMSP_SUCCESS;
    if (!x || !y) return ARMSP_ERROR_INVALID_VALUE;
    const int64_t total = (int64_t)M * N;
    hipLaunchKernelGGL(k_rows_to_bf16_2d, grid_for(total), dim3(256), 0, as_stream(stream), total, N, x, ldx, y, ldy);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_quantize_fp8_rows(int K, int N, const float* B, int ldb, uint8_t* Bq, float* scale, void* stream) {
    using namespace armsp;
    if (K < 0 || N < 0 || ldb < N) return ARMSP_ERROR_INVALID_VALUE;
    if (K == 0 || N == 0) return ARMSP_SUCCESS;
    if (!B || !Bq || !scale) return ARMSP_ERROR_INVALID_VALUE;
    hipLaunchKernelGGL(k_quant_fp8, dim3((unsigned)((K + 3) / 4)), dim3(256), 0, as_stream(stream), K, N, B, ldb, Bq,
                       scale);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_dequantize_fp8_rows(int K, int N, const uint8_t* Bq, int ldq, const float* scale, float* out, int ldo,
                              void* stream) {
    using namespace armsp;
    if (K < 0 || N < 0 || ldq < N || ldo < N) return ARMSP_ERROR_INVALID_VALUE;
    if (K == 0 || N == 0) return ARMSP_SUCCESS;
    if (!Bq || !out) return ARMSP_ERROR_INVALID_VALUE;
    hipLaunchKernelGGL(k_dequant_fp8, grid_for((int64_t)K * ((N + 3) / 4)), dim3(256), 0, as_stream(stream), K, N, Bq,
                       ldq, scale, out, ldo);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

}  // extern "C"
