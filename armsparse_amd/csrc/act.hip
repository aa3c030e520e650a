// The GCN hidden layer's epilogue in one pass: bias, ReLU, dropout and the
// bfloat16 cast that the next layer's x·W reads.
//
// Reference: Net.forward (gcn_custom.py:84-86) applies F.relu to conv1's
// output and then F.dropout, and GCNConv adds its bias last (op.py:122-124):
// three passes over the [nodes × hidden] activations, plus a cast where the
// next layer runs in bf16.  Here the forward reads the fp32 pre-activation
// once and writes the bf16 activation and one mask bit per element (ReLU
// passed and dropout kept); the backward reads the bf16 gradient and the
// mask once, writes the bf16 gradient of the pre-activation (the operand of
// the weight-gradient GEMM) and sums the bias gradient on the way, in a fixed
// order (deterministic).
//
// Dropout draws one 24-bit uniform per element from a counter-based hash of
// (seed, element index) — the SplitMix64 output function (Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014) — and
// keeps the element when it is at least p·2^24.  The seed lives in device
// memory and the call advances it after its kernel, so every replay of a
// captured HIP graph draws a fresh mask.
#include "common.hpp"

namespace armsp {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// one item = 8 consecutive columns of one row (N % 8 == 0): two float4 loads,
// one 16-byte store of bf16, one mask byte
__global__ __launch_bounds__(256) void k_bias_relu_dropout(int64_t items, int N, const float* __restrict__ Y,
                                                           int64_t ldy, const float* __restrict__ bias,
                                                           uint32_t thresh, float scale,
                                                           const uint64_t* __restrict__ seedp,
                                                           uint16_t* __restrict__ H, uint8_t* __restrict__ mask) {
    const uint64_t seed = thresh ? *seedp : 0;
    const int ipr = N >> 3;
    for (int64_t it = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; it < items;
         it += (int64_t)gridDim.x * blockDim.x) {
        const int64_t m = it / ipr;
        const int n0 = (int)(it - m * ipr) * 8;
        const float4* src = reinterpret_cast<const float4*>(Y + m * ldy + n0);
        const float4 a = src[0], b = src[1];
        float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        if (bias) {
            const float4* bp = reinterpret_cast<const float4*>(bias + n0);
            const float4 c = bp[0], d = bp[1];
            v[0] += c.x; v[1] += c.y; v[2] += c.z; v[3] += c.w;
            v[4] += d.x; v[5] += d.y; v[6] += d.z; v[7] += d.w;
        }
        uint32_t bits = 0;
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            bool pass = v[j] > 0.0f;
            if (thresh) {
                const uint64_t e = (uint64_t)it * 8 + j;  // = m·N + n0 + j
                pass = pass && (uint32_t)(mix64(seed + e * 0x9e3779b97f4a7c15ull) >> 40) >= thresh;
            }
            const uint32_t hb = f32_to_bf16_rne(pass ? v[j] * scale : 0.0f);
            if (j & 1)
                o[j >> 1] |= hb << 16;
            else
                o[j >> 1] = hb;
            bits |= (uint32_t)pass << j;
        }
        *reinterpret_cast<uint4*>(H + m * N + n0) = make_uint4(o[0], o[1], o[2], o[3]);
        mask[it] = (uint8_t)bits;
    }
}

__global__ void k_seed_advance(uint64_t* seed) { *seed += 1; }

__global__ void k_zero(int n, float* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0f;
}

// Backward over rows [b·R, (b+1)·R): 8 columns per thread, the block's threads
// in P = 256 / (N/8) row phases; gy = pass ? g·scale : 0 in bf16, and per column
// the phases' fp32 sums joined in LDS in phase order into part[b][·].
__global__ __launch_bounds__(256) void k_act_backward(int M, int N, int R, const uint16_t* __restrict__ G,
                                                      const uint8_t* __restrict__ mask, float scale,
                                                      uint16_t* __restrict__ GY, int64_t ldgy,
                                                      float* __restrict__ part) {
    extern __shared__ float s_acc[];  // [P][N]
    const int ipr = N >> 3;
    const int P = 256 / ipr;
    const int t = threadIdx.x;
    const int cg = t % ipr, ph = t / ipr;
    const int64_t r0 = (int64_t)blockIdx.x * R;
    const int64_t r1 = r0 + R < M ? r0 + R : M;
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (ph < P) {
        for (int64_t m = r0 + ph; m < r1; m += P) {
            const int64_t it = m * ipr + cg;
            const uint4 gv = *reinterpret_cast<const uint4*>(G + it * 8);
            const uint32_t bits = mask[it];
            const uint32_t w[4] = {gv.x, gv.y, gv.z, gv.w};
            uint32_t o[4];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint16_t gb = (uint16_t)(j & 1 ? w[j >> 1] >> 16 : w[j >> 1] & 0xffffu);
                const float r = (bits >> j) & 1u ? bf16_to_f32(gb) * scale : 0.0f;
                acc[j] += r;
                const uint32_t hb = f32_to_bf16_rne(r);
                if (j & 1)
                    o[j >> 1] |= hb << 16;
                else
                    o[j >> 1] = hb;
            }
            *reinterpret_cast<uint4*>(GY + m * ldgy + cg * 8) = make_uint4(o[0], o[1], o[2], o[3]);
        }
        if (part) {
#pragma unroll
            for (int j = 0; j < 8; ++j) s_acc[ph * N + cg * 8 + j] = acc[j];
        }
    }
    if (!part) return;
    __syncthreads();
    for (int n = t; n < N; n += 256) {
        float sum = 0.0f;
        for (int q = 0; q < P; ++q) sum += s_acc[q * N + n];
        part[(int64_t)blockIdx.x * N + n] = sum;
    }
}

// out[n] = Σ_b part[b][n]: 16 columns × 16 phases per block, phases joined in
// order; 8 loads in flight ahead of each thread's in-order adds.
__global__ __launch_bounds__(256) void k_parts_sum(int nb, int N, const float* __restrict__ part,
                                                   float* __restrict__ out) {
    __shared__ float s[16][16];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int n = blockIdx.x * 16 + tx;
    float acc = 0.0f;
    if (n < N) {
#pragma unroll 8
        for (int b = ty; b < nb; b += 16) acc += part[(int64_t)b * N + n];
    }
    s[ty][tx] = acc;
    __syncthreads();
    if (ty == 0 && n < N) {
        float sum = 0.0f;
        for (int q = 0; q < 16; ++q) sum += s[q][tx];
        out[n] = sum;
    }
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// keep threshold of dropout probability p: an element is kept when its 24-bit
// uniform is at least p·2^24 (p = 0: no dropout)
inline uint32_t keep_threshold(float p) {
    const double t = (double)p * 16777216.0 + 0.5;
    return t >= 16777216.0 ? 16777215u : (uint32_t)t;
}

}  // namespace
}  // namespace armsp

extern "C" {

int armsp_bias_relu_dropout_bf16(int M, int N, const float* Y, int64_t ldy, const float* bias, float p,
                                 uint64_t* seed, uint16_t* H, uint8_t* mask, void* stream) {
    using namespace armsp;
    if (M < 0 || N < 0 || !(p >= 0.0f && p < 1.0f)) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0) return ARMSP_SUCCESS;
    if (N % 8 || ldy < N || ldy % 4 || !Y || !H || !mask || (p > 0.0f && !seed)) return ARMSP_ERROR_INVALID_VALUE;
    if (!aligned16(Y) || !aligned16(H) || (bias && !aligned16(bias))) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    const uint32_t thresh = keep_threshold(p);
    const float scale = thresh ? 1.0f / (1.0f - p) : 1.0f;
    const int64_t items = (int64_t)M * (N / 8);
    int64_t blocks = (items + 255) / 256;
    const int64_t cap = 16 * (int64_t)device_cus();
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_bias_relu_dropout, dim3((unsigned)blocks), dim3(256), 0, s, items, N, Y, ldy, bias, thresh,
                       scale, seed, H, mask);
    ARMSP_LAUNCHED();
    if (thresh) {
        hipLaunchKernelGGL(k_seed_advance, dim3(1), dim3(1), 0, s, seed);
        ARMSP_LAUNCHED();
    }
    return ARMSP_SUCCESS;
}

int armsp_bias_relu_dropout_backward_bf16(int M, int N, const uint16_t* G, const uint8_t* mask, float p,
                                          uint16_t* GY, int64_t ldgy, float* dbias, void* stream) {
    using namespace armsp;
    if (M < 0 || N < 0 || !(p >= 0.0f && p < 1.0f)) return ARMSP_ERROR_INVALID_VALUE;
    if (N == 0) return ARMSP_SUCCESS;
    if (N % 8 || N > 2048 || ldgy < N || ldgy % 8) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    if (M == 0) {  // a fill kernel, not a memset node, so that the call stays capturable
        if (dbias) {
            hipLaunchKernelGGL(k_zero, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, dbias);
            ARMSP_LAUNCHED();
        }
        return ARMSP_SUCCESS;
    }
    if (!G || !mask || !GY || !aligned16(G) || !aligned16(GY)) return ARMSP_ERROR_INVALID_VALUE;
    const uint32_t thresh = keep_threshold(p);
    const float scale = thresh ? 1.0f / (1.0f - p) : 1.0f;
    int R = (int)(((int64_t)M + 1023) / 1024);
    if (R < 64) R = 64;
    const int nb = (int)(((int64_t)M + R - 1) / R);
    const int P = 256 / (N / 8);
    const size_t lds = dbias ? (size_t)P * N * sizeof(float) : 0;
    void* ws = nullptr;
    if (dbias) {
        const int st = ws_alloc(&ws, (size_t)nb * N * sizeof(float), s);
        if (st) return st;
    }
    float* part = static_cast<float*>(ws);
    hipLaunchKernelGGL(k_act_backward, dim3((unsigned)nb), dim3(256), lds, s, M, N, R, G, mask, scale, GY, ldgy,
                       part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && dbias) {
        hipLaunchKernelGGL(k_parts_sum, dim3((unsigned)((N + 15) / 16)), dim3(256), 0, s, nb, N, part, dbias);
        e = hipGetLastError();
    }
    if (ws) ws_free(ws, s);
    return e == hipSuccess ? ARMSP_SUCCESS : hip_fail(e);
}

}  // extern "C"
