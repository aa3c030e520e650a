// Kernels of the learned schedule selector: sparsity features of a CSR
// matrix, and the forward / backward / update / inference kernels of a small
// fp32 MLP.  No reference counterpart: the reference picks SpMM kernels by a
// fixed N threshold (gcn/pytorch-custom/spmm_kernel.cu:437-457).
#include <cfloat>

#include "common.hpp"

namespace armsp {
namespace {

// ------------------------------------------------------------ row stats
// hist[0] counts empty rows, hist[b] (b ≥ 1) rows with 2^(b-1) ≤ len < 2^b,
// the last bin everything longer; stats[0] = max row length.
__global__ void k_zero_int(int n, int* __restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(256) void k_row_stats(int M, const int* __restrict__ rowptr, int nbins,
                                                   int* __restrict__ hist, int* __restrict__ maxlen) {
    __shared__ int h[64];
    __shared__ int mx;
    if (threadIdx.x < 64) h[threadIdx.x] = 0;
    if (threadIdx.x == 0) mx = 0;
    __syncthreads();
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < M; r += gridDim.x * blockDim.x) {
        const int len = rowptr[r + 1] - rowptr[r];
        const int b = len == 0 ? 0 : min(nbins - 1, 1 + (31 - __clz(len)));
        atomicAdd(&h[b], 1);
        atomicMax(&mx, len);
    }
    __syncthreads();
    if (threadIdx.x < nbins && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
    if (threadIdx.x == 0) atomicMax(maxlen, mx);
}

// ------------------------------------------------------------ fp32 GEMM
// C = op(A)·op(B) (+ beta·C) (+ bias[n]) (then ReLU).  64×64 tile, 16×16
// threads each owning a 4×4 block, K in steps of 16 through LDS.
constexpr int TM = 64, TN = 64, TK = 16;

__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, const float* __restrict__ A, int64_t sam,
                                               int64_t sak, const float* __restrict__ B, int64_t sbk, int64_t sbn,
                                               float* __restrict__ C, int64_t ldc, float beta,
                                               const float* __restrict__ bias, int relu) {
    __shared__ float sA[TK][TM + 1];
    __shared__ float sB[TK][TN + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += TK) {
        for (int i = threadIdx.x; i < TK * TM; i += 256) {
            const int kk = i / TM, mm = i % TM;
            const int gm = m0 + mm, gk = k0 + kk;
            sA[kk][mm] = (gm < M && gk < K) ? A[gm * sam + gk * sak] : 0.0f;
        }
        for (int i = threadIdx.x; i < TK * TN; i += 256) {
            const int kk = i / TN, nn = i % TN;
            const int gn = n0 + nn, gk = k0 + kk;
            sB[kk][nn] = (gn < N && gk < K) ? B[gk * sbk + gn * sbn] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sA[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = sB[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gm = m0 + ty * 4 + i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gn = n0 + tx * 4 + j;
            if (gn >= N) continue;
            float v = acc[i][j];
            float* p = C + gm * ldc + gn;
            if (beta != 0.0f) v += beta * *p;
            if (bias) v += bias[gn];
            if (relu) v = fmaxf(v, 0.0f);
            *p = v;
        }
    }
}

// ------------------------------------------------- loss and elementwise
// One block: per-row softmax and mean cross-entropy against hard labels or
// soft target rows; dlogits = (softmax - target) / rows.
__global__ __launch_bounds__(256) void k_softmax_xent(int rows, int classes, const float* __restrict__ logits,
                                                      const int* __restrict__ labels,
                                                      const float* __restrict__ targets, float* __restrict__ dlogits,
                                                      float* __restrict__ loss) {
    __shared__ float part[256];
    float l = 0.0f;
    for (int r = threadIdx.x; r < rows; r += 256) {
        const float* z = logits + (int64_t)r * classes;
        float mx = -FLT_MAX;
        for (int c = 0; c < classes; ++c) mx = fmaxf(mx, z[c]);
        float s = 0.0f;
        for (int c = 0; c < classes; ++c) s += expf(z[c] - mx);
        const float lse = logf(s);
        for (int c = 0; c < classes; ++c) {
            const float q = targets ? targets[(int64_t)r * classes + c] : (c == labels[r] ? 1.0f : 0.0f);
            if (q != 0.0f) l -= q * (z[c] - mx - lse);
            if (dlogits) dlogits[(int64_t)r * classes + c] = (expf(z[c] - mx) / s - q) / rows;
        }
    }
    part[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = part[0] / rows;
}

__global__ void k_relu_backward(int64_t n, const float* __restrict__ act, float* __restrict__ grad) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n && act[i] <= 0.0f) grad[i] = 0.0f;
}

// out[n] = Σ_m X[m][n], one thread per column, rows summed in order.
__global__ void k_colsum(int M, int N, const float* __restrict__ X, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
#pragma unroll 8
    for (int m = 0; m < M; ++m) s += X[(int64_t)m * N + n];  // loads run ahead of the in-order adds
    out[n] = s;
}

// Tall matrices: block b sums rows [b·R, (b+1)·R) — 64 columns by 4 row
// phases, the phases joined in LDS in a fixed order — into part[b][·];
// k_colsum_parts then adds the blocks' sums the same way (block phases, fixed
// join).  Deterministic.  R is small (128 rows) so that a short matrix still
// spreads over many blocks, and every loop keeps 8 loads in flight ahead of
// its in-order adds.
__global__ __launch_bounds__(256) void k_colsum_blocks(int M, int N, int R, const float* __restrict__ X,
                                                       float* __restrict__ part) {
    __shared__ float s[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int n = blockIdx.y * 64 + tx;
    const int64_t r0 = (int64_t)blockIdx.x * R;
    const int64_t r1 = r0 + R < M ? r0 + R : M;
    float acc = 0.0f;
    if (n < N) {
#pragma unroll 8
        for (int64_t m = r0 + ty; m < r1; m += 4) acc += X[m * N + n];
    }
    s[ty][tx] = acc;
    __syncthreads();
    if (ty == 0 && n < N) part[(int64_t)blockIdx.x * N + n] = ((s[0][tx] + s[1][tx]) + s[2][tx]) + s[3][tx];
}

__global__ __launch_bounds__(256) void k_colsum_parts(int nb, int N, const float* __restrict__ part,
                                                      float* __restrict__ out) {
    __shared__ float s[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + tx;
    float acc = 0.0f;
    if (n < N) {
#pragma unroll 8
        for (int b = ty; b < nb; b += 4) acc += part[(int64_t)b * N + n];
    }
    s[ty][tx] = acc;
    __syncthreads();
    if (ty == 0 && n < N) out[n] = ((s[0][tx] + s[1][tx]) + s[2][tx]) + s[3][tx];
}

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float wd, float c1, float c2) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    m[i] = b1 * m[i] + (1.0f - b1) * gi;
    v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
    p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
}

// ------------------------------------------------------------ inference
// One workgroup per feature vector: standardise, three dense layers
// (ReLU, ReLU, linear), argmax.  H ≤ 256, C ≤ 256.
__global__ __launch_bounds__(256) void k_mlp_infer(int nvec, int F, int H, int Cc, const float* __restrict__ x,
                                                   const float* __restrict__ mu, const float* __restrict__ sd,
                                                   const float* __restrict__ W1, const float* __restrict__ b1,
                                                   const float* __restrict__ W2, const float* __restrict__ b2,
                                                   const float* __restrict__ W3, const float* __restrict__ b3,
                                                   float* __restrict__ logits, int* __restrict__ cls) {
    __shared__ float xs[256], h1[256], h2[256], z[256];
    const int v = blockIdx.x, t = threadIdx.x;
    if (v >= nvec) return;
    if (t < F) xs[t] = (x[(int64_t)v * F + t] - mu[t]) / sd[t];
    __syncthreads();
    if (t < H) {
        float s = b1[t];
        for (int k = 0; k < F; ++k) s = fmaf(xs[k], W1[k * H + t], s);
        h1[t] = fmaxf(s, 0.0f);
    }
    __syncthreads();
    if (t < H) {
        float s = b2[t];
        for (int k = 0; k < H; ++k) s = fmaf(h1[k], W2[k * H + t], s);
        h2[t] = fmaxf(s, 0.0f);
    }
    __syncthreads();
    if (t < Cc) {
        float s = b3[t];
        for (int k = 0; k < H; ++k) s = fmaf(h2[k], W3[k * Cc + t], s);
        z[t] = s;
        if (logits) logits[(int64_t)v * Cc + t] = s;
    }
    __syncthreads();
    if (t == 0) {
        int best = 0;
        for (int c = 1; c < Cc; ++c)
            if (z[c] > z[best]) best = c;
        cls[v] = best;
    }
}

inline unsigned blocks_for(int64_t n, int tpb = 256) { return (unsigned)((n + tpb - 1) / tpb); }

}  // namespace
}  // namespace armsp

extern "C" {

int armsp_row_stats(int M, const int* rowptr, int nbins, int* hist, int* maxlen, void* stream) {
    using namespace armsp;
    if (M < 0 || nbins < 1 || nbins > 64 || !hist || !maxlen || (M > 0 && !rowptr)) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, nbins, hist);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, 1, maxlen);
    if (M > 0) {
        const unsigned g = min(blocks_for(M), 1024u);
        hipLaunchKernelGGL(k_row_stats, dim3(g), dim3(256), 0, s, M, rowptr, nbins, hist, maxlen);
    }
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_sgemm_f32(int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk,
                    int64_t sbn, float* C, int64_t ldc, float beta, const float* bias, int relu, void* stream) {
    using namespace armsp;
    if (M < 0 || N < 0 || K < 0 || ldc < N) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0) return ARMSP_SUCCESS;
    if (!C || (K > 0 && (!A || !B))) return ARMSP_ERROR_INVALID_VALUE;
    dim3 grid((M + TM - 1) / TM, (N + TN - 1) / TN);
    if (grid.y > 65535) return ARMSP_ERROR_NOT_SUPPORTED;
    hipLaunchKernelGGL(k_sgemm, grid, dim3(256), 0, as_stream(stream), M, N, K, A, sam, sak, B, sbk, sbn, C, ldc,
                       beta, bias, relu);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_softmax_xent(int rows, int classes, const float* logits, const int* labels, const float* targets,
                       float* dlogits, float* loss, void* stream) {
    using namespace armsp;
    if (rows <= 0 || classes <= 0 || !logits || (!labels && !targets) || !loss) return ARMSP_ERROR_INVALID_VALUE;
    hipLaunchKernelGGL(k_softmax_xent, dim3(1), dim3(256), 0, as_stream(stream), rows, classes, logits, labels,
                       targets, dlogits, loss);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_relu_backward(int64_t n, const float* act, float* grad, void* stream) {
    using namespace armsp;
    if (n < 0 || (n > 0 && (!act || !grad))) return ARMSP_ERROR_INVALID_VALUE;
    if (n == 0) return ARMSP_SUCCESS;
    hipLaunchKernelGGL(k_relu_backward, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), n, act, grad);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_colsum_f32(int M, int N, const float* X, float* out, void* stream) {
    using namespace armsp;
    if (M < 0 || N < 0 || (N > 0 && (!X || !out))) return ARMSP_ERROR_INVALID_VALUE;
    if (N == 0) return ARMSP_SUCCESS;
    hipStream_t s = as_stream(stream);
    if (M <= 4096 || N > 65535 * 64) {  // short columns: a thread per column
        hipLaunchKernelGGL(k_colsum, dim3(blocks_for(N)), dim3(256), 0, s, M, N, X, out);
        ARMSP_LAUNCHED();
        return ARMSP_SUCCESS;
    }
    const int64_t nb64 = ((int64_t)M + 127) / 128;
    const int nb = (int)(nb64 < 4096 ? nb64 : 4096);
    const int R = (int)(((int64_t)M + nb - 1) / nb);
    void* ws = nullptr;
    int st = ws_alloc(&ws, (size_t)nb * N * sizeof(float), s);
    if (st) return st;
    float* part = static_cast<float*>(ws);
    hipLaunchKernelGGL(k_colsum_blocks, dim3((unsigned)nb, (unsigned)((N + 63) / 64)), dim3(256), 0, s, M, N, R, X,
                       part);
    hipLaunchKernelGGL(k_colsum_parts, dim3((unsigned)((N + 63) / 64)), dim3(256), 0, s, nb, N, part, out);
    const hipError_t e = hipGetLastError();
    ws_free(ws, s);
    return e == hipSuccess ? ARMSP_SUCCESS : hip_fail(e);
}

int armsp_adam_step(int64_t n, float* param, const float* grad, float* m, float* v, float lr, float beta1,
                    float beta2, float eps, float weight_decay, int step, void* stream) {
    using namespace armsp;
    if (n < 0 || step < 1 || (n > 0 && (!param || !grad || !m || !v))) return ARMSP_ERROR_INVALID_VALUE;
    if (n == 0) return ARMSP_SUCCESS;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), n, param, grad, m, v, lr, beta1,
                       beta2, eps, weight_decay, c1, c2);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

int armsp_mlp_infer(int nvec, int F, int H, int C, const float* x, const float* mu, const float* sd, const float* W1,
                    const float* b1, const float* W2, const float* b2, const float* W3, const float* b3,
                    float* logits, int* cls, void* stream) {
    using namespace armsp;
    if (nvec < 0 || F < 1 || F > 256 || H < 1 || H > 256 || C < 1 || C > 256) return ARMSP_ERROR_INVALID_VALUE;
    if (nvec == 0) return ARMSP_SUCCESS;
    if (!x || !mu || !sd || !W1 || !b1 || !W2 || !b2 || !W3 || !b3 || !cls) return ARMSP_ERROR_INVALID_VALUE;
    hipLaunchKernelGGL(k_mlp_infer, dim3(nvec), dim3(256), 0, as_stream(stream), nvec, F, H, C, x, mu, sd, W1, b1, W2,
                       b2, W3, b3, logits, cls);
    ARMSP_LAUNCHED();
    return ARMSP_SUCCESS;
}

}  // extern "C"
