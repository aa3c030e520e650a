// bf16 GEMM on gfx950 MFMA (v_mfma_f32_32x32x16_bf16), fp32 accumulation.
//
// Used for the dense feature transform of the GCN layer, x @ W
// (gcn/pytorch-custom/op.py:114), and for its weight gradient xᵀ @ dY.  Both
// operands take arbitrary element strides so transposed views need no copy.
// Block tile 64×64×32, four waves each owning a 32×32 quarter; tiles are
// staged in LDS k-contiguous so each lane fetches its 8 operand values with one
// 16-byte LDS read.  When the output grid is too small to fill the chip the K
// range is split; partial tiles go to a workspace and are summed in split
// order by a second kernel (deterministic).
#include "common.hpp"

namespace armsp {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;  // LDS row pitch in bf16

__device__ __forceinline__ uint16_t ld_elem(const uint16_t* __restrict__ p, int64_t off) { return p[off]; }

// C_part[z] = op(A)[m0:m0+64, kz0:kz1] · op(B)[kz0:kz1, n0:n0+64]
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, int kchunk, const uint16_t* __restrict__ A,
                                                   int64_t sam, int64_t sak, const uint16_t* __restrict__ B,
                                                   int64_t sbk, int64_t sbn, float* __restrict__ C, int64_t ldc,
                                                   int64_t zstride, int accumulate) {
    __shared__ __attribute__((aligned(16))) uint16_t sA[BM * LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sB[BN * LDK];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kz0 = blockIdx.z * kchunk;
    const int kz1 = min(K, kz0 + kchunk);
    const int wm = (wid & 1) * 32, wn = (wid >> 1) * 32;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

    // each thread stages 8 elements of each operand per K step
    const bool a_mcontig = (sam == 1 && sak != 1);
    const bool b_ncontig = (sbn == 1 && sbk != 1);
    for (int k0 = kz0; k0 < kz1; k0 += BK) {
        // A tile: 64 rows × 32 k
        if (!a_mcontig) {
            const int r = tid >> 2, kk = (tid & 3) * 8;  // row r, k in [kk, kk+8)
            const int gm = m0 + r;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int gk = k0 + kk + i;
                sA[r * LDK + kk + i] = (gm < M && gk < kz1) ? ld_elem(A, (int64_t)gm * sam + (int64_t)gk * sak) : 0;
            }
        } else {
            const int kk = tid >> 3, r = (tid & 7) * 8;  // k kk, rows [r, r+8)
            const int gk = k0 + kk;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int gm = m0 + r + i;
                sA[(r + i) * LDK + kk] = (gm < M && gk < kz1) ? ld_elem(A, (int64_t)gm * sam + (int64_t)gk * sak) : 0;
            }
        }
        // B tile stored n-major: sB[n][k]
        if (!b_ncontig) {
            const int c = tid >> 2, kk = (tid & 3) * 8;
            const int gn = n0 + c;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int gk = k0 + kk + i;
                sB[c * LDK + kk + i] = (gn < N && gk < kz1) ? ld_elem(B, (int64_t)gk * sbk + (int64_t)gn * sbn) : 0;
            }
        } else {
            const int kk = tid >> 3, c = (tid & 7) * 8;
            const int gk = k0 + kk;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int gn = n0 + c + i;
                sB[(c + i) * LDK + kk] = (gn < N && gk < kz1) ? ld_elem(B, (int64_t)gk * sbk + (int64_t)gn * sbn) : 0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            const int koff = ks + (lane >> 5) * 8;
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[(wm + (lane & 31)) * LDK + koff]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[(wn + (lane & 31)) * LDK + koff]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // output layout: register i → row 8*(i/4) + 4*(lane/32) + i%4, column lane%32
    float* Cz = C + (int64_t)blockIdx.z * zstride;
    const int gn = n0 + wn + (lane & 31);
    if (gn < N) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int gm = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
            if (gm < M) {
                float* p = Cz + (int64_t)gm * ldc + gn;
                *p = accumulate ? *p + acc[i] : acc[i];
            }
        }
    }
}

__global__ void k_splitk_reduce(int M, int N, int splits, const float* __restrict__ part, float* __restrict__ C,
                                int64_t ldc, int accumulate) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)M * N) return;
    const int m = (int)(idx / N), n = (int)(idx % N);
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[(int64_t)z * M * N + idx];
    float* p = C + (int64_t)m * ldc + n;
    *p = accumulate ? *p + s : s;
}

}  // namespace
}  // namespace armsp

extern "C" int armsp_gemm_bf16(int M, int N, int Kd, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
                               int64_t sbk, int64_t sbn, float* C, int64_t ldc, int accumulate, void* stream) {
    using namespace armsp;
    if (M < 0 || N < 0 || Kd < 0 || ldc < N) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0) return ARMSP_SUCCESS;
    if (!A || !B || !C) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    const int gx = (M + BM - 1) / BM, gy = (N + BN - 1) / BN;
    if (gy > 65535) return ARMSP_ERROR_NOT_SUPPORTED;
    // split K when the tile grid is small and K is long
    int splits = 1;
    const int tiles = gx * gy, target = 2 * device_cus();
    while (tiles * splits < target && Kd / (splits * 2) >= 4 * BK && splits < 64) splits *= 2;
    int kchunk = (Kd + splits - 1) / splits;
    kchunk = (kchunk + BK - 1) / BK * BK;
    if (kchunk == 0) kchunk = BK;
    splits = Kd > 0 ? (Kd + kchunk - 1) / kchunk : 1;
    TimedLaunch t("gemm", s);
    if (splits == 1) {
        hipLaunchKernelGGL(k_gemm_bf16, dim3(gx, gy, 1), dim3(256), 0, s, M, N, Kd, kchunk, A, sam, sak, B, sbk, sbn, C,
                           ldc, (int64_t)0, accumulate);
        ARMSP_LAUNCHED();
        t.stop();
        return ARMSP_SUCCESS;
    }
    void* ws = nullptr;
    int st = ws_alloc(&ws, (size_t)splits * M * N * sizeof(float), s);
    if (st) return st;
    float* part = static_cast<float*>(ws);
    hipLaunchKernelGGL(k_gemm_bf16, dim3(gx, gy, splits), dim3(256), 0, s, M, N, Kd, kchunk, A, sam, sak, B, sbk, sbn,
                       part, (int64_t)N, (int64_t)M * N, 0);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        const int64_t n = (int64_t)M * N;
        hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, M, N, splits, part, C,
                           ldc, accumulate);
        e = hipGetLastError();
    }
    ws_free(ws, s);
    if (e != hipSuccess) return hip_fail(e);
    t.stop();
    return ARMSP_SUCCESS;
}
