// Sparse × sparse → dense for gfx950: C[M×N] += A·B, A and B CSR.
//
// Reference: sparse_matrix_mul3 (SpMM.cpp:309-327) — per row m of A, a
// Gustavson scatter C[m, colB[t]] += valA[idx]·valB[t] directly into the dense
// output row.  Here one wave owns (row m, a tile of up to TN columns): the C
// tile is staged in LDS, the nonzeros of A's row are taken in CSR order and
// each B row is scattered by the lanes (its columns are distinct, so one pass
// has no conflicts), then the tile is written back.  A window of 64 A nonzeros
// loads its B-row bounds at once, and each B row's entries are loaded one step
// before they are scattered, so the walk does not wait on memory per nonzero.  Every output element
// sees the same fmaf sequence as the reference, so the result is bit-exact.
#include "common.hpp"

namespace armsp {
namespace {

constexpr int kTN = 4096;  // columns per wave tile (16 KiB of LDS per wave; 1024 and 2048 time the same)

__global__ __launch_bounds__(256) void k_spgemm_dense(int M, int N, int ntiles, const int* __restrict__ rpA,
                                                      const int* __restrict__ ciA, const float* __restrict__ vA,
                                                      const int* __restrict__ rpB, const int* __restrict__ ciB,
                                                      const float* __restrict__ vB, float* __restrict__ C, int ldc,
                                                      int vec4) {
    __shared__ __attribute__((aligned(16))) float tile[4][kTN];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int item = uniform((int)(blockIdx.x * 4u + wid));
    if (item >= M * ntiles) return;
    const int m = item / ntiles;
    const int c0 = (item % ntiles) * kTN;
    const int tn = min(kTN, N - c0);
    float* t = tile[wid];
    float* crow = C + (size_t)m * ldc + c0;
    if (vec4) {  // C rows 16-byte aligned, tn a multiple of 4
        for (int k = lane * 4; k < tn; k += 256)
            *reinterpret_cast<float4*>(&t[k]) = *reinterpret_cast<const float4*>(&crow[k]);
    } else {
        for (int k = lane; k < tn; k += 64) t[k] = crow[k];
    }
    wave_sync_lds();

    const int beg = rpA[m], end = rpA[m + 1];
    for (int wbase = beg; wbase < end; wbase += 64) {
        // this window's A nonzeros and the bounds of their B rows, one per lane,
        // all loaded at once; the steps below take them in CSR order
        const int n = min(64, end - wbase);
        int wk = 0, bl = 0, el = 0;
        float wa = 0.0f;
        if (lane < n) {
            wk = ciA[wbase + lane];
            wa = vA ? vA[wbase + lane] : 1.0f;
            bl = rpB[wk];
            el = rpB[wk + 1];
        }
        // the first 64 entries of B row q are loaded one step ahead, so a step
        // scatters while the next step's loads are in flight
        auto fetch = [&](int q, int& j, float& v) {
            const int bb = read_lane(bl, q), be = read_lane(el, q);
            const int p = bb + lane;
            j = -1;
            v = 0.0f;
            if (p < be && lane < 64) {
                j = ciB[p];
                v = vB ? vB[p] : 1.0f;
            }
        };
        int jc;
        float vc;
        fetch(0, jc, vc);
        for (int q = 0; q < n; ++q) {
            const float a = read_lane(wa, q);
            int jn = -1;
            float vn = 0.0f;
            if (q + 1 < n) fetch(q + 1, jn, vn);
            // B row q's columns are distinct, so the lanes' updates do not collide
            {
                const int j = jc - c0;
                if (jc >= 0 && j >= 0 && j < tn) t[j] = fmaf(a, vc, t[j]);
            }
            const int bb = read_lane(bl, q), be = read_lane(el, q);
            for (int p = bb + 64 + lane; p < be; p += 64) {  // the rest of a B row over 64 entries
                const int j = ciB[p] - c0;
                if (j >= 0 && j < tn) t[j] = fmaf(a, vB ? vB[p] : 1.0f, t[j]);
            }
            wave_sync_lds();  // step q's updates land before step q+1 touches the same elements
            jc = jn;
            vc = vn;
        }
    }
    if (vec4) {
        for (int k = lane * 4; k < tn; k += 256)
            *reinterpret_cast<float4*>(&crow[k]) = *reinterpret_cast<const float4*>(&t[k]);
    } else {
        for (int k = lane; k < tn; k += 64) crow[k] = t[k];
    }
}

}  // namespace
}  // namespace armsp

extern "C" int armsp_spgemm_dense_f32(int M, int K, int N, int nnzA, const int* rowptrA, const int* colindA,
                                      const float* valA, int nnzB, const int* rowptrB, const int* colindB,
                                      const float* valB, float* C, int ldc, void* stream) {
    using namespace armsp;
    if (M < 0 || K < 0 || N < 0 || nnzA < 0 || nnzB < 0) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0 || nnzA == 0 || nnzB == 0) return ARMSP_SUCCESS;
    if (!rowptrA || !colindA || !rowptrB || !colindB || !C || ldc < N) return ARMSP_ERROR_INVALID_VALUE;
    const int ntiles = (N + kTN - 1) / kTN;
    const int64_t items = (int64_t)M * ntiles;
    if (items > 0x7fffffffLL - 4) return ARMSP_ERROR_NOT_SUPPORTED;
    hipStream_t s = as_stream(stream);
    TimedLaunch t("spgemm", s);
    // 16-byte row staging when every tile row starts 16-byte aligned and is whole float4s
    const int vec4 = ((uintptr_t)C % 16 == 0 && ldc % 4 == 0 && N % 4 == 0) ? 1 : 0;
    hipLaunchKernelGGL(k_spgemm_dense, dim3((unsigned)((items + 3) / 4)), dim3(256), 0, s, M, N, ntiles, rowptrA,
                       colindA, valA, rowptrB, colindB, valB, C, ldc, vec4);
    ARMSP_LAUNCHED();
    t.stop();
    return ARMSP_SUCCESS;
}
