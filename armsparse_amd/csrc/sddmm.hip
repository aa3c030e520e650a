// SDDMM for gfx950: out[e] = s_e · <U[row(e)], V[col(e)]> on a CSR or COO pattern.
//
// Reference semantics: gcn/pytorch-custom/sddmm.cu:7-424 (COO kernels with 16
// edges per block and an 8/16/32-lane shuffle all-reduce per edge; CSR kernels
// that binary-search the row of every edge, computeUtil.h:11-28).  Here each
// wave owns a contiguous range of nonzeros; a group of G lanes computes one
// dot product with W-element vector loads, each group takes G consecutive
// nonzeros, and a transpose-reduce (G-1 shuffles for G dot products) leaves
// lane L holding nonzero base+L, so the output store is one coalesced line
// per wave.  The row of a CSR nonzero is found by a per-lane binary search
// restricted to the wave's row range.
#include <climits>

#include "common.hpp"

namespace armsp {
namespace {

template <typename T, int W>
__device__ __forceinline__ void load_vec(const T* __restrict__ p, float (&o)[W]) {
    if constexpr (sizeof(T) == 4) {
        if constexpr (W == 4) {
            float4 t = *reinterpret_cast<const float4*>(p);
            o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
        } else {
#pragma unroll
            for (int i = 0; i < W; ++i) o[i] = p[i];
        }
    } else {
        if constexpr (W == 8) {
            uint4 t = *reinterpret_cast<const uint4*>(p);
            uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                o[2 * i] = __uint_as_float(w[i] << 16);
                o[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
            }
        } else if constexpr (W == 4) {
            uint2 t = *reinterpret_cast<const uint2*>(p);
            o[0] = __uint_as_float(t.x << 16); o[1] = __uint_as_float(t.x & 0xffff0000u);
            o[2] = __uint_as_float(t.y << 16); o[3] = __uint_as_float(t.y & 0xffff0000u);
        } else {
#pragma unroll
            for (int i = 0; i < W; ++i) o[i] = bf16_to_f32(p[i]);
        }
    }
}

// G lanes per nonzero, W elements per load, T loads per lane (Kd ≤ G·W·T).
template <typename T, int G, int W, bool COO>
__global__ __launch_bounds__(256) void k_sddmm(int M, int Kd, int nnz, int chunk, const int* __restrict__ rowidx,
                                               const int* __restrict__ colind, const float* __restrict__ scale,
                                               const T* __restrict__ U, int ldu, const T* __restrict__ V, int ldv,
                                               float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int g = lane / G, l = lane & (G - 1);
    const int wave = uniform((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    const int64_t j0l = (int64_t)wave * chunk;
    if (j0l >= nnz) return;
    const int j0 = (int)j0l;
    const int j1 = (int)(j0l + chunk < nnz ? j0l + chunk : nnz);
    int rlo = 0, rhi = 0;
    if constexpr (!COO) {
        rlo = uniform(rows_ending_by_wave(j0, rowidx, M));
        rhi = uniform(rows_ending_by_wave(j1 - 1, rowidx, M));
    }
    const int steps = (Kd + G * W - 1) / (G * W);
    const bool full = (Kd % (G * W)) == 0;

    for (int base = j0; base < j1; base += 64) {
        const int e = base + lane;
        int myrow = 0, mycol = 0;
        if (e < j1) {
            mycol = colind[e];
            if constexpr (COO) {
                myrow = rowidx[e];
            } else {
                int lo = rlo, hi = rhi;  // row r with rowptr[r] <= e < rowptr[r+1]
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (rowidx[mid] <= e)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                myrow = lo;
            }
        }
        if constexpr (!COO) {
            if (base + 64 < j1) rlo = uniform(__shfl(myrow, 63));  // later nonzeros are in rows ≥ this one
        }
        float p[G];
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const int src = g * G + q;
            const int r = __shfl(myrow, src);
            const int c = __shfl(mycol, src);
            float acc = 0.0f;
            if (base + src < j1) {
                const T* up = U + (size_t)r * ldu;
                const T* vp = V + (size_t)c * ldv;
                for (int t = 0; t < steps; ++t) {
                    const int k = (t * G + l) * W;
                    float a[W], b[W];
                    if (full || k + W <= Kd) {
                        load_vec<T, W>(up + k, a);
                        load_vec<T, W>(vp + k, b);
                    } else {
#pragma unroll
                        for (int i = 0; i < W; ++i) {
                            const bool in = k + i < Kd;
                            if constexpr (sizeof(T) == 4) {
                                a[i] = in ? (float)up[k + i] : 0.0f;
                                b[i] = in ? (float)vp[k + i] : 0.0f;
                            } else {
                                a[i] = in ? bf16_to_f32(up[k + i]) : 0.0f;
                                b[i] = in ? bf16_to_f32(vp[k + i]) : 0.0f;
                            }
                        }
                    }
#pragma unroll
                    for (int i = 0; i < W; ++i) acc = fmaf(a[i], b[i], acc);
                }
            }
            p[q] = acc;
        }
        // transpose-reduce: after the stage with mask m, lane keeps the half of
        // its G partials selected by (l & m); finally lane l holds nonzero l.
#pragma unroll
        for (int m = G / 2; m >= 1; m >>= 1) {
            const bool upper = (l & m) != 0;
#pragma unroll
            for (int i = 0; i < m; ++i) {
                const float send = upper ? p[i] : p[i + m];
                const float keep = upper ? p[i + m] : p[i];
                p[i] = keep + __shfl_xor(send, m);
            }
        }
        if (e < j1) out[e] = scale ? scale[e] * p[0] : p[0];
    }
}

// ------------------------------------------------ staged, 32-bit offsets
// Same lane mapping and transpose-reduce as k_sddmm, with the index work
// moved out of the dot-product loop: the wave first stages, in LDS, the byte
// offsets of the U row and the V row of every nonzero of its chunk (CSR: the
// chunk is a merge-path range of rows + nonzeros, so it holds at most chunk+1
// row ends and each nonzero's row is a binary search in LDS).  The loop then
// issues buffer loads at (row offset + lane offset) with no 64-bit address
// arithmetic and no predication: lanes past the chunk read row 0 and their
// results are never stored.
template <typename T, int W>
__device__ __forceinline__ void gather_vec(__amdgpu_buffer_rsrc_t rs, uint32_t off, float (&o)[W]) {
    constexpr int BYTES = W * (int)sizeof(T);
    uint32_t w[BYTES >= 4 ? BYTES / 4 : 1];
    if constexpr (BYTES == 16) {
        auto t = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
        w[0] = t[0]; w[1] = t[1]; w[2] = t[2]; w[3] = t[3];
    } else if constexpr (BYTES == 8) {
        auto t = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
        w[0] = t[0]; w[1] = t[1];
    } else if constexpr (BYTES == 4) {
        w[0] = __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
    } else {
        w[0] = __builtin_amdgcn_raw_buffer_load_b16(rs, off, 0, 0);
    }
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int i = 0; i < W; ++i) o[i] = __uint_as_float(w[i]);
    } else {
#pragma unroll
        for (int i = 0; i < W; ++i) o[i] = __uint_as_float((i & 1) ? (w[i >> 1] & 0xffff0000u) : (w[i >> 1] << 16));
    }
}

// S = 2: the dot products are split into two column halves (Kd is the half),
// and blocks are numbered so that the 8 XCDs (block b runs on XCD b % 8)
// alternate halves: each XCD's L2 then caches half of every hot V row.  Half 0
// goes to out, half 1 to out2; k_sddmm_combine adds them and scales.
template <typename T, int G, int W, bool COO, bool FULL>
__global__ __launch_bounds__(256) void k_sddmm_stream(int M, int Kd, int nnz, int chunk,
                                                      const int* __restrict__ rowidx, const int* __restrict__ colind,
                                                      const float* __restrict__ scale, const T* __restrict__ U,
                                                      int ldu, const T* __restrict__ V, int ldv,
                                                      const int* __restrict__ part, float* __restrict__ out, int S,
                                                      int ngroups, float* __restrict__ out2) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int g = lane / G, l = lane & (G - 1);
    int grp = (int)blockIdx.x;
    if (S == 2) {
        const int x = (int)(blockIdx.x & 7u);
        grp = (int)(blockIdx.x >> 3) * 4 + (x >> 1);
        if (grp >= ngroups) return;
        if (x & 1) {  // the upper half of every row of U and V
            U += Kd;
            V += Kd;
            out = out2;
        }
        scale = nullptr;
    }
    const int wave = uniform(grp * 4 + wid);
    const int64_t total = COO ? (int64_t)nnz : (int64_t)M + nnz;
    const int64_t d0l = (int64_t)wave * chunk;
    if (d0l >= total) return;
    const int d0 = (int)d0l;
    const int d1 = (int)min(d0l + chunk, total);
    int i0 = 0, j0 = d0, j1 = d1;
    if constexpr (!COO) {
        i0 = uniform(partition_at(part, wave, d0, M, nnz));  // merge_partition
        const int i1 = uniform(partition_at(part, wave + 1, d1, M, nnz));
        j0 = d0 - i0;
        j1 = d1 - i1;
    }
    const int nz = j1 - j0;
    const size_t wstride = ((size_t)chunk * 8 + (size_t)(chunk + 1) * 4 + 15) & ~(size_t)15;
    uint2* s_off = reinterpret_cast<uint2*>(smem + wid * wstride);
    int* s_re = reinterpret_cast<int*>(smem + wid * wstride + (size_t)chunk * 8);
    const uint32_t urow = (uint32_t)ldu * (uint32_t)sizeof(T), vrow = (uint32_t)ldv * (uint32_t)sizeof(T);
    int nre = 0;
    if constexpr (!COO) {
        nre = d1 - d0 - nz + 1;  // row ends of rows i0..i1
        for (int k = lane; k < nre; k += 64) s_re[k] = (i0 + k < M) ? rowidx[i0 + k + 1] : INT_MAX;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const int nzr = (nz + 63) & ~63;
    for (int k = lane; k < nzr; k += 64) {
        uint2 o = make_uint2(0u, 0u);
        if (k < nz) {
            const int e = j0 + k;
            int r;
            if constexpr (COO) {
                r = __builtin_nontemporal_load(rowidx + e);
            } else {  // rows i0 + (number of row ends ≤ e)
                int lo = 0, hi = nre - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_re[mid] <= e)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                r = i0 + lo;
            }
            o = make_uint2((uint32_t)r * urow, (uint32_t)__builtin_nontemporal_load(colind + e) * vrow);
        }
        s_off[k] = o;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const __amdgpu_buffer_rsrc_t ru = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(U), (short)0, (int)0xffffffff,
                                                                         0x00020000);
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(V), (short)0, (int)0xffffffff,
                                                                         0x00020000);
    const int steps = (Kd + G * W - 1) / (G * W);  // FULL: Kd = steps·G·W; else Kd % W == 0
    // dot products whose operands are fetched together: 16 floats of each operand per lane
    constexpr int QB = (16 / W) < G ? (16 / W) : G;

    for (int base = 0; base < nz; base += 64) {
        float p[G];
#pragma unroll
        for (int q = 0; q < G; ++q) p[q] = 0.0f;
        for (int t = 0; t < steps; ++t) {
            const int k = (t * G + l) * W;
            const bool in = FULL || k < Kd;  // lanes past Kd in the last step read column 0, use 0
            const uint32_t lo = in ? (uint32_t)(k * (int)sizeof(T)) : 0u;
#pragma unroll
            for (int qb = 0; qb < G; qb += QB) {
                float a[QB][W], b[QB][W];
#pragma unroll
                for (int q = 0; q < QB; ++q) {
                    const uint2 o = s_off[base + g * G + qb + q];
                    gather_vec<T, W>(ru, o.x + lo, a[q]);
                    gather_vec<T, W>(rv, o.y + lo, b[q]);
                    if (!FULL && !in) {
#pragma unroll
                        for (int i = 0; i < W; ++i) a[q][i] = 0.0f;
                    }
                }
#pragma unroll
                for (int q = 0; q < QB; ++q)
#pragma unroll
                    for (int i = 0; i < W; ++i) p[qb + q] = fmaf(a[q][i], b[q][i], p[qb + q]);
                // keep the next batch's loads below this point: bounded registers, more waves
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int m = G / 2; m >= 1; m >>= 1) {
            const bool upper = (l & m) != 0;
#pragma unroll
            for (int i = 0; i < m; ++i) {
                const float send = upper ? p[i] : p[i + m];
                const float keep = upper ? p[i + m] : p[i];
                p[i] = keep + __shfl_xor(send, m);
            }
        }
        const int e = base + lane;
        if (e < nz) {
            const int ge = j0 + e;
            out[ge] = scale ? scale[ge] * p[0] : p[0];
        }
    }
}

// out[e] = scale[e] · (out[e] + out2[e]): the two column halves of the split kernel.
__global__ __launch_bounds__(256) void k_sddmm_combine(int nnz, const float* __restrict__ scale,
                                                       float* __restrict__ out, const float* __restrict__ out2) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const float d = out[e] + out2[e];
    out[e] = scale ? scale[e] * d : d;
}

inline size_t sddmm_lds_bytes(int chunk) {
    return 4 * (((size_t)chunk * 8 + (size_t)(chunk + 1) * 4 + 15) & ~(size_t)15);
}

template <typename T, int G, int W>
void launch_sddmm_stream(bool coo, dim3 grid, hipStream_t s, int M, int Kd, int nnz, int chunk, const int* ri,
                         const int* ci, const float* sc, const T* U, int ldu, const T* V, int ldv, const int* part,
                         float* out, int S, int ngroups, float* out2) {
    const size_t lds = sddmm_lds_bytes(chunk);
    const bool full = Kd % (G * W) == 0;
#define ARMSP_SDDMM_ST(CO, FU)                                                                                 \
    hipLaunchKernelGGL((k_sddmm_stream<T, G, W, CO, FU>), grid, dim3(256), lds, s, M, Kd, nnz, chunk, ri, ci, sc, U, \
                       ldu, V, ldv, part, out, S, ngroups, out2)
    if (coo) {
        if (full) ARMSP_SDDMM_ST(true, true); else ARMSP_SDDMM_ST(true, false);
    } else {
        if (full) ARMSP_SDDMM_ST(false, true); else ARMSP_SDDMM_ST(false, false);
    }
#undef ARMSP_SDDMM_ST
}

template <typename T, int G, int W>
void launch_sddmm(bool coo, dim3 grid, hipStream_t s, int M, int Kd, int nnz, int chunk, const int* ri, const int* ci,
                  const float* sc, const T* U, int ldu, const T* V, int ldv, float* out) {
    if (coo)
        hipLaunchKernelGGL((k_sddmm<T, G, W, true>), grid, dim3(256), 0, s, M, Kd, nnz, chunk, ri, ci, sc, U, ldu, V,
                           ldv, out);
    else
        hipLaunchKernelGGL((k_sddmm<T, G, W, false>), grid, dim3(256), 0, s, M, Kd, nnz, chunk, ri, ci, sc, U, ldu, V,
                           ldv, out);
}

inline bool aligned(const void* p, int bytes) { return ((uintptr_t)p % (uintptr_t)bytes) == 0; }

template <typename T>
int sddmm_impl(bool coo, int M, int K, int Kd, int nnz, const int* rowidx, const int* colind, const float* scale,
               const T* U, int ldu, const T* V, int ldv, float* out, const armsp_config* cfg, void* stream) {
    if (M < 0 || K < 0 || Kd < 0 || nnz < 0) return ARMSP_ERROR_INVALID_VALUE;
    if (nnz == 0) return ARMSP_SUCCESS;
    if (!rowidx || !colind || !out || Kd == 0 || !U || !V || ldu < Kd || ldv < Kd) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    armsp_config c = cfg ? *cfg : armsp_config{0, 0, 0, 0, 0};
    if (!coo && c.schedule == 0 && c.group == 0 && c.vec == 0 && c.chunk == 0 && c.unroll == 0) {
        armsp_config sel;  // left to the library: the installed schedule selector, if any
        if (selector_config(1, M, nnz, Kd, (int)sizeof(T), rowidx, s, true, &sel)) c = sel;
    }
    if (c.schedule != ARMSP_SCHED_AUTO && c.schedule != ARMSP_SCHED_NNZ && c.schedule != ARMSP_SCHED_MERGE)
        return ARMSP_ERROR_NOT_SUPPORTED;
    // vector width: 16-byte loads when every row start is aligned
    constexpr int WMAX = sizeof(T) == 4 ? 4 : 8;
    int W = c.vec > 0 ? c.vec : WMAX;
    if (W != 1 && W != WMAX) W = 1;
    if (W > 1 && (ldu % W || ldv % W || !aligned(U, 16) || !aligned(V, 16))) W = 1;
    // default chunk: 128 items per wave (the best single configuration over the
    // selector corpus; within 2% of the best on the 1M R-MAT at Kd = 32 and 128)
    const int chunk = c.chunk > 0 ? c.chunk : 128;
    if (chunk % 64) return ARMSP_ERROR_INVALID_VALUE;
    // staged kernel: 32-bit row offsets into U and V, at most 2048 items per wave in
    // LDS, Kd a multiple of the vector width (schedule NNZ asks for the unstaged one)
    const bool staged = c.schedule != ARMSP_SCHED_NNZ && chunk <= 2048 && Kd % W == 0 &&
                        (uint64_t)M * (uint64_t)ldu * sizeof(T) <= 0xffffffffull &&
                        (uint64_t)K * (uint64_t)ldv * sizeof(T) <= 0xffffffffull;
    int G = c.group;
    if (G <= 0) {
        // lanes per dot product: the staged kernel keeps G partials per lane, so it
        // stops at 8 and loops over Kd instead (bf16, 1M R-MAT, Kd = 128: 8 lanes
        // 578 us, 4 lanes 926 us, 16 lanes 618 us)
        const int gmax = staged ? 8 : 16;
        G = 4;
        while (G < gmax && G * W < Kd) G <<= 1;
    }
    if (G != 4 && G != 8 && G != 16 && G != 32) return ARMSP_ERROR_INVALID_VALUE;
    const int64_t items = (staged && !coo) ? (int64_t)M + nnz : (int64_t)nnz;
    const int64_t nwaves = (items + chunk - 1) / chunk;
    if (nwaves > INT_MAX / 4) return ARMSP_ERROR_NOT_SUPPORTED;
    // unroll = 2: split the dot products into column halves spread over the XCDs
    // (each XCD's L2 then holds half of every hot V row), plus one pass that adds
    // the halves — a win on skewed columns whose V rows outgrow the L2 (the 1M
    // R-MAT at Kd = 128: 13%), a loss where columns are uniform or local, so it is
    // the selector's choice, not AUTO's.  Where it cannot apply (COO, the unstaged
    // kernel, odd halves) the call runs unsplit.
    if (c.unroll < 0 || c.unroll > 2) return ARMSP_ERROR_INVALID_VALUE;
    const int KdS = Kd / 2;
    const bool split = c.unroll == 2 && staged && !coo && KdS > 0 && Kd % 2 == 0 && KdS % W == 0 &&
                       aligned(U + KdS, W * (int)sizeof(T)) && aligned(V + KdS, W * (int)sizeof(T));
    if (split && c.group <= 0) {
        G = 4;
        while (G < 8 && G * W < KdS) G <<= 1;
    }
    const int ngroups = (int)((nwaves + 3) / 4);
    const dim3 grid(split ? (unsigned)(((int64_t)ngroups + 3) / 4 * 8) : (unsigned)ngroups);
    // the staged CSR kernel starts each wave from a precomputed merge-path partition
    void* ws = nullptr;
    int* part = nullptr;
    float* out2 = nullptr;
    if (staged && !coo) {
        const size_t pb = ((size_t)(nwaves + 1) * sizeof(int) + 255) & ~(size_t)255;
        int st = ws_alloc(&ws, pb + (split ? (size_t)nnz * sizeof(float) : 0), s);
        if (st) return st;
        part = static_cast<int*>(ws);
        if (split) out2 = reinterpret_cast<float*>(static_cast<char*>(ws) + pb);
        TimedLaunch tp("merge_partition", s);
        st = merge_partition(M, nnz, chunk, nwaves, rowidx, part, s);
        if (st) {
            ws_free(ws, s);
            return st;
        }
        tp.stop();
    }
    TimedLaunch t("sddmm", s);
#define ARMSP_SDDMM_G(WW)                                                                                    \
    if (staged) {                                                                                            \
        switch (G) {                                                                                         \
            case 4: launch_sddmm_stream<T, 4, WW>(coo, grid, s, M, split ? KdS : Kd, nnz, chunk, rowidx, colind, scale, U, ldu, V, ldv, part, out, split ? 2 : 1, ngroups, out2); break; \
            case 8: launch_sddmm_stream<T, 8, WW>(coo, grid, s, M, split ? KdS : Kd, nnz, chunk, rowidx, colind, scale, U, ldu, V, ldv, part, out, split ? 2 : 1, ngroups, out2); break; \
            case 16: launch_sddmm_stream<T, 16, WW>(coo, grid, s, M, split ? KdS : Kd, nnz, chunk, rowidx, colind, scale, U, ldu, V, ldv, part, out, split ? 2 : 1, ngroups, out2); break; \
            default: launch_sddmm_stream<T, 32, WW>(coo, grid, s, M, split ? KdS : Kd, nnz, chunk, rowidx, colind, scale, U, ldu, V, ldv, part, out, split ? 2 : 1, ngroups, out2); break; \
        }                                                                                                    \
    } else {                                                                                                 \
        switch (G) {                                                                                         \
            case 4: launch_sddmm<T, 4, WW>(coo, grid, s, M, Kd, nnz, chunk, rowidx, colind, scale, U, ldu, V, ldv, out); break; \
            case 8: launch_sddmm<T, 8, WW>(coo, grid, s, M, Kd, nnz, chunk, rowidx, colind, scale, U, ldu, V, ldv, out); break; \
            case 16: launch_sddmm<T, 16, WW>(coo, grid, s, M, Kd, nnz, chunk, rowidx, colind, scale, U, ldu, V, ldv, out); break; \
            default: launch_sddmm<T, 32, WW>(coo, grid, s, M, Kd, nnz, chunk, rowidx, colind, scale, U, ldu, V, ldv, out); break; \
        }                                                                                                    \
    }
    if (W == 1) {
        ARMSP_SDDMM_G(1)
    } else {
        ARMSP_SDDMM_G(WMAX)
    }
#undef ARMSP_SDDMM_G
    if (split)
        hipLaunchKernelGGL(k_sddmm_combine, dim3((unsigned)(((int64_t)nnz + 255) / 256)), dim3(256), 0, s, nnz,
                           scale, out, out2);
    const hipError_t e = hipGetLastError();
    if (ws) ws_free(ws, s);
    if (e != hipSuccess) return hip_fail(e);
    t.stop();
    return ARMSP_SUCCESS;
}

}  // namespace
}  // namespace armsp

extern "C" {

int armsp_sddmm_csr_f32(int M, int K, int Kd, int nnz, const int* rowptr, const int* colind, const float* scale,
                        const float* U, int ldu, const float* V, int ldv, float* out, const armsp_config* cfg,
                        void* stream) {
    return armsp::sddmm_impl<float>(false, M, K, Kd, nnz, rowptr, colind, scale, U, ldu, V, ldv, out, cfg, stream);
}

int armsp_sddmm_coo_f32(int M, int K, int Kd, int nnz, const int* rowind, const int* colind, const float* scale,
                        const float* U, int ldu, const float* V, int ldv, float* out, void* stream) {
    return armsp::sddmm_impl<float>(true, M, K, Kd, nnz, rowind, colind, scale, U, ldu, V, ldv, out, nullptr, stream);
}

int armsp_sddmm_csr_bf16(int M, int K, int Kd, int nnz, const int* rowptr, const int* colind, const float* scale,
                         const uint16_t* U, int ldu, const uint16_t* V, int ldv, float* out, const armsp_config* cfg,
                         void* stream) {
    return armsp::sddmm_impl<uint16_t>(false, M, K, Kd, nnz, rowptr, colind, scale, U, ldu, V, ldv, out, cfg, stream);
}

}  // extern "C"
