// CSR SpMM for gfx950: C (=|+=) A·B with B in fp32, bf16 or fp8-e4m3.
//
// The reference kernel family (SDMM.cpp:21-180, 422-448;
// gcn/pytorch-custom/spmm.cpp:49-93) runs an OpenMP loop over rows and, per
// nonzero, an axpy of a B row into the C row (num_vector_neon, SDMM.cpp:449).
// Here the C row lives in registers instead: lanes cover the columns, and the
// nonzeros are streamed in CSR order, so the per-element fmaf sequence is the
// reference's one and the C row is written once.
//
// Schedules:
//  ROW   — a group of G lanes owns a row (G = 64 for wide N, down to 4 for
//          narrow N with 64/G rows per wave).  Bit-exact with the CPU order.
//  NNZ   — each wave owns a fixed range of nonzeros.
//  MERGE — each wave owns a fixed range of the merge path over
//          (row ends, nonzeros), so empty rows and power-law hub rows cost
//          the same per item (Merrill & Garland's merge-based CSR).
//  NNZ and MERGE keep the partial sum of a row cut by a wave boundary in a
//  carry buffer; a second kernel adds each run of carries into its row in
//  wave order, so results are deterministic run to run.
#include <climits>
#include <mutex>

#include "common.hpp"

namespace armsp {
namespace {

// ----------------------------------------------------------- dense loads
template <typename TB, int V>
__device__ __forceinline__ void load_b(const TB* __restrict__ p, float (&o)[V]) {
    if constexpr (sizeof(TB) == 4) {
        if constexpr (V == 1) {
            o[0] = p[0];
        } else if constexpr (V == 2) {
            float2 t = *reinterpret_cast<const float2*>(p);
            o[0] = t.x; o[1] = t.y;
        } else if constexpr (V == 4) {
            float4 t = *reinterpret_cast<const float4*>(p);
            o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
        }
    } else if constexpr (sizeof(TB) == 2) {  // bf16 bits
        if constexpr (V == 1) {
            o[0] = bf16_to_f32(p[0]);
        } else if constexpr (V == 2) {
            uint32_t t = *reinterpret_cast<const uint32_t*>(p);
            o[0] = __uint_as_float(t << 16); o[1] = __uint_as_float(t & 0xffff0000u);
        } else if constexpr (V == 4) {
            uint2 t = *reinterpret_cast<const uint2*>(p);
            o[0] = __uint_as_float(t.x << 16); o[1] = __uint_as_float(t.x & 0xffff0000u);
            o[2] = __uint_as_float(t.y << 16); o[3] = __uint_as_float(t.y & 0xffff0000u);
        } else if constexpr (V == 8) {
            uint4 t = *reinterpret_cast<const uint4*>(p);
            uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                o[2 * i] = __uint_as_float(w[i] << 16);
                o[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
            }
        }
    } else {  // fp8-e4m3 bytes
        uint32_t w[2] = {0u, 0u};
        if constexpr (V == 1) w[0] = p[0];
        else if constexpr (V == 2) w[0] = *reinterpret_cast<const uint16_t*>(p);
        else if constexpr (V == 4) w[0] = *reinterpret_cast<const uint32_t*>(p);
        else if constexpr (V == 8) {
            uint2 t = *reinterpret_cast<const uint2*>(p);
            w[0] = t.x; w[1] = t.y;
        }
#pragma unroll
        for (int i = 0; i < V; ++i) o[i] = e4m3_to_f32((w[i >> 2] >> (8 * (i & 3))) & 0xffu);
    }
}

template <int V>
__device__ __forceinline__ void load_c(const float* __restrict__ p, float (&o)[V]) {
    load_b<float, V>(p, o);
}

template <int V>
__device__ __forceinline__ void store_c(float* __restrict__ p, const float (&o)[V]) {
    if constexpr (V == 1) {
        __builtin_nontemporal_store(o[0], p);
    } else if constexpr (V == 2) {
        __builtin_nontemporal_store(o[0], p);
        __builtin_nontemporal_store(o[1], p + 1);
    } else {
#pragma unroll
        for (int i = 0; i < V; ++i) __builtin_nontemporal_store(o[i], p + i);
    }
}

// Optional operand scales and output epilogue (armsp_epilogue + the fp8
// per-row scale of B).  coefficient of nonzero (r, c): val · bscale[c] ·
// cscale[c]; output row r: rscale[r] · (A·B)[r] + bias, then ReLU.
struct Epi {
    const float* bscale;  // fp8 B row scales
    const float* cscale;  // column scale of A
    const float* rscale;  // output row scale
    const float* bias;    // per-column bias
    int relu;
};

__device__ __forceinline__ float coef(float a, const Epi& ep, int c) {
    if (ep.bscale) a *= ep.bscale[c];
    if (ep.cscale) a *= ep.cscale[c];
    return a;
}

// out = rscale·acc + bias (+ C_old) (ReLU when `final`)
template <int V>
__device__ __forceinline__ void finish(float (&o)[V], const float (&acc)[V], const Epi& ep, int r, int col,
                                       const float (&cold)[V], bool has_cold, bool final) {
#pragma unroll
    for (int v = 0; v < V; ++v) o[v] = acc[v];
    if (ep.rscale) {
        const float rs = ep.rscale[r];
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] *= rs;
    }
    if (ep.bias) {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] += ep.bias[col + v];
    }
    if (has_cold) {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] += cold[v];
    }
    if (final && ep.relu) {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = fmaxf(o[v], 0.0f);
    }
}

// LDS of one wave of the staged merge kernels: per nonzero (row byte offset,
// coefficient), per row its end relative to the chunk, and with a row-scale
// epilogue per row its scale (so writing a row needs no memory load).
__host__ __device__ inline size_t merge_wave_lds(int chunk, bool row_scales) {
    return ((size_t)chunk * 8 + (size_t)(chunk + 1) * 4 * (row_scales ? 2 : 1) + 15) & ~(size_t)15;
}

// Writes one finished row of a merge kernel: acc·rs + bias (+ C when
// accumulating), ReLU when the row is final.  `rs` is the row's scale (staged
// in LDS), `bc` the lane's bias columns, loaded once per wave.  The C load and its use sit in one branch:
// were the loaded value merged with a default after the branch, the wait for
// it would run on every row and, the memory counter being in order, drain the
// B gathers already in flight for the next rows.
template <int V>
__device__ __forceinline__ void write_row(float* cp, const float (&acc)[V], const Epi& ep, float rs,
                                          const float (&bc)[V], bool acc_mode, bool final) {
    float o[V];
#pragma unroll
    for (int v = 0; v < V; ++v) o[v] = acc[v];
    if (ep.rscale) {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] *= rs;
    }
    if (ep.bias) {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] += bc[v];
    }
    const bool relu = final && ep.relu;
    if (acc_mode) {
        float cold[V];
        load_c<V>(cp, cold);
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] += cold[v];
        if (relu) {
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = fmaxf(o[v], 0.0f);
        }
        store_c<V>(cp, o);
    } else {
        if (relu) {
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = fmaxf(o[v], 0.0f);
        }
        store_c<V>(cp, o);
    }
}

// The lane's bias columns, loaded once per wave.  The empty asm statements
// consume them, so their load completes here, before the gather loop: a load
// still pending at the loop entry would otherwise make the compiler wait for
// it (and, the memory counter being in order, for every gather) at each row.
template <int V>
__device__ __forceinline__ void load_bias(float (&bc)[V], const Epi& ep, int col, bool active) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
        bc[v] = (ep.bias && active) ? ep.bias[col + v] : 0.0f;
        asm volatile("" ::"v"(bc[v]));
    }
}

// ------------------------------------------------------------- ROW split
template <typename TB, int V, int G>
__global__ __launch_bounds__(256) void k_spmm_row(int M, int N, const int* __restrict__ rowptr,
                                                  const int* __restrict__ colind, const float* __restrict__ val,
                                                  const TB* __restrict__ B, int ldb, Epi ep,
                                                  float* __restrict__ C, int ldc, int flags) {
    constexpr int U = G < 4 ? G : 4;
    const int lane = threadIdx.x & (G - 1);
    int row = (int)((blockIdx.x * 256u + threadIdx.x) / G);
    if constexpr (G == 64) row = uniform(row);
    if (row >= M) return;
    const int col = (blockIdx.y * G + lane) * V;
    const bool active = col < N;
    const int beg = rowptr[row];
    int end = rowptr[row + 1];
    if (flags & ARMSP_TRUNC4) end = beg + ((end - beg) & ~3);
    // Without an epilogue the accumulation starts from C (the reference's
    // fmaf sequence, bit-exact); with one it starts from 0 and C is added last.
    const bool plain = !ep.rscale && !ep.bias && !ep.relu;
    float acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;
    if (plain && (flags & ARMSP_ACCUMULATE) && active) load_c<V>(C + (size_t)row * ldc + col, acc);

    for (int base = beg; base < end; base += G) {
        const int n = min(G, end - base);
        int mc = 0;
        float mv = 0.0f;
        if (lane < n) {
            mc = colind[base + lane];
            mv = coef(val ? val[base + lane] : 1.0f, ep, mc);
        }
        for (int j = 0; j < n; j += U) {
            float b[U][V];
            float a[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int c;
                if constexpr (G == 64) {
                    c = read_lane(mc, j + u);
                    a[u] = read_lane(mv, j + u);
                } else {
                    c = __shfl(mc, j + u, G);
                    a[u] = __shfl(mv, j + u, G);
                }
#pragma unroll
                for (int v = 0; v < V; ++v) b[u][v] = 0.0f;
                if (j + u < n && active) load_b<TB, V>(B + (size_t)c * ldb + col, b[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < n) {
#pragma unroll
                    for (int v = 0; v < V; ++v) acc[v] = fmaf(a[u], b[u][v], acc[v]);
                }
        }
    }
    if (!active) return;
    float* cp = C + (size_t)row * ldc + col;
    if (plain) {
        store_c<V>(cp, acc);
    } else {
        float o[V], cold[V];
        const bool accm = (flags & ARMSP_ACCUMULATE) != 0;
        if (accm) load_c<V>(cp, cold);
        finish<V>(o, acc, ep, row, col, cold, accm, true);
        store_c<V>(cp, o);
    }
}

// ----------------------------------------------------- NNZ / MERGE split
// One wave per `chunk` items.  Lanes cover 64·V columns of the tile
// blockIdx.y.  colind/val of the wave's range are read 64 at a time into a
// register window and broadcast with readlane; U dense rows are fetched
// before any of them is consumed so U gathers are in flight per wave.
template <typename TB, int V, int U, bool MERGE>
__global__ __launch_bounds__(256) void k_spmm_split(int M, int N, int nnz, int chunk, const int* __restrict__ rowptr,
                                                    const int* __restrict__ colind, const float* __restrict__ val,
                                                    const TB* __restrict__ B, int ldb, Epi ep,
                                                    float* __restrict__ C, int ldc, int flags,
                                                    int* __restrict__ carry_row, float* __restrict__ carry_val) {
    const int lane = threadIdx.x & 63;
    const int wave = uniform((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    const int64_t total = MERGE ? (int64_t)M + nnz : (int64_t)nnz;
    const int64_t d0l = (int64_t)wave * chunk;
    if (d0l >= total) return;
    const int d0 = (int)d0l;
    const int d1 = (int)min(d0l + chunk, total);
    int i0, j0, i1, j1;
    if constexpr (MERGE) {
        i0 = uniform(merge_path_search_wave(d0, rowptr, M, nnz));
        i1 = uniform(merge_path_search_wave(d1, rowptr, M, nnz));
        j0 = d0 - i0;
        j1 = d1 - i1;
    } else {
        j0 = d0;
        j1 = d1;
        i0 = d0 == 0 ? 0 : uniform(rows_ending_by_wave(j0, rowptr, M));
        i1 = uniform(rows_ending_by_wave(j1, rowptr, M));
    }
    const bool acc_mode = (flags & ARMSP_ACCUMULATE) != 0;
    const int col = (int)(blockIdx.y * 64u + lane) * V;
    const bool active = col < N;

    // Register windows, double-buffered: nonzeros [jw, jw+64) and the next 64,
    // row ends of rows [rw, rw+64) and the next 64.  The next window is
    // requested as soon as the current one is taken, 64 items ahead of use.
    // The CSR streams are read once: non-temporal loads keep them from
    // evicting hot dense rows from the L2.
    auto load_nz = [&](int base, int& c, float& a) {
        c = 0;
        a = 0.0f;
        if (base + lane < j1) {
            c = __builtin_nontemporal_load(colind + base + lane);
            a = coef(val ? __builtin_nontemporal_load(val + base + lane) : 1.0f, ep, c);
        }
    };
    auto load_re = [&](int base) {
        return (base + lane < M) ? __builtin_nontemporal_load(rowptr + base + 1 + lane) : INT_MAX;
    };
    int jw = j0, wc, nc;
    float wv, nv;
    load_nz(jw, wc, wv);
    load_nz(jw + 64, nc, nv);
    int r = i0, rw = i0;
    int wre = load_re(rw), nre = load_re(rw + 64);
    int re = read_lane(wre, 0);

    float acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;

    // a row whose first nonzeros belong to an earlier wave gets its carries
    // (and its ReLU) from the fixup kernel
    const int carry_in_row = (i0 < M && rowptr[i0] < j0) ? i0 : -1;
    auto emit = [&]() {
        if (active) {
            float* cp = C + (size_t)r * ldc + col;
            float o[V], cold[V];
            if (acc_mode) load_c<V>(cp, cold);
            finish<V>(o, acc, ep, r, col, cold, acc_mode, r != carry_in_row);
            store_c<V>(cp, o);
        }
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = 0.0f;
        ++r;
        if (r - rw == 64) {
            rw = r;
            wre = nre;
            nre = load_re(rw + 64);
        }
        re = read_lane(wre, r - rw);
    };

    for (int jb = j0; jb < j1; jb += U) {
        if (jb - jw == 64) {
            jw = jb;
            wc = nc;
            wv = nv;
            load_nz(jw + 64, nc, nv);
        }
        float b[U][V];
        float a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = read_lane(wc, jb - jw + u);
            a[u] = read_lane(wv, jb - jw + u);
#pragma unroll
            for (int v = 0; v < V; ++v) b[u][v] = 0.0f;
            if (jb + u < j1 && active) load_b<TB, V>(B + (size_t)c * ldb + col, b[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int jj = jb + u;
            if (jj < j1) {
                while (r < i1 && re <= jj) emit();
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(a[u], b[u][v], acc[v]);
            }
        }
    }
    while (r < i1) emit();

    // partial sum of row i1 (nonzeros [max(j0, rowptr[i1]), j1)) goes to the carry
    bool has = false;
    if (i1 < M) {
        const int rs = rowptr[i1];
        has = j1 > (rs > j0 ? rs : j0);
    }
    if (blockIdx.y == 0 && lane == 0) carry_row[wave] = has ? i1 : -1;
    if (has && active) {
        float* cp = carry_val + (size_t)wave * N + col;
#pragma unroll
        for (int v = 0; v < V; ++v) cp[v] = acc[v];
    }
}

// ------------------------------------------ MERGE split, 32-bit offsets
// The merge-path split of k_spmm_split with a leaner per-nonzero stream, used
// whenever every dense-row byte offset fits in 32 bits:
//  - the wave first stages its chunk in LDS: per nonzero the dense-row BYTE
//    offset and the coefficient, per row its end; the nonzero loop then reads
//    them with wave-uniform LDS loads (a broadcast, no bank conflicts);
//  - a gather is one buffer load, address = row offset + lane column offset;
//  - lanes past N load column 0 of the tile (always in bounds) and never
//    store, so no load is predicated;
//  - U gathers form a batch and one scalar compare per batch decides whether a
//    row ends inside it; only such batches walk nonzero by nonzero;
//  - two batches are in flight: batch k+1 is issued before batch k is used.
template <typename TB, int V>
__device__ __forceinline__ void gather_b(__amdgpu_buffer_rsrc_t rs, uint32_t voff, float (&o)[V]) {
    constexpr int BYTES = V * (int)sizeof(TB);
    constexpr int W = BYTES >= 4 ? BYTES / 4 : 1;
    uint32_t w[W];
    if constexpr (BYTES == 16) {
        auto t = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 0, 0);
        w[0] = t[0]; w[1] = t[1]; w[2] = t[2]; w[3] = t[3];
    } else if constexpr (BYTES == 8) {
        auto t = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, 0, 0);
        w[0] = t[0]; w[1] = t[1];
    } else if constexpr (BYTES == 4) {
        w[0] = __builtin_amdgcn_raw_buffer_load_b32(rs, voff, 0, 0);
    } else if constexpr (BYTES == 2) {
        w[0] = __builtin_amdgcn_raw_buffer_load_b16(rs, voff, 0, 0);
    } else {
        w[0] = __builtin_amdgcn_raw_buffer_load_b8(rs, voff, 0, 0);
    }
    if constexpr (sizeof(TB) == 4) {
#pragma unroll
        for (int i = 0; i < V; ++i) o[i] = __uint_as_float(w[i]);
    } else if constexpr (sizeof(TB) == 2) {
#pragma unroll
        for (int i = 0; i < V; ++i) o[i] = __uint_as_float((i & 1) ? (w[i >> 1] & 0xffff0000u) : (w[i >> 1] << 16));
    } else {  // packed fp8 conversion, two codes per instruction
#pragma unroll
        for (int i = 0; i < (V + 3) / 4; ++i) {
            float f[4];
            e4m3x4_to_f32(w[i], f);
#pragma unroll
            for (int q = 0; q < 4 && 4 * i + q < V; ++q) o[4 * i + q] = f[q];
        }
    }
}

template <typename TB, int V, int U>
__global__ __launch_bounds__(256) void k_spmm_stream(int M, int N, int nnz, int chunk,
                                                     const int* __restrict__ rowptr, const int* __restrict__ colind,
                                                     const float* __restrict__ val, const TB* __restrict__ B,
                                                     int ldb, Epi ep, float* __restrict__ C, int ldc, int flags,
                                                     const int* __restrict__ part, int* __restrict__ carry_row,
                                                     float* __restrict__ carry_val) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int wave = uniform((int)(blockIdx.x * 4u + wid));
    const int64_t total = (int64_t)M + nnz;
    const int64_t d0l = (int64_t)wave * chunk;
    if (d0l >= total) return;
    const int d0 = (int)d0l;
    const int d1 = (int)min(d0l + chunk, total);
    const int i0 = uniform(partition_at(part, wave, d0, M, nnz));  // merge_partition
    const int i1 = uniform(partition_at(part, wave + 1, d1, M, nnz));
    const int j0 = d0 - i0, j1 = d1 - i1;
    const int nz = j1 - j0;      // ≤ chunk
    const int nr = i1 - i0 + 1;  // rows i0..i1, ≤ chunk + 1

    // stage the chunk: (row byte offset, coefficient) per nonzero, end per row
    const uint32_t rowbytes = (uint32_t)ldb * (uint32_t)sizeof(TB);
    const bool has_rs = ep.rscale != nullptr;
    const size_t wstride = merge_wave_lds(chunk, has_rs);
    uint2* s_nz = reinterpret_cast<uint2*>(smem + wid * wstride);
    int* s_re = reinterpret_cast<int*>(smem + wid * wstride + (size_t)chunk * 8);
    float* s_rs = reinterpret_cast<float*>(smem + wid * wstride + (size_t)chunk * 8 + (size_t)(chunk + 1) * 4);
    for (int k = lane; k < nz; k += 64) {
        const int c = __builtin_nontemporal_load(colind + j0 + k);
        const float a = coef(val ? __builtin_nontemporal_load(val + j0 + k) : 1.0f, ep, c);
        s_nz[k] = make_uint2((uint32_t)c * rowbytes, __float_as_uint(a));
    }
    for (int k = lane; k < nr; k += 64) {
        s_re[k] = (i0 + k < M) ? rowptr[i0 + k + 1] - j0 : INT_MAX;
        if (has_rs) s_rs[k] = (i0 + k < M) ? ep.rscale[i0 + k] : 1.0f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const bool acc_mode = (flags & ARMSP_ACCUMULATE) != 0;
    const int tile0 = (int)blockIdx.y * 64 * V;
    const int col = tile0 + lane * V;
    const bool active = col < N;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<TB*>(B + tile0), (short)0, (int)0xffffffff, 0x00020000);
    const uint32_t voff = active ? (uint32_t)(lane * V * (int)sizeof(TB)) : 0u;

    float acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;
    // rows are emitted in order; re = end of row i0 + rr, relative to j0
    int rr = 0;
    int re = uniform(s_re[0]);
    const int carry_in_row = (i0 < M && rowptr[i0] < j0) ? i0 : -1;
    float bc[V];
    load_bias<V>(bc, ep, col, active);
    auto emit = [&]() {
        const int r = i0 + rr;
        if (active)
            write_row<V>(C + (size_t)r * ldc + col, acc, ep, has_rs ? s_rs[rr] : 1.0f, bc, acc_mode,
                         r != carry_in_row);
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = 0.0f;
        ++rr;
        re = rr < nr ? uniform(s_re[rr]) : INT_MAX;  // never past the staged rows
    };
    auto issue = [&](int k, float (&b)[U][V], float (&a)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint2 e = s_nz[k + u];
            a[u] = __uint_as_float(e.y);
            gather_b<TB, V>(rs, voff + e.x, b[u]);
        }
    };
    auto consume = [&](int k, const float (&b)[U][V], const float (&a)[U]) {
        if (re >= k + U) {
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(a[u], b[u][v], acc[v]);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                while (re <= k + u) emit();
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(a[u], b[u][v], acc[v]);
            }
        }
    };

    const int kend = (nz / U) * U;  // end of the full batches
    float bA[U][V], bB[U][V];
    float aA[U], aB[U];
    int k = 0;
    if (k < kend) issue(k, bA, aA);
    while (k < kend) {
        if (k + U < kend) issue(k + U, bB, aB);
        consume(k, bA, aA);
        k += U;
        if (k >= kend) break;
        if (k + U < kend) issue(k + U, bA, aA);
        consume(k, bB, aB);
        k += U;
    }
#pragma unroll
    for (int u = 0; u < U - 1; ++u) {  // fewer than U nonzeros left
        if (kend + u < nz) {
            const uint2 e = s_nz[kend + u];
            float b[V];
            gather_b<TB, V>(rs, voff + e.x, b);
            while (re <= kend + u) emit();
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] = fmaf(__uint_as_float(e.y), b[v], acc[v]);
        }
    }
    while (rr < nr - 1) emit();

    // partial sum of row i1 (nonzeros [max(j0, rowptr[i1]), j1)) goes to the carry
    bool has = false;
    if (i1 < M) {
        const int rs0 = rowptr[i1];
        has = j1 > (rs0 > j0 ? rs0 : j0);
    }
    if (blockIdx.y == 0 && lane == 0) carry_row[wave] = has ? i1 : -1;
    if (has && active) {
        float* cp = carry_val + (size_t)wave * N + col;
#pragma unroll
        for (int v = 0; v < V; ++v) cp[v] = acc[v];
    }
}

// ------------------------------------- MERGE split over 128-byte column slices
// The stream kernel with the dense columns cut into slices of 128 bytes (one
// cache line of a B row: 32 fp32, 64 bf16 or 128 fp8 columns).  Half a wave
// covers a slice (4 bytes per lane), so each load instruction gathers two
// nonzeros, and the two halves' partial sums of a row meet at the row end
// through one cross-half shuffle.  Blocks are numbered so that, with the
// round-robin assignment of workgroups to the 8 XCDs, every XCD serves one
// slice (when the slice count divides 8): each XCD's L2 then holds one line
// per B row instead of all of them, i.e. several times more distinct rows of
// the hot set.  The numbering only affects speed, never results.
template <typename TB, int V, int U>
__global__ __launch_bounds__(256) void k_spmm_slice(int M, int N, int nnz, int chunk, int S, int G,
                                                    const int* __restrict__ rowptr, const int* __restrict__ colind,
                                                    const float* __restrict__ val, const TB* __restrict__ B, int ldb,
                                                    Epi ep, float* __restrict__ C, int ldc, int flags,
                                                    const int* __restrict__ part, int* __restrict__ carry_row,
                                                    float* __restrict__ carry_val) {
    // V columns per lane (4, 8 or 16 bytes); a slice is the 32·V columns of a half wave
    constexpr int SW = 32 * V;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int h = lane >> 5, l = lane & 31;
    int g, sl;
    if (8 % S == 0) {  // XCD x (= block % 8) serves slice x % S
        const int x = (int)(blockIdx.x & 7u);
        sl = x % S;
        g = (int)(blockIdx.x >> 3) * (8 / S) + x / S;
    } else {
        sl = (int)(blockIdx.x % (unsigned)S);
        g = (int)(blockIdx.x / (unsigned)S);
    }
    if (g >= G) return;
    const int wave = uniform(g * 4 + wid);  // keeps every chunk bound and loop test scalar
    const int64_t total = (int64_t)M + nnz;
    const int64_t d0l = (int64_t)wave * chunk;
    if (d0l >= total) return;
    const int d0 = (int)d0l;
    const int d1 = (int)min(d0l + chunk, total);
    const int i0 = uniform(partition_at(part, wave, d0, M, nnz));  // merge_partition
    const int i1 = uniform(partition_at(part, wave + 1, d1, M, nnz));
    const int j0 = d0 - i0, j1 = d1 - i1;
    const int nz = j1 - j0;
    const int nr = i1 - i0 + 1;

    const uint32_t rowbytes = (uint32_t)ldb * (uint32_t)sizeof(TB);
    const bool has_rs = ep.rscale != nullptr;
    const size_t wstride = merge_wave_lds(chunk, has_rs);
    uint2* s_nz = reinterpret_cast<uint2*>(smem + wid * wstride);
    int* s_re = reinterpret_cast<int*>(smem + wid * wstride + (size_t)chunk * 8);
    float* s_rs = reinterpret_cast<float*>(smem + wid * wstride + (size_t)chunk * 8 + (size_t)(chunk + 1) * 4);
    for (int k = lane; k < nz; k += 64) {
        const int c = __builtin_nontemporal_load(colind + j0 + k);
        const float a = coef(val ? __builtin_nontemporal_load(val + j0 + k) : 1.0f, ep, c);
        s_nz[k] = make_uint2((uint32_t)c * rowbytes, __float_as_uint(a));
    }
    for (int k = lane; k < nr; k += 64) {
        s_re[k] = (i0 + k < M) ? rowptr[i0 + k + 1] - j0 : INT_MAX;
        if (has_rs) s_rs[k] = (i0 + k < M) ? ep.rscale[i0 + k] : 1.0f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const bool acc_mode = (flags & ARMSP_ACCUMULATE) != 0;
    const int col = sl * SW + l * V;
    const bool active = col < N;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<TB*>(B + sl * SW), (short)0, (int)0xffffffff, 0x00020000);
    const uint32_t voff = active ? (uint32_t)(l * V * (int)sizeof(TB)) : 0u;

    float acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;
    int rr = 0;
    int re = uniform(s_re[0]);
    const int carry_in_row = (i0 < M && rowptr[i0] < j0) ? i0 : -1;
    float bc[V];
    load_bias<V>(bc, ep, col, active);
    auto emit = [&]() {  // called by the whole wave
        const int r = i0 + rr;
        float sum[V];
#pragma unroll
        for (int v = 0; v < V; ++v) sum[v] = half_sum(acc[v]);
        if (h == 0 && active)
            write_row<V>(C + (size_t)r * ldc + col, sum, ep, has_rs ? s_rs[rr] : 1.0f, bc, acc_mode,
                         r != carry_in_row);
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = 0.0f;
        ++rr;
        re = rr < nr ? uniform(s_re[rr]) : INT_MAX;  // never past the staged rows
    };
    auto issue = [&](int kb, float (&b)[U][V], float (&a)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint2 e = s_nz[kb + 2 * u + h];
            a[u] = __uint_as_float(e.y);
            gather_b<TB, V>(rs, voff + e.x, b[u]);
        }
    };
    // nonzeros k2 (lower half) and k2+1 (upper half, when k2+1 < nz).  Rows are
    // emitted only up to the pair's last existing nonzero `last`: a row ending at
    // nz is either emitted after the loop or is the carry row i1.
    auto step = [&](int k2, int last, const float (&b)[V], float a, bool has) {
        const int kk = k2 + h;
        bool done = !has;
        while (re <= last) {  // a row ends at or before the pair's last nonzero
            if (!done && kk < re) {
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(a, b[v], acc[v]);
                done = true;
            }
            emit();
        }
        if (!done) {
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] = fmaf(a, b[v], acc[v]);
        }
    };
    auto consume = [&](int kb, const float (&b)[U][V], const float (&a)[U]) {
        if (re >= kb + 2 * U) {
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(a[u], b[u][v], acc[v]);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) step(kb + 2 * u, kb + 2 * u + 1, b[u], a[u], true);
        }
    };

    const int kend = (nz / (2 * U)) * (2 * U);
    float bA[U][V], bB[U][V];
    float aA[U], aB[U];
    int k = 0;
    if (k < kend) issue(k, bA, aA);
    while (k < kend) {
        if (k + 2 * U < kend) issue(k + 2 * U, bB, aB);
        consume(k, bA, aA);
        k += 2 * U;
        if (k >= kend) break;
        if (k + 2 * U < kend) issue(k + 2 * U, bA, aA);
        consume(k, bB, aB);
        k += 2 * U;
    }
    for (int k2 = kend; k2 < nz; k2 += 2) {  // fewer than 2U nonzeros left
        const bool has = k2 + h < nz;
        float b[V];
        float a = 0.0f;
#pragma unroll
        for (int v = 0; v < V; ++v) b[v] = 0.0f;
        if (has) {
            const uint2 e = s_nz[k2 + h];
            a = __uint_as_float(e.y);
            gather_b<TB, V>(rs, voff + e.x, b);
        }
        step(k2, k2 + 1 < nz ? k2 + 1 : k2, b, a, has);
    }
    while (rr < nr - 1) emit();

    bool has = false;
    if (i1 < M) {
        const int rs0 = rowptr[i1];
        has = j1 > (rs0 > j0 ? rs0 : j0);
    }
    if (sl == 0 && lane == 0) carry_row[wave] = has ? i1 : -1;
    if (has) {
        float sum[V];
#pragma unroll
        for (int v = 0; v < V; ++v) sum[v] = half_sum(acc[v]);
        if (h == 0 && active) {
            float* cp = carry_val + (size_t)wave * N + col;
#pragma unroll
            for (int v = 0; v < V; ++v) cp[v] = sum[v];
        }
    }
}

// ---------------------------------------------- MERGE split for narrow N
// For N ≤ 4 a wave per nonzero would leave most lanes idle, and a row per
// lane group stalls on hub rows.  Here the wave first stages the products
// a·B[c, 0..N) of its tile's nonzeros in LDS (every lane issues the gathers of
// IPT nonzeros at once), then each lane walks IPT consecutive items of the
// tile's merge path (as the SpMV merge kernel does), keeping the N outputs of
// its current row in registers: rows that start and end inside the lane are
// written directly; the first row a lane finishes also gets the partial sums
// of the lanes before it (a segmented scan over the wave); the partial of the
// tile's last row goes to the carry.
template <typename TB, int NV, int IPT>
__global__ __launch_bounds__(256) void k_spmm_narrow(int M, int N, int nnz, const int* __restrict__ rowptr,
                                                     const int* __restrict__ colind, const float* __restrict__ val,
                                                     const TB* __restrict__ B, int ldb, Epi ep, float* __restrict__ C,
                                                     int ldc, int flags, const int* __restrict__ part,
                                                     int* __restrict__ carry_row, float* __restrict__ carry_val) {
    constexpr int TILE = 64 * IPT;
    __shared__ float s_prod[4][TILE * NV];
    __shared__ int s_re[4][TILE + 1];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int wave = uniform((int)(blockIdx.x * 4u + wid));
    const int64_t total = (int64_t)M + nnz;
    const int64_t d0l = (int64_t)wave * TILE;
    if (d0l >= total) return;
    const int d0 = (int)d0l;
    const int d1 = (int)min(d0l + TILE, total);
    const int i0 = uniform(partition_at(part, wave, d0, M, nnz));  // merge_partition
    const int i1 = uniform(partition_at(part, wave + 1, d1, M, nnz));
    const int j0 = d0 - i0, j1 = d1 - i1;
    const int nz = j1 - j0, nrows = i1 - i0;  // rows i0..i1-1 end inside the tile
    const uint32_t rowbytes = (uint32_t)ldb * (uint32_t)sizeof(TB);
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<TB*>(B), (short)0, (int)0xffffffff, 0x00020000);
    float* prod = s_prod[wid];
    int* re = s_re[wid];
    {
        // nonzero k = lane + 64·q; all IPT gathers of a lane are in flight together
        int cs[IPT];
        float as[IPT];
#pragma unroll
        for (int q = 0; q < IPT; ++q) {
            const int k = lane + 64 * q;
            cs[q] = 0;
            as[q] = 0.0f;
            if (k < nz) {
                cs[q] = __builtin_nontemporal_load(colind + j0 + k);
                as[q] = coef(val ? __builtin_nontemporal_load(val + j0 + k) : 1.0f, ep, cs[q]);
            }
        }
        float bv[IPT][NV];
        if (sizeof(TB) == 4 && NV >= 2 && N >= 2 && N >= NV - 1) {
            // fp32 rows of 2, 3 or 4 columns: one dword-aligned 8-, 12- or 16-byte load per
            // nonzero instead of one load per column
#pragma unroll
            for (int q = 0; q < IPT; ++q) {
                const uint32_t off = (uint32_t)cs[q] * rowbytes;
#pragma unroll
                for (int n = 0; n < NV; ++n) bv[q][n] = 0.0f;
                if (NV == 2) {
                    const auto t = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
                    bv[q][0] = __uint_as_float(t[0]);
                    bv[q][1 % NV] = __uint_as_float(t[1]);
                } else if (N == 3) {
                    const auto t = __builtin_amdgcn_raw_buffer_load_b96(rs, off, 0, 0);
                    bv[q][0] = __uint_as_float(t[0]);
                    bv[q][1 % NV] = __uint_as_float(t[1]);
                    bv[q][2 % NV] = __uint_as_float(t[2]);
                } else {
                    const auto t = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
                    bv[q][0] = __uint_as_float(t[0]);
                    bv[q][1 % NV] = __uint_as_float(t[1]);
                    bv[q][2 % NV] = __uint_as_float(t[2]);
                    bv[q][3 % NV] = __uint_as_float(t[3]);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < IPT; ++q)
#pragma unroll
                for (int n = 0; n < NV; ++n) {
                    bv[q][n] = 0.0f;
                    if (n < N) {
                        float b[1];
                        gather_b<TB, 1>(rs, (uint32_t)cs[q] * rowbytes + (uint32_t)(n * (int)sizeof(TB)), b);
                        bv[q][n] = b[0];
                    }
                }
        }
#pragma unroll
        for (int q = 0; q < IPT; ++q) {
            const int k = lane + 64 * q;
            if (k < nz) {
#pragma unroll
                for (int n = 0; n < NV; ++n) prod[k * NV + n] = as[q] * bv[q][n];
            }
        }
    }
    for (int k = lane; k < nrows; k += 64) re[k] = rowptr[i0 + k + 1] - j0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const bool acc_mode = (flags & ARMSP_ACCUMULATE) != 0;
    const int carry_in_row = (i0 < M && rowptr[i0] < j0) ? i0 : -1;
    // this lane's start on the tile's merge path
    const int items = nrows + nz;
    const int ld = lane * IPT < items ? lane * IPT : items;
    int lo = ld - nz > 0 ? ld - nz : 0, hi = ld < nrows ? ld : nrows;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (re[mid] <= ld - mid - 1)
            lo = mid + 1;
        else
            hi = mid;
    }
    int ri = lo, rj = ld - lo;
    const int ri_start = ri;
    float acc[NV], first[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) acc[n] = first[n] = 0.0f;
    auto write_row = [&](int r, const float (&v)[NV]) {
        float* cp = C + (size_t)r * ldc;
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            if (n < N) {
                float o = v[n];
                if (ep.rscale) o *= ep.rscale[r];
                if (ep.bias) o += ep.bias[n];
                if (acc_mode) o += cp[n];
                if (ep.relu && r != carry_in_row) o = fmaxf(o, 0.0f);
                cp[n] = o;
            }
        }
    };
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
        if (ri + rj < items) {
            if (ri < nrows && re[ri] <= rj) {  // row end
                if (ri == ri_start) {
#pragma unroll
                    for (int n = 0; n < NV; ++n) first[n] = acc[n];
                } else {
                    write_row(i0 + ri, acc);
                }
#pragma unroll
                for (int n = 0; n < NV; ++n) acc[n] = 0.0f;
                ++ri;
            } else {
#pragma unroll
                for (int n = 0; n < NV; ++n) acc[n] += prod[rj * NV + n];
                ++rj;
            }
        }
    }
    // segmented inclusive scan of the open-row partials: key = the row each lane ends in
    const int key = ri;
    float S[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) S[n] = acc[n];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int ok = __shfl_up(key, d);
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            const float os = __shfl_up(S[n], d);
            if (lane >= d && ok == key) S[n] += os;
        }
    }
    // the first row this lane finished began in earlier lanes when ri_start is
    // the key the lane before ended in: add their partials
    const int prev_key = __shfl_up(key, 1);
    float prev[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) prev[n] = __shfl_up(S[n], 1);
    if (ri > ri_start) {
        if (lane > 0 && prev_key == ri_start) {
#pragma unroll
            for (int n = 0; n < NV; ++n) first[n] += prev[n];
        }
        write_row(i0 + ri_start, first);
    }
    // partial of row i1 (the scan total of the lanes that end in it) → carry
    float tail[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) tail[n] = read_lane(S[n], 63);
    if (lane == 0) {
        bool has = false;
        if (i1 < M) {
            const int rs0 = rowptr[i1];
            has = j1 > (rs0 > j0 ? rs0 : j0);
        }
        carry_row[wave] = has ? i1 : -1;
        if (has) {
#pragma unroll
            for (int n = 0; n < NV; ++n)
                if (n < N) carry_val[(size_t)wave * N + n] = tail[n];
        }
    }
}

inline size_t stream_lds_bytes(int chunk, bool row_scales) { return 4 * merge_wave_lds(chunk, row_scales); }

// Adds each run of carries (consecutive waves cut inside the same row) into
// that row: a wave per carry, and the one that heads a run sums it.  One load
// round reads the 64 row tags from the previous wave's on (head test and run
// length together); the lanes cover the columns with lpr lanes a row, so
// 64 / lpr carries share a load instruction, 8 loads in flight per lane, and
// a butterfly over the sub-rows joins them in a fixed order — a hub row cut
// into hundreds of chunks costs a few load round trips, and the sum does not
// depend on timing.
template <int V>
__global__ __launch_bounds__(256) void k_carry_fixup(int nwaves, int N, const int* __restrict__ carry_row,
                                                     const float* __restrict__ carry_val, float* __restrict__ C,
                                                     int ldc, Epi ep) {
    constexpr int UF = 8;
    const int lane = threadIdx.x & 63;
    const int w = uniform((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    if (w >= nwaves) return;
    const int idx = w - 1 + lane;  // lane 0: the previous wave, lane 1: this one
    const int cr = (idx >= 0 && idx < nwaves) ? carry_row[idx] : -2;
    const int r = read_lane(cr, 1);
    if (r < 0 || read_lane(cr, 0) == r) return;
    const uint64_t m = __ballot(cr == r) >> 1;  // bit k: wave w + k carries row r, k < 63
    int run = __builtin_ctzll(~m);
    if (run == 63) {
        for (int base = w + 63;; base += 64) {
            const int c2 = (base + lane < nwaves) ? carry_row[base + lane] : -2;
            const uint64_t m2 = __ballot(c2 == r);
            const int cnt = (~m2 == 0ull) ? 64 : __builtin_ctzll(~m2);
            run += cnt;
            if (cnt < 64) break;
        }
    }
    const int cols = (N + V - 1) / V;  // lanes a row needs
    int lpr = 1;
    while (lpr < 64 && lpr < cols) lpr <<= 1;
    const int P = 64 / lpr;  // carries per load instruction
    const int p = lane / lpr, cl = lane % lpr;
    const float* cv = carry_val + (size_t)w * N;
    for (int col0 = 0; col0 < N; col0 += lpr * V) {
        const int col = col0 + cl * V;
        const bool cok = col < N;
        float* cp = C + (size_t)r * ldc + col;
        float o[V];  // C's row, loaded together with the first carries
        if (p == 0 && cok) load_c<V>(cp, o);
        float s[V];
#pragma unroll
        for (int v = 0; v < V; ++v) s[v] = 0.0f;
        if (cok) {
            int k = p;
            for (; k + (UF - 1) * P < run; k += UF * P) {
                float t[UF][V];
#pragma unroll
                for (int q = 0; q < UF; ++q) load_c<V>(cv + (size_t)(k + q * P) * N + col, t[q]);
#pragma unroll
                for (int q = 0; q < UF; ++q)
#pragma unroll
                    for (int v = 0; v < V; ++v) s[v] += t[q][v];
            }
            for (; k < run; k += P) {
                float t[V];
                load_c<V>(cv + (size_t)k * N + col, t);
#pragma unroll
                for (int v = 0; v < V; ++v) s[v] += t[v];
            }
        }
        for (int o = lpr; o < 64; o <<= 1)
#pragma unroll
            for (int v = 0; v < V; ++v) s[v] += __shfl_xor(s[v], o);
        if (p == 0 && cok) {
            const float rs = ep.rscale ? ep.rscale[r] : 1.0f;
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] += rs * s[v];
            if (ep.relu) {
#pragma unroll
                for (int v = 0; v < V; ++v) o[v] = fmaxf(o[v], 0.0f);
            }
#pragma unroll
            for (int v = 0; v < V; ++v) cp[v] = o[v];
        }
    }
}

// ------------------------------------------------------------- dispatch
template <typename TB>
constexpr int max_vec() {
    return sizeof(TB) == 4 ? 4 : 8;
}

inline bool aligned(const void* p, int bytes) { return ((uintptr_t)p % (uintptr_t)bytes) == 0; }

template <typename TB>
int pick_vec(int N, int ldb, int ldc, const TB* B, const float* C, int want, int lanes) {
    int v = want;
    if (v <= 0) {
        v = 1;
        while (v < max_vec<TB>() && lanes * v < N) v <<= 1;
    }
    while (v > 1 && (N % v || ldb % v || ldc % v || !aligned(B, v * (int)sizeof(TB)) || !aligned(C, v * 4))) v >>= 1;
    return v;
}

template <typename TB, int V>
int launch_row_v(int G, dim3 grid, hipStream_t s, int M, int N, const int* rp, const int* ci, const float* va,
                 const TB* B, int ldb, const Epi& bs, float* C, int ldc, int flags) {
    switch (G) {
#define ARMSP_ROW_CASE(GG)                                                                                   \
    case GG:                                                                                                 \
        hipLaunchKernelGGL((k_spmm_row<TB, V, GG>), grid, dim3(256), 0, s, M, N, rp, ci, va, B, ldb, bs, C, ldc, \
                           flags);                                                                           \
        break;
        ARMSP_ROW_CASE(4)
        ARMSP_ROW_CASE(8)
        ARMSP_ROW_CASE(16)
        ARMSP_ROW_CASE(32)
        ARMSP_ROW_CASE(64)
#undef ARMSP_ROW_CASE
        default: return ARMSP_ERROR_INVALID_VALUE;
    }
    return ARMSP_SUCCESS;
}

template <typename TB, int V, bool MERGE>
int launch_split_vu(int U, dim3 grid, hipStream_t s, int M, int N, int nnz, int chunk, const int* rp, const int* ci,
                    const float* va, const TB* B, int ldb, const Epi& bs, float* C, int ldc, int flags, int* crow,
                    float* cval) {
    switch (U) {
#define ARMSP_SPLIT_CASE(UU)                                                                                \
    case UU:                                                                                                \
        hipLaunchKernelGGL((k_spmm_split<TB, V, UU, MERGE>), grid, dim3(256), 0, s, M, N, nnz, chunk, rp, ci, va, \
                           B, ldb, bs, C, ldc, flags, crow, cval);                                          \
        break;
        ARMSP_SPLIT_CASE(4)
        ARMSP_SPLIT_CASE(8)
        ARMSP_SPLIT_CASE(16)
#undef ARMSP_SPLIT_CASE
        default: return ARMSP_ERROR_INVALID_VALUE;
    }
    return ARMSP_SUCCESS;
}

template <typename TB, int V>
int launch_stream_vu(int U, dim3 grid, hipStream_t s, int M, int N, int nnz, int chunk, const int* rp, const int* ci,
                     const float* va, const TB* B, int ldb, const Epi& bs, float* C, int ldc, int flags,
                     const int* part, int* crow, float* cval) {
    const size_t lds = stream_lds_bytes(chunk, bs.rscale != nullptr);
    switch (U) {
#define ARMSP_STREAM_CASE(UU)                                                                                  \
    case UU:                                                                                                   \
        hipLaunchKernelGGL((k_spmm_stream<TB, V, UU>), grid, dim3(256), lds, s, M, N, nnz, chunk, rp, ci, va, B, ldb, \
                           bs, C, ldc, flags, part, crow, cval);                                               \
        break;
        ARMSP_STREAM_CASE(4)
        ARMSP_STREAM_CASE(8)
        ARMSP_STREAM_CASE(16)
#undef ARMSP_STREAM_CASE
        default: return ARMSP_ERROR_INVALID_VALUE;
    }
    return ARMSP_SUCCESS;
}

template <typename TB, int V>
int launch_slice_v(int U, int S, int G, hipStream_t s, int M, int N, int nnz, int chunk, const int* rp,
                   const int* ci, const float* va, const TB* B, int ldb, const Epi& bs, float* C, int ldc, int flags,
                   const int* part, int* crow, float* cval) {
    const size_t lds = stream_lds_bytes(chunk, bs.rscale != nullptr);
    const int64_t nblk = (8 % S == 0) ? ((int64_t)(G + 8 / S - 1) / (8 / S)) * 8 : (int64_t)G * S;
    if (nblk > INT_MAX) return ARMSP_ERROR_NOT_SUPPORTED;
    const dim3 grid((unsigned)nblk);
    switch (U) {
#define ARMSP_SLICE_CASE(UU)                                                                                     \
    case UU:                                                                                                     \
        hipLaunchKernelGGL((k_spmm_slice<TB, V, UU>), grid, dim3(256), lds, s, M, N, nnz, chunk, S, G, rp, ci, va, B, \
                           ldb, bs, C, ldc, flags, part, crow, cval);                                            \
        break;
        ARMSP_SLICE_CASE(4)
        ARMSP_SLICE_CASE(8)
        ARMSP_SLICE_CASE(16)
#undef ARMSP_SLICE_CASE
        default: return ARMSP_ERROR_INVALID_VALUE;
    }
    return ARMSP_SUCCESS;
}

// lane width sv columns: 4, 8 or 16 bytes per lane
template <typename TB>
int launch_slice(int sv, int U, int S, int G, hipStream_t s, int M, int N, int nnz, int chunk, const int* rp,
                 const int* ci, const float* va, const TB* B, int ldb, const Epi& bs, float* C, int ldc, int flags,
                 const int* part, int* crow, float* cval) {
    switch (sv * (int)sizeof(TB)) {
        case 4:
            return launch_slice_v<TB, 4 / sizeof(TB)>(U, S, G, s, M, N, nnz, chunk, rp, ci, va, B, ldb, bs, C, ldc,
                                                      flags, part, crow, cval);
        case 8:
            return launch_slice_v<TB, 8 / sizeof(TB)>(U, S, G, s, M, N, nnz, chunk, rp, ci, va, B, ldb, bs, C, ldc,
                                                      flags, part, crow, cval);
        case 16:
            return launch_slice_v<TB, 16 / sizeof(TB)>(U, S, G, s, M, N, nnz, chunk, rp, ci, va, B, ldb, bs, C, ldc,
                                                       flags, part, crow, cval);
        default: return ARMSP_ERROR_INVALID_VALUE;
    }
}

template <typename TB, int NV>
int launch_narrow_nv(int ipt, int64_t nwaves, hipStream_t s, int M, int N, int nnz, const int* rp, const int* ci,
                     const float* va, const TB* B, int ldb, const Epi& bs, float* C, int ldc, int flags,
                     const int* part, int* crow, float* cval) {
    const dim3 grid((unsigned)((nwaves + 3) / 4));
    switch (ipt) {
#define ARMSP_NARROW_CASE(II)                                                                                   \
    case II:                                                                                                    \
        hipLaunchKernelGGL((k_spmm_narrow<TB, NV, II>), grid, dim3(256), 0, s, M, N, nnz, rp, ci, va, B, ldb, bs, C, \
                           ldc, flags, part, crow, cval);                                                       \
        break;
        ARMSP_NARROW_CASE(4)
        ARMSP_NARROW_CASE(8)
        ARMSP_NARROW_CASE(16)
#undef ARMSP_NARROW_CASE
        default: return ARMSP_ERROR_INVALID_VALUE;
    }
    return ARMSP_SUCCESS;
}

template <typename TB>
int launch_narrow(int ipt, int64_t nwaves, hipStream_t s, int M, int N, int nnz, const int* rp, const int* ci,
                  const float* va, const TB* B, int ldb, const Epi& bs, float* C, int ldc, int flags,
                  const int* part, int* crow, float* cval) {
    if (N <= 1) return launch_narrow_nv<TB, 1>(ipt, nwaves, s, M, N, nnz, rp, ci, va, B, ldb, bs, C, ldc, flags, part, crow, cval);
    if (N <= 2) return launch_narrow_nv<TB, 2>(ipt, nwaves, s, M, N, nnz, rp, ci, va, B, ldb, bs, C, ldc, flags, part, crow, cval);
    return launch_narrow_nv<TB, 4>(ipt, nwaves, s, M, N, nnz, rp, ci, va, B, ldb, bs, C, ldc, flags, part, crow, cval);
}

// MERGE gathers through 32-bit buffer offsets unless a dense-row offset could exceed them
template <typename TB, int V, bool MERGE>
int launch_split(bool wide, int U, dim3 grid, hipStream_t s, int M, int N, int nnz, int chunk, const int* rp,
                 const int* ci, const float* va, const TB* B, int ldb, const Epi& bs, float* C, int ldc, int flags,
                 const int* part, int* crow, float* cval) {
    if (wide || !MERGE || chunk > 2048)  // the LDS staging holds at most 2048 items per wave
        return launch_split_vu<TB, V, MERGE>(U, grid, s, M, N, nnz, chunk, rp, ci, va, B, ldb, bs, C, ldc, flags, crow,
                                             cval);
    return launch_stream_vu<TB, V>(U, grid, s, M, N, nnz, chunk, rp, ci, va, B, ldb, bs, C, ldc, flags, part, crow,
                                   cval);
}

// Device memory a call may use instead of allocating its workspace (the entry
// point that allocates nothing): `bytes` at `p`, 256-byte aligned.
struct Scratch {
    void* p;
    size_t bytes;
};

template <typename TB>
int spmm_impl(int M, int K, int N, int nnz, const int* rowptr, const int* colind, const float* val, const TB* B,
              const float* b_row_scale, int ldb, float* C, int ldc, int flags, const armsp_epilogue* epi,
              const armsp_config* cfg, void* stream, const Scratch* scratch = nullptr) {
    if (M < 0 || K < 0 || N < 0 || nnz < 0) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0) return ARMSP_SUCCESS;
    if (!rowptr || !C || ldc < N) return ARMSP_ERROR_INVALID_VALUE;
    if (nnz > 0 && (!colind || !B || ldb < N || K == 0)) return ARMSP_ERROR_INVALID_VALUE;
    if (sizeof(TB) == 1 && nnz > 0 && !b_row_scale) return ARMSP_ERROR_INVALID_VALUE;
    Epi bscale{sizeof(TB) == 1 ? b_row_scale : nullptr, epi ? epi->col_scale : nullptr,
               epi ? epi->row_scale : nullptr, epi ? epi->bias : nullptr, epi ? epi->relu : 0};
    if (flags & ~(ARMSP_ACCUMULATE | ARMSP_TRUNC4)) return ARMSP_ERROR_INVALID_VALUE;
    if ((int64_t)M + nnz > INT_MAX) return ARMSP_ERROR_NOT_SUPPORTED;
    hipStream_t s = as_stream(stream);

    armsp_config c = cfg ? *cfg : armsp_config{0, 0, 0, 0, 0};
    // left to the library: the installed schedule selector, if any (never on the
    // allocation-free path unless this matrix's choice is already known)
    if (c.schedule == 0 && c.group == 0 && c.vec == 0 && c.chunk == 0 && c.unroll == 0 && !(flags & ARMSP_TRUNC4) &&
        !epi && nnz > 0) {
        armsp_config sel;
        if (selector_config(0, M, nnz, N, (int)sizeof(TB), rowptr, s, scratch == nullptr, &sel)) c = sel;
    }
    int sched = c.schedule;
    // AUTO: the merge path at every width — N ≤ 4 runs the per-lane item walk
    // (k_spmm_narrow), wider N a whole wave or half waves per nonzero.  (A row per
    // lane group is often faster on regular rows, but a hub row stalls its group:
    // 20x slower on the 1M R-MAT at N = 8; the schedule selector picks it where it
    // pays.)
    if (sched == ARMSP_SCHED_AUTO) sched = ARMSP_SCHED_MERGE;
    if (flags & ARMSP_TRUNC4) sched = ARMSP_SCHED_ROW;  // truncation is per row
    if (nnz == 0) sched = ARMSP_SCHED_ROW;               // writes (or keeps) every row

    if (sched == ARMSP_SCHED_ROW) {
        int G = c.group;
        int V = pick_vec<TB>(N, ldb, ldc, B, C, c.vec, G > 0 ? G : 64);
        if (G <= 0) {
            G = 4;
            while (G < 64 && G * V < N) G <<= 1;
        }
        if (G < 4 || G > 64 || (G & (G - 1))) return ARMSP_ERROR_INVALID_VALUE;
        dim3 grid((unsigned)(((int64_t)M * G + 255) / 256), (unsigned)((N + G * V - 1) / (G * V)));
        if (grid.y > 65535) return ARMSP_ERROR_NOT_SUPPORTED;
        TimedLaunch t("spmm", s);
        int st;
        switch (V) {
            case 1: st = launch_row_v<TB, 1>(G, grid, s, M, N, rowptr, colind, val, B, ldb, bscale, C, ldc, flags); break;
            case 2: st = launch_row_v<TB, 2>(G, grid, s, M, N, rowptr, colind, val, B, ldb, bscale, C, ldc, flags); break;
            case 4: st = launch_row_v<TB, 4>(G, grid, s, M, N, rowptr, colind, val, B, ldb, bscale, C, ldc, flags); break;
            case 8:
                if constexpr (sizeof(TB) < 4) {
                    st = launch_row_v<TB, 8>(G, grid, s, M, N, rowptr, colind, val, B, ldb, bscale, C, ldc, flags);
                    break;
                }
                [[fallthrough]];
            default: return ARMSP_ERROR_INVALID_VALUE;
        }
        if (st) return st;
        ARMSP_LAUNCHED();
        t.stop();
        return ARMSP_SUCCESS;
    }
    if (sched != ARMSP_SCHED_NNZ && sched != ARMSP_SCHED_MERGE) return ARMSP_ERROR_INVALID_VALUE;

    const bool merge = sched == ARMSP_SCHED_MERGE;
    const int V = pick_vec<TB>(N, ldb, ldc, B, C, c.vec, 64);
    int U = c.unroll > 0 ? c.unroll : (V <= 2 ? 8 : 4);
    if (U != 4 && U != 8 && U != 16) return ARMSP_ERROR_INVALID_VALUE;
    int chunk = c.chunk > 0 ? c.chunk : 512;
    if (chunk % 64) return ARMSP_ERROR_INVALID_VALUE;
    // dense-row byte offsets in 32 bits: (K-1)·ldb + N elements from B
    const bool wide = c.group == 1 || ((uint64_t)K * (uint64_t)ldb * sizeof(TB)) > 0xffffffffull;
    // N ≤ 4 (group 0): lanes walk runs of 4, 8 or 16 merge items (k_spmm_narrow); the
    // wave tile is 64 × that run, so the chunk is rounded to 256, 512 or 1024.  From
    // N = 5 up a whole wave per nonzero is faster (1M R-MAT, N = 8: 238 vs 347 us)
    const bool narrow = merge && !wide && N <= 4 && c.group == 0;
    // AUTO column slices (see below) only where they pay: slices start on 128-byte
    // lines only when every row of B does — otherwise the line a slice boundary cuts
    // is fetched by two XCDs (1M R-MAT: bf16 N=100 688 vs 490 us, fp32 N=200 1840 vs
    // 1707 us against the stream kernel) — and the row must fill its slices (bf16
    // N=100 in 256-byte rows: 4.9 ms with a 72-byte last slice, 4.6 ms with the
    // stream kernel on a 126M-edge graph)
    const int esz = (int)sizeof(TB);
    const bool line_rows = ((int64_t)ldb * esz) % 128 == 0 && aligned(B, 128);
    const bool slice_rows = line_rows && ((N * esz >= 512 && (N * esz) % 256 == 0) ||
                                          (esz == 2 && N * esz >= 128 && N * esz < 512 && (N * esz) % 128 == 0));
    // half-wave slices, and a whole wave per nonzero on narrow rows of B (N < 32), run
    // best with 256 items per wave and 4 gathers per batch
    if (merge && !wide && !narrow && c.group == 0 && (slice_rows || N < 32)) {
        if (c.chunk <= 0) chunk = 256;
        if (c.unroll <= 0) U = 4;
    }
    if (narrow) {  // default: runs of 8 items for N ≤ 2, of 4 above (the staged products fill the LDS)
        if (c.chunk <= 0) chunk = N <= 2 ? 512 : 256;
        chunk = chunk <= 256 ? 256 : (chunk <= 512 ? 512 : 1024);
    }
    const int64_t total = merge ? (int64_t)M + nnz : (int64_t)nnz;
    if (merge && !narrow && c.chunk <= 0) {
        // small inputs: shorter chunks until 16 waves per CU are in flight (PubMed-size
        // graphs otherwise leave half the CUs idle while each wave walks its chunk)
        const int64_t want = 16 * (int64_t)device_cus();
        while (chunk > 64 && (total + chunk - 1) / chunk < want) chunk >>= 1;
    }
    const int64_t nwaves = (total + chunk - 1) / chunk;
    if (nwaves > INT_MAX / 4) return ARMSP_ERROR_NOT_SUPPORTED;
    const unsigned tiles = (unsigned)((N + 64 * V - 1) / (64 * V));
    if (tiles > 65535) return ARMSP_ERROR_NOT_SUPPORTED;

    // the LDS-staged merge kernels (stream, slice, narrow) start from a precomputed
    // merge-path partition: nwaves + 1 row boundaries
    const bool staged = merge && !wide && chunk <= 2048;
    void* ws = nullptr;
    const size_t row_bytes = ((size_t)nwaves * sizeof(int) + 255) & ~(size_t)255;
    const size_t val_bytes = ((size_t)nwaves * N * sizeof(float) + 255) & ~(size_t)255;
    const size_t part_bytes = staged ? (size_t)(nwaves + 1) * sizeof(int) : 0;
    int st;
    if (scratch) {  // no allocation: the caller's region, or a clear refusal
        if (row_bytes + val_bytes + part_bytes > scratch->bytes) return ARMSP_ERROR_NOT_SUPPORTED;
        ws = scratch->p;
    } else {
        st = ws_alloc(&ws, row_bytes + val_bytes + part_bytes, s);
        if (st) return st;
    }
    // a caller's region is never freed here
    auto release = [&](void* w, hipStream_t st_) {
        if (!scratch) ws_free(w, st_);
    };
    int* crow = static_cast<int*>(ws);
    float* cval = reinterpret_cast<float*>(static_cast<char*>(ws) + row_bytes);
    int* part = staged ? reinterpret_cast<int*>(static_cast<char*>(ws) + row_bytes + val_bytes) : nullptr;
    if (staged) {
        TimedLaunch t("merge_partition", s);
        st = merge_partition(M, nnz, chunk, nwaves, rowptr, part, s);
        if (st) {
            release(ws, s);
            return st;
        }
        t.stop();
    }

    // group 32: half a wave per nonzero over column slices of B (k_spmm_slice), each lane
    // 4 bytes of B (vec 0) or vec columns of 4, 8 or 16 bytes.
    // AUTO (group 0): when a row of B spans two or more 256-byte slices, half a wave per
    // nonzero with 8-byte lanes, the slices spread over the XCDs — on power-law matrices
    // each XCD's L2 then holds half (or less) of every hot row (measured: 18–25% faster
    // on the 1M R-MAT at N=128/256 fp32, 8% at N=256 bf16, within 6% elsewhere)
    // bf16 rows of 128-511 bytes: half-wave slices with 4-byte lanes (1M R-MAT: N = 128
    // 637 -> 574 us, N = 64 336 -> 315 us)
    const int alb = N * esz >= 512 ? 8 : 4;  // AUTO's lane bytes
    const int asv = alb / esz;
    const bool auto_slice = merge && !wide && !narrow && c.group == 0 && chunk <= 2048 && slice_rows &&
                            N % asv == 0 && ldb % asv == 0 && ldc % asv == 0 && aligned(B, alb) &&
                            aligned(C, 4 * asv);
    const int SV = auto_slice ? asv : (c.vec > 0 ? c.vec : 4 / esz);
    const int slb = SV * esz;
    const bool slice = auto_slice ||
                       (merge && !wide && c.group == 32 && chunk <= 2048 && (slb == 4 || slb == 8 || slb == 16) &&
                        N % SV == 0 && ldb % SV == 0 && ldc % SV == 0 && aligned(B, slb) && aligned(C, 4 * SV));
    if (c.group == 32 && !slice) {
        release(ws, s);
        return ARMSP_ERROR_NOT_SUPPORTED;
    }
    dim3 grid((unsigned)((nwaves + 3) / 4), tiles);
    if (narrow) {
        TimedLaunch t("spmm", s);
        st = launch_narrow<TB>(chunk / 64, nwaves, s, M, N, nnz, rowptr, colind, val, B, ldb, bscale, C, ldc, flags,
                               part, crow, cval);
        hipError_t e = st ? hipSuccess : hipGetLastError();
        if (st || e != hipSuccess) {
            release(ws, s);
            return st ? st : hip_fail(e);
        }
        t.stop();
    } else if (slice) {
        TimedLaunch t("spmm", s);
        const int S = (N + 32 * SV - 1) / (32 * SV);
        st = launch_slice<TB>(SV, U, S, (int)((nwaves + 3) / 4), s, M, N, nnz, chunk, rowptr, colind, val, B, ldb,
                              bscale, C, ldc, flags, part, crow, cval);
        hipError_t e = st ? hipSuccess : hipGetLastError();
        if (st || e != hipSuccess) {
            release(ws, s);
            return st ? st : hip_fail(e);
        }
        t.stop();
    } else {
        TimedLaunch t("spmm", s);
#define ARMSP_SPLIT_V(VV)                                                                                          \
    case VV:                                                                                                       \
        st = merge ? launch_split<TB, VV, true>(wide, U, grid, s, M, N, nnz, chunk, rowptr, colind, val, B, ldb,    \
                                                bscale, C, ldc, flags, part, crow, cval)                           \
                   : launch_split<TB, VV, false>(wide, U, grid, s, M, N, nnz, chunk, rowptr, colind, val, B, ldb,   \
                                                 bscale, C, ldc, flags, part, crow, cval);                         \
        break;
        switch (V) {
            ARMSP_SPLIT_V(1)
            ARMSP_SPLIT_V(2)
            ARMSP_SPLIT_V(4)
            case 8:
                if constexpr (sizeof(TB) < 4) {
                    st = merge ? launch_split<TB, 8, true>(wide, U, grid, s, M, N, nnz, chunk, rowptr, colind, val, B,
                                                           ldb, bscale, C, ldc, flags, part, crow, cval)
                               : launch_split<TB, 8, false>(wide, U, grid, s, M, N, nnz, chunk, rowptr, colind, val, B,
                                                            ldb, bscale, C, ldc, flags, part, crow, cval);
                    break;
                }
                [[fallthrough]];
            default: st = ARMSP_ERROR_INVALID_VALUE;
        }
#undef ARMSP_SPLIT_V
        if (st) {
            release(ws, s);
            return st;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            release(ws, s);
            return hip_fail(e);
        }
        t.stop();
    }
    {
        TimedLaunch t("spmm_fixup", s);
        const int fv = (N % 4 == 0 && ldc % 4 == 0 && aligned(C, 16)) ? 4 : 1;
        dim3 fgrid((unsigned)((nwaves + 3) / 4), 1);
        if (fv == 4)
            hipLaunchKernelGGL((k_carry_fixup<4>), fgrid, dim3(256), 0, s, (int)nwaves, N, crow, cval, C, ldc, bscale);
        else
            hipLaunchKernelGGL((k_carry_fixup<1>), fgrid, dim3(256), 0, s, (int)nwaves, N, crow, cval, C, ldc, bscale);
        hipError_t e = hipGetLastError();
        release(ws, s);
        if (e != hipSuccess) return hip_fail(e);
        t.stop();
    }
    return ARMSP_SUCCESS;
}

// ------------------------------------------------- allocation-free entry point
// The merge-path partition and carry buffers of armsp_spmm_f32_noalloc live in
// this file-scope array: each device holds its own copy for the life of the
// process.  Host code reaches a device's copy only through the address
// hipGetSymbolAddress returns with that device current (the array's name on the
// host is not a device address).
constexpr size_t kNoAllocScratchBytes = (size_t)ARMSP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

// The calling thread's current device's copy of `scratch`, looked up once per device.
int scratch_address(void** out) {
    int dev = 0;
    ARMSP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return ARMSP_ERROR_NOT_SUPPORTED;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e);
        if (!p) return ARMSP_ERROR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return ARMSP_SUCCESS;
}

}  // namespace
}  // namespace armsp

extern "C" {

int armsp_spmm_f32_noalloc(int M, int K, int N, int nnz, const int* rowptr, const int* colind, const float* val,
                           const float* B, int ldb, float* C, int ldc, int flags, const armsp_config* cfg,
                           void* stream) {
    void* p = nullptr;
    const int st = armsp::scratch_address(&p);
    if (st) return st;
    const armsp::Scratch sc{p, armsp::kNoAllocScratchBytes};
    armsp::TimingPause quiet;  // no timing events either
    return armsp::spmm_impl<float>(M, K, N, nnz, rowptr, colind, val, B, nullptr, ldb, C, ldc, flags, nullptr, cfg,
                                   stream, &sc);
}

int armsp_spmm_f32(int M, int K, int N, int nnz, const int* rowptr, const int* colind, const float* val,
                   const float* B, int ldb, float* C, int ldc, int flags, const armsp_config* cfg, void* stream) {
    return armsp::spmm_impl<float>(M, K, N, nnz, rowptr, colind, val, B, nullptr, ldb, C, ldc, flags, nullptr, cfg,
                                   stream);
}

int armsp_spmm_bf16(int M, int K, int N, int nnz, const int* rowptr, const int* colind, const float* val,
                    const uint16_t* B, int ldb, float* C, int ldc, int flags, const armsp_config* cfg, void* stream) {
    return armsp::spmm_impl<uint16_t>(M, K, N, nnz, rowptr, colind, val, B, nullptr, ldb, C, ldc, flags, nullptr, cfg,
                                      stream);
}

int armsp_spmm_fp8(int M, int K, int N, int nnz, const int* rowptr, const int* colind, const float* val,
                   const uint8_t* Bq, const float* b_row_scale, int ldb, float* C, int ldc, int flags,
                   const armsp_config* cfg, void* stream) {
    return armsp::spmm_impl<uint8_t>(M, K, N, nnz, rowptr, colind, val, Bq, b_row_scale, ldb, C, ldc, flags, nullptr,
                                     cfg, stream);
}

int armsp_spmm_ex(int M, int K, int N, int nnz, const int* rowptr, const int* colind, const float* val, int b_dtype,
                  const void* B, const float* b_row_scale, int ldb, float* C, int ldc, int flags,
                  const armsp_epilogue* epilogue, const armsp_config* cfg, void* stream) {
    switch (b_dtype) {
        case ARMSP_DTYPE_F32:
            return armsp::spmm_impl<float>(M, K, N, nnz, rowptr, colind, val, static_cast<const float*>(B), nullptr,
                                           ldb, C, ldc, flags, epilogue, cfg, stream);
        case ARMSP_DTYPE_BF16:
            return armsp::spmm_impl<uint16_t>(M, K, N, nnz, rowptr, colind, val, static_cast<const uint16_t*>(B),
                                              nullptr, ldb, C, ldc, flags, epilogue, cfg, stream);
        case ARMSP_DTYPE_FP8_E4M3:
            return armsp::spmm_impl<uint8_t>(M, K, N, nnz, rowptr, colind, val, static_cast<const uint8_t*>(B),
                                             b_row_scale, ldb, C, ldc, flags, epilogue, cfg, stream);
        default:
            return ARMSP_ERROR_INVALID_VALUE;
    }
}

}  // extern "C"
