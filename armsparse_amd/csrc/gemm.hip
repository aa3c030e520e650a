// bf16 GEMM on gfx950 MFMA (v_mfma_f32_32x32x16_bf16), fp32 accumulation.
//
// Used for the dense feature transform of the GCN layer, x @ W
// (gcn/pytorch-custom/op.py:114), and for its weight gradient xᵀ @ dY.  Both
// operands take arbitrary element strides so transposed views need no copy.
// Three kernels: a tall one for M ≫ K, N (B resident in LDS, A streamed once),
// a reduction one for K ≫ M, N (per-block K slabs into partials summed in
// block order), and a general 64×64×32 tile kernel, split along K when its
// tile grid is too small (partials again summed in split order).  Operand
// tiles sit in LDS k-contiguous so each lane fetches its 8 MFMA operand
// values with one 16-byte LDS read.
#include "common.hpp"

namespace armsp {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;  // LDS row pitch in bf16

// Staging of a 64×32 operand tile (rows r0..r0+63 of M or N, k0..k0+31): each
// thread carries 8 consecutive elements, packed in a uint4.  Tile modes: 0 — the
// operand is k-contiguous (stride 1 along k, 16-byte aligned rows), one 16-byte
// load along k; 1 — row-contiguous (stride 1 along the row dimension, 16-byte
// aligned k slices), one 16-byte load along the rows, transposed into LDS by
// eight 2-byte stores with consecutive lanes on consecutive k (bank-conflict
// free); 2 — any strides, eight element loads along k.  Out-of-range elements
// (rows ≥ R, k ≥ ke) are zero.
__device__ __forceinline__ uint32_t pack2(uint16_t lo, uint16_t hi) { return lo | ((uint32_t)hi << 16); }

__device__ __forceinline__ uint4 tile_fetch(const uint16_t* __restrict__ P, int64_t sr, int64_t sk, int mode, int R,
                                            int r0, int k0, int ke, int tid) {
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    uint16_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (mode == 1) {
        const int r = r0 + (tid >> 5) * 8, gk = k0 + (tid & 31);
        if (gk < ke && r < R) {
            const uint16_t* q = P + (int64_t)r + (int64_t)gk * sk;
            if (r + 8 <= R) return *reinterpret_cast<const uint4*>(q);
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (r + i < R) t[i] = q[i];
        }
    } else {
        const int r = r0 + (tid >> 2), gk = k0 + (tid & 3) * 8;
        if (r < R && gk < ke) {
            const uint16_t* q = P + (int64_t)r * sr + (int64_t)gk * sk;
            if (mode == 0 && gk + 8 <= ke) return *reinterpret_cast<const uint4*>(q);
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (gk + i < ke) t[i] = q[(int64_t)i * sk];
        }
    }
    v.x = pack2(t[0], t[1]);
    v.y = pack2(t[2], t[3]);
    v.z = pack2(t[4], t[5]);
    v.w = pack2(t[6], t[7]);
    return v;
}

// tile in LDS, k-contiguous rows of pitch LDK
__device__ __forceinline__ void tile_commit(uint16_t* __restrict__ S, int mode, uint4 v, int tid) {
    if (mode == 1) {
        const int r = (tid >> 5) * 8, kk = tid & 31;
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 8; ++i) S[(r + i) * LDK + kk] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
    } else {
        *reinterpret_cast<uint4*>(&S[(tid >> 2) * LDK + (tid & 3) * 8]) = v;
    }
}

// tile mode of an operand whose element (r, k) sits at P[r·sr + k·sk]
int tile_mode(const void* P, int64_t sr, int64_t sk) {
    const bool al = ((uintptr_t)P % 16) == 0;
    if (al && sk == 1 && sr % 8 == 0) return 0;
    if (al && sr == 1 && sk % 8 == 0) return 1;
    return 2;
}

// C_part[z] = op(A)[m0:m0+64, kz0:kz1] · op(B)[kz0:kz1, n0:n0+64]; the next k
// step's tiles are loaded into registers while the current one is multiplied.
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, int kchunk, const uint16_t* __restrict__ A,
                                                   int64_t sam, int64_t sak, int amode,
                                                   const uint16_t* __restrict__ B, int64_t sbk, int64_t sbn,
                                                   int bmode, float* __restrict__ C, int64_t ldc, int64_t zstride,
                                                   int accumulate) {
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
    uint4 va = tile_fetch(A, sam, sak, amode, M, m0, kz0, kz1, tid);
    uint4 vb = tile_fetch(B, sbn, sbk, bmode, N, n0, kz0, kz1, tid);
    for (int k0 = kz0; k0 < kz1; k0 += BK) {
        tile_commit(sA, amode, va, tid);
        tile_commit(sB, bmode, vb, tid);
        __syncthreads();
        if (k0 + BK < kz1) {
            va = tile_fetch(A, sam, sak, amode, M, m0, k0 + BK, kz1, tid);
            vb = tile_fetch(B, sbn, sbk, bmode, N, n0, k0 + BK, kz1, tid);
        }
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

// ------------------------------------------------------------ tall GEMM
// C[M×N] (=|+=) A[M×K]·B[K×N] for M ≫ K, N (the GCN layer's x·W and the
// gradient back to x): each block keeps all of B in LDS (n-major, k
// contiguous) and loops over row tiles of A, so A is read once and B once per
// block.  Four waves as WR row strips of 32 × (4/WR) column groups of 32·NTW;
// a row tile is 32·WR rows, staged in LDS 32 k at a time with vector loads
// along k when A's rows allow (va = 8, 4 or 1 elements).  B is staged with
// 16-byte loads when it is contiguous along n (x·W) or along k (a Wᵀ view).
constexpr int TK = 32, TKP = TK + 8;  // A stage: 32 k, row pitch 40 bf16
constexpr int RTK = 64, RTKP = RTK + 8;  // reduction stage: 64 k, row pitch 72 bf16

template <int WR, int NTW, bool OUTBF>
__global__ __launch_bounds__(256) void k_gemm_tall(int M, int N, int K, int KP, const uint16_t* __restrict__ A,
                                                   int64_t sam, int64_t sak, const uint16_t* __restrict__ B,
                                                   int64_t sbk, int64_t sbn, void* __restrict__ Cv, int64_t ldc,
                                                   int accumulate, int va) {
    constexpr int TR = 32 * WR;               // rows per tile
    constexpr int NP = 32 * NTW * (4 / WR);  // columns held
    extern __shared__ __attribute__((aligned(16))) uint16_t lds[];
    uint16_t* sB = lds;                      // [NP][KP + 8]
    uint16_t* sA = lds + NP * (KP + 8);      // [TR][TKP], at least 4096 elements
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wr = wid % WR, wc = wid / WR;
    const int bp = KP + 8;
    const bool bal = ((uintptr_t)B % 16) == 0;
    if (bal && sbn == 1 && sbk % 8 == 0) {  // rows of B contiguous in n: 16-byte loads, consecutive lanes on k
        for (int idx = tid; idx < (NP / 8) * KP; idx += 256) {
            const int g = idx / KP, k = idx - g * KP, n = g * 8;
            uint16_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (k < K && n < N) {
                const uint16_t* q = B + (int64_t)k * sbk + n;
                if (n + 8 <= N) {
                    const uint4 v = *reinterpret_cast<const uint4*>(q);
                    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int i = 0; i < 8; ++i) t[i] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
                } else {
                    for (int i = 0; n + i < N; ++i) t[i] = q[i];
                }
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) sB[(n + i) * bp + k] = t[i];
        }
    } else if (bal && sbk == 1 && sbn % 8 == 0) {  // columns of B contiguous in k: 16-byte loads along k
        const int per = KP / 8;
        for (int idx = tid; idx < NP * per; idx += 256) {
            const int n = idx / per, k = (idx - n * per) * 8;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (n < N && k < K) {
                const uint16_t* q = B + (int64_t)n * sbn + k;
                if (k + 8 <= K) {
                    v = *reinterpret_cast<const uint4*>(q);
                } else {
                    uint16_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    for (int i = 0; k + i < K; ++i) t[i] = q[i];
                    v = make_uint4(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16),
                                   t[4] | ((uint32_t)t[5] << 16), t[6] | ((uint32_t)t[7] << 16));
                }
            }
            *reinterpret_cast<uint4*>(&sB[n * bp + k]) = v;
        }
    } else {
        for (int idx = tid; idx < NP * KP; idx += 256) {
            const int n = idx / KP, k = idx - n * KP;
            sB[n * bp + k] = (n < N && k < K) ? B[(int64_t)k * sbk + (int64_t)n * sbn] : (uint16_t)0;
        }
    }
    __syncthreads();
    const int ntiles = (M + TR - 1) / TR;
    // MFMAs of one staged k step
    auto mma_step = [&](f32x16 (&acc)[NTW], int k0) {
#pragma unroll
        for (int ks = 0; ks < TK; ks += 16) {
            const int koff = ks + (lane >> 5) * 8;
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[(wr * 32 + (lane & 31)) * TKP + koff]);
#pragma unroll
            for (int j = 0; j < NTW; ++j) {
                const int n = (wc * NTW + j) * 32 + (lane & 31);
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[n * bp + k0 + koff]);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[j], 0, 0, 0);
            }
        }
    };
    // the tile's results.  bf16 output with 16-byte aligned rows: each 32×32
    // block goes through the wave's 2 KB of LDS (the A stage, free once the last k
    // step is done) and leaves as 16-byte stores, two per lane, 64 bytes of a row
    // per lane pair; otherwise 32 consecutive columns of one row per half wave.
    const bool vout = OUTBF && ((uintptr_t)Cv % 16) == 0 && ldc % 8 == 0;
    auto store_tile = [&](const f32x16 (&acc)[NTW], int m0) {
        if (vout) {
            uint16_t* stg = sA + wid * 1024;  // [32][32] bf16
            const int orow = lane >> 1, ocol = (lane & 1) * 16;
            const int gm = m0 + wr * 32 + orow;
#pragma unroll
            for (int j = 0; j < NTW; ++j) {
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    stg[(8 * (i >> 2) + 4 * (lane >> 5) + (i & 3)) * 32 + (lane & 31)] = f32_to_bf16_rne(acc[j][i]);
                wave_sync_lds();
                const uint4 v0 = *reinterpret_cast<const uint4*>(&stg[orow * 32 + ocol]);
                const uint4 v1 = *reinterpret_cast<const uint4*>(&stg[orow * 32 + ocol + 8]);
                const int gn = (wc * NTW + j) * 32 + ocol;
                if (gm < M) {
                    uint16_t* q = static_cast<uint16_t*>(Cv) + (int64_t)gm * ldc + gn;
                    if (gn + 16 <= N) {
                        *reinterpret_cast<uint4*>(q) = v0;
                        *reinterpret_cast<uint4*>(q + 8) = v1;
                    } else {
                        const uint32_t w[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
                        for (int c = 0; c < 16 && gn + c < N; ++c) q[c] = (uint16_t)(w[c >> 1] >> (16 * (c & 1)));
                    }
                }
                wave_sync_lds();  // the next block reuses the wave's staging area
            }
            __syncthreads();  // the staging areas are the next tile's A stage
            return;
        }
#pragma unroll
        for (int j = 0; j < NTW; ++j) {
            const int gn = (wc * NTW + j) * 32 + (lane & 31);
            if (gn < N) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int gm = m0 + wr * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                    if (gm < M) {
                        if constexpr (OUTBF) {  // bf16 output (round to nearest even), overwrite
                            static_cast<uint16_t*>(Cv)[(int64_t)gm * ldc + gn] = f32_to_bf16_rne(acc[j][i]);
                        } else {
                            float* p = static_cast<float*>(Cv) + (int64_t)gm * ldc + gn;
                            *p = accumulate ? *p + acc[j][i] : acc[j][i];
                        }
                    }
                }
            }
        }
    };
    // Rows of A contiguous in k with 16-byte rows and at most PF k steps: every
    // k step of the next tile is loaded into registers while this tile runs (the
    // register of step s is refilled as soon as step s is staged), so a tile
    // waits on one load latency, not on one per k step.
    constexpr int PF = 8;          // k steps held in registers (K ≤ 256)
    constexpr int NVW = TR / 64;   // 16-byte vectors per thread per k step
    const int nsteps = KP / TK;
    if (sak == 1 && va == 8 && nsteps <= PF) {
        uint4 pre[PF][NVW];
        const int kk = (tid & 3) * 8;  // this thread's k offset in every step (idx & 3 is tid & 3)
        // rows of the tile being fetched, at this thread's k offset; rows past M read
        // row M-1 (their results are never stored)
        const uint16_t* rp[NVW];
        auto rows_of = [&](int m0) {
#pragma unroll
            for (int v = 0; v < NVW; ++v) {
                const int gm = m0 + ((tid + v * 256) >> 2);
                rp[v] = A + (int64_t)(gm < M ? gm : M - 1) * sam + kk;
            }
        };
        // every step but the last lies below K (KP is K rounded up to TK)
        auto fetch = [&](int st, uint4 (&dst)[NVW]) {
            const int gk = st * TK + kk;
#pragma unroll
            for (int v = 0; v < NVW; ++v) {
                if (st + 1 < nsteps || gk + 8 <= K) {
                    dst[v] = *reinterpret_cast<const uint4*>(rp[v] + st * TK);
                } else {  // the k tail: zero past K
                    uint16_t t[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) t[i] = gk + i < K ? rp[v][st * TK + i] : (uint16_t)0;
                    dst[v] = make_uint4(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16),
                                        t[4] | ((uint32_t)t[5] << 16), t[6] | ((uint32_t)t[7] << 16));
                }
            }
        };
        if (blockIdx.x < (unsigned)ntiles) {
            rows_of((int)blockIdx.x * TR);
#pragma unroll
            for (int st = 0; st < PF; ++st)
                if (st < nsteps) fetch(st, pre[st]);
        }
        for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
            const int m0 = tile * TR;
            const int next = tile + (int)gridDim.x;
            const bool more = next < ntiles;
            if (more) rows_of(next * TR);
            f32x16 acc[NTW];
#pragma unroll
            for (int j = 0; j < NTW; ++j)
#pragma unroll
                for (int i = 0; i < 16; ++i) acc[j][i] = 0.0f;
#pragma unroll
            for (int st = 0; st < PF; ++st) {
                if (st < nsteps) {
#pragma unroll
                    for (int v = 0; v < NVW; ++v)
                        *reinterpret_cast<uint4*>(&sA[((tid + v * 256) >> 2) * TKP + kk]) = pre[st][v];
                    if (more) fetch(st, pre[st]);
                    __syncthreads();
                    mma_step(acc, st * TK);
                    __syncthreads();
                }
            }
            store_tile(acc, m0);
        }
        return;
    }
    // Longer rows (a multiple of PF k steps) with one column group per wave: the k
    // steps of the block's tiles form one sequence and the step PF ahead is loaded
    // into the register slot of the step just staged, so every step waits on a load
    // issued PF steps earlier (x·W at 19,717 × 500 → 64: one tile of 16 steps per
    // block).  Only NTW = 1 carries this path: a second accumulator set would cost
    // the wider kernels their occupancy.
    if constexpr (NTW == 1) if (sak == 1 && va == 8 && nsteps % PF == 0) {
        uint4 pre[PF][NVW];
        const int kk = (tid & 3) * 8;
        const int my_tiles = (int)blockIdx.x < ntiles ? (ntiles - 1 - (int)blockIdx.x) / (int)gridDim.x + 1 : 0;
        const int nitems = my_tiles * nsteps;
        // item i: k step i % nsteps of the block's tile i / nsteps (rows past M read row
        // M-1, never stored; only a tile's last step reaches past K, zero there)
        auto fetch = [&](int i, uint4 (&dst)[NVW]) {
            const int tq = i / nsteps, st = i - tq * nsteps;
            const int m0 = ((int)blockIdx.x + tq * (int)gridDim.x) * TR;
            const int gk = st * TK + kk;
#pragma unroll
            for (int v = 0; v < NVW; ++v) {
                const int gm = m0 + ((tid + v * 256) >> 2);
                const uint16_t* rp = A + (int64_t)(gm < M ? gm : M - 1) * sam + gk;
                if (st + 1 < nsteps || gk + 8 <= K) {
                    dst[v] = *reinterpret_cast<const uint4*>(rp);
                } else {
                    uint16_t t[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) t[e] = gk + e < K ? rp[e] : (uint16_t)0;
                    dst[v] = make_uint4(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16),
                                        t[4] | ((uint32_t)t[5] << 16), t[6] | ((uint32_t)t[7] << 16));
                }
            }
        };
#pragma unroll
        for (int j = 0; j < PF; ++j)
            if (j < nitems) fetch(j, pre[j]);
        for (int tq = 0; tq < my_tiles; ++tq) {
            f32x16 acc[NTW];
#pragma unroll
            for (int q = 0; q < NTW; ++q)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[q][e] = 0.0f;
            // a tile's steps come in groups of PF: slot j holds step base + j
            for (int base = 0; base < nsteps; base += PF) {
#pragma unroll
                for (int j = 0; j < PF; ++j) {
                    const int i = tq * nsteps + base + j;
#pragma unroll
                    for (int v = 0; v < NVW; ++v)
                        *reinterpret_cast<uint4*>(&sA[((tid + v * 256) >> 2) * TKP + kk]) = pre[j][v];
                    if (i + PF < nitems) fetch(i + PF, pre[j]);
                    __syncthreads();
                    mma_step(acc, (base + j) * TK);
                    __syncthreads();
                }
            }
            store_tile(acc, ((int)blockIdx.x + tq * (int)gridDim.x) * TR);
        }
        return;
    }
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int m0 = tile * TR;
        f32x16 acc[NTW];
#pragma unroll
        for (int j = 0; j < NTW; ++j)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[j][i] = 0.0f;
        for (int k0 = 0; k0 < KP; k0 += TK) {
            if (sak == 1 && va > 1) {  // rows of A contiguous in k: va-wide loads
                const int per_row = TK / va;
                for (int idx = tid; idx < TR * per_row; idx += 256) {
                    const int r = idx / per_row, kk = (idx - r * per_row) * va;
                    const int gm = m0 + r, gk = k0 + kk;
                    uint16_t* dst = &sA[r * TKP + kk];
                    const uint16_t* src = A + (int64_t)gm * sam + gk;
                    if (gm < M && gk + va <= K) {
                        if (va == 8)
                            *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(src);
                        else
                            *reinterpret_cast<uint2*>(dst) = *reinterpret_cast<const uint2*>(src);
                    } else {
                        for (int i = 0; i < va; ++i) dst[i] = (gm < M && gk + i < K) ? src[i] : (uint16_t)0;
                    }
                }
            } else {  // any strides: element loads, k fastest when sak == 1, else m fastest
                for (int idx = tid; idx < TR * TK; idx += 256) {
                    int r, kk;
                    if (sak == 1) {
                        r = idx / TK;
                        kk = idx - r * TK;
                    } else {
                        kk = idx / TR;
                        r = idx - kk * TR;
                    }
                    const int gm = m0 + r, gk = k0 + kk;
                    sA[r * TKP + kk] = (gm < M && gk < K) ? A[(int64_t)gm * sam + (int64_t)gk * sak] : (uint16_t)0;
                }
            }
            __syncthreads();
            mma_step(acc, k0);
            __syncthreads();
        }
        store_tile(acc, m0);
    }
}

// fp32 [M×N] (pitch N) → bf16 rows of pitch ldc
__global__ void k_rows_to_bf16(int M, int N, const float* __restrict__ X, uint16_t* __restrict__ Y, int64_t ldc) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)M * N) return;
    const int m = (int)(idx / N), n = (int)(idx % N);
    Y[(int64_t)m * ldc + n] = f32_to_bf16_rne(X[idx]);
}

// ------------------------------------------------------- reduction GEMM
// C[M×N] = Σ_k A(m,k)·B(k,n) over a long K with M, N ≤ 256 (the weight
// gradient xᵀ·dY): block b reduces k ∈ [b·ks, (b+1)·ks) into part[b] with the
// whole M×N output in MFMA accumulators (TPW 32×32 tiles per wave).  Operand
// slabs of 32 k go through LDS; each thread fetches its share of the next slab
// as 8-byte vectors along the operand's contiguous dimension (m for A = xᵀ,
// n for B = dY, or k) into registers while the current slab is multiplied,
// so the loads of one slab are all in flight together and overlap the MFMAs.
// The partials are summed in block order afterwards (deterministic).
//
// vector along the contiguous dimension: 4 bf16 (8 bytes), or 8 (16 bytes) in mode 3
struct Slab {
    int along;     // 0: vectors along the row dimension (m or n), 1: along k, 2: element loads,
                   // 3: as 0 with 16-byte vectors
    int64_t srow;  // stride of the row dimension
    int64_t sk;    // stride of k
};


// mode 3: elements (r..r+7, k) of a row-contiguous operand, zero past R or ke
__device__ __forceinline__ uint4 load8(const uint16_t* __restrict__ P, const Slab& sl, int r, int gk, int R, int ke) {
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (r < R && gk < ke) {
        const uint16_t* p = P + r + (int64_t)gk * sl.sk;
        if (r + 8 <= R) {
            v = *reinterpret_cast<const uint4*>(p);
        } else {
            uint16_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = 0; i < 8 && r + i < R; ++i) t[i] = p[i];
            v = make_uint4(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16), t[4] | ((uint32_t)t[5] << 16),
                           t[6] | ((uint32_t)t[7] << 16));
        }
    }
    return v;
}

// element (r, k) of a slab operand, 4 consecutive along the vector dimension
__device__ __forceinline__ uint2 load4(const uint16_t* __restrict__ P, const Slab& sl, int r, int gk, int R, int ke) {
    uint2 v = make_uint2(0u, 0u);
    if (sl.along == 2) {  // any strides: four element loads along the row dimension
        uint16_t t[4] = {0, 0, 0, 0};
        if (gk < ke) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (r + i < R) t[i] = P[(int64_t)(r + i) * sl.srow + (int64_t)gk * sl.sk];
        }
        return make_uint2(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16));
    }
    if (sl.along == 0) {  // r..r+3 contiguous (srow == 1)
        if (r < R && gk < ke) {
            const uint16_t* p = P + r + (int64_t)gk * sl.sk;
            if (r + 4 <= R) {
                v = *reinterpret_cast<const uint2*>(p);
            } else {
                uint16_t t[4] = {0, 0, 0, 0};
                for (int i = 0; i < 4 && r + i < R; ++i) t[i] = p[i];
                v = make_uint2(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16));
            }
        }
    } else {  // k..k+3 contiguous (sk == 1)
        if (r < R && gk < ke) {
            const uint16_t* p = P + (int64_t)r * sl.srow + gk;
            if (gk + 4 <= ke) {
                v = *reinterpret_cast<const uint2*>(p);
            } else {
                uint16_t t[4] = {0, 0, 0, 0};
                for (int i = 0; i < 4 && gk + i < ke; ++i) t[i] = p[i];
                v = make_uint2(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16));
            }
        }
    }
    return v;
}

// slab vector q of this thread: (r, kk) of its first element
template <bool V16>
__device__ __forceinline__ void slab_pos(const Slab& sl, int v, int RP, int& r, int& kk) {
    if constexpr (V16) {  // mode 3: as mode 0, 8 rows per vector
        kk = v % RTK;
        r = (v / RTK) * 8;
        return;
    }
    if (sl.along == 0) {
        // consecutive lanes take consecutive k: the four transposing 2-byte LDS stores
        // of a wave then each fill one 128-byte row segment, bank-conflict free (lanes
        // along the row dimension made every store a 16-way conflict: xᵀ·dY with
        // 100×256 outputs over 2.4M rows 1498 → 952 us, and 916 us with 64-k stages)
        kk = v % RTK;
        r = (v / RTK) * 4;
    } else if (sl.along == 2) {  // RP/4 vectors per k
        kk = v / (RP / 4);
        r = (v - kk * (RP / 4)) * 4;
    } else {  // TK/4 vectors per row
        r = v / (RTK / 4);
        kk = (v - r * (RTK / 4)) * 4;
    }
}

// mode 3: eight transposing 2-byte stores (consecutive lanes on consecutive k)
__device__ __forceinline__ void store8(uint16_t* __restrict__ S, int r, int kk, uint4 v) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) S[(r + i) * RTKP + kk] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
}

__device__ __forceinline__ void store4(uint16_t* __restrict__ S, const Slab& sl, int r, int kk, uint2 v) {
    const uint16_t e[4] = {(uint16_t)(v.x & 0xffffu), (uint16_t)(v.x >> 16), (uint16_t)(v.y & 0xffffu),
                           (uint16_t)(v.y >> 16)};
    if (sl.along != 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) S[(r + i) * RTKP + kk] = e[i];
    } else {
        *reinterpret_cast<uint2*>(&S[r * RTKP + kk]) = v;
    }
}

constexpr int RW = 16;  // waves per reduction block: ≤ 4 accumulator tiles per wave for 256×256

template <int TPW, bool VA, bool VB>
__global__ __launch_bounds__(64 * RW) void k_gemm_reduce(int M, int N, int K, int ks, const uint16_t* __restrict__ A,
                                                     Slab sa, const uint16_t* __restrict__ B, Slab sb,
                                                     float* __restrict__ part) {
    extern __shared__ __attribute__((aligned(16))) uint16_t lds[];
    const int MT = (M + 31) / 32, NT = (N + 31) / 32;
    const int MP = 32 * MT, NP = 32 * NT;
    uint16_t* sA = lds;              // [MP][RTKP]
    uint16_t* sB = lds + MP * RTKP;  // [NP][RTKP]
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int kb = blockIdx.x * ks, ke = min(K, kb + ks);
    // vectors per slab: MP·RTK/4 = 512·MT for A, 512·NT for B → ≤ 4 per thread each;
    // an operand in mode 3 (VA, VB) takes 16-byte vectors, half as many → ≤ 2 per thread
    constexpr int NT_ = 64 * RW;
    constexpr int NQA = VA ? 2 : 4, NQB = VB ? 2 : 4;
    using VecA = typename std::conditional<VA, uint4, uint2>::type;
    using VecB = typename std::conditional<VB, uint4, uint2>::type;
    const int nva = (VA ? 256 : 512) * MT, nvb = (VB ? 256 : 512) * NT;
    VecA ra[NQA];
    VecB rb[NQB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NQA; ++q) {
            const int v = tid + NT_ * q;
            if (v < nva) {
                int r, kk;
                slab_pos<VA>(sa, v, MP, r, kk);
                if constexpr (VA)
                    ra[q] = load8(A, sa, r, k0 + kk, M, ke);
                else
                    ra[q] = load4(A, sa, r, k0 + kk, M, ke);
            }
        }
#pragma unroll
        for (int q = 0; q < NQB; ++q) {
            const int v = tid + NT_ * q;
            if (v < nvb) {
                int r, kk;
                slab_pos<VB>(sb, v, NP, r, kk);
                if constexpr (VB)
                    rb[q] = load8(B, sb, r, k0 + kk, N, ke);
                else
                    rb[q] = load4(B, sb, r, k0 + kk, N, ke);
            }
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int q = 0; q < NQA; ++q) {
            const int v = tid + NT_ * q;
            if (v < nva) {
                int r, kk;
                slab_pos<VA>(sa, v, MP, r, kk);
                if constexpr (VA)
                    store8(sA, r, kk, ra[q]);
                else
                    store4(sA, sa, r, kk, ra[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < NQB; ++q) {
            const int v = tid + NT_ * q;
            if (v < nvb) {
                int r, kk;
                slab_pos<VB>(sb, v, NP, r, kk);
                if constexpr (VB)
                    store8(sB, r, kk, rb[q]);
                else
                    store4(sB, sb, r, kk, rb[q]);
            }
        }
    };
    f32x16 acc[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[j][i] = 0.0f;
    if (kb < ke) fetch(kb);
    for (int k0 = kb; k0 < ke; k0 += RTK) {
        commit();
        __syncthreads();
        if (k0 + RTK < ke) fetch(k0 + RTK);  // next slab in flight during the MFMAs
#pragma unroll
        for (int kq = 0; kq < RTK; kq += 16) {
            const int koff = kq + (lane >> 5) * 8;
#pragma unroll
            for (int j = 0; j < TPW; ++j) {
                const int t = wid * TPW + j;
                if (t < MT * NT) {
                    const int tm = t / NT, tn = t - tm * NT;
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[(tm * 32 + (lane & 31)) * RTKP + koff]);
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[(tn * 32 + (lane & 31)) * RTKP + koff]);
                    acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    float* P = part + (int64_t)blockIdx.x * M * N;
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        const int t = wid * TPW + j;
        if (t < MT * NT) {
            const int tm = t / NT, tn = t - tm * NT;
            const int gn = tn * 32 + (lane & 31);
            if (gn < N) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int gm = tm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                    if (gm < M) P[(int64_t)gm * N + gn] = acc[j][i];
                }
            }
        }
    }
}

// C[m, n] (=|+=) Σ_z part[z][m·N + n].  A workgroup takes OPB consecutive
// outputs; 256/OPB lanes share each output, lane phase q summing partials
// q, q + 256/OPB, ... in order, and the phases' sums are then added in phase
// order (deterministic for given splits and OPB).  A lane alone per output
// would chain all `splits` loads: 62–512 of them in the GCN's products.
template <int OPB>
__global__ __launch_bounds__(256) void k_splitk_reduce(int M, int N, int splits, const float* __restrict__ part,
                                                       float* __restrict__ C, int64_t ldc, int accumulate) {
    constexpr int PH = 256 / OPB;
    __shared__ float sum[PH][OPB];
    const int o = threadIdx.x % OPB, q = threadIdx.x / OPB;
    const int64_t MN = (int64_t)M * N;
    const int64_t idx = blockIdx.x * (int64_t)OPB + o;
    float acc = 0.0f;
    if (idx < MN) {
#pragma unroll 4
        for (int z = q; z < splits; z += PH) acc += part[(int64_t)z * MN + idx];
    }
    sum[q][o] = acc;
    __syncthreads();
    if (q == 0 && idx < MN) {
        float t = sum[0][o];
#pragma unroll
        for (int r = 1; r < PH; ++r) t += sum[r][o];
        float* p = C + (idx / N) * ldc + idx % N;
        *p = accumulate ? *p + t : t;
    }
}

// the split-K sum of `splits` partial [M×N] products into C: 16 outputs per
// workgroup (16 lanes each) until the outputs alone fill 2 workgroups per CU,
// then 64 (4 lanes each)
void splitk_reduce(int M, int N, int splits, const float* part, float* C, int64_t ldc, int accumulate,
                   hipStream_t s) {
    const int64_t n = (int64_t)M * N;
    if (n / 64 >= 2 * device_cus())
        hipLaunchKernelGGL(k_splitk_reduce<64>, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s, M, N, splits, part,
                           C, ldc, accumulate);
    else
        hipLaunchKernelGGL(k_splitk_reduce<16>, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, s, M, N, splits, part,
                           C, ldc, accumulate);
}

}  // namespace
}  // namespace armsp

namespace armsp {
namespace {
int gemm_f32_out(int M, int N, int Kd, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                 int64_t sbn, float* C, int64_t ldc, int accumulate, hipStream_t s);
int gemm_tall(int M, int N, int Kd, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
              int64_t sbn, void* C, bool outbf, int64_t ldc, int accumulate, hipStream_t s, bool& done);
}  // namespace
}  // namespace armsp

extern "C" int armsp_gemm_bf16(int M, int N, int Kd, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
                               int64_t sbk, int64_t sbn, float* C, int64_t ldc, int accumulate, void* stream) {
    using namespace armsp;
    if (M < 0 || N < 0 || Kd < 0 || ldc < N) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0) return ARMSP_SUCCESS;
    if (!A || !B || !C) return ARMSP_ERROR_INVALID_VALUE;
    return gemm_f32_out(M, N, Kd, A, sam, sak, B, sbk, sbn, C, ldc, accumulate, as_stream(stream));
}

extern "C" int armsp_gemm_bf16_ex(int M, int N, int Kd, const uint16_t* A, int64_t sam, int64_t sak,
                                  const uint16_t* B, int64_t sbk, int64_t sbn, void* C, int c_dtype, int64_t ldc,
                                  int accumulate, void* stream) {
    using namespace armsp;
    if (c_dtype == ARMSP_DTYPE_F32)
        return armsp_gemm_bf16(M, N, Kd, A, sam, sak, B, sbk, sbn, static_cast<float*>(C), ldc, accumulate, stream);
    if (c_dtype != ARMSP_DTYPE_BF16 || accumulate) return ARMSP_ERROR_INVALID_VALUE;
    if (M < 0 || N < 0 || Kd < 0 || ldc < N) return ARMSP_ERROR_INVALID_VALUE;
    if (M == 0 || N == 0) return ARMSP_SUCCESS;
    if (!A || !B || !C) return ARMSP_ERROR_INVALID_VALUE;
    hipStream_t s = as_stream(stream);
    bool done = false;
    int st = gemm_tall(M, N, Kd, A, sam, sak, B, sbk, sbn, C, true, ldc, 0, s, done);
    if (st || done) return st;
    // other shapes: fp32 into a workspace, then rounded into C
    void* ws = nullptr;
    st = ws_alloc(&ws, (size_t)M * N * sizeof(float), s);
    if (st) return st;
    st = gemm_f32_out(M, N, Kd, A, sam, sak, B, sbk, sbn, static_cast<float*>(ws), N, 0, s);
    if (!st) {
        const int64_t n = (int64_t)M * N;
        hipLaunchKernelGGL(k_rows_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, M, N,
                           static_cast<const float*>(ws), static_cast<uint16_t*>(C), ldc);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = hip_fail(e);
    }
    ws_free(ws, s);
    return st;
}

namespace armsp {
namespace {
// tall: M ≫ K, N — B resident in LDS, A streamed once; done = false when the shape does not fit
int gemm_tall(int M, int N, int Kd, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
              int64_t sbn, void* C, bool outbf, int64_t ldc, int accumulate, hipStream_t s, bool& done) {
    done = false;
    if (!(M >= 4096 && N <= 256 && Kd <= 1024 && Kd > 0)) return ARMSP_SUCCESS;  // and B fits in LDS (below)
    const int KP = (Kd + TK - 1) / TK * TK;
    int WR = 2, NTW = 4;
    if (N <= 32) {
        WR = 4;
        NTW = 1;
    } else if (N <= 64) {
        NTW = 1;
    } else if (N <= 128) {
        NTW = 2;
    }
    const int NP = 32 * NTW * (4 / WR);
    // the A stage doubles as the bf16 output staging: 2 KB per wave
    const size_t stage = (size_t)32 * WR * TKP > 4096 ? (size_t)32 * WR * TKP : 4096;
    const size_t lds = ((size_t)NP * (KP + 8) + stage) * sizeof(uint16_t);
    if (lds > 96 * 1024) return ARMSP_SUCCESS;
    int va = 1;
    if (sak == 1 && sam % 8 == 0 && ((uintptr_t)A % 16) == 0)
        va = 8;
    else if (sak == 1 && sam % 4 == 0 && ((uintptr_t)A % 8) == 0)
        va = 4;
    const int ntiles = (M + 32 * WR - 1) / (32 * WR);
    const int nblk = ntiles < 2 * device_cus() ? ntiles : 2 * device_cus();
    TimedLaunch t("gemm", s);
#define ARMSP_TALL(WW, NN)                                                                                         \
    do {                                                                                                           \
        if (outbf)                                                                                                 \
            hipLaunchKernelGGL((k_gemm_tall<WW, NN, true>), dim3(nblk), dim3(256), lds, s, M, N, Kd, KP, A, sam, sak, \
                               B, sbk, sbn, C, ldc, accumulate, va);                                               \
        else                                                                                                       \
            hipLaunchKernelGGL((k_gemm_tall<WW, NN, false>), dim3(nblk), dim3(256), lds, s, M, N, Kd, KP, A, sam,  \
                               sak, B, sbk, sbn, C, ldc, accumulate, va);                                          \
    } while (0)
    if (WR == 4)
        ARMSP_TALL(4, 1);
    else if (NTW == 1)
        ARMSP_TALL(2, 1);
    else if (NTW == 2)
        ARMSP_TALL(2, 2);
    else
        ARMSP_TALL(2, 4);
#undef ARMSP_TALL
    ARMSP_LAUNCHED();
    t.stop();
    done = true;
    return ARMSP_SUCCESS;
}

int gemm_f32_out(int M, int N, int Kd, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                 int64_t sbn, float* C, int64_t ldc, int accumulate, hipStream_t s) {
    {
        bool done = false;
        const int st = gemm_tall(M, N, Kd, A, sam, sak, B, sbk, sbn, C, false, ldc, accumulate, s, done);
        if (st || done) return st;
    }
    // operand slabs as 8-byte vectors: along the row dimension when it is contiguous
    // (A = xᵀ: sam == 1; B = dY: sbn == 1), else along k; 8-byte aligned
    auto slab = [](const uint16_t* P, int64_t srow, int64_t sk, Slab& sl) {
        const bool al = ((uintptr_t)P % 8) == 0;
        if (((uintptr_t)P % 16) == 0 && srow == 1 && sk % 8 == 0)
            sl = Slab{3, srow, sk};
        else if (al && srow == 1 && sk % 4 == 0)
            sl = Slab{0, srow, sk};
        else if (al && sk == 1 && srow % 4 == 0)
            sl = Slab{1, srow, sk};
        else
            sl = Slab{2, srow, sk};  // element loads
        return true;
    };
    Slab sa{}, sb{};
    if (Kd >= 65536 && M <= 256 && N <= 256 && slab(A, sam, sak, sa) && slab(B, sbn, sbk, sb)) {
        const int MT = (M + 31) / 32, NT = (N + 31) / 32;
        const int tpw = (MT * NT + RW - 1) / RW;
        if (tpw <= 4) {
            int nblk = 2 * device_cus();
            int ks = (Kd + nblk - 1) / nblk;
            ks = (ks + RTK - 1) / RTK * RTK;
            nblk = (Kd + ks - 1) / ks;
            const size_t lds = (size_t)32 * (MT + NT) * RTKP * sizeof(uint16_t);
            void* ws = nullptr;
            int st = ws_alloc(&ws, (size_t)nblk * M * N * sizeof(float), s);
            if (st) return st;
            float* part = static_cast<float*>(ws);
            TimedLaunch t("gemm", s);
            // 16-byte vectors for an operand that is row-contiguous with 16-byte aligned rows
            const bool va16 = sa.along == 3, vb16 = sb.along == 3;
#define ARMSP_RED(TT, AA, BB)                                                                                      \
    hipLaunchKernelGGL((k_gemm_reduce<TT, AA, BB>), dim3(nblk), dim3(64 * RW), lds, s, M, N, Kd, ks, A, sa, B, sb, \
                       part)
#define ARMSP_RED_T(AA, BB)       \
    do {                          \
        if (tpw <= 1)             \
            ARMSP_RED(1, AA, BB); \
        else if (tpw <= 2)        \
            ARMSP_RED(2, AA, BB); \
        else                      \
            ARMSP_RED(4, AA, BB); \
    } while (0)
            if (va16 && vb16)
                ARMSP_RED_T(true, true);
            else if (va16)
                ARMSP_RED_T(true, false);
            else if (vb16)
                ARMSP_RED_T(false, true);
            else
                ARMSP_RED_T(false, false);
#undef ARMSP_RED_T
#undef ARMSP_RED
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) {
                splitk_reduce(M, N, nblk, part, C, ldc, accumulate, s);
                e = hipGetLastError();
            }
            ws_free(ws, s);
            if (e != hipSuccess) return hip_fail(e);
            t.stop();
            return ARMSP_SUCCESS;
        }
    }
    const int gx = (M + BM - 1) / BM, gy = (N + BN - 1) / BN;
    if (gy > 65535) return ARMSP_ERROR_NOT_SUPPORTED;
    // split K when the tile grid is small and K is long: up to 4 workgroups per CU,
    // each at least 4 k steps (a workgroup's steps run one global-load latency
    // apart; 19,717-deep PubMed products: 5 steps each instead of 10)
    int splits = 1;
    const int tiles = gx * gy, target = 4 * device_cus();
    while (tiles * splits < target && Kd / (splits * 2) >= 4 * BK && splits < 128) splits *= 2;
    int kchunk = (Kd + splits - 1) / splits;
    kchunk = (kchunk + BK - 1) / BK * BK;
    if (kchunk == 0) kchunk = BK;
    splits = Kd > 0 ? (Kd + kchunk - 1) / kchunk : 1;
    const int amode = tile_mode(A, sam, sak), bmode = tile_mode(B, sbn, sbk);
    TimedLaunch t("gemm", s);
    if (splits == 1) {
        hipLaunchKernelGGL(k_gemm_bf16, dim3(gx, gy, 1), dim3(256), 0, s, M, N, Kd, kchunk, A, sam, sak, amode, B, sbk,
                           sbn, bmode, C, ldc, (int64_t)0, accumulate);
        ARMSP_LAUNCHED();
        t.stop();
        return ARMSP_SUCCESS;
    }
    void* ws = nullptr;
    int st = ws_alloc(&ws, (size_t)splits * M * N * sizeof(float), s);
    if (st) return st;
    float* part = static_cast<float*>(ws);
    hipLaunchKernelGGL(k_gemm_bf16, dim3(gx, gy, splits), dim3(256), 0, s, M, N, Kd, kchunk, A, sam, sak, amode, B, sbk,
                       sbn, bmode, part, (int64_t)N, (int64_t)M * N, 0);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        splitk_reduce(M, N, splits, part, C, ldc, accumulate, s);
        e = hipGetLastError();
    }
    ws_free(ws, s);
    if (e != hipSuccess) return hip_fail(e);
    t.stop();
    return ARMSP_SUCCESS;
}
}  // namespace
}  // namespace armsp
