// Internal helpers shared by the HIP translation units of libarmsparse.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "armsp.h"

namespace armsp {

constexpr int kWave = 64;

// ------------------------------------------------------------------ runtime
// Records the HIP error for armsp_last_hip_error() and maps it to a status.
int hip_fail(hipError_t e);
#define ARMSP_HIP(call)                                 \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return ::armsp::hip_fail(e_); \
    } while (0)
#define ARMSP_LAUNCHED()                                         \
    do {                                                         \
        hipError_t e_ = hipGetLastError();                       \
        if (e_ != hipSuccess) return ::armsp::hip_fail(e_);      \
    } while (0)

// The stream of an entry point.  Every entry point calls this before its
// launches; it also clears an error that an earlier runtime call on this
// thread left behind (hipGetLastError reports the last error of any call), so
// the launch checks below report only this entry point's launches.
inline hipStream_t as_stream(void* s) {
    (void)hipGetLastError();
    return reinterpret_cast<hipStream_t>(s);
}

// True when ARMSP_FAIL_ALLOC asks every device allocation to fail (tests of
// the failure paths, also without a GPU).
bool alloc_injected();

// Stream-ordered workspace from the device's default memory pool.  The pool's
// release threshold is raised at first use so freed blocks stay cached.
int ws_alloc(void** p, size_t bytes, hipStream_t s);
void ws_free(void* p, hipStream_t s);
int workspace_reserve(size_t bytes, size_t* reserved);

// Optional per-launch timing (armsp_timing_enable).  `TimedLaunch t(family, s)`
// records a start event; t.stop() records the stop event.
bool timing_on();
struct TimedLaunch {
    void* rec = nullptr;
    hipStream_t stream_ = nullptr;
    TimedLaunch(const char* family, hipStream_t s);
    void stop();
};

// While alive, TimedLaunch records nothing on this thread (entry points that
// must not create HIP objects).
struct TimingPause {
    TimingPause();
    ~TimingPause();
};

// The installed schedule selector's configuration for an automatic call of
// operation op (0 SpMM, 1 SDDMM), if it has one (true: *out set).  may_compute:
// a first choice for this matrix may allocate and synchronise `s`.
bool selector_config(int op, int M, int nnz, int N, int esz, const int* rowptr, hipStream_t s, bool may_compute,
                     armsp_config* out);

// Number of compute units on the current device (cached).
int device_cus();

// part[w] = rows consumed by the merge path of (row ends, nonzeros) at diagonal
// min(w·chunk, M + nnz), w = 0..nwaves (nwaves + 1 entries), on stream s.
int merge_partition(int M, int nnz, int chunk, int64_t nwaves, const int* rowptr, int* part, hipStream_t s);

// ------------------------------------------------------------ device utils
__device__ __forceinline__ int wave_lane() { return __lane_id(); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int read_lane(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ float read_lane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// OCP fp8 e4m3 (e4m3fn: bias 7, no infinities, 0x7f/0xff = NaN) → fp32,
// built from integer fields so it is exact for normals and subnormals.
__device__ __forceinline__ float e4m3_to_f32(uint32_t b) {
    uint32_t sign = (b & 0x80u) << 24;
    uint32_t e = (b >> 3) & 0xfu, m = b & 7u;
    float mag;
    if (e == 0) {
        mag = (float)m * 0x1p-9f;
    } else if (e == 15 && m == 7) {
        mag = __uint_as_float(0x7fc00000u);
    } else {
        mag = __uint_as_float(((e + 120u) << 23) | (m << 20));
    }
    return __uint_as_float(__float_as_uint(mag) | sign);
}

// Four OCP e4m3 codes (bytes of w, lowest first) → fp32 with the gfx950
// packed conversion (two codes per instruction).
__device__ __forceinline__ void e4m3x4_to_f32(uint32_t w, float (&o)[4]) {
    const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)w, false);
    const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)w, true);
    o[0] = lo[0];
    o[1] = lo[1];
    o[2] = hi[0];
    o[3] = hi[1];
}

// x[l] + x[l ^ 32] in every lane l: v_permlane32_swap of x with itself gives
// one copy whose upper half holds the lower lanes and one whose lower half holds
// the upper lanes, so the two results always add to the pair's sum (a VALU op;
// a 32-lane shuffle would go through the LDS crossbar).
__device__ __forceinline__ float half_sum(float x) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Merge-path search: for diagonal d over (row ends rowptr[1..M], nnz 0..nnz-1)
// returns the number of rows consumed; nnz consumed is d - result.
__device__ __forceinline__ int merge_path_search(int d, const int* __restrict__ rowptr, int M, int nnz) {
    int lo = d - nnz > 0 ? d - nnz : 0;
    int hi = d < M ? d : M;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (rowptr[mid + 1] <= d - mid - 1)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// The same search done by a whole wave: each round the 64 lanes probe 64
// evenly spaced candidates and a ballot narrows the range 64-fold, so a search
// over a million rows takes 4 rounds of loads instead of 20 dependent ones.
// Every lane of the wave must call it with the same d; all get the result.
__device__ __forceinline__ int merge_path_search_wave(int d, const int* __restrict__ rowptr, int M, int nnz) {
    const int lane = __lane_id();
    int lo = d - nnz > 0 ? d - nnz : 0;  // answer in [lo, hi]
    int hi = d < M ? d : M;
    while (lo < hi) {
        const int span = hi - lo;
        const int step = (span + 63) >> 6;  // lanes probe lo + l·step, l·step < span
        const int i = lo + lane * step;
        // P(i): row i ends at or before the path reaches it; true on a prefix of i
        const bool p = (lane * step < span) && rowptr[i + 1] <= d - i - 1;
        const int t = __popcll(__ballot(p));
        if (step == 1) return lo + t;
        if (t == 0) {
            hi = lo;  // P(lo) false
        } else {
            const int nlo = lo + (t - 1) * step + 1;
            hi = lo + t * step < hi ? lo + t * step : hi;
            lo = nlo;
        }
    }
    return lo;
}

// LDS writes of a wave visible to its own later LDS reads (no workgroup barrier).
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row boundary w of a precomputed merge-path partition (merge_partition) for
// diagonal d, clamped to the interval a search could return so that a
// malformed rowptr cannot carry a kernel outside its chunk.
__device__ __forceinline__ int partition_at(const int* __restrict__ part, int w, int d, int M, int nnz) {
    const int lo = d - nnz > 0 ? d - nnz : 0;
    const int hi = d < M ? d : M;
    const int i = part[w];
    return i < lo ? lo : (i > hi ? hi : i);
}

// Number of rows whose end is ≤ j (rows fully before nonzero j).
__device__ __forceinline__ int rows_ending_by(int j, const int* __restrict__ rowptr, int M) {
    int lo = 0, hi = M;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (rowptr[mid + 1] <= j)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// rows_ending_by done by a whole wave, 64 probes per round.
__device__ __forceinline__ int rows_ending_by_wave(int j, const int* __restrict__ rowptr, int M) {
    const int lane = __lane_id();
    int lo = 0, hi = M;  // answer in [lo, hi]
    while (lo < hi) {
        const int span = hi - lo;
        const int step = (span + 63) >> 6;
        const int i = lo + lane * step;
        const bool p = (lane * step < span) && rowptr[i + 1] <= j;
        const int t = __popcll(__ballot(p));
        if (step == 1) return lo + t;
        if (t == 0) {
            hi = lo;
        } else {
            const int nlo = lo + (t - 1) * step + 1;
            hi = lo + t * step < hi ? lo + t * step : hi;
            lo = nlo;
        }
    }
    return lo;
}

}  // namespace armsp
