// Merge-path partition of a CSR matrix for the chunked merge kernels (SpMM,
// SpMV, SDDMM): part[w] = the rows consumed at diagonal min(w·chunk, M + nnz)
// of the merge of the M row ends with the nnz nonzeros, w = 0..nwaves.
//
// Row i's end is consumed at diagonal e(i) = rowptr[i+1] + i + 1, which grows
// strictly with i, so part[w] = #{i : e(i) <= w·chunk} and row i owns exactly
// the boundaries w with e(i-1) <= w·chunk < e(i) (e(-1) = 0; the extra thread
// i = M owns the rest, up to w = nwaves).  One thread per row reads two row
// pointers and writes the boundaries it owns: a streaming pass over rowptr
// instead of a binary search per boundary, and the same values
// merge_path_search gives.
#include "common.hpp"

namespace armsp {
namespace {

__global__ __launch_bounds__(256) void k_merge_partition(int M, int nnz, int chunk, int64_t nwaves,
                                                         const int* __restrict__ rowptr, int* __restrict__ part) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i > M) return;
    const int64_t lo = i == 0 ? 0 : (int64_t)rowptr[i] + i;
    const int64_t hi = i < M ? (int64_t)rowptr[i + 1] + i + 1 : nwaves * chunk + 1;
    // clamped so that a malformed rowptr cannot write outside part[0..nwaves]
    int64_t wb = lo <= 0 ? 0 : (lo + chunk - 1) / chunk;
    int64_t we = hi <= 0 ? 0 : (hi + chunk - 1) / chunk;
    if (we > nwaves + 1) we = nwaves + 1;
    for (int64_t w = wb; w < we; ++w) part[w] = (int)i;
}

}  // namespace

int merge_partition(int M, int nnz, int chunk, int64_t nwaves, const int* rowptr, int* part, hipStream_t s) {
    (void)nnz;
    const int64_t n = (int64_t)M + 1;
    hipLaunchKernelGGL(k_merge_partition, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, M, nnz, chunk, nwaves,
                       rowptr, part);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ARMSP_SUCCESS : hip_fail(e);
}

}  // namespace armsp
