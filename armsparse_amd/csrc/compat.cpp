// Host-pointer drop-ins with the reference's names (see armsparse_compat.h).
// Each shim validates the inputs on the host, stages them on the device,
// runs the library kernel on a private stream, and copies the result back.
#include "armsparse_compat.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "armsp.h"
#include "armsp_gen.h"
#include "common.hpp"

namespace {

[[noreturn]] void die(const char* fn, const char* why) {
    std::fprintf(stderr, "armsparse: %s: %s\n", fn, why);
    std::exit(EXIT_FAILURE);
}

void hip_or_die(const char* fn, hipError_t e) {
    if (e != hipSuccess) die(fn, hipGetErrorString(e));
}

void status_or_die(const char* fn, int st) {
    if (st != ARMSP_SUCCESS) die(fn, armsp_status_string(st));
}

// RAII device buffer
struct Dev {
    void* p = nullptr;
    Dev(const char* fn, size_t bytes) {
        if (bytes && armsp::alloc_injected()) die(fn, "device allocation failed");
        if (bytes) hip_or_die(fn, hipMalloc(&p, bytes));
    }
    ~Dev() {
        if (p) hipFree(p);
    }
    template <typename T>
    T* as() {
        return static_cast<T*>(p);
    }
};

// The reference kernels take any CSR whose row pointers start at 0 and never
// decrease and whose column indices are in range: toCSR writes each row's
// columns in increasing order, toCSR2 (SDMM.cpp:371-394) every odd row in
// decreasing order, and the kernels walk a row's entries in storage order,
// whatever it is.  `distinct` also rejects a column repeated within a row (the
// sparse B of sparse_matrix_mul3, whose row entries 64 lanes scatter at once).
void check_csr(const char* fn, int M, int K, const int* rowPtr, const int* colInd, bool distinct = false) {
    if (M < 0 || K < 0) die(fn, "negative dimension");
    if (!rowPtr) die(fn, "null rowPtr");
    if (rowPtr[0] != 0) die(fn, "rowPtr[0] != 0");
    for (int r = 0; r < M; ++r)
        if (rowPtr[r + 1] < rowPtr[r]) die(fn, "rowPtr is decreasing");
    const int nnz = rowPtr[M];
    if (nnz > 0 && !colInd) die(fn, "null colInd");
    for (int p = 0; p < nnz; ++p)
        if (colInd[p] < 0 || colInd[p] >= K) die(fn, "column index out of range");
    if (distinct && nnz > 0) {
        std::vector<int> seen((size_t)K, -1);  // the last row each column appeared in
        for (int r = 0; r < M; ++r)
            for (int p = rowPtr[r]; p < rowPtr[r + 1]; ++p) {
                if (seen[(size_t)colInd[p]] == r) die(fn, "column index repeated within a row");
                seen[(size_t)colInd[p]] = r;
            }
    }
}

int max_col(int nnz, const int* colInd) {
    int k = -1;
    for (int i = 0; i < nnz; ++i)
        if (colInd[i] > k) k = colInd[i];
    return k;
}

// Shared body of the CSR × dense shims.  ncols: columns updated (≤ N).
void spmm_shim(const char* fn, int* rowPtr, int* colInd, float* val, float* B, float* C, int M, int K, int N,
               int ncols, int flags) {
    if (N < 0) die(fn, "negative dimension");
    check_csr(fn, M, K, rowPtr, colInd);
    if (M == 0 || N == 0 || ncols == 0) return;
    if (!B || !C) die(fn, "null dense operand");
    const int nnz = rowPtr[M];
    Dev drp(fn, (size_t)(M + 1) * 4), dci(fn, (size_t)nnz * 4), dva(fn, (size_t)nnz * 4);
    Dev dB(fn, (size_t)K * N * 4), dC(fn, (size_t)M * N * 4);
    hipStream_t s;
    hip_or_die(fn, hipStreamCreate(&s));
    hip_or_die(fn, hipMemcpyAsync(drp.p, rowPtr, (size_t)(M + 1) * 4, hipMemcpyHostToDevice, s));
    if (nnz) {
        hip_or_die(fn, hipMemcpyAsync(dci.p, colInd, (size_t)nnz * 4, hipMemcpyHostToDevice, s));
        if (val) hip_or_die(fn, hipMemcpyAsync(dva.p, val, (size_t)nnz * 4, hipMemcpyHostToDevice, s));
        hip_or_die(fn, hipMemcpyAsync(dB.p, B, (size_t)K * N * 4, hipMemcpyHostToDevice, s));
    }
    hip_or_die(fn, hipMemcpyAsync(dC.p, C, (size_t)M * N * 4, hipMemcpyHostToDevice, s));
    armsp_config cfg{ARMSP_SCHED_ROW, 0, 0, 0, 0};
    status_or_die(fn, armsp_spmm_f32(M, K, ncols, nnz, drp.as<int>(), dci.as<int>(), val ? dva.as<float>() : nullptr,
                                     dB.as<float>(), N, dC.as<float>(), N, flags, &cfg, s));
    hip_or_die(fn, hipMemcpyAsync(C, dC.p, (size_t)M * N * 4, hipMemcpyDeviceToHost, s));
    hip_or_die(fn, hipStreamSynchronize(s));
    hipStreamDestroy(s);
}

}  // namespace

void rowBaseSpmm(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int) {
    spmm_shim("rowBaseSpmm", rowPtr, colInd, val, B, C, M, K, N, N & ~3, ARMSP_ACCUMULATE | ARMSP_TRUNC4);
}

void rowBaseSpmm2(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int) {
    spmm_shim("rowBaseSpmm2", rowPtr, colInd, val, B, C, M, K, N, N, ARMSP_ACCUMULATE);
}

void rowBaseSpmm3(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int) {
    spmm_shim("rowBaseSpmm3", rowPtr, colInd, val, B, C, M, K, N, N, ARMSP_ACCUMULATE);
}

void rowWiseBlocking(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int) {
    if (M % 8 || N % 8) die("rowWiseBlocking", "requires M % 8 == 0 and N % 8 == 0");
    spmm_shim("rowWiseBlocking", rowPtr, colInd, val, B, C, M, K, N, N, ARMSP_ACCUMULATE);
}

void rowWiseRegisterBlocking(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int) {
    if (N % 4) die("rowWiseRegisterBlocking", "requires N % 4 == 0");
    spmm_shim("rowWiseRegisterBlocking", rowPtr, colInd, val, B, C, M, K, N, N, ARMSP_ACCUMULATE);
}

void irregular(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int) {
    spmm_shim("irregular", rowPtr, colInd, val, B, C, M, K, N, N & ~31, ARMSP_ACCUMULATE);
}

void origin(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int) {
    spmm_shim("origin", rowPtr, colInd, val, B, C, M, K, N, N, 0);
}

void sparse_matrix_mul3(int rowPtr1[], int colInd1[], float val1[], int rowPtr2[], int colInd2[], float val2[],
                        float* matrixC, int M, int K, int N, int) {
    const char* fn = "sparse_matrix_mul3";
    if (N < 0) die(fn, "negative dimension");
    check_csr(fn, M, K, rowPtr1, colInd1);
    check_csr(fn, K, N, rowPtr2, colInd2, true);
    if (M == 0 || N == 0) return;
    if (!matrixC) die(fn, "null output");
    const int nnz1 = rowPtr1[M], nnz2 = rowPtr2[K];
    hipStream_t s;
    hip_or_die(fn, hipStreamCreate(&s));
    Dev r1(fn, (size_t)(M + 1) * 4), c1(fn, (size_t)nnz1 * 4), v1(fn, (size_t)nnz1 * 4);
    Dev r2(fn, (size_t)(K + 1) * 4), c2(fn, (size_t)nnz2 * 4), v2(fn, (size_t)nnz2 * 4);
    Dev dC(fn, (size_t)M * N * 4);
    hip_or_die(fn, hipMemcpyAsync(r1.p, rowPtr1, (size_t)(M + 1) * 4, hipMemcpyHostToDevice, s));
    hip_or_die(fn, hipMemcpyAsync(r2.p, rowPtr2, (size_t)(K + 1) * 4, hipMemcpyHostToDevice, s));
    if (nnz1) {
        hip_or_die(fn, hipMemcpyAsync(c1.p, colInd1, (size_t)nnz1 * 4, hipMemcpyHostToDevice, s));
        if (val1) hip_or_die(fn, hipMemcpyAsync(v1.p, val1, (size_t)nnz1 * 4, hipMemcpyHostToDevice, s));
    }
    if (nnz2) {
        hip_or_die(fn, hipMemcpyAsync(c2.p, colInd2, (size_t)nnz2 * 4, hipMemcpyHostToDevice, s));
        if (val2) hip_or_die(fn, hipMemcpyAsync(v2.p, val2, (size_t)nnz2 * 4, hipMemcpyHostToDevice, s));
    }
    hip_or_die(fn, hipMemcpyAsync(dC.p, matrixC, (size_t)M * N * 4, hipMemcpyHostToDevice, s));
    status_or_die(fn, armsp_spgemm_dense_f32(M, K, N, nnz1, r1.as<int>(), c1.as<int>(), val1 ? v1.as<float>() : nullptr,
                                             nnz2, r2.as<int>(), c2.as<int>(), val2 ? v2.as<float>() : nullptr,
                                             dC.as<float>(), N, s));
    hip_or_die(fn, hipMemcpyAsync(matrixC, dC.p, (size_t)M * N * 4, hipMemcpyDeviceToHost, s));
    hip_or_die(fn, hipStreamSynchronize(s));
    hipStreamDestroy(s);
}

void sparse_dense_vector(float* B, float* C, int rowPtr[], int colInd[], float val[], int M, int) {
    const char* fn = "sparse_dense_vector";
    if (M < 0 || !rowPtr) die(fn, "bad row pointer");
    if (M == 0) return;
    const int nnz = rowPtr[M];
    if (nnz % 4) die(fn, "blocked CSR length is not a multiple of 4");
    for (int r = 0; r < M; ++r)
        if (rowPtr[r] > rowPtr[r + 1] || (rowPtr[r] & 3)) die(fn, "row pointers are not 4-aligned and monotone");
    for (int i = 0; i < nnz; i += 4)
        if (colInd[i] < 0 || (colInd[i] & 3)) die(fn, "block start column is not a non-negative multiple of 4");
    if (!C || (nnz && (!B || !val))) die(fn, "null operand");
    const int K = nnz ? max_col(nnz, colInd) + 1 : 0;
    hipStream_t s;
    hip_or_die(fn, hipStreamCreate(&s));
    Dev drp(fn, (size_t)(M + 1) * 4), dci(fn, (size_t)nnz * 4), dva(fn, (size_t)nnz * 4);
    Dev dx(fn, (size_t)K * 4), dy(fn, (size_t)M * 4);
    hip_or_die(fn, hipMemcpyAsync(drp.p, rowPtr, (size_t)(M + 1) * 4, hipMemcpyHostToDevice, s));
    if (nnz) {
        hip_or_die(fn, hipMemcpyAsync(dci.p, colInd, (size_t)nnz * 4, hipMemcpyHostToDevice, s));
        hip_or_die(fn, hipMemcpyAsync(dva.p, val, (size_t)nnz * 4, hipMemcpyHostToDevice, s));
        hip_or_die(fn, hipMemcpyAsync(dx.p, B, (size_t)K * 4, hipMemcpyHostToDevice, s));
    }
    hip_or_die(fn, hipMemcpyAsync(dy.p, C, (size_t)M * 4, hipMemcpyHostToDevice, s));
    status_or_die(fn, armsp_spmv_vcsr_f32(M, nnz, drp.as<int>(), dci.as<int>(), dva.as<float>(), dx.as<float>(),
                                          dy.as<float>(), 0, s));
    hip_or_die(fn, hipMemcpyAsync(C, dy.p, (size_t)M * 4, hipMemcpyDeviceToHost, s));
    hip_or_die(fn, hipStreamSynchronize(s));
    hipStreamDestroy(s);
}

void csr_spmm_kernel(int A_nrows, int B_ncols, int* A_csrRowPtr, int* A_csrColInd, float* A_csrVal, float* B_dnVal,
                     float* C_dnVal) {
    const char* fn = "csr_spmm_kernel";
    if (A_nrows < 0 || !A_csrRowPtr) die(fn, "bad row pointer");
    const int nnz = A_nrows ? A_csrRowPtr[A_nrows] : 0;
    const int K = nnz > 0 && A_csrColInd ? max_col(nnz, A_csrColInd) + 1 : 0;
    spmm_shim(fn, A_csrRowPtr, A_csrColInd, A_csrVal, B_dnVal, C_dnVal, A_nrows, K, B_ncols, B_ncols,
              ARMSP_ACCUMULATE);
}

void DataReuseSpmm(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int,
                   bool* flag) {
    const char* fn = "DataReuseSpmm";
    if (M < 0 || K < 0 || N < 0 || !rowPtr) die(fn, "bad dimensions or row pointer");
    if (M == 0 || N == 0) return;
    const int nnz = rowPtr[M];
    if (rowPtr[0] != 0) die(fn, "rowPtr[0] != 0");
    for (int r = 0; r < M; ++r)
        if (rowPtr[r + 1] < rowPtr[r]) die(fn, "rowPtr is decreasing");
    if (nnz > 0 && (!colInd || !val || !flag || !B)) die(fn, "null operand");
    if (!C) die(fn, "null C");
    for (int i = 0; i < nnz; ++i)
        if (colInd[i] < 0 || colInd[i] >= K) die(fn, "column index out of range");
    std::vector<uint8_t> fl((size_t)nnz);
    for (int i = 0; i < nnz; ++i) fl[i] = flag[i] ? 1 : 0;
    hipStream_t s;
    hip_or_die(fn, hipStreamCreate(&s));
    Dev drp(fn, (size_t)(M + 1) * 4), dci(fn, (size_t)nnz * 4), dva(fn, (size_t)nnz * 4), dfl(fn, (size_t)nnz);
    Dev dB(fn, (size_t)K * N * 4), dC(fn, (size_t)M * N * 4);
    hip_or_die(fn, hipMemcpyAsync(drp.p, rowPtr, (size_t)(M + 1) * 4, hipMemcpyHostToDevice, s));
    if (nnz) {
        hip_or_die(fn, hipMemcpyAsync(dci.p, colInd, (size_t)nnz * 4, hipMemcpyHostToDevice, s));
        hip_or_die(fn, hipMemcpyAsync(dva.p, val, (size_t)nnz * 4, hipMemcpyHostToDevice, s));
        hip_or_die(fn, hipMemcpyAsync(dfl.p, fl.data(), (size_t)nnz, hipMemcpyHostToDevice, s));
        hip_or_die(fn, hipMemcpyAsync(dB.p, B, (size_t)K * N * 4, hipMemcpyHostToDevice, s));
    }
    hip_or_die(fn, hipMemcpyAsync(dC.p, C, (size_t)M * N * 4, hipMemcpyHostToDevice, s));
    status_or_die(fn, armsp_spmm_datareuse_f32(M, K, N, nnz, drp.as<int>(), dci.as<int>(), dva.as<float>(),
                                               dfl.as<uint8_t>(), dB.as<float>(), N, dC.as<float>(), N,
                                               ARMSP_ACCUMULATE, s));
    hip_or_die(fn, hipMemcpyAsync(C, dC.p, (size_t)M * N * 4, hipMemcpyDeviceToHost, s));
    hip_or_die(fn, hipStreamSynchronize(s));
    hipStreamDestroy(s);
}
