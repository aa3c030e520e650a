/*
 * armsparse_compat.h — drop-in replacements for the ARMSparse CPU kernels,
 * with the reference's own names, C++ linkage and host-pointer signatures.
 *
 * Link a reference driver against libarmsparse.so instead of compiling the
 * kernel bodies, and these run on the MI355X: host arrays are copied to the
 * device, the kernel runs with the row-split schedule (which keeps the
 * reference's per-element fmaf order, so results are bit-exact with the
 * reference's FMA arithmetic), and C is copied back.  `threads` is accepted
 * and ignored.  Preconditions the reference leaves undefined are checked
 * (rowPtr[0] = 0 and non-decreasing, column indices in range, the blocking
 * divisibility rules; a row's columns may come in any order, as toCSR2,
 * SDMM.cpp:371-394, writes them); on any failure the shim
 * prints "armsparse: <function>: <reason>" to stderr and exits with
 * EXIT_FAILURE, as the reference's own error paths do
 * (gcn/pytorch-custom/spmm_kernel.cu:5-19, numaAwareSDMM.cpp:353-357).
 *
 * Helpers the reference drivers define themselves (toCSR, prepareSparse,
 * getHighResolutionTime, num_vector_neon) are deliberately not exported.
 */
#ifndef ARMSPARSE_COMPAT_H
#define ARMSPARSE_COMPAT_H

/* C += A·B over the first 4·floor(nnz_row/4) nonzeros of each row and the
 * first 4·floor(N/4) columns (SDMM.cpp:21-64, tails dropped as there). */
void rowBaseSpmm(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int threads);
/* C += A·B (SDMM.cpp:65-89, cblas_saxpy per nonzero). */
void rowBaseSpmm2(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int threads);
/* C += A·B (SDMM.cpp:90-114, ARMSparse's timed kernel). */
void rowBaseSpmm3(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int threads);
/* C += A·B; requires M % 8 == 0 and N % 8 == 0 (SDMM.cpp:115-135). */
void rowWiseBlocking(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int threads);
/* C += A·B; requires N % 4 == 0 (SDMM.cpp:136-160). */
void rowWiseRegisterBlocking(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N,
                             int threads);
/* C += A·B on the first 32·floor(N/32) columns (SDMM.cpp:161-180). */
void irregular(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int threads);
/* C = A·B (SDMM.cpp:422-448). */
void origin(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int threads);

/* Dense C[M×N] += A[M×K]·B[K×N], A and B CSR (SpMM.cpp:309-327); a row of B
 * must not hold a column twice. */
void sparse_matrix_mul3(int rowPtr1[], int colInd1[], float val1[], int rowPtr2[], int colInd2[], float val2[],
                        float* matrixC, int M, int K, int N, int threads);

/* C[m] += Σ over 4-blocks, on the zero-padded blocked CSR of
 * toVectorAwareCSR; B is the dense vector x, C the output (SpMV.cpp:155-179). */
void sparse_dense_vector(float* B, float* C, int rowPtr[], int colInd[], float val[], int M, int threads);

/* C_dnVal[A_nrows×B_ncols] += A·B (gcn/pytorch-custom/spmm.cpp:49-71). */
void csr_spmm_kernel(int A_nrows, int B_ncols, int* A_csrRowPtr, int* A_csrColInd, float* A_csrVal, float* B_dnVal,
                     float* C_dnVal);

/* C[m + flag] += val·B[colInd] for every entry of row m, on the 2-row
 * DataReuse format of toDataReuseFormat (suitesparse_sdmm.cpp:25-43). */
void DataReuseSpmm(int rowPtr[], int colInd[], float val[], float B[], float C[], int M, int K, int N, int threads,
                   bool* flag);

#endif /* ARMSPARSE_COMPAT_H */
