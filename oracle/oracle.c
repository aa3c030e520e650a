/*
 * oracle.c — CPU restatement of the ARMSparse hot-path kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (armsparse_amd/, include/)
 * links, loads or calls this file; it is used by tests/ as the parity checker,
 * by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg
 * as the timed CPU baseline.
 *
 * The reference kernels need <arm_neon.h> and OpenBLAS and do not compile on
 * x86 (the only hot-path file that does is numaAware/numaAwareSpMM.cpp, built
 * separately under oracle/_ref).  They are restated here operation for
 * operation:
 *   - vfmaq_n_f32 / vfmaq_laneq_f32 (fused multiply-add) → fmaf
 *   - the scalar tails `C[i] = C[i] + A[i]*b` and `res += val*B` are contracted
 *     to fused multiply-adds by GCC -O3 (default -ffp-contract=fast on
 *     aarch64), so they are fmaf as well
 *   - cblas_saxpy (y += a·x elementwise, OpenBLAS FMA kernels) → fmaf
 *   - vmlaq_f32 (ACLE: multiply then add, not fused) → separate multiply and
 *     add; the fused reading is kept behind a flag because some compilers
 *     contract it.
 * This file is compiled with -ffp-contract=off so that only the explicit
 * fmaf calls are fused.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* num_vector_neon (SDMM.cpp:449-469, also gcn/pytorch-custom/spmm.cpp:73-93):
 * C[0:len] += A[0:len] * (*B), 4-wide FMA body and an FMA-contracted tail. */
void orc_num_vector(const float* A, const float* B, float* C, int len) {
    const float b = *B;
    for (int i = 0; i < len; ++i) C[i] = fmaf(A[i], b, C[i]);
}

/* rowBaseSpmm3 (SDMM.cpp:90-114): rows in parallel, per nonzero one
 * num_vector_neon of the B row into the C row.  rowBaseSpmm2 (SDMM.cpp:65-89,
 * cblas_saxpy) and csr_spmm_kernel (gcn/pytorch-custom/spmm.cpp:49-71) have
 * the same per-element arithmetic. */
void orc_rowBaseSpmm3(const int* rowPtr, const int* colInd, const float* val, const float* B, float* C, int M, int K,
                      int N, int threads) {
    (void)K;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int m = 0; m < M; m++) {
        for (int idx = rowPtr[m]; idx < rowPtr[m + 1]; idx++) {
            const int k = colInd[idx];
            const float one = 1.0f;
            orc_num_vector(B + (size_t)k * N, val ? val + idx : &one, C + (size_t)m * N, N);
        }
    }
}

/* rowBaseSpmm (SDMM.cpp:21-64): groups of 4 nonzeros (tail of the row
 * dropped), column blocks of 4 (tail columns dropped); inside a group the four
 * vfmaq_laneq_f32 are applied in nonzero order. */
void orc_rowBaseSpmm(const int* rowPtr, const int* colInd, const float* val, const float* B, float* C, int M, int K,
                     int N, int threads) {
    (void)K;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int m = 0; m < M; m++) {
        const int row_end = rowPtr[m + 1];
        for (int idx = rowPtr[m]; idx + 4 <= row_end; idx += 4) {
            for (int j = 0; j + 4 <= N; j += 4) {
                for (int t = 0; t < 4; ++t) {
                    float c = C[(size_t)m * N + j + t];
                    for (int q = 0; q < 4; ++q) c = fmaf(B[(size_t)colInd[idx + q] * N + j + t], val[idx + q], c);
                    C[(size_t)m * N + j + t] = c;
                }
            }
        }
    }
}

/* rowWiseBlocking (SDMM.cpp:115-135): 8-row × 8-column blocks, inside a block
 * the rows in parallel and num_vector_neon over 8 columns.  Requires
 * M % 8 == 0 and N % 8 == 0 (the reference indexes past the end otherwise). */
void orc_rowWiseBlocking(const int* rowPtr, const int* colInd, const float* val, const float* B, float* C, int M,
                         int K, int N, int threads) {
    (void)K;
    const int bsize = 8;
    for (int R = 0; R < M; R += bsize)
        for (int Col = 0; Col < N; Col += bsize) {
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int m = R; m < R + bsize; m++)
                for (int idx = rowPtr[m]; idx < rowPtr[m + 1]; idx++)
                    orc_num_vector(B + (size_t)colInd[idx] * N + Col, val + idx, C + (size_t)m * N + Col, bsize);
        }
}

/* rowWiseRegisterBlocking (SDMM.cpp:136-160): per row and 4-column block, an
 * accumulator loaded from C, FMA over the row's nonzeros, stored back.
 * Requires N % 4 == 0. */
void orc_rowWiseRegisterBlocking(const int* rowPtr, const int* colInd, const float* val, const float* B, float* C,
                                 int M, int K, int N, int threads) {
    (void)K;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int m = 0; m < M; m++)
        for (int kk = 0; kk < N; kk += 4) {
            float acc[4];
            for (int t = 0; t < 4; ++t) acc[t] = C[(size_t)m * N + kk + t];
            for (int idx = rowPtr[m]; idx < rowPtr[m + 1]; idx++)
                for (int t = 0; t < 4; ++t) acc[t] = fmaf(B[(size_t)colInd[idx] * N + kk + t], val[idx], acc[t]);
            for (int t = 0; t < 4; ++t) C[(size_t)m * N + kk + t] = acc[t];
        }
}

/* irregular (SDMM.cpp:161-180): per nonzero, column blocks of 32 in parallel;
 * only the first 32·floor(N/32) columns are touched. */
void orc_irregular(const int* rowPtr, const int* colInd, const float* val, const float* B, float* C, int M, int K,
                   int N, int threads) {
    (void)K;
    const int bsize = 32, Cols = N / bsize;
    for (int m = 0; m < M; m++)
        for (int idx = rowPtr[m]; idx < rowPtr[m + 1]; ++idx) {
            const int k = colInd[idx];
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int Col = 0; Col < Cols; ++Col)
                orc_num_vector(B + (size_t)k * N + Col * bsize, val + idx, C + (size_t)m * N + Col * bsize, bsize);
        }
}

/* origin (SDMM.cpp:422-448): C = A·B, per element res = 0; res += val*B in CSR
 * order (contracted to fmaf). */
void orc_origin(const int* rowPtr, const int* colInd, const float* val, const float* B, float* C, int M, int K, int N,
                int threads) {
    (void)K;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            float res = 0.0f;
            for (int ptr = rowPtr[i]; ptr < rowPtr[i + 1]; ptr++)
                res = fmaf(val[ptr], B[(size_t)colInd[ptr] * N + j], res);
            C[(size_t)i * N + j] = res;
        }
}

/* sparse_matrix_mul3 (SpMM.cpp:309-327): dense C[M×N] += A·B with A, B CSR;
 * C[m, colInd2[t]] += val1[idx]*val2[t] (contracted to fmaf). */
void orc_sparse_matrix_mul3(const int* rowPtr1, const int* colInd1, const float* val1, const int* rowPtr2,
                            const int* colInd2, const float* val2, float* C, int M, int K, int N, int threads) {
    (void)K;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int m = 0; m < M; m++)
        for (int idx = rowPtr1[m]; idx < rowPtr1[m + 1]; idx++) {
            const int k = colInd1[idx];
            for (int t = rowPtr2[k]; t < rowPtr2[k + 1]; t++)
                C[(size_t)m * N + colInd2[t]] = fmaf(val1[idx], val2[t], C[(size_t)m * N + colInd2[t]]);
        }
}

/* sparse_dense_vector (SpMV.cpp:155-179): per row four lanes
 * s_l = s_l + val[idx+l]*x[colInd[idx]+l] over the row's 4-blocks
 * (vmlaq_f32; unfused unless `fused`), then
 * vadd_f32(high, low) = (s2+s0, s3+s1), vpadd lane 0 = (s2+s0)+(s3+s1),
 * and C[m] += that. */
void orc_sparse_dense_vector(const float* x, float* y, const int* rowPtr, const int* colInd, const float* val, int M,
                             int threads, int fused) {
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int m = 0; m < M; m++) {
        float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int idx = rowPtr[m]; idx < rowPtr[m + 1]; idx += 4) {
            const int k = colInd[idx];
            for (int l = 0; l < 4; ++l) {
                if (fused) {
                    s[l] = fmaf(val[idx + l], x[k + l], s[l]);
                } else {
                    const float p = val[idx + l] * x[k + l];
                    s[l] = s[l] + p;
                }
            }
        }
        const float r0 = s[2] + s[0], r1 = s[3] + s[1];
        y[m] = y[m] + (r0 + r1);
    }
}

/* toCSR (SDMM.cpp:355-370): row-major scan of a dense matrix, nonzeros kept. */
int orc_toCSR(const float* dense, int rows, int cols, int* rowPtr, int* colInd, float* val) {
    int size = 0;
    for (int row = 0; row < rows; row++) {
        rowPtr[row] = size;
        for (int col = 0; col < cols; col++) {
            const float num = dense[(size_t)row * cols + col];
            if (num != 0) {
                if (val) val[size] = num;
                if (colInd) colInd[size] = col;
                size++;
            }
        }
    }
    rowPtr[rows] = size;
    return size;
}

/* toDataReuseFormat (suitesparse_sdmm.cpp:150-179, datasets/suitesparse.cpp:
 * 229-263): row-major scan; after each nonzero (row, col) of an even row, the
 * nonzero (row+1, col) of the next row, if any, follows it with flag 1 and is
 * erased from the matrix (max_tile = 2), so an odd row keeps only the columns
 * its even partner lacks.  Works on (and, like the reference, erases entries
 * of) `dense`; returns the entry count. */
int orc_toDataReuseFormat(float* dense, int rows, int cols, int* rowPtr, int* colInd, float* val,
                          unsigned char* flag) {
    int size = 0;
    const int max_tile = 2;
    for (int row = 0; row < rows; row++) {
        rowPtr[row] = size;
        for (int col = 0; col < cols; col++) {
            const float num = dense[(size_t)row * cols + col];
            if (num != 0) {
                val[size] = num;
                colInd[size] = col;
                flag[size] = 0;
                size++;
                for (int cur = row + 1; cur % max_tile != 0; cur++) {
                    if (cur < rows && dense[(size_t)cur * cols + col] != 0) {
                        val[size] = dense[(size_t)cur * cols + col];
                        colInd[size] = col;
                        flag[size] = 1;
                        size++;
                        dense[(size_t)cur * cols + col] = 0;
                    } else {
                        break;
                    }
                }
            }
        }
    }
    rowPtr[rows] = size;
    return size;
}

/* DataReuseSpmm (suitesparse_sdmm.cpp:25-43): each entry of row m is a
 * cblas_saxpy of its B row into C row m + flag, rows taken in pairs by one
 * thread (schedule(static, 2)), so C row m+1 gets the flagged entries of row
 * m first, then row m+1's own. */
void orc_DataReuseSpmm(const int* rowPtr, const int* colInd, const float* val, const unsigned char* flag,
                       const float* B, float* C, int M, int K, int N, int threads) {
    (void)K;
#pragma omp parallel for num_threads(threads) schedule(static, 2)
    for (int m = 0; m < M; m++) {
        for (int idx = rowPtr[m]; idx < rowPtr[m + 1]; idx++) {
            const int k = colInd[idx];
            float* c = C + (size_t)(m + (flag[idx] ? 1 : 0)) * N;
            const float* b = B + (size_t)k * N;
            for (int i = 0; i < N; i++) c[i] = fmaf(val[idx], b[i], c[i]);
        }
    }
}

/* toCSC (SpMM.cpp:80-95): column-major scan of a dense matrix. */
int orc_toCSC(const float* dense, int rows, int cols, int* colPtr, int* rowInd, float* val) {
    int size = 0;
    for (int col = 0; col < cols; col++) {
        colPtr[col] = size;
        for (int row = 0; row < rows; row++) {
            const float num = dense[(size_t)row * cols + col];
            if (num != 0) {
                if (val) val[size] = num;
                if (rowInd) rowInd[size] = row;
                size++;
            }
        }
    }
    colPtr[cols] = size;
    return size;
}

/* toVectorAwareCSR (SpMV.cpp:15-38): for each aligned 4-column block with any
 * nonzero (|v| > threshold when threshold > 0, numaAware/numaAwareSpMV.cpp:48),
 * store all four values and column ids.  Requires cols % 4 == 0 (the reference
 * reads past the row end otherwise).  Pointers may be NULL to count only. */
int orc_toVectorAwareCSR(const float* dense, int rows, int cols, int* rowPtr, int* colInd, float* val,
                         float threshold) {
    int size = 0;
    for (int row = 0; row < rows; row++) {
        if (rowPtr) rowPtr[row] = size;
        for (int col = 0; col < cols; col += 4) {
            int load = 0;
            for (int t = col; t < col + 4 && t < cols; t++) {
                const float v = dense[(size_t)row * cols + t];
                if (threshold > 0 ? fabsf(v) > threshold : v != 0) {
                    load = 1;
                    break;
                }
            }
            if (load) {
                for (int t = col; t < col + 4; t++) {
                    if (val) val[size] = dense[(size_t)row * cols + t];
                    if (colInd) colInd[size] = t;
                    size++;
                }
            }
        }
    }
    if (rowPtr) rowPtr[rows] = size;
    return size;
}

/* numaAwareSDMM (numaAware/numaAwareSDMM.cpp:58-75, 183-226): A split into
 * `nodes` column ranges [node·K/4, (node+1)·(K/4)) (the reference's bounds,
 * generalised to `nodes`), each node accumulates its own C_node with one
 * saxpy per nonzero, then C_0 += C_1, += C_2, ... (sumUp).  Result in C
 * (overwritten); work holds (nodes-1)·M·N floats. */
void orc_numa_sdmm(const int* rowPtr, const int* colInd, const float* val, const float* B, float* C, float* work,
                   int M, int K, int N, int nodes) {
    const int part = K / nodes;
    memset(C, 0, sizeof(float) * (size_t)M * N);
    memset(work, 0, sizeof(float) * (size_t)(nodes - 1) * M * N);
    for (int node = 0; node < nodes; ++node) {
        float* Cn = node == 0 ? C : work + (size_t)(node - 1) * M * N;
        const int lo = node * K / nodes, hi = (node + 1) * part;
        for (int m = 0; m < M; m++)
            for (int idx = rowPtr[m]; idx < rowPtr[m + 1]; idx++) {
                const int k = colInd[idx];
                if (k < lo || k >= hi) continue;
                orc_num_vector(B + (size_t)k * N, val + idx, Cn + (size_t)m * N, N);
            }
    }
    for (int node = 1; node < nodes; ++node) {
        const float* Cn = work + (size_t)(node - 1) * M * N;
        for (size_t i = 0; i < (size_t)M * N; ++i) C[i] = C[i] + Cn[i];
    }
}

/* ----------------------------------------------------------- FP64 oracle */

/* C = beta·C_in + A·B in double; Bd is the (dequantised) dense operand. */
void orc_spmm_f64(const int* rowPtr, const int* colInd, const float* val, const float* B, int ldb, const float* Cin,
                  double* C, int M, int N, int threads) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64)
    for (int m = 0; m < M; m++) {
        double* c = C + (size_t)m * N;
        for (int j = 0; j < N; ++j) c[j] = Cin ? (double)Cin[(size_t)m * N + j] : 0.0;
        for (int idx = rowPtr[m]; idx < rowPtr[m + 1]; idx++) {
            const double a = val ? (double)val[idx] : 1.0;
            const float* b = B + (size_t)colInd[idx] * ldb;
            for (int j = 0; j < N; ++j) c[j] += a * (double)b[j];
        }
    }
}

void orc_spmv_f64(const int* rowPtr, const int* colInd, const float* val, const float* x, const float* yin, double* y,
                  int M, int threads) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
    for (int m = 0; m < M; m++) {
        double s = yin ? (double)yin[m] : 0.0;
        for (int idx = rowPtr[m]; idx < rowPtr[m + 1]; idx++) s += (val ? (double)val[idx] : 1.0) * (double)x[colInd[idx]];
        y[m] = s;
    }
}

/* out[e] = s_e·<U[row(e)], V[col(e)]> (gcn/pytorch-custom/sddmm.cu:255-424
 * semantics, optional A_ij scale). */
void orc_sddmm_f64(const int* rowPtr, const int* colInd, const float* scale, const float* U, int ldu, const float* V,
                   int ldv, int Kd, double* out, int M, int threads) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
    for (int m = 0; m < M; m++)
        for (int e = rowPtr[m]; e < rowPtr[m + 1]; e++) {
            const float* u = U + (size_t)m * ldu;
            const float* v = V + (size_t)colInd[e] * ldv;
            double s = 0.0;
            for (int k = 0; k < Kd; ++k) s += (double)u[k] * (double)v[k];
            out[e] = scale ? (double)scale[e] * s : s;
        }
}

/* CSR → CSC by a stable counting sort (rows ascend inside a column): the
 * structure toCSC (SpMM.cpp:80-95) produces from the dense form. */
void orc_csr_to_csc(const int* rowPtr, const int* colInd, const float* val, int M, int K, int* colPtr, int* rowInd,
                    float* valT) {
    memset(colPtr, 0, sizeof(int) * (size_t)(K + 1));
    const int nnz = rowPtr[M];
    for (int i = 0; i < nnz; ++i) colPtr[colInd[i] + 1]++;
    for (int c = 0; c < K; ++c) colPtr[c + 1] += colPtr[c];
    int* cur = (int*)malloc(sizeof(int) * (size_t)(K > 0 ? K : 1));
    memcpy(cur, colPtr, sizeof(int) * (size_t)K);
    for (int m = 0; m < M; ++m)
        for (int i = rowPtr[m]; i < rowPtr[m + 1]; ++i) {
            const int p = cur[colInd[i]]++;
            rowInd[p] = m;
            if (valT) valT[p] = val[i];
        }
    free(cur);
}

/* Dense fp64 product of two CSR matrices, C = A·B (M×N), for SpGEMM checks. */
void orc_spgemm_f64(const int* rowPtr1, const int* colInd1, const float* val1, const int* rowPtr2, const int* colInd2,
                    const float* val2, const float* Cin, double* C, int M, int N) {
    for (size_t i = 0; i < (size_t)M * N; ++i) C[i] = Cin ? (double)Cin[i] : 0.0;
    for (int m = 0; m < M; m++)
        for (int idx = rowPtr1[m]; idx < rowPtr1[m + 1]; idx++) {
            const int k = colInd1[idx];
            for (int t = rowPtr2[k]; t < rowPtr2[k + 1]; t++)
                C[(size_t)m * N + colInd2[t]] += (double)val1[idx] * (double)val2[t];
        }
}
