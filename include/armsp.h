/*
 * armsp.h — C ABI of libarmsparse.so, the MI355X (gfx950) sparse-kernel library.
 *
 * Every entry point takes plain device pointers, plain sizes and a HIP stream
 * (passed as `void*`, NULL = the null stream), launches asynchronously on that
 * stream and returns an armsp_status.  No torch or HIP types appear in the
 * signatures.  Index arrays are int32 CSR/COO as in the reference; row-major
 * dense operands carry a leading dimension.
 *
 * Which reference interface each entry point replaces is stated beside it
 * (paths are relative to the ARMSparse source tree).  The C++-linkage drop-ins
 * with the reference's own names and host-pointer signatures are declared in
 * armsparse_compat.h; they are thin shims over these functions.
 *
 * Workspace: entry points whose schedule needs a carry buffer (nnz-split and
 * merge-path SpMM / SpMV, split-K GEMM, transpose) take it from the library's
 * own stream-ordered memory pool on the caller's stream
 * (hipMallocFromPoolAsync / hipFreeAsync), so concurrent calls on different
 * streams never share it and the device's default pool is left untouched.
 * While the stream is being captured into a HIP graph they take it from the
 * arena reserved by armsp_workspace_reserve instead (nothing may allocate
 * during capture).
 */
#ifndef ARMSP_H
#define ARMSP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARMSP_VERSION 100 /* 1.0.0 */

typedef enum armsp_status {
    ARMSP_SUCCESS = 0,
    ARMSP_ERROR_INVALID_VALUE = 1, /* bad size, null pointer, malformed CSR metadata */
    ARMSP_ERROR_NOT_SUPPORTED = 2, /* valid request this build has no kernel for */
    ARMSP_ERROR_HIP = 3,           /* a HIP runtime call failed (see armsp_last_hip_error) */
    ARMSP_ERROR_ALLOC = 4,         /* device workspace allocation failed */
    ARMSP_ERROR_NO_DEVICE = 5      /* no usable gfx950 device */
} armsp_status;

/* Work decomposition of a sparse kernel. */
typedef enum armsp_schedule {
    ARMSP_SCHED_AUTO = 0,  /* library heuristic (or the learned selector, set by the host layer) */
    ARMSP_SCHED_ROW = 1,   /* row-split: a group of `group` lanes owns one row, CSR order kept */
    ARMSP_SCHED_NNZ = 2,   /* nnz-split: each wave owns `chunk` nonzeros, carries fixed up after */
    ARMSP_SCHED_MERGE = 3  /* merge-path: each wave owns `chunk` items of (rows + nnz), carries fixed up */
} armsp_schedule;

/* Schedule and tile shape.  Zero fields mean "choose automatically". */
typedef struct armsp_config {
    int32_t schedule; /* armsp_schedule */
    int32_t group;    /* lanes per row (ROW) or per nonzero (SDDMM): 1,2,4,...,64.
                         SpMM MERGE: 0 automatic (lane walks for N <= 4, half-wave
                         slices with 8-byte lanes for B rows of 512+ bytes and with
                         4-byte lanes for bf16 rows of 128-511 bytes, when ldb
                         spans a multiple of 128 bytes and the row fills whole
                         slices; else a whole wave per nonzero); 64 a whole wave per nonzero; 32 half
                         a wave per nonzero over column slices of 32·vec elements
                         spread over the XCDs; 1 forces the 64-bit-address kernel
                         (otherwise used only when a dense-row offset needs more than
                         32 bits) */
    int32_t vec;      /* dense elements per lane: 1, 2, 4 (8 for bf16/fp8 B) */
    int32_t chunk;    /* items per wave for NNZ / MERGE (multiple of 64) */
    int32_t unroll;   /* SpMM: nonzeros whose dense rows are fetched together: 4, 8, 16.
                         SDDMM (staged CSR kernel): 2 splits each dot product into
                         column halves spread over the XCDs plus a combine pass (pays
                         on skewed columns whose V rows outgrow the L2); 0/1 none */
} armsp_config;

/* flags for the SpMM / SpMV entry points */
#define ARMSP_ACCUMULATE 0x1   /* C += A·B (reference rowBaseSpmm2/3, csr_spmm_kernel); else C = A·B (origin) */
#define ARMSP_TRUNC4     0x2   /* rowBaseSpmm semantics: only the first 4*floor(nnz_row/4) nonzeros of each row */
#define ARMSP_FUSED_MLA  0x4   /* VCSR SpMV: fused multiply-add instead of the unfused vmlaq_f32 reading */
#define ARMSP_VCSR_BALANCED 0x8 /* VCSR SpMV: merge-path schedule over the padded blocks (load-balanced on
                                   skewed rows; within 1e-5 per row, not bit-exact) */

/* ---------------------------------------------------------------- runtime */
int armsp_version(void);
const char* armsp_status_string(int status);
/* Last HIP error code seen by the library on this thread (0 if none). */
int armsp_last_hip_error(void);
/* 1 if a gfx950 device is visible, else 0.  Never fails. */
int armsp_device_available(void);
/* Reserves the device's capture arena: the workspace entry points use while
 * their stream is being captured into a HIP graph.  Its size is at least
 * `bytes` and at least the largest workspace any call on this device has
 * asked for so far, so reserve after a warm-up of the work to be captured.
 * The arena is handed to one call at a time: the captured work must run on
 * one stream, and graphs that share it must not replay concurrently.  An arena
 * lives until the process ends (a captured graph keeps its address); a larger
 * reservation retires it.  *reserved (may be NULL) receives the size.  No
 * counterpart in the reference (it has no device memory). */
int armsp_workspace_reserve(size_t bytes, size_t* reserved);

/* Kernel timing: when enabled, every main kernel launch is bracketed by HIP
 * events on its own stream; totals are read per kernel family. */
int armsp_timing_enable(int on);
int armsp_timing_reset(void);
/* Synchronises the recorded events.  family: "spmm", "spmm_fixup", "spmv",
 * "spmv_fixup", "sddmm", "spgemm", "gemm", "transpose", ... */
int armsp_timing_read(const char* family, double* total_ms, int64_t* launches);

/* ------------------------------------------------------------------- SpMM */
/* C[M×N] (=|+=) A[M×K] · B[K×N].  A: CSR rowptr[M+1], colind[nnz], val[nnz]
 * (val may be NULL: all ones — the value-less aggregation of
 * gcn/pytorch-custom/spmm_kernel.cu:31-207).  B row-major with ldb, C with ldc.
 * Replaces rowBaseSpmm3/rowBaseSpmm2/rowWiseBlocking/rowWiseRegisterBlocking
 * (SDMM.cpp:65-160) with flags=ACCUMULATE, origin (SDMM.cpp:422-448) with
 * flags=0, rowBaseSpmm (SDMM.cpp:21-64) with ACCUMULATE|TRUNC4 and N
 * rounded down to a multiple of 4, irregular (SDMM.cpp:161-180) with N
 * rounded down to a multiple of 32, and csr_spmm_kernel
 * (gcn/pytorch-custom/spmm.cpp:49-71) on a zeroed C.  ROW schedule keeps the
 * reference's per-element fmaf order (bit-exact); NNZ/MERGE re-associate
 * across the carry boundary. */
int armsp_spmm_f32(int M, int K, int N, int nnz,
                   const int* rowptr, const int* colind, const float* val,
                   const float* B, int ldb, float* C, int ldc,
                   int flags, const armsp_config* cfg, void* stream);

/* armsp_spmm_f32 without any allocation: the same kernels and results (bit for
 * bit, for the same arguments), with the merge-path partition and carry buffers
 * in a device-resident array of ARMSP_NOALLOC_SCRATCH_BYTES that the library
 * holds on every GPU it runs on until the process ends, instead of workspace
 * from the memory pool.  No hipMalloc, hipFree, pool, event or other allocating
 * call is made.  A call whose workspace would exceed the array returns
 * ARMSP_ERROR_NOT_SUPPORTED before launching anything.  The merge-path
 * schedules need about 4·W·(N + 2) bytes for W = ceil((M + nnz) / chunk) waves,
 * chunk being the items per wave the call runs with (cfg->chunk, else 256 to
 * 1024 chosen by shape); the 1M×1M R-MAT with 2^24 nnz at N = 128 needs 36 MB.
 * The row schedule needs none.  Every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (concurrent streams, host threads)
 * race: serialise them.  The array keeps the last call's contents; nothing in it
 * needs clearing between calls.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
#define ARMSP_NOALLOC_SCRATCH_BYTES (64u << 20)
int armsp_spmm_f32_noalloc(int M, int K, int N, int nnz,
                           const int* rowptr, const int* colind, const float* val,
                           const float* B, int ldb, float* C, int ldc,
                           int flags, const armsp_config* cfg, void* stream);

/* Same with B stored as bfloat16 bits; fp32 accumulate, fp32 C.  No
 * reference counterpart (the reference is fp32 only, spmm.cpp:112). */
int armsp_spmm_bf16(int M, int K, int N, int nnz,
                    const int* rowptr, const int* colind, const float* val,
                    const uint16_t* B, int ldb, float* C, int ldc,
                    int flags, const armsp_config* cfg, void* stream);

/* Same with B stored as OCP fp8-e4m3 bytes and one fp32 scale per row of B:
 * B[k][n] = scale[k] * e4m3(Bq[k][n]).  No reference counterpart. */
int armsp_spmm_fp8(int M, int K, int N, int nnz,
                   const int* rowptr, const int* colind, const float* val,
                   const uint8_t* Bq, const float* b_row_scale, int ldb,
                   float* C, int ldc, int flags, const armsp_config* cfg, void* stream);

/* Dense operand element types for armsp_spmm_ex. */
#define ARMSP_DTYPE_F32      0
#define ARMSP_DTYPE_BF16     1
#define ARMSP_DTYPE_FP8_E4M3 2

/* Fused operand scales and output epilogue of a SpMM:
 *   C[r,:] (=|+=) row_scale[r] · Σ_e val_e · col_scale[c_e] · B[c_e,:] + bias[:],
 * then ReLU when relu != 0 (applied to the final value of the row).  Any
 * pointer may be NULL.  This is the GCN layer D^-½·A·D^-½·(XW) + b of
 * GCNConv (gcn/pytorch-custom/op.py:143-149, there four separate tensor ops)
 * in one launch; its backward D^-½·Aᵀ·D^-½·dY is the same call on Aᵀ. */
typedef struct armsp_epilogue {
    const float* col_scale;
    const float* row_scale;
    const float* bias;
    int32_t relu;
} armsp_epilogue;

/* SpMM with B of type b_dtype (b_row_scale: fp8 per-row scales, else
 * ignored) and an optional epilogue.  With epilogue == NULL it computes
 * exactly what armsp_spmm_f32 / _bf16 / _fp8 compute. */
int armsp_spmm_ex(int M, int K, int N, int nnz,
                  const int* rowptr, const int* colind, const float* val,
                  int b_dtype, const void* B, const float* b_row_scale, int ldb,
                  float* C, int ldc, int flags, const armsp_epilogue* epilogue,
                  const armsp_config* cfg, void* stream);

/* ------------------------------------------------------------------- SpMV */
/* y[M] (=|+=) A·x on plain CSR.  Replaces the row split of
 * numaAware/numaAwareSpMV.cpp:102-125 (one launch, no per-node copies). */
int armsp_spmv_f32(int M, int K, int nnz,
                   const int* rowptr, const int* colind, const float* val,
                   const float* x, float* y, int flags, const armsp_config* cfg, void* stream);

/* y[M] += A·x on the zero-padded 1×4-blocked "vector-aware" CSR built by
 * toVectorAwareCSR (SpMV.cpp:15-38): every row holds whole aligned 4-blocks,
 * colind[idx] of a block start is a multiple of 4.  Replaces
 * sparse_dense_vector (SpMV.cpp:155-179): per row four partial sums
 * s_l += val*x (unfused, or fused with ARMSP_FUSED_MLA) reduced as
 * (s0+s2)+(s1+s3), bit-exact with the reference's order.  Rows of up to 16
 * blocks get a 4-lane group each, rows of up to 1024 blocks a wave each, longer
 * rows (ARMSP_VCSR_LONG_BLOCKS in the environment moves the 1024) a workgroup
 * each when val and x are 16-byte aligned, all in one launch; every walker
 * still adds each row's blocks in order, so the longest row's chain of adds
 * bounds the call.  ARMSP_VCSR_BALANCED runs
 * the merge-path SpMV over the same arrays instead (rows of any length spread
 * over the device; the merge path's order of additions).  x must hold
 * K rounded up to a multiple of 4 entries. */
int armsp_spmv_vcsr_f32(int M, int nnz_padded,
                        const int* rowptr, const int* colind, const float* val,
                        const float* x, float* y, int flags, void* stream);

/* ------------------------------------------------------------------ SDDMM */
/* out[e] = s_e · <U[row(e),:Kd], V[colind[e],:Kd]> over the pattern of A.
 * scale NULL means s_e = 1 (the semantics of csr_sddmm,
 * gcn/pytorch-custom/sddmm.cpp:42-61 / sddmm.cu:255-424); scale = A's values
 * gives the A_ij-scaled form.  U is M×Kd (ldu), V is K×Kd (ldv), row-major. */
int armsp_sddmm_csr_f32(int M, int K, int Kd, int nnz,
                        const int* rowptr, const int* colind, const float* scale,
                        const float* U, int ldu, const float* V, int ldv,
                        float* out, const armsp_config* cfg, void* stream);
/* COO form: replaces coo_sddmm (sddmm.cpp:21-40 / sddmm.cu:7-252). */
int armsp_sddmm_coo_f32(int M, int K, int Kd, int nnz,
                        const int* rowind, const int* colind, const float* scale,
                        const float* U, int ldu, const float* V, int ldv,
                        float* out, void* stream);
/* bfloat16 U and V, fp32 accumulate and output.  No reference counterpart. */
int armsp_sddmm_csr_bf16(int M, int K, int Kd, int nnz,
                         const int* rowptr, const int* colind, const float* scale,
                         const uint16_t* U, int ldu, const uint16_t* V, int ldv,
                         float* out, const armsp_config* cfg, void* stream);

/* ------------------------------------------------------- SpGEMM to dense */
/* C[M×N] += A[M×K]·B[K×N] with A and B both CSR and C dense row-major.
 * Replaces sparse_matrix_mul3 (SpMM.cpp:309-327, numaAware/numaAwareSpMM.cpp
 * :452-470): per output element the fmaf sequence follows A's CSR order, as
 * in the reference, so the result is bit-exact. */
int armsp_spgemm_dense_f32(int M, int K, int N,
                           int nnzA, const int* rowptrA, const int* colindA, const float* valA,
                           int nnzB, const int* rowptrB, const int* colindB, const float* valB,
                           float* C, int ldc, void* stream);

/* ------------------------------------------------------- DataReuse SpMM */
/* C (=|+= with ARMSP_ACCUMULATE) A·B on the 2-row DataReuse format
 * (armsp_csr_to_datareuse): every entry of row m goes to C row m + flag[p].
 * Replaces DataReuseSpmm (suitesparse_sdmm.cpp:25-43, datasets/suitesparse.cpp
 * :23-50); a row pair shares the B row of a flagged entry and its partner, and
 * each C row gets its additions in the reference's order (bit-identical to
 * the sequential kernel).  Flags of odd rows are treated as 0. */
int armsp_spmm_datareuse_f32(int M, int K, int N, int nnz,
                             const int* rowptr, const int* colind, const float* val, const uint8_t* flag,
                             const float* B, int ldb, float* C, int ldc, int flags, void* stream);

/* ------------------------------------------------------- max aggregation */
/* C[r][j] = max(init, max over the row's nonzeros p of val[p]·B[colind[p]][j]),
 * folded in CSR order with `acc > x ? acc : x`: the DGL max reducer of
 * gcn/dgl-custom/binary_reduce_max.cu (:151-166 and the tiled kernels
 * :26-149, whose init is max_init() = -10000, :22-24).  val may be NULL (copy
 * of the source features, as the reference).  argmax (NULL, or M×N int32,
 * row-major) receives the winning position p, or -1 where init won (a later
 * equal value wins). */
int armsp_spmm_max_f32(int M, int K, int N, int nnz,
                       const int* rowptr, const int* colind, const float* val,
                       const float* B, int ldb, float init, float* C, int ldc, int* argmax, void* stream);

/* Backward of armsp_spmm_max_f32: dB[colind[p]][j] += val[p]·dC[r][j] for
 * p = argmax[r][j] ≥ 0 (atomic adds; the reference has no backward for its
 * custom max kernel).  dB is accumulated into, not cleared. */
int armsp_spmm_max_backward_f32(int M, int K, int N, int nnz, const int* colind, const float* val,
                                const float* dC, int ldc, const int* argmax, float* dB, int ldb, void* stream);

/* -------------------------------------------------------------- transpose */
/* CSC of A (= CSR of Aᵀ): colptr[K+1], rowind[nnz], valT[nnz] (valT and val
 * may be NULL together).  Stable: rows ascend inside each column, exactly as
 * toCSC (SpMM.cpp:80-95) / scipy tocsc produce; replaces the cuSPARSE call in
 * csr2cscKernel (gcn/pytorch-custom/spmm_kernel.cu:381-423). */
int armsp_csr_transpose(int M, int K, int nnz,
                        const int* rowptr, const int* colind, const float* val,
                        int* colptr, int* rowind, float* valT, void* stream);

/* Canonical CSR from nnz unordered COO triplets (rows[e], cols[e], vals[e]):
 * entries in (row, column) order, duplicates summed in input order, and with
 * drop_zeros the entries whose value (after summing) is 0 left out — the
 * structure toCSR builds from a dense matrix (SDMM.cpp:355-370,
 * SpMM.cpp:64-79), on the device.  Outputs: rowptr[M+1], colind[nnz] and
 * val[nnz] (capacity nnz; vals and val may be NULL together for a pattern),
 * and *nnz_out (a device int) = the entries written.  Every index is checked
 * (ARMSP_ERROR_INVALID_VALUE when out of range); the check synchronises the
 * stream once. */
int armsp_coo_to_csr(int M, int K, int nnz,
                     const int* rows, const int* cols, const float* vals, int drop_zeros,
                     int* rowptr, int* colind, float* val, int* nnz_out, void* stream);

/* CSR → the zero-padded 1×4-blocked CSR of toVectorAwareCSR (SpMV.cpp:15-38),
 * on the device: the same arrays armsp_csr_to_vcsr builds on the host, bit for
 * bit.  A block [4b, 4b+4) of a row is stored iff one of the row's values in it
 * triggers (v != 0; with threshold > 0, |v| > threshold as in
 * numaAware/numaAwareSpMV.cpp:48; val == NULL: every value is 1), as four
 * slots (the row's values, 0 where absent) with column indices 4b..4b+3, so
 * x needs ceil(K/4)·4 entries.  Always writes out_rowptr[M+1] and *padded (a
 * host int64: the stored entries); with out_colind and out_val (both or
 * neither) of capacity >= *padded it also writes the blocks.  Call once with
 * NULL outputs to size them.  The input must be canonical CSR: rowptr[0] = 0,
 * non-decreasing, rowptr[M] = nnz, columns in [0, K) strictly increasing in a
 * row (else ARMSP_ERROR_INVALID_VALUE, with nothing written).  The check
 * synchronises the stream once; the blocks are written after it, in stream
 * order. */
int armsp_csr_to_vcsr_device(int M, int K, int nnz, const int* rowptr, const int* colind, const float* val,
                             float threshold, int* out_rowptr, int* out_colind, float* out_val, int64_t capacity,
                             int64_t* padded, void* stream);

/* A CSR matrix split into nblocks (1..63) blocks of columns [x·W, (x+1)·W),
 * W = ceil(K / nblocks), each a CSR matrix over all M rows whose rows keep
 * their entries' order: block x's row pointers are
 * out_rowptr[x·(M+1) .. x·(M+1)+M] (from 0), its entries
 * out_colind/out_val[block_start[x] .. block_start[x+1]) (block_start: host,
 * nblocks+1 entries).  A product whose gathered rows of B outgrow the L2 then
 * runs block after block with a block's rows of B cached (the product over the
 * blocks, accumulated, is A·B; the sums are ordered per block).  No counterpart
 * in the reference: it is the column split of the NUMA drivers
 * (numaAware/numaAwareSDMM.cpp:100-113 DistributetoCSR), taken in time instead
 * of across nodes.  The input must be canonical CSR (rowptr from 0 to nnz,
 * columns in [0, K) strictly increasing in a row; else
 * ARMSP_ERROR_INVALID_VALUE).  Synchronises the stream. */
int armsp_csr_column_split(int M, int K, int nnz, const int* rowptr, const int* colind, const float* val,
                           int nblocks, int* out_rowptr, int* out_colind, float* out_val, int64_t* block_start,
                           void* stream);

/* -------------------------------------------------------- dense helpers */
/* C[M×N] (=|+=) op(A)[M×Kd] · op(B)[Kd×N] in bfloat16 on MFMA with fp32
 * accumulation; element (i,k) of op(A) is A[i*sam + k*sak], element (k,j) of
 * op(B) is B[k*sbk + j*sbn].  Replaces torch.matmul(x, weight) of GCNConv
 * (gcn/pytorch-custom/op.py:114) and its weight gradient. */
int armsp_gemm_bf16(int M, int N, int Kd,
                    const uint16_t* A, int64_t sam, int64_t sak,
                    const uint16_t* B, int64_t sbk, int64_t sbn,
                    float* C, int64_t ldc, int accumulate, void* stream);
/* The same product written as C of type c_dtype: ARMSP_DTYPE_F32 (as
 * armsp_gemm_bf16) or ARMSP_DTYPE_BF16 (rounded to nearest even; accumulate
 * must be 0), so a bf16 layer output needs no separate conversion pass. */
int armsp_gemm_bf16_ex(int M, int N, int Kd,
                       const uint16_t* A, int64_t sam, int64_t sak,
                       const uint16_t* B, int64_t sbk, int64_t sbn,
                       void* C, int c_dtype, int64_t ldc, int accumulate, void* stream);
/* fp32 → bf16 (round to nearest even), n elements. */
int armsp_convert_f32_to_bf16(int64_t n, const float* x, uint16_t* y, void* stream);
/* The same for an M×N matrix: rows of pitch ldx in, rows of pitch ldy out (a
 * bf16 operand laid out for the SpMM gathers, e.g. an odd number of 128-byte
 * lines per row). */
int armsp_convert_rows_f32_to_bf16(int M, int N, const float* x, int64_t ldx, uint16_t* y, int64_t ldy,
                                   void* stream);
/* Per-row fp8-e4m3 quantisation: scale[k] = max|B[k,:]| / 448 (1 if the row
 * is zero), Bq = RNE(B / scale) saturated to ±448. */
int armsp_quantize_fp8_rows(int K, int N, const float* B, int ldb,
                            uint8_t* Bq, float* scale, void* stream);

/* Inverse of armsp_quantize_fp8_rows: out[k][j] = e4m3(Bq[k][j])·scale[k]
 * (scale NULL: 1), decoded by the gfx950 packed fp8 conversion that the fp8
 * SpMM uses.  Extension (the reference is fp32 only, spmm.cpp:112). */
int armsp_dequantize_fp8_rows(int K, int N, const uint8_t* Bq, int ldq, const float* scale,
                              float* out, int ldo, void* stream);
/* Counter-based fills (same streams as the host generators in armsp_gen.h). */
int armsp_fill_uniform_f32(int64_t n, uint64_t seed, uint64_t stream_id,
                           float lo, float hi, float* out, void* stream);
int armsp_fill_normal_f32(int64_t n, uint64_t seed, uint64_t stream_id,
                          float* out, void* stream);

/* ------------------------------------------- learned schedule selector */
/* No reference counterpart: the reference picks its SpMM kernel by a fixed N
 * threshold (gcn/pytorch-custom/spmm_kernel.cu:437-457).  These kernels
 * compute the selector's sparsity features and train / run its MLP. */
/* Row-length histogram: hist[0] empty rows, hist[b] rows with
 * 2^(b-1) <= len < 2^b, hist[nbins-1] everything longer; *maxlen = longest
 * row.  hist and maxlen are device int arrays (overwritten). */
int armsp_row_stats(int M, const int* rowptr, int nbins, int* hist, int* maxlen, void* stream);
/* C = op(A)·op(B) (+ beta·C) (+ bias[n]) (then ReLU if relu != 0), fp32,
 * strided operands as in armsp_gemm_bf16. */
int armsp_sgemm_f32(int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk,
                    int64_t sbn, float* C, int64_t ldc, float beta, const float* bias, int relu, void* stream);
/* Mean softmax cross-entropy of logits[rows×classes] into *loss (device)
 * against hard labels[rows] or, when targets is non-NULL, soft target rows
 * targets[rows×classes]; dlogits (nullable) = (softmax - target) / rows. */
int armsp_softmax_xent(int rows, int classes, const float* logits, const int* labels, const float* targets,
                       float* dlogits, float* loss, void* stream);
/* grad[i] = 0 where act[i] <= 0. */
int armsp_relu_backward(int64_t n, const float* act, float* grad, void* stream);
/* out[n] = sum over m of X[m][n]: rows summed in order up to 4096 rows; taller
 * matrices in blocks of 128 rows or more (at most 4096 blocks, each summed in
 * four interleaved row phases) whose sums are added the same way in a fixed
 * order.  Deterministic; also the bias gradient
 * of the GCN layer (a torch reduction there could record memset nodes into a
 * captured graph). */
int armsp_colsum_f32(int M, int N, const float* X, float* out, void* stream);
/* GCN hidden-layer epilogue in one pass (replaces the bias add of GCNConv,
 * op.py:122-124, and the F.relu / F.dropout of Net.forward, gcn_custom.py:84-86):
 * H[m,n] = bf16(keep·max(Y[m,n] + bias[n], 0)/(1-p)) with H contiguous [M×N]
 * bf16, and mask[m·N/8 + n/8] bit n%8 = (Y + bias > 0 and kept).  keep draws a
 * 24-bit uniform per element from a counter-based hash of (*seed, m·N + n);
 * with p > 0 the call then advances *seed (device memory) by one on the same
 * stream, so a captured graph draws a new mask per replay.  N % 8 == 0,
 * ldy % 4 == 0, Y, H and bias 16-byte aligned; bias may be NULL; 0 <= p < 1. */
int armsp_bias_relu_dropout_bf16(int M, int N, const float* Y, int64_t ldy, const float* bias, float p,
                                 uint64_t* seed, uint16_t* H, uint8_t* mask, void* stream);
/* Its backward: GY = bf16(mask ? G/(1-p) : 0) (G contiguous [M×N] bf16, GY
 * rows of pitch ldgy, a multiple of 8, so that the gradient can be written in
 * the layout the next SpMM gathers best) and, when dbias is not NULL,
 * dbias[n] = Σ_m of the same fp32 values, summed in a fixed order
 * (deterministic).  N % 8 == 0, N <= 2048. */
int armsp_bias_relu_dropout_backward_bf16(int M, int N, const uint16_t* G, const uint8_t* mask, float p,
                                          uint16_t* GY, int64_t ldgy, float* dbias, void* stream);
/* One Adam step (L2 weight decay added to the gradient); step counts from 1. */
int armsp_adam_step(int64_t n, float* param, const float* grad, float* m, float* v, float lr, float beta1,
                    float beta2, float eps, float weight_decay, int step, void* stream);
/* MLP inference: x[nvec×F] standardised by (mu, sd), h1 = relu(x·W1 + b1),
 * h2 = relu(h1·W2 + b2), logits = h2·W3 + b3 (W row-major [in×out]);
 * cls[v] = argmax; logits (nullable) written.  F, H, C <= 256. */
int armsp_mlp_infer(int nvec, int F, int H, int C, const float* x, const float* mu, const float* sd, const float* W1,
                    const float* b1, const float* W2, const float* b2, const float* W3, const float* b3,
                    float* logits, int* cls, void* stream);

/* Measurement probe (no reference counterpart): the B-row gathers of the SpMM
 * kernels alone.  For i in [0, n) it reads the column slice of row idx[i] of B
 * (K rows of N fp32 columns, row stride ldb) that its lane group covers:
 * `group` lanes (8, 16, 32, 64) of `vec` floats (2 or 4) each; the N / (group·vec)
 * slices (a divisor of 8) are spread over the XCDs as in the SpMM slice kernel.
 * Writes per-lane sums to out (at least 64·S·ceil(n/512) floats, S the slice
 * count, rounded up to whole 4-wave groups).  idx values must lie in [0, K):
 * like the SpMM kernels' column indices they are not range-checked on the
 * device.  Its rate on a matrix's column indices against relabelled or uniform
 * ones shows what the gather pattern costs by itself. */
int armsp_probe_gather_f32(int n, const int* idx, const float* B, int K, int ldb, int N, int group, int vec,
                           float* out, void* stream);

/* The learned selector as the library's own AUTO: after armsp_selector_set,
 * every SpMM (fp32 B, no epilogue) or CSR SDDMM (fp32) call whose armsp_config
 * is NULL or all zero runs the configuration the network picks for its matrix
 * and width.  The first such call on a matrix computes the features on the
 * device (armsp_row_stats, a features kernel), runs armsp_mlp_infer and reads
 * the chosen class back (one synchronisation of the call's stream; the
 * network's weights are copied to the device at first use); the choice is then
 * cached per (op, rowptr address, M, nnz, width, element size) until the next
 * armsp_selector_set or armsp_selector_clear.  A caller that reuses a rowptr
 * address for another structure may get a slower schedule, never a different
 * result.  Under HIP graph capture, and in armsp_spmm_f32_noalloc, only cached
 * choices are used (else the heuristic).  op: ARMSP_SELECTOR_SPMM or
 * ARMSP_SELECTOR_SDDMM; F must be 17 (the features above: log2 of M, nnz + 1,
 * nnz/M + 1, max row length + 1 and the width, then the fraction of rows in
 * each of 12 log-scale row-length bins, as armsp_row_stats counts them); the
 * arrays are host memory, W row-major [in×out]; classes[c] is the configuration
 * of class c (all zero: the library's heuristic).  No reference counterpart
 * (the reference picks its kernel by a fixed N threshold,
 * gcn/pytorch-custom/spmm_kernel.cu:437-457). */
#define ARMSP_SELECTOR_SPMM 0
#define ARMSP_SELECTOR_SDDMM 1
int armsp_selector_set(int op, int F, int H, int C, const float* mu, const float* sd, const float* W1,
                       const float* b1, const float* W2, const float* b2, const float* W3, const float* b3,
                       const armsp_config* classes);
/* Removes op's network: its AUTO calls use the heuristic again. */
int armsp_selector_clear(int op);
/* *out = the configuration an AUTO call of op on this matrix and width would
 * run (all zero: the heuristic); computes and caches the choice like a call. */
int armsp_selector_choose(int op, int M, int nnz, int N, int element_bytes, const int* rowptr,
                          armsp_config* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ARMSP_H */
