/*
 * armsp_gen.h — seeded, CSR-native input generators and host format helpers
 * of libarmsparse.so.  Pure host code: callable without a GPU.
 *
 * The reference draws its inputs with srand(time(NULL)) into a dense M×K
 * array and converts (SDMM.cpp:329-370, SpMV.cpp:39-50), which is neither
 * reproducible nor able to reach 1M×1M.  These generators replace
 * prepareSparse/toCSR for the benchmark shapes: every value is a pure
 * function of (seed, stream, counter), so host and device harnesses rebuild
 * identical inputs and no large data is stored.
 *
 * Counter-based RNG:  key(seed, s) = mix64(seed + G·(s+1)),
 *                     r(seed, s, i) = mix64(key + G·(i+1)),
 * G = 0x9E3779B97F4A7C15 and mix64 the splitmix64 finaliser.
 * U(lo,hi) from r: lo + (hi-lo)·((r>>40)·2^-24)   (exact for U(-1,1)).
 *
 * Canonical CSR everywhere: rowptr[0]=0, non-decreasing, rowptr[M]=nnz,
 * column indices strictly increasing inside a row, all < K — the structure
 * toCSR (SDMM.cpp:355-370) produces from a dense matrix.
 */
#ifndef ARMSP_GEN_H
#define ARMSP_GEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* RNG streams used by the generators (documented so other harnesses can
 * regenerate the same data). */
#define ARMSP_STREAM_POS    1  /* uniform positions */
#define ARMSP_STREAM_VAL    2  /* nonzero values, indexed by final CSR position */
#define ARMSP_STREAM_RMAT   3  /* R-MAT quadrant draws: edge i uses draws 8i..8i+7, 16 bits per level */
#define ARMSP_STREAM_GRAPH  4  /* Chung–Lu endpoint draws */
#define ARMSP_STREAM_DENSE  16 /* first stream for dense operands (B, x, U, V ...) */

uint64_t armsp_rng_u64(uint64_t seed, uint64_t stream, uint64_t i);
float armsp_rng_uniform(uint64_t seed, uint64_t stream, uint64_t i, float lo, float hi);

/* Dense fills on the host.  normal: Box–Muller from draws 2i, 2i+1. */
int armsp_host_fill_uniform(int64_t n, uint64_t seed, uint64_t stream, float lo, float hi, float* out);
int armsp_host_fill_normal(int64_t n, uint64_t seed, uint64_t stream, float* out);

/* Exactly `nnz` distinct uniformly placed positions in M×K (first `nnz`
 * distinct draws of the position stream), values U(-1,1).  val may be NULL. */
int armsp_gen_uniform_csr(int M, int K, int64_t nnz, uint64_t seed,
                          int* rowptr, int* colind, float* val);

/* R-MAT on 2^scale × 2^scale with quadrant probabilities a, b, c (d = 1-a-b-c):
 * exactly `nnz` distinct edges (first distinct draws), values U(-1,1). */
int armsp_gen_rmat_csr(int scale, int64_t nnz, double a, double b, double c, uint64_t seed,
                       int* rowptr, int* colind, float* val);

/* Symmetric Chung–Lu power-law graph on n nodes with exactly `pairs`
 * distinct undirected non-loop edges; node weights w_i ∝ (i+1)^(-1/(gamma-1)).
 * Output: CSR of the directed adjacency (2·pairs entries, values omitted). */
int armsp_gen_powerlaw_graph(int n, int64_t pairs, double gamma, uint64_t seed,
                             int* rowptr, int* colind);

/* The same graph with a self loop on every node, as the GCN data path builds
 * A + I (gcn/pytorch-custom/gcn_custom.py:38-58): 2·pairs + n entries. */
int armsp_gen_gcn_graph(int n, int64_t pairs, double gamma, uint64_t seed,
                        int* rowptr, int* colind);

/* Row-wise generators for the schedule-selector corpus (square M×M).  They
 * fill caller arrays of capacity `cap` and return the nonzero count, or -1 if
 * `cap` is too small.  Duplicates inside a row are merged, so the count is
 * data dependent; an upper bound is M·(ceil(avg_row)+1) for kinds 0-2 and
 * 32·M·param for kind 3.  kind: 0 uniform, 1 R-MAT (M a power of two;
 * a = param, b = c = 0.4(1-a), d = 0.2(1-a)), 2 banded (half-bandwidth =
 * param), 3 block-diagonal with 32×32 dense blocks (param = blocks per
 * block-row, the diagonal block plus param-1 random ones).  Values U(-1,1). */
int64_t armsp_gen_corpus_csr(int kind, int M, double avg_row, double param, uint64_t seed,
                             int64_t cap, int* rowptr, int* colind, float* val);

/* CSR → zero-padded 1×4-blocked CSR (toVectorAwareCSR, SpMV.cpp:15-38; with
 * threshold > 0 the |v| > threshold trigger of numaAware/numaAwareSpMV.cpp:48).
 * A block [4b, 4b+4) of a row is stored iff some stored value in it triggers;
 * its four slots hold the row's values (0 where absent).  Columns ≥ K are
 * padded with zeros (the reference reads past the row end there).  Call with
 * out_colind == NULL to get the padded count (out_rowptr filled). */
int64_t armsp_csr_to_vcsr(int M, int K, const int* rowptr, const int* colind, const float* val,
                          float threshold, int* out_rowptr, int* out_colind, float* out_val);

/* Canonical CSR → the 2-row DataReuse format of toDataReuseFormat
 * (suitesparse_sdmm.cpp:150-179, datasets/suitesparse.cpp:229-263), built
 * from CSR instead of a dense matrix: after each entry (r, c) of an even row
 * r, the entry (r+1, c) of the next row, if any, follows with flag 1; an odd
 * row keeps only the columns its even partner lacks.  Same entry count as the
 * CSR (explicit zeros of the CSR stay entries; a dense matrix has none);
 * outputs out_rowptr[M+1], out_colind/out_val/out_flag[nnz] (val may be NULL:
 * values 1).  Returns 0, 1 bad arguments, 2 not canonical CSR. */
int armsp_csr_to_datareuse(int M, int K, const int* rowptr, const int* colind, const float* val,
                           int* out_rowptr, int* out_colind, float* out_val, uint8_t* out_flag);

/* Validates canonical CSR metadata on the host: returns 0 if valid, else a
 * positive code (1 rowptr[0]!=0, 2 decreasing rowptr, 3 rowptr[M]!=nnz,
 * 4 column out of range, 5 columns not strictly increasing). */
int armsp_csr_check(int M, int K, int64_t nnz, const int* rowptr, const int* colind);

/* Matrix Market coordinate file → canonical CSR (host).  Replaces the
 * SuiteSparse drivers' reader (suitesparse_spmm.cpp:90-121,
 * suitesparse_sdmm.cpp:63-90, suitesparse_spmv.cpp:73-100: dense scatter, a
 * later entry overwrites an earlier one, toCSR keeps nonzeros) and readMtx
 * (gcn/util/util.hpp:286-333: pattern entries 1.0, symmetric files mirrored).
 * Same rules here: last write wins, zeros dropped, pattern = 1.0, symmetric
 * files mirrored (skew-symmetric negated) unless ARMSP_MTX_STORED_TRIANGLE.
 * Call once with rowptr = colind = val = NULL for *M, *K, *nnz, then with
 * rowptr[M+1], colind[nnz], val[nnz] (val may stay NULL) and *nnz unchanged.
 * Returns 0, or 1 cannot open, 2 not a real/integer/pattern coordinate
 * matrix, 3 malformed size line or entry (or fewer entries than declared),
 * 4 index out of range, 5 bad arguments. */
#define ARMSP_MTX_STORED_TRIANGLE 0x1
int armsp_read_mtx(const char* path, int flags, int* M, int* K, int64_t* nnz,
                   int* rowptr, int* colind, float* val);

/* CSR values are never exactly zero (toCSR drops zeros, SDMM.cpp:362): a drawn
 * 0 is replaced by 2^-24. */

#ifdef __cplusplus
}
#endif
#endif /* ARMSP_GEN_H */
