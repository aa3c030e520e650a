// CSR → CSC transpose (placeholder until the radix-sort kernels land).
#include "common.hpp"

extern "C" int armsp_csr_transpose(int M, int K, int nnz, const int* rowptr, const int* colind, const float* val,
                                   int* colptr, int* rowind, float* valT, void* stream) {
    (void)M; (void)K; (void)nnz; (void)rowptr; (void)colind; (void)val; (void)colptr; (void)rowind; (void)valT;
    (void)stream;
    return ARMSP_ERROR_NOT_SUPPORTED;
}
