// C-linkage entry for the reference's own sparse_matrix_mul3
// (numaAware/numaAwareSpMM.cpp:452-470), compiled from the reference tree by
// oracle/Makefile.  Test infrastructure only.
void sparse_matrix_mul3(int rowPtr1[], int colInd1[], float val1[], int rowPtr2[], int colInd2[], float val2[],
                        float* matrixC, int M, int K, int N, int threads);

extern "C" void ref_sparse_matrix_mul3(int* rowPtr1, int* colInd1, float* val1, int* rowPtr2, int* colInd2,
                                       float* val2, float* matrixC, int M, int K, int N, int threads) {
    sparse_matrix_mul3(rowPtr1, colInd1, val1, rowPtr2, colInd2, val2, matrixC, M, K, N, threads);
}
