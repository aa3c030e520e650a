// Host-side code of the library and the oracle under AddressSanitizer and
// UndefinedBehaviorSanitizer: the seeded generators, the format converters and
// the CSR check (armsparse_amd/csrc/gen.cpp), the Matrix Market reader
// (armsparse_amd/csrc/mtx.cpp) on the reference's data files and on malformed
// ones, and the oracle's restated kernels (oracle/oracle.c) against plain loops.
// Exit status 0 when every check passes; a sanitizer report aborts the run.
//   host_checks <dir for temporary files> <file.mtx>...
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "armsp_gen.h"

extern "C" {
void orc_rowBaseSpmm3(const int*, const int*, const float*, const float*, float*, int, int, int, int);
void orc_rowBaseSpmm(const int*, const int*, const float*, const float*, float*, int, int, int, int);
void orc_origin(const int*, const int*, const float*, const float*, float*, int, int, int, int);
void orc_irregular(const int*, const int*, const float*, const float*, float*, int, int, int, int);
void orc_sparse_matrix_mul3(const int*, const int*, const float*, const int*, const int*, const float*, float*, int,
                            int, int, int);
void orc_sparse_dense_vector(const float*, float*, const int*, const int*, const float*, int, int, int);
int orc_toCSR(const float*, int, int, int*, int*, float*);
int orc_toCSC(const float*, int, int, int*, int*, float*);
int orc_toVectorAwareCSR(const float*, int, int, int*, int*, float*, float);
int orc_toDataReuseFormat(float*, int, int, int*, int*, float*, unsigned char*);
void orc_DataReuseSpmm(const int*, const int*, const float*, const unsigned char*, const float*, float*, int, int, int,
                       int);
void orc_csr_to_csc(const int*, const int*, const float*, int, int, int*, int*, float*);
void orc_spmm_f64(const int*, const int*, const float*, const float*, int, const float*, double*, int, int, int);
void orc_sddmm_f64(const int*, const int*, const float*, const float*, int, const float*, int, int, double*, int, int);
}

static int failures = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

struct Csr {
    int M = 0, K = 0;
    std::vector<int> rp, ci;
    std::vector<float> va;
};

static void generators() {
    Csr a;
    a.M = 300;
    a.K = 200;
    a.rp.resize(301);
    a.ci.resize(3000);
    a.va.resize(3000);
    CHECK(armsp_gen_uniform_csr(300, 200, 3000, 7, a.rp.data(), a.ci.data(), a.va.data()) == 0);
    CHECK(armsp_csr_check(300, 200, 3000, a.rp.data(), a.ci.data()) == 0);
    std::vector<int> rp(1025), ci(8192);
    std::vector<float> va(8192);
    CHECK(armsp_gen_rmat_csr(10, 8192, 0.57, 0.19, 0.19, 3, rp.data(), ci.data(), va.data()) == 0);
    CHECK(armsp_csr_check(1024, 1024, 8192, rp.data(), ci.data()) == 0);
    std::vector<int> grp(513), gci(2 * 1500 + 512);
    CHECK(armsp_gen_gcn_graph(512, 1500, 3.0, 5, grp.data(), gci.data()) == 0);
    CHECK(armsp_csr_check(512, 512, grp[512], grp.data(), gci.data()) == 0);
    for (int kind = 0; kind < 4; ++kind) {
        const int M = 1024;
        const int64_t cap = kind == 3 ? 32 * (int64_t)M * 2 : (int64_t)M * 20;
        std::vector<int> crp(M + 1), cci(cap);
        std::vector<float> cva(cap);
        const double param = kind == 1 ? 0.6 : kind == 2 ? 16.0 : kind == 3 ? 2.0 : 0.0;
        const int64_t n = armsp_gen_corpus_csr(kind, M, kind == 3 ? 64.0 : 8.0, param, 11 + kind, cap, crp.data(),
                                               cci.data(), cva.data());
        CHECK(n >= 0 && n <= cap);
        if (n >= 0) CHECK(armsp_csr_check(M, M, n, crp.data(), cci.data()) == 0);
    }
    // converters
    std::vector<int> vrp(301);
    const int64_t np = armsp_csr_to_vcsr(300, 200, a.rp.data(), a.ci.data(), a.va.data(), 0.0f, vrp.data(), nullptr,
                                         nullptr);
    CHECK(np > 0 && np % 4 == 0);
    std::vector<int> vci(np);
    std::vector<float> vva(np);
    CHECK(armsp_csr_to_vcsr(300, 200, a.rp.data(), a.ci.data(), a.va.data(), 0.0f, vrp.data(), vci.data(),
                            vva.data()) == np);
    std::vector<int> drp(301), dci(3000);
    std::vector<float> dva(3000);
    std::vector<uint8_t> dfl(3000);
    CHECK(armsp_csr_to_datareuse(300, 200, a.rp.data(), a.ci.data(), a.va.data(), drp.data(), dci.data(), dva.data(),
                                 dfl.data()) == 0);
    // malformed CSR
    int bad_rp[] = {0, 2, 1}, bad_ci[] = {0, 1};
    CHECK(armsp_csr_check(2, 2, 2, bad_rp, bad_ci) == 2);
    int rp2[] = {0, 1, 2}, ci2[] = {0, 9};
    CHECK(armsp_csr_check(2, 2, 2, rp2, ci2) == 4);
}

static void write_file(const std::string& path, const char* text) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return;
    std::fputs(text, f);
    std::fclose(f);
}

static void matrix_market(const std::string& tmp, int nfiles, char** files) {
    for (int i = 0; i < nfiles; ++i) {
        int M = 0, K = 0;
        int64_t nnz = 0;
        CHECK(armsp_read_mtx(files[i], 0, &M, &K, &nnz, nullptr, nullptr, nullptr) == 0);
        std::vector<int> rp(M + 1), ci(nnz);
        std::vector<float> va(nnz);
        CHECK(armsp_read_mtx(files[i], 0, &M, &K, &nnz, rp.data(), ci.data(), va.data()) == 0);
        CHECK(armsp_csr_check(M, K, nnz, rp.data(), ci.data()) == 0);
    }
    const char* bad[] = {
        "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n",             // fewer entries than declared
        "%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n",             // row out of range
        "%%MatrixMarket matrix coordinate real general\n3 3 1\n1 0 1.0\n",             // column 0
        "%%MatrixMarket matrix coordinate real general\n-3 3 1\n1 1 1.0\n",            // negative size
        "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",                 // dense format
        "not a matrix market file\n",
        "%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1 abc\n",             // bad value
    };
    for (size_t i = 0; i < sizeof(bad) / sizeof(bad[0]); ++i) {
        const std::string p = tmp + "/bad" + std::to_string(i) + ".mtx";
        write_file(p, bad[i]);
        int M = 0, K = 0;
        int64_t nnz = 0;
        CHECK(armsp_read_mtx(p.c_str(), 0, &M, &K, &nnz, nullptr, nullptr, nullptr) != 0);
    }
    int M = 0, K = 0;
    int64_t nnz = 0;
    CHECK(armsp_read_mtx((tmp + "/does_not_exist.mtx").c_str(), 0, &M, &K, &nnz, nullptr, nullptr, nullptr) != 0);
}

static void oracle_kernels() {
    const int M = 64, K = 48, N = 24;
    std::vector<float> dense((size_t)M * K, 0.0f), B((size_t)K * N), C((size_t)M * N, 0.0f);
    for (int i = 0; i < M * K; ++i)
        if (armsp_rng_uniform(1, 1, i, 0.0f, 1.0f) < 0.15f) dense[i] = armsp_rng_uniform(1, 2, i, -1.0f, 1.0f);
    for (int i = 0; i < K * N; ++i) B[i] = armsp_rng_uniform(1, 3, i, -1.0f, 1.0f);
    std::vector<int> rp(M + 1), ci(M * K);
    std::vector<float> va(M * K);
    const int nnz = orc_toCSR(dense.data(), M, K, rp.data(), ci.data(), va.data());
    CHECK(nnz > 0 && rp[M] == nnz);
    orc_rowBaseSpmm3(rp.data(), ci.data(), va.data(), B.data(), C.data(), M, K, N, 2);
    std::vector<double> ref((size_t)M * N);
    orc_spmm_f64(rp.data(), ci.data(), va.data(), B.data(), N, nullptr, ref.data(), M, N, 2);
    for (int i = 0; i < M * N; ++i) CHECK(std::fabs(C[i] - ref[i]) <= 1e-5 * (1.0 + std::fabs(ref[i])));
    std::vector<float> C2((size_t)M * N, 0.0f);
    orc_origin(rp.data(), ci.data(), va.data(), B.data(), C2.data(), M, K, N, 2);
    orc_rowBaseSpmm(rp.data(), ci.data(), va.data(), B.data(), C2.data(), M, K, N, 2);
    orc_irregular(rp.data(), ci.data(), va.data(), B.data(), C2.data(), M, K, N, 2);
    std::vector<int> cp(K + 1), ri(nnz);
    std::vector<float> vt(nnz);
    orc_csr_to_csc(rp.data(), ci.data(), va.data(), M, K, cp.data(), ri.data(), vt.data());
    std::vector<int> cp2(K + 1), ri2(M * K);
    std::vector<float> vt2(M * K);
    CHECK(orc_toCSC(dense.data(), M, K, cp2.data(), ri2.data(), vt2.data()) == nnz);
    for (int i = 0; i <= K; ++i) CHECK(cp[i] == cp2[i]);
    // blocked CSR and its SpMV
    std::vector<int> vrp(M + 1), vci(M * K);
    std::vector<float> vva(M * K), x(K, 1.0f), y(M, 0.0f);
    const int np = orc_toVectorAwareCSR(dense.data(), M, K, vrp.data(), vci.data(), vva.data(), 0.0f);
    CHECK(np % 4 == 0);
    orc_sparse_dense_vector(x.data(), y.data(), vrp.data(), vci.data(), vva.data(), M, 2, 0);
    // sparse × sparse
    std::vector<float> Cd((size_t)M * K, 0.0f);
    std::vector<int> rpb(K + 1), cib(K * K);
    std::vector<float> vab(K * K), db((size_t)K * K, 0.0f);
    for (int i = 0; i < K; ++i) db[(size_t)i * K + (i * 7) % K] = 1.0f;
    orc_toCSR(db.data(), K, K, rpb.data(), cib.data(), vab.data());
    orc_sparse_matrix_mul3(rp.data(), ci.data(), va.data(), rpb.data(), cib.data(), vab.data(), Cd.data(), M, K, K, 2);
    // DataReuse format and SpMM
    std::vector<float> d2 = dense;
    std::vector<int> drp(M + 1), dci(M * K);
    std::vector<float> dva(M * K), C3((size_t)M * N, 0.0f);
    std::vector<unsigned char> dfl(M * K);
    const int dn = orc_toDataReuseFormat(d2.data(), M, K, drp.data(), dci.data(), dva.data(), dfl.data());
    CHECK(dn == nnz);
    orc_DataReuseSpmm(drp.data(), dci.data(), dva.data(), dfl.data(), B.data(), C3.data(), M, K, N, 2);
    for (int i = 0; i < M * N; ++i) CHECK(std::fabs(C3[i] - ref[i]) <= 1e-5 * (1.0 + std::fabs(ref[i])));
    // SDDMM in FP64
    std::vector<float> U((size_t)M * 8, 0.5f), V((size_t)K * 8, 0.25f);
    std::vector<double> out(nnz);
    orc_sddmm_f64(rp.data(), ci.data(), nullptr, U.data(), 8, V.data(), 8, 8, out.data(), M, 2);
    for (int i = 0; i < nnz; ++i) CHECK(std::fabs(out[i] - 1.0) < 1e-12);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: host_checks <tmpdir> <file.mtx>...\n");
        return 2;
    }
    generators();
    matrix_market(argv[1], argc - 2, argv + 2);
    oracle_kernels();
    if (failures) std::fprintf(stderr, "%d checks failed\n", failures);
    return failures ? 1 : 0;
}
