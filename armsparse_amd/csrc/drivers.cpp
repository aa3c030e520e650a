// Command-line drivers with the reference's names, arguments and output lines,
// running on the MI355X through libarmsparse.  One binary; the behaviour is
// picked by the name it is invoked as (armsparse_amd/bin/<name> symlinks):
//
//   SDMM M K N threads ratio        (SDMM.cpp:181-327)
//   SpMM M K N threads              (SpMM.cpp:105-211, sparse × sparse → dense)
//   SpMV M K threads                (SpMV.cpp:52-127, blocked CSR, ratio 0.6)
//   numaAwareSDMM M K N             (numaAware/numaAwareSDMM.cpp:265-397)
//   numaAwareSPDM M K N             (numaAware/numaAwareSPDM.cpp)
//   balanceNUMAAwareSPDM M K N      (numaAware/balanceNUMAAwareSPDM.cpp)
//   numaAwareSpMV M K threads       (numaAware/numaAwareSpMV.cpp:126-)
//   numaAwareSpMM M K N threads     (numaAware/numaAwareSpMM.cpp)
//   suitesparse_sdmm threads [file] (suitesparse_sdmm.cpp:63-147: rowBaseSpmm and DataReuseSpmm, N = 1024)
//   suitesparse_spmm threads [file] (suitesparse_spmm.cpp:90-172: sparse_matrix_mul3, A·A)
//   suitesparse_spmv threads [file] (suitesparse_spmv.cpp:73-132: blocked-CSR SpMV, x = 5)
//
// The suitesparse drivers read a Matrix Market file (default: the reference's
// hard-coded datasets/ path) as the reference does — stored entries only, a
// later entry overwriting an earlier one, zeros dropped — and print its lines.
//
// Inputs come from the seeded CSR-native generators instead of
// srand(time(NULL)) + a dense M×K array (the reference's densities and value
// ranges are kept: (rand%20000-10000)/1000 becomes U(-10,10), rand%10 becomes
// U(0,10)).  Times are HIP-event kernel times in seconds with the operands
// resident on the device; the NUMA drivers run as one "node" (one GPU: the
// merge-path schedule balances the whole matrix, so there is no node split and
// no sumUp).  `threads` is accepted and ignored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "armsp.h"
#include "armsp_gen.h"

namespace {

[[noreturn]] void die(const char* what) {
    std::fprintf(stderr, "armsparse: %s\n", what);
    std::exit(EXIT_FAILURE);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "armsparse: %s: %s\n", what, hipGetErrorString(e));
        std::exit(EXIT_FAILURE);
    }
}

void st_check(int st, const char* what) {
    if (st != ARMSP_SUCCESS) {
        std::fprintf(stderr, "armsparse: %s: %s\n", what, armsp_status_string(st));
        std::exit(EXIT_FAILURE);
    }
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

long arg(int argc, char** argv, int i, const char* usage) {
    if (i >= argc) die(usage);
    char* end = nullptr;
    long v = std::strtol(argv[i], &end, 10);
    if (!end || *end || v < 0) die(usage);
    return v;
}

template <typename T>
T* upload(const std::vector<T>& h) {
    T* d = nullptr;
    hip_check(hipMalloc(&d, std::max<size_t>(h.size(), 1) * sizeof(T)), "hipMalloc");
    if (!h.empty()) hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    return d;
}

struct Csr {
    int M = 0, K = 0;
    std::vector<int> rp, ci;
    std::vector<float> va;
    int nnz() const { return rp.empty() ? 0 : rp.back(); }
};

// nnz distinct uniform positions, values U(lo, hi) (seeded)
Csr gen_uniform(int M, int K, double ratio, uint64_t seed, float lo, float hi) {
    Csr a;
    a.M = M;
    a.K = K;
    const int64_t nnz = std::min<int64_t>((int64_t)std::llround(ratio * (double)M * (double)K), (int64_t)M * K);
    a.rp.resize(M + 1);
    a.ci.resize(nnz);
    a.va.resize(nnz);
    if (armsp_gen_uniform_csr(M, K, nnz, seed, a.rp.data(), a.ci.data(), a.va.data())) die("generator failed");
    for (auto& v : a.va) v = lo + (hi - lo) * (0.5f * (v + 1.0f));  // U(-1,1) → U(lo,hi)
    for (auto& v : a.va)
        if (v == 0.0f) v = hi;  // CSR keeps nonzeros only
    return a;
}

std::vector<float> dense_uniform(int64_t n, uint64_t seed, float lo, float hi) {
    std::vector<float> v(n);
    if (armsp_host_fill_uniform(n, seed, ARMSP_STREAM_DENSE, lo, hi, v.data())) die("fill failed");
    return v;
}

// Event-timed device launches (mean seconds over `ite` after one warm-up).
template <class F>
double time_kernel(F launch, int ite) {
    launch();
    hip_check(hipDeviceSynchronize(), "warm-up");
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
    hip_check(hipEventRecord(a, nullptr), "event");
    for (int i = 0; i < ite; ++i) launch();
    hip_check(hipEventRecord(b, nullptr), "event");
    hip_check(hipEventSynchronize(b), "event");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, a, b), "event");
    hipEventDestroy(a);
    hipEventDestroy(b);
    return ms / 1e3 / ite;
}

// Result check against an FP64 product on the host: the reference's validate()
// tolerance (|gpu - exact| <= 0.1 everywhere, numaAwareSDMM.cpp:227-264) on the
// result of one call and, stricter, a norm-wise relative error of at most 1e-5 per
// row.  Prints both
// figures and "Correct Result" / "False Result"; returns 1, 0, or -1 when the
// output is too large to check on the host (nothing printed).
struct Check {
    double max_abs = 0.0, max_rel = 0.0;
    // got holds `scale` accumulated calls: the per-call result got/scale is compared
    void row(const double* exact, const float* got, int n, double scale) {
        double num = 0.0, den = 0.0;
        for (int j = 0; j < n; ++j) {
            const double d = (double)got[j] / scale - exact[j];
            max_abs = std::max(max_abs, std::fabs(d));
            num += d * d;
            den += exact[j] * exact[j];
        }
        if (den > 0.0) max_rel = std::max(max_rel, std::sqrt(num / den));
        else if (num > 0.0) max_rel = INFINITY;
    }
    int report() const {
        std::printf("max |C - FP64|: %g  max row rel err vs FP64: %.2e\n", max_abs, max_rel);
        const bool ok = max_abs <= 0.1 && max_rel <= 1e-5;
        std::printf(ok ? "Correct Result\n" : "False Result\n");
        return ok ? 1 : 0;
    }
};

int validate_spmm(const Csr& a, const std::vector<float>& B, int N, const std::vector<float>& C, double scale) {
    if ((int64_t)a.M * N > (1 << 26)) return -1;
    Check ck;
    std::vector<double> acc(N);
    for (int m = 0; m < a.M; ++m) {
        std::fill(acc.begin(), acc.end(), 0.0);
        for (int p = a.rp[m]; p < a.rp[m + 1]; ++p)
            for (int j = 0; j < N; ++j) acc[j] += (double)a.va[p] * B[(size_t)a.ci[p] * N + j];
        ck.row(acc.data(), &C[(size_t)m * N], N, scale);
    }
    return ck.report();
}

// ------------------------------------------------------------- CSR × dense
int run_spmm(const char* name, int M, int K, int N, double ratio, float vlo, float vhi, float blo, float bhi,
             bool numa_lines) {
    double t0 = now();
    Csr a = gen_uniform(M, K, ratio, 0xA5A50000ULL, vlo, vhi);
    const double to_csr = now() - t0;
    std::vector<float> B = dense_uniform((int64_t)K * N, 0xA5A50000ULL, blo, bhi);
    int* rp = upload(a.rp);
    int* ci = upload(a.ci);
    float* va = upload(a.va);
    t0 = now();
    float* dB = upload(B);
    const double memcopy = now() - t0;
    float* dC = nullptr;
    hip_check(hipMalloc(&dC, std::max<size_t>((size_t)M * N, 1) * 4), "hipMalloc");
    hip_check(hipMemset(dC, 0, (size_t)M * N * 4), "memset");
    armsp_config row{ARMSP_SCHED_ROW, 0, 0, 0, 0}, merge{ARMSP_SCHED_MERGE, 0, 0, 0, 0};
    const int ite = numa_lines ? 1 : 1;
    // rowBaseSpmm3 semantics: C += A·B, the reference's row order (bit-exact)
    const double take = time_kernel(
        [&] { st_check(armsp_spmm_f32(M, K, N, a.nnz(), rp, ci, va, dB, N, dC, N, ARMSP_ACCUMULATE, &row, nullptr),
                       "spmm"); }, ite);
    const double take2 = time_kernel(
        [&] { st_check(armsp_spmm_f32(M, K, N, a.nnz(), rp, ci, va, dB, N, dC, N, ARMSP_ACCUMULATE, &merge, nullptr),
                       "spmm"); }, ite);
    // the reference's phases on one device: time1 = launching the work (its
    // pthread_create loop), time2 = waiting for it (the joins), sumUp = adding the
    // partial results — here the merge path's carry fixup, the pass that adds the
    // partial sums of rows cut between chunks (event-timed)
    double time1 = 0.0, time2 = 0.0, sumup = 0.0;
    if (numa_lines) {
        armsp_timing_enable(1);
        armsp_timing_reset();
        const double h0 = now();
        st_check(armsp_spmm_f32(M, K, N, a.nnz(), rp, ci, va, dB, N, dC, N, ARMSP_ACCUMULATE, &merge, nullptr),
                 "spmm");
        const double h1 = now();
        hip_check(hipDeviceSynchronize(), "spmm");
        const double h2 = now();
        time1 = h1 - h0;
        time2 = h2 - h1;
        double ms = 0.0;
        int64_t n = 0;
        armsp_timing_read("spmm_fixup", &ms, &n);
        sumup = ms / 1e3;
        armsp_timing_enable(0);
    }
    std::vector<float> C((size_t)M * N);
    hip_check(hipMemcpy(C.data(), dC, C.size() * 4, hipMemcpyDeviceToHost), "download");
    if (!numa_lines) {
        std::printf("toCSRTime:%g\n", to_csr);
        std::printf("m k n ratio:%d %d %d %g  Time take:%g\n", M, K, N, ratio, take);
        std::printf("blas_timeTake:%g\n", take2);
    } else {
        std::printf("Running parallelSum------------------\n");
        std::printf("0 time %lf\n", take2);
        std::printf("TimeTake %lf s\n", take2);
        std::printf("time1 %lf\n", time1);
        std::printf("time2 %lf\n", time2);
        std::printf("--memCopyTime: %lf s   sumUpTime: %lf s\n", memcopy, sumup);
    }
    // C now holds the sum of every call: 2·(1+ite) (+1 for the phase-timed call) times A·B,
    // as the reference's C = (1+ite)·A·B
    const int ok = validate_spmm(a, B, N, C, 2.0 * (1 + ite) + (numa_lines ? 1.0 : 0.0));
    std::printf("nnz:%d GB/s:%.2f (%s, row-order kernel %.2f)\n", a.nnz(),
                (4.0 * (M + 1) + 8.0 * a.nnz() + 4.0 * K * N + 8.0 * M * N) / take2 / 1e9, name,
                (4.0 * (M + 1) + 8.0 * a.nnz() + 4.0 * K * N + 8.0 * M * N) / take / 1e9);
    hipFree(rp); hipFree(ci); hipFree(va); hipFree(dB); hipFree(dC);
    return ok == 0 ? EXIT_FAILURE : EXIT_SUCCESS;
}

// ----------------------------------------------------- sparse × sparse
int run_spgemm(int M, int K, int N, double ratio, bool numa_lines) {
    Csr a = gen_uniform(M, K, ratio, 0xA5A50000ULL, 0.0f, 10.0f);
    Csr b = gen_uniform(K, N, ratio, 0xA5A50001ULL, 0.0f, 10.0f);
    int *rp1 = upload(a.rp), *ci1 = upload(a.ci), *rp2 = upload(b.rp), *ci2 = upload(b.ci);
    float *v1 = upload(a.va), *v2 = upload(b.va);
    float* dC = nullptr;
    hip_check(hipMalloc(&dC, std::max<size_t>((size_t)M * N, 1) * 4), "hipMalloc");
    hip_check(hipMemset(dC, 0, (size_t)M * N * 4), "memset");
    const int ite = 30;
    const double take = time_kernel(
        [&] { st_check(armsp_spgemm_dense_f32(M, K, N, a.nnz(), rp1, ci1, v1, b.nnz(), rp2, ci2, v2, dC, N, nullptr),
                       "spgemm"); }, ite);
    if (numa_lines) {
        // one device, no K split: there are no partial matrices to sum up
        std::printf("Running parallelSum------------------\n");
        std::printf("compute time %lf s\n", take);
        std::printf("sumup time %lf s\n", 0.0);
        std::printf("total %lf\n", take);
    } else {
        std::printf("M k N sparity: %d %d %d %g\n", M, K, N, ratio);
        std::printf("timeTake: %g\n", take);
    }
    // C = (1 + ite)·A·B (warm-up included) against an FP64 A·B
    int ok = -1;
    if ((int64_t)M * N <= (1 << 26)) {
        std::vector<float> C((size_t)M * N);
        hip_check(hipMemcpy(C.data(), dC, C.size() * 4, hipMemcpyDeviceToHost), "download");
        Check ck;
        std::vector<double> acc(N);
        for (int m = 0; m < M; ++m) {
            std::fill(acc.begin(), acc.end(), 0.0);
            for (int p = a.rp[m]; p < a.rp[m + 1]; ++p)
                for (int t = b.rp[a.ci[p]]; t < b.rp[a.ci[p] + 1]; ++t) acc[b.ci[t]] += (double)a.va[p] * b.va[t];
            ck.row(acc.data(), &C[(size_t)m * N], N, 1.0 + ite);
        }
        ok = ck.report();
    }
    hipFree(rp1); hipFree(ci1); hipFree(rp2); hipFree(ci2); hipFree(v1); hipFree(v2); hipFree(dC);
    return ok == 0 ? EXIT_FAILURE : EXIT_SUCCESS;
}

// ------------------------------------------------------------------ SpMV
int run_spmv(int M, int K, double ratio, float threshold, bool numa_lines) {
    Csr a = gen_uniform(M, K, ratio, 0xA5A50000ULL, 0.0f, 10.0f);
    std::vector<int> vrp(M + 1);
    double t0 = now();
    const int64_t np = armsp_csr_to_vcsr(M, K, a.rp.data(), a.ci.data(), a.va.data(), threshold, vrp.data(), nullptr,
                                         nullptr);
    std::vector<int> vci(np);
    std::vector<float> vva(np);
    armsp_csr_to_vcsr(M, K, a.rp.data(), a.ci.data(), a.va.data(), threshold, vrp.data(), vci.data(), vva.data());
    const double to_vcsr = now() - t0;
    const int Kp = (K + 3) / 4 * 4;
    std::vector<float> x(Kp, 5.0f), y(M, 0.0f);  // B[i] = 5, C[i] = 0 (SpMV.cpp:400-405)
    int *rp = upload(vrp), *ci = upload(vci);
    float *va = upload(vva), *dx = upload(x), *dy = upload(y);
    std::printf("%d\n", (int)np);
    const int ite = numa_lines ? 1 : 30;
    const double take = time_kernel(
        [&] { st_check(armsp_spmv_vcsr_f32(M, (int)np, rp, ci, va, dx, dy, 0, nullptr), "spmv"); }, ite);
    if (numa_lines) {
        std::printf("Running parallelSum------------------\n");
        std::printf("TimeTake %lf s\n", take);
    } else {
        std::printf("m k ratio: %d %d %g", M, K, ratio);
        std::printf("time take: %g\n", take);
        std::printf("ToTCSRTime: %g\n", to_vcsr);
    }
    // y = (1 + ite)·A·x (warm-up included) against FP64 over the padded blocks
    std::vector<float> yg(M);
    hip_check(hipMemcpy(yg.data(), dy, (size_t)M * 4, hipMemcpyDeviceToHost), "download");
    Check ck;
    for (int m = 0; m < M; ++m) {
        double e = 0.0;
        for (int p = vrp[m]; p < vrp[m + 1]; ++p) e += (double)vva[p] * x[(size_t)vci[p - (p & 3)] + (p & 3)];
        ck.row(&e, &yg[m], 1, 1.0 + ite);
    }
    const int ok = ck.report();
    hipFree(rp); hipFree(ci); hipFree(va); hipFree(dx); hipFree(dy);
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}

// ------------------------------------------------------ Matrix Market inputs
Csr read_mtx_file(const char* path, int* declared) {
    Csr a;
    int64_t nnz = 0;
    int st = armsp_read_mtx(path, ARMSP_MTX_STORED_TRIANGLE, &a.M, &a.K, &nnz, nullptr, nullptr, nullptr);
    if (st) {
        std::fprintf(stderr, "armsparse: %s: Matrix Market read error %d\n", path, st);
        std::exit(EXIT_FAILURE);
    }
    if (nnz > 0x7fffffff) die("matrix too large");
    a.rp.resize(a.M + 1);
    a.ci.resize(nnz);
    a.va.resize(nnz);
    st = armsp_read_mtx(path, ARMSP_MTX_STORED_TRIANGLE, &a.M, &a.K, &nnz, a.rp.data(), a.ci.data(), a.va.data());
    if (st) die("Matrix Market read failed");
    // the header's count, which the reference prints
    *declared = (int)nnz;
    if (FILE* f = std::fopen(path, "r")) {
        char line[4096];
        while (std::fgets(line, sizeof line, f))
            if (line[0] != '%') {
                long long m = 0, k = 0, n = 0;
                if (std::sscanf(line, "%lld %lld %lld", &m, &k, &n) == 3) *declared = (int)n;
                break;
            }
        std::fclose(f);
    }
    return a;
}

int run_suitesparse(const std::string& name, const char* path) {
    int declared = 0;
    Csr a = read_mtx_file(path, &declared);
    std::printf("%d %d %d\n", a.M, a.K, declared);
    int *rp = upload(a.rp), *ci = upload(a.ci);
    float* va = upload(a.va);
    const int ite = name == "suitesparse_spmm" ? 10 : 30;
    if (name == "suitesparse_sdmm") {  // rowBaseSpmm: whole groups of 4 nonzeros and columns
        const int N = 1024;
        std::vector<float> B = dense_uniform((int64_t)a.K * N, 0xA5A50000ULL, 0.0f, 10.0f);
        float* dB = upload(B);
        float* dC = nullptr;
        hip_check(hipMalloc(&dC, std::max<size_t>((size_t)a.M * N, 1) * 4), "hipMalloc");
        hip_check(hipMemset(dC, 0, (size_t)a.M * N * 4), "memset");
        armsp_config row{ARMSP_SCHED_ROW, 0, 0, 0, 0};
        const double t = time_kernel([&] {
            st_check(armsp_spmm_f32(a.M, a.K, N, a.nnz(), rp, ci, va, dB, N, dC, N, ARMSP_ACCUMULATE | ARMSP_TRUNC4,
                                    &row, nullptr), "spmm");
        }, ite);
        // the DataReuse format the reference builds beside it (toDataReuseFormat) and its SpMM
        std::vector<int> drp(a.M + 1), dci(a.nnz());
        std::vector<float> dva(a.nnz());
        std::vector<uint8_t> dfl(a.nnz());
        if (armsp_csr_to_datareuse(a.M, a.K, a.rp.data(), a.ci.data(), a.va.data(), drp.data(), dci.data(),
                                   dva.data(), dfl.data()))
            die("DataReuse conversion failed");
        int *xrp = upload(drp), *xci = upload(dci);
        float* xva = upload(dva);
        uint8_t* xfl = upload(dfl);
        const double t2 = time_kernel([&] {
            st_check(armsp_spmm_datareuse_f32(a.M, a.K, N, a.nnz(), xrp, xci, xva, xfl, dB, N, dC, N,
                                              ARMSP_ACCUMULATE, nullptr), "datareuse");
        }, ite);
        std::printf("m k n ratio:%d %d %d  nnz:%d\n", a.M, a.K, N, declared);
        std::printf("rowBaseSpmmTime %g\n", t);
        std::printf("DataReuseSpmmTime %g\n", t2);
        hipFree(xrp); hipFree(xci); hipFree(xva); hipFree(xfl);
        hipFree(dB);
        hipFree(dC);
    } else if (name == "suitesparse_spmm") {  // A·A into a dense C, then toCSR of C
        const int N = a.M;
        if (a.K != a.M) die("suitesparse_spmm multiplies the matrix by itself: it must be square");
        float* dC = nullptr;
        hip_check(hipMalloc(&dC, std::max<size_t>((size_t)a.M * N, 1) * 4), "hipMalloc");
        hip_check(hipMemset(dC, 0, (size_t)a.M * N * 4), "memset");
        const double t = time_kernel([&] {
            st_check(armsp_spgemm_dense_f32(a.M, a.K, N, a.nnz(), rp, ci, va, a.nnz(), rp, ci, va, dC, N, nullptr),
                     "spgemm");
        }, ite);
        std::vector<float> C((size_t)a.M * N);
        hip_check(hipMemcpy(C.data(), dC, C.size() * 4, hipMemcpyDeviceToHost), "download");
        const double t0 = now();
        int64_t cnnz = 0;
        for (float v : C) cnnz += v != 0.0f;
        const double tocsr = now() - t0;
        std::printf("tocsrTime: %g\n", tocsr);
        std::printf("m k n ratio:%d %d %d  nnz:%d\n", a.M, a.K, N, declared);
        std::printf("SpMMTime %g\n", t);
        std::printf("C nnz:%lld\n", (long long)cnnz);
        hipFree(dC);
    } else {  // blocked CSR (toVectorAwareCSR), x = 5
        std::vector<int> vrp(a.M + 1);
        const int64_t np = armsp_csr_to_vcsr(a.M, a.K, a.rp.data(), a.ci.data(), a.va.data(), 0.0f, vrp.data(),
                                             nullptr, nullptr);
        std::vector<int> vci(np);
        std::vector<float> vva(np);
        armsp_csr_to_vcsr(a.M, a.K, a.rp.data(), a.ci.data(), a.va.data(), 0.0f, vrp.data(), vci.data(), vva.data());
        const int Kp = (a.K + 3) / 4 * 4;
        std::vector<float> x(Kp, 5.0f), y(a.M, 0.0f);
        int *vr = upload(vrp), *vc = upload(vci);
        float *vv = upload(vva), *dx = upload(x), *dy = upload(y);
        std::printf("123");
        const double t = time_kernel(
            [&] { st_check(armsp_spmv_vcsr_f32(a.M, (int)np, vr, vc, vv, dx, dy, 0, nullptr), "spmv"); }, ite);
        std::printf("m k: %d %d \n", a.M, a.K);
        std::printf("time take: %g\n", t);
        hipFree(vr); hipFree(vc); hipFree(vv); hipFree(dx); hipFree(dy);
    }
    hipFree(rp); hipFree(ci); hipFree(va);
    return EXIT_SUCCESS;
}

std::string base_name(const char* p) {
    const char* s = std::strrchr(p, '/');
    return s ? s + 1 : p;
}

}  // namespace

int main(int argc, char** argv) {
    std::string name = base_name(argv[0]);
    if (name == "armsparse_driver" && argc > 1) {  // armsparse_driver <name> args...
        name = argv[1];
        ++argv;
        --argc;
    }
    if (!armsp_device_available()) die("no gfx950 device");
    if (name == "SDMM") {
        const char* u = "usage: SDMM M K N threads ratio";
        const int M = (int)arg(argc, argv, 1, u), K = (int)arg(argc, argv, 2, u), N = (int)arg(argc, argv, 3, u);
        arg(argc, argv, 4, u);
        if (argc < 6) die(u);
        const double ratio = std::strtod(argv[5], nullptr);
        return run_spmm("SDMM", M, K, N, ratio, -10.0f, 10.0f, -10.0f, 10.0f, false);
    }
    if (name == "numaAwareSDMM" || name == "numaAwareSPDM" || name == "balanceNUMAAwareSPDM") {
        const char* u = "usage: numaAwareSDMM M K N";
        const int M = (int)arg(argc, argv, 1, u), K = (int)arg(argc, argv, 2, u), N = (int)arg(argc, argv, 3, u);
        const double ratio = name == "numaAwareSDMM" ? 0.01 : 0.1;
        return run_spmm(name.c_str(), M, K, N, ratio, 0.0f, 10.0f, 0.0f, 10.0f, true);
    }
    if (name == "SpMM" || name == "numaAwareSpMM") {
        const char* u = "usage: SpMM M K N threads";
        const int M = (int)arg(argc, argv, 1, u), K = (int)arg(argc, argv, 2, u), N = (int)arg(argc, argv, 3, u);
        arg(argc, argv, 4, u);
        return run_spgemm(M, K, N, name == "SpMM" ? 0.01 : 0.1, name != "SpMM");
    }
    if (name == "SpMV" || name == "numaAwareSpMV") {
        const char* u = "usage: SpMV M K threads";
        const int M = (int)arg(argc, argv, 1, u), K = (int)arg(argc, argv, 2, u);
        arg(argc, argv, 3, u);
        return run_spmv(M, K, 0.6, name == "SpMV" ? 0.0f : 0.1f, name != "SpMV");
    }
    if (name == "suitesparse_sdmm" || name == "suitesparse_spmm" || name == "suitesparse_spmv") {
        const char* u = "usage: suitesparse_{sdmm,spmm,spmv} threads [file.mtx]";
        arg(argc, argv, 1, u);
        const char* dflt = name == "suitesparse_sdmm" ? "datasets/wang3.mtx"
                           : name == "suitesparse_spmm" ? "datasets/big.mtx"
                                                        : "datasets/mouse_gene.mtx";
        return run_suitesparse(name, argc > 2 ? argv[2] : dflt);
    }
    die("unknown driver name (SDMM, SpMM, SpMV, numaAwareSDMM, numaAwareSPDM, balanceNUMAAwareSPDM, "
        "numaAwareSpMV, numaAwareSpMM, suitesparse_sdmm, suitesparse_spmm, suitesparse_spmv)");
}
