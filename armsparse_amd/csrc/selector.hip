// The learned schedule selector inside the library: when an SpMM or SDDMM call
// leaves its configuration to the library (cfg NULL or all zero), an installed
// network picks it.  Per matrix: the row-length histogram (armsp_row_stats), a
// features kernel, the MLP's inference kernel (armsp_mlp_infer), and one 4-byte
// read of the chosen class; the choice is cached per (operation, rowptr
// address, M, nnz, dense width, element size).  Without a network, or when a
// choice would need a device round trip where none is allowed (HIP graph
// capture, the entry point that allocates nothing), the library's heuristic
// decides.  The reference has no counterpart: it picks its SpMM kernel by a
// fixed N threshold (gcn/pytorch-custom/spmm_kernel.cu:437-457).
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"

namespace armsp {
namespace {

constexpr int kOps = 2;      // ARMSP_SELECTOR_SPMM, ARMSP_SELECTOR_SDDMM
constexpr int kBins = 12;    // row-length histogram bins of the features
constexpr int kFeat = 5 + kBins;
constexpr int kDevs = 64;

// Features of the selector (log2 of M, nnz+1, mean row length+1, max row
// length+1 and N; then the fraction of rows in each histogram bin).
__global__ void k_selector_features(int M, int nnz, int N, const int* __restrict__ stats, float* __restrict__ x) {
    const int t = threadIdx.x;
    const double m = M > 1 ? (double)M : 1.0;
    if (t == 0) x[0] = (float)log2(m);
    if (t == 1) x[1] = (float)log2((double)nnz + 1.0);
    if (t == 2) x[2] = (float)log2((double)nnz / m + 1.0);
    if (t == 3) x[3] = (float)log2((double)stats[kBins] + 1.0);
    if (t == 4) x[4] = (float)log2(N > 1 ? (double)N : 1.0);
    if (t >= 5 && t < kFeat) x[t] = (float)((double)stats[t - 5] / m);
}

struct Net {
    int F = 0, H = 0, C = 0;
    std::vector<float> w;                 // mu, sd, W1, b1, W2, b2, W3, b3, packed
    std::vector<armsp_config> classes;    // class c → configuration (AUTO: the heuristic)
    uint64_t version = 0;
};

struct DevNet {
    float* w = nullptr;   // the packed weights, then 17 features
    int* stats = nullptr; // histogram bins, max row length, chosen class
    uint64_t version = 0;
};

std::mutex g_mu;
Net g_net[kOps];
DevNet g_dev[kOps][kDevs];
uint64_t g_version = 0;
using Key = std::tuple<int, uintptr_t, int, int, int, int>;
std::map<Key, armsp_config> g_cache;

size_t packed_floats(int F, int H, int C) { return 2 * (size_t)F + (size_t)F * H + H + (size_t)H * H + H + (size_t)H * C + C; }

}  // namespace

// Configuration the installed network picks for an automatic call, if any
// (true: *out set).  may_compute: a device round trip (allocation on first use,
// a synchronisation of `s`) is allowed.
bool selector_config(int op, int M, int nnz, int N, int esz, const int* rowptr, hipStream_t s, bool may_compute,
                     armsp_config* out) {
    if (op < 0 || op >= kOps || M <= 0 || nnz <= 0 || !rowptr) return false;
    std::lock_guard<std::mutex> g(g_mu);
    const Net& net = g_net[op];
    if (net.version == 0 || esz != 4) return false;  // trained on fp32 operands only
    const Key key{op, (uintptr_t)rowptr, M, nnz, N, esz};
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
        *out = it->second;
        return true;
    }
    if (!may_compute) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevs) return false;
    DevNet& d = g_dev[op][dev];
    const size_t nw = packed_floats(net.F, net.H, net.C);
    if (d.version != net.version) {
        if (alloc_injected()) return false;
        if (d.w) hipFree(d.w);
        if (d.stats) hipFree(d.stats);
        d.w = nullptr;
        d.stats = nullptr;
        if (hipMalloc(&d.w, (nw + kFeat) * sizeof(float)) != hipSuccess ||
            hipMalloc(&d.stats, (kBins + 2) * sizeof(int)) != hipSuccess ||
            hipMemcpy(d.w, net.w.data(), nw * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            d.version = 0;
            return false;
        }
        d.version = net.version;
    }
    const int F = net.F, H = net.H, C = net.C;
    const float* mu = d.w;
    const float* sd = mu + F;
    const float* W1 = sd + F;
    const float* b1 = W1 + (size_t)F * H;
    const float* W2 = b1 + H;
    const float* b2 = W2 + (size_t)H * H;
    const float* W3 = b2 + H;
    const float* b3 = W3 + (size_t)H * C;
    float* x = d.w + nw;
    int* cls = d.stats + kBins + 1;
    if (armsp_row_stats(M, rowptr, kBins, d.stats, d.stats + kBins, s)) return false;
    hipLaunchKernelGGL(k_selector_features, dim3(1), dim3(64), 0, s, M, nnz, N, d.stats, x);
    if (armsp_mlp_infer(1, F, H, C, x, mu, sd, W1, b1, W2, b2, W3, b3, nullptr, cls, s)) return false;
    int c = -1;
    if (hipMemcpyAsync(&c, cls, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess || c < 0 || c >= C) {
        (void)hipGetLastError();
        return false;
    }
    g_cache[key] = net.classes[(size_t)c];
    *out = net.classes[(size_t)c];
    return true;
}

}  // namespace armsp

extern "C" {

int armsp_selector_set(int op, int F, int H, int C, const float* mu, const float* sd, const float* W1,
                       const float* b1, const float* W2, const float* b2, const float* W3, const float* b3,
                       const armsp_config* classes) {
    using namespace armsp;
    if (op < 0 || op >= kOps || F != kFeat || H < 1 || H > 256 || C < 1 || C > 256) return ARMSP_ERROR_INVALID_VALUE;
    if (!mu || !sd || !W1 || !b1 || !W2 || !b2 || !W3 || !b3 || !classes) return ARMSP_ERROR_INVALID_VALUE;
    Net n;
    n.F = F;
    n.H = H;
    n.C = C;
    n.w.reserve(packed_floats(F, H, C));
    auto put = [&](const float* p, size_t k) { n.w.insert(n.w.end(), p, p + k); };
    put(mu, F);
    put(sd, F);
    put(W1, (size_t)F * H);
    put(b1, H);
    put(W2, (size_t)H * H);
    put(b2, H);
    put(W3, (size_t)H * C);
    put(b3, C);
    n.classes.assign(classes, classes + C);
    std::lock_guard<std::mutex> g(g_mu);
    n.version = ++g_version;
    g_net[op] = std::move(n);
    g_cache.clear();
    return ARMSP_SUCCESS;
}

int armsp_selector_clear(int op) {
    using namespace armsp;
    if (op < 0 || op >= kOps) return ARMSP_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> g(g_mu);
    g_net[op] = Net{};
    g_cache.clear();
    return ARMSP_SUCCESS;
}

int armsp_selector_choose(int op, int M, int nnz, int N, int element_bytes, const int* rowptr, armsp_config* out,
                          void* stream) {
    using namespace armsp;
    if (!out || op < 0 || op >= kOps) return ARMSP_ERROR_INVALID_VALUE;
    armsp_config c{0, 0, 0, 0, 0};
    selector_config(op, M, nnz, N, element_bytes, rowptr, as_stream(stream), true, &c);
    *out = c;
    return ARMSP_SUCCESS;
}

}  // extern "C"
