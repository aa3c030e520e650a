// Adam step of the GCN training loop over all of a model's float32 parameters
// in two launches.
//
// Reference: gcn_custom.py:116-119 builds torch.optim.Adam over two parameter
// groups (weight decay 5e-4 on the first layer, none on the second).  A
// PubMed-size model holds 32K parameters; a multi-tensor optimizer that gives
// each 64K-element chunk one workgroup then runs the whole update on two
// workgroups, about a tenth of a 0.2 ms training step.  Here every 256
// elements are a workgroup of their own, one element per lane.
//
// The update is Adam (Kingma and Ba, "Adam: A Method for Stochastic
// Optimization", ICLR 2015, Algorithm 1) with torch.optim.Adam's conventions:
// L2 weight decay added to the gradient, the bias corrections applied to the
// step size and to the square root of the second moment, eps added after that
// square root.  For element i of a tensor at step t (its counter after the
// increment):
//   g  = grad + wd·param
//   m  = β1·m + (1 − β1)·g
//   v  = β2·v + (1 − β2)·g²
//   param −= (lr / (1 − β1^t)) · m / (√v / √(1 − β2^t) + eps)
// Each tensor carries its own step counter in device memory (as a capturable
// torch optimizer does), so the call can sit in a captured HIP graph: the
// first launch increments every counter, the second reads them.
#include "armsp_optim.h"
#include "common.hpp"

namespace armsp {
namespace {

constexpr int ADAM_MAX = ARMSP_ADAM_MAX_TENSORS;
constexpr int ADAM_BLOCK = 256;

struct AdamList {
    float* p[ADAM_MAX];
    const float* g[ADAM_MAX];
    float* m[ADAM_MAX];
    float* v[ADAM_MAX];
    const float* step[ADAM_MAX];
    int64_t n[ADAM_MAX];
    float wd[ADAM_MAX];
    int64_t first[ADAM_MAX + 1];  // first workgroup of each tensor, then the total
    int count;
};

struct AdamSteps {
    float* step[ADAM_MAX];
    int count;
};

__global__ void k_adam_count(AdamSteps s) {
    const int t = threadIdx.x;
    if (t < s.count) *s.step[t] += 1.0f;
}

// om1 = 1 − β1 and om2 = 1 − β2 come rounded from double, like the scalars of
// torch's update; the bias corrections 1 − β^t = −expm1(t·log1p(−(1 − β)))
// keep full float precision where 1 − powf(β, t) would cancel (β2 = 0.999).
__global__ __launch_bounds__(ADAM_BLOCK) void k_adam(AdamList L, float lr, float beta1, float beta2, float om1,
                                                     float om2, float eps) {
    const int64_t b = blockIdx.x;
    int t = 0;
    while (t + 1 < L.count && b >= L.first[t + 1]) ++t;
    const int64_t i = (b - L.first[t]) * ADAM_BLOCK + threadIdx.x;
    if (i >= L.n[t]) return;
    const float st = *L.step[t];
    const float bc1 = -expm1f(st * log1pf(-om1));
    const float bc2_sqrt = sqrtf(-expm1f(st * log1pf(-om2)));
    const float step_size = lr / bc1;
    float* __restrict__ P = L.p[t];
    float* __restrict__ Mo = L.m[t];
    float* __restrict__ Vo = L.v[t];
    const float p = P[i];
    float g = L.g[t][i];
    if (L.wd[t] != 0.0f) g += L.wd[t] * p;
    const float m = beta1 * Mo[i] + om1 * g;
    const float v = beta2 * Vo[i] + om2 * g * g;
    Mo[i] = m;
    Vo[i] = v;
    P[i] = p - step_size * m / (sqrtf(v) / bc2_sqrt + eps);
}

}  // namespace
}  // namespace armsp

extern "C" {

int armsp_adam_f32(const armsp_adam_tensor* tensors, int count, double lr, double beta1, double beta2, double eps,
                   void* stream) {
    using namespace armsp;
    if (count < 0 || count > ADAM_MAX || (count > 0 && !tensors)) return ARMSP_ERROR_INVALID_VALUE;
    if (!(lr >= 0.0) || !(eps >= 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0))
        return ARMSP_ERROR_INVALID_VALUE;
    AdamList L{};
    AdamSteps S{};
    int64_t blocks = 0;
    for (int k = 0; k < count; ++k) {
        const armsp_adam_tensor& a = tensors[k];
        if (a.n < 0 || !(a.weight_decay >= 0.0f)) return ARMSP_ERROR_INVALID_VALUE;
        if (!a.step) return ARMSP_ERROR_INVALID_VALUE;
        if (a.n > 0 && (!a.param || !a.grad || !a.exp_avg || !a.exp_avg_sq)) return ARMSP_ERROR_INVALID_VALUE;
        for (int j = 0; j < k; ++j)  // a shared counter would be advanced twice
            if (tensors[j].step == a.step) return ARMSP_ERROR_INVALID_VALUE;
        S.step[k] = a.step;
        L.p[k] = a.param;
        L.g[k] = a.grad;
        L.m[k] = a.exp_avg;
        L.v[k] = a.exp_avg_sq;
        L.step[k] = a.step;
        L.n[k] = a.n;
        L.wd[k] = a.weight_decay;
        L.first[k] = blocks;
        blocks += (a.n + ADAM_BLOCK - 1) / ADAM_BLOCK;
        if (blocks > INT_MAX) return ARMSP_ERROR_NOT_SUPPORTED;
    }
    if (count == 0) return ARMSP_SUCCESS;
    L.first[count] = blocks;
    L.count = count;
    S.count = count;
    hipStream_t s = as_stream(stream);
    TimedLaunch tl("adam", s);
    hipLaunchKernelGGL(k_adam_count, dim3(1), dim3(64), 0, s, S);
    ARMSP_LAUNCHED();
    if (blocks > 0) {
        hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks), dim3(ADAM_BLOCK), 0, s, L, (float)lr, (float)beta1,
                           (float)beta2, (float)(1.0 - beta1), (float)(1.0 - beta2), (float)eps);
        ARMSP_LAUNCHED();
    }
    tl.stop();
    return ARMSP_SUCCESS;
}

}  // extern "C"
