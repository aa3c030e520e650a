/*
 * armsp_optim.h — the optimizer step of the GCN training loop in
 * libarmsparse.so (device code, same conventions as armsp.h: plain device
 * pointers, a HIP stream passed as `void*`, an armsp_status returned).
 *
 * The reference trains its GCN with torch.optim.Adam over two parameter
 * groups, weight decay 5e-4 on the first layer and none on the second
 * (gcn/pytorch-custom/gcn_custom.py:116-119).  armsp_adam_f32 is that step
 * for float32 parameters: every tensor of the model in one call, two kernel
 * launches, the step counters in device memory so that the call can be
 * captured in a HIP graph with the rest of the training step.
 */
#ifndef ARMSP_OPTIM_H
#define ARMSP_OPTIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A float32 tensor of an Adam step: parameter, gradient, first and second
 * moments (n elements each, device memory), its L2 weight decay and its step
 * counter (one float in device memory, 0 before the first step). */
typedef struct armsp_adam_tensor {
    float* param;
    const float* grad;
    float* exp_avg;
    float* exp_avg_sq;
    float* step;
    int64_t n;
    float weight_decay;
} armsp_adam_tensor;

#define ARMSP_ADAM_MAX_TENSORS 16

/* One Adam step over up to ARMSP_ADAM_MAX_TENSORS tensors, replacing
 * torch.optim.Adam.step (gcn_custom.py:116-119, called at :126).  Each
 * tensor's counter is incremented on the device first; then, with t the new
 * count,
 *   g = grad + wd·param,  m = β1·m + (1−β1)·g,  v = β2·v + (1−β2)·g²,
 *   param −= (lr / (1−β1^t)) · m / (√v / √(1−β2^t) + eps)
 * — torch.optim.Adam's update with amsgrad and maximize off; the scalars lr,
 * β, 1 − β and eps are formed in double and rounded to float, as torch's are.
 * Every tensor needs a counter of its own; 0 <= β1, β2 < 1; lr, eps, wd >= 0.
 * count == 0 is a no-op. */
int armsp_adam_f32(const armsp_adam_tensor* tensors, int count, double lr, double beta1, double beta2, double eps,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ARMSP_OPTIM_H */
