// Host-side launch wrappers of the far-field MLP kernels (cv_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cv_mlp.h"

namespace cv {

struct AdamStep {
    float lr, beta1, beta2, eps;
    float bc1, bc2;  // bias corrections 1 - beta^t
    int apply;       // 0: reduce gradients only
};

void launch_mlp_pack(const float* master, uint16_t* fwd, uint16_t* bwd, float* bias, hipStream_t s);
void launch_mlp_forward(const MlpDev& w, const float* X, int n, float* Y, hipStream_t s);
int mlp_train_groups(int n);
void launch_mlp_train(const MlpDev& w, const float* X, const float* T, int n, float* partial, float* loss_part,
                      hipStream_t s);
void launch_mlp_adam(const float* partial, int nwg, const float* loss_part, float* master, float* m1, float* m2,
                     float* grad_out, float* loss_out, const AdamStep& st, hipStream_t s);

}  // namespace cv
