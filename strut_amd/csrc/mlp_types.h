// Kernel argument blocks of the predictor's training kernels (predictor.hip).
#pragma once

#include <stdint.h>

#include "mlp_device.h"

namespace sgpu {

struct MlpPrepArgs {
    MlpDev m;
    int32_t batch;        // padded minibatch rows (multiple of 16)
    int32_t rows;         // valid rows (<= batch); the rest are zero
    int64_t first;        // first entry of perm (or first sample when perm is null)
    const int32_t *perm;  // sample order, or null for consecutive samples
    const float *x;       // raw features [samples][n_in]
    const float *y;       // raw targets [samples][n_out], or null
    uint16_t *xb;         // bf16 [batch][k0]
    uint16_t *xt;         // bf16 [k0][batch]
    float *t;             // normalised targets [batch][n5], or null
};

struct MlpTrainArgs {
    MlpDev m;
    int32_t batch, rows;
    int32_t backward;                // 0: forward only
    float grad_scale;                // d(loss)/dy factor: 2 / (rows * n_out)
    const uint16_t *xb, *xt;
    const float *t;
    uint16_t *ht[kMlpLayers - 1];    // bf16 [256][batch] hidden activations
    uint16_t *dzt[kMlpLayers];       // bf16 [N_l][batch] pre-activation gradients
    float *y_raw;                    // de-normalised outputs [rows][n_out], or null
    float *loss_acc;                 // per 16-row tile: running sum of squared errors
    float *g;                        // flat FP32 gradients (layout of the master weights)
};

struct MlpTiles {
    int32_t base[kMlpLayers + 1];    // first wgrad tile of each layer
    int64_t off_w[kMlpLayers], off_b[kMlpLayers];
};

struct MlpAdamArgs {
    float *p, *g, *mom, *vel;
    uint16_t *w, *wt;
    int64_t n;
    float lr_t, b1, b2, eps_t;       // bias-corrected step size and epsilon
    int64_t off_w[kMlpLayers], off_bf[kMlpLayers];
    int32_t N[kMlpLayers], K[kMlpLayers];
};

}  // namespace sgpu
