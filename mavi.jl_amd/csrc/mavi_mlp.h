// mavi_mlp.h -- network shape, parameter layout and the device-side state of
// the learned propagator's MLP (mavi_mlp.hip, mavi_learned.hip).
#pragma once
#include "mavi_internal.h"

constexpr int kMlpIn = MAVI_MLP_IN;       // features per particle
constexpr int kMlpHid = MAVI_MLP_HIDDEN;  // width of the 3 hidden layers
constexpr int kMlpOut = MAVI_MLP_OUT;     // predicted displacement (2)
constexpr int kMlpOutPad = 16;            // output layer padded to one MFMA block

// Flat FP32 parameter layout: W0^T [128][160], b0, W1^T [128][128], b1,
// W2^T [128][128], b2, W3^T [2][128], b3.
constexpr int64_t kOffW0 = 0;
constexpr int64_t kOffB0 = kOffW0 + (int64_t)kMlpHid * kMlpIn;
constexpr int64_t kOffW1 = kOffB0 + kMlpHid;
constexpr int64_t kOffB1 = kOffW1 + (int64_t)kMlpHid * kMlpHid;
constexpr int64_t kOffW2 = kOffB1 + kMlpHid;
constexpr int64_t kOffB2 = kOffW2 + (int64_t)kMlpHid * kMlpHid;
constexpr int64_t kOffW3 = kOffB2 + kMlpHid;
constexpr int64_t kOffB3 = kOffW3 + (int64_t)kMlpOut * kMlpHid;
constexpr int64_t kMlpParams = kOffB3 + kMlpOut;
static_assert(kMlpParams == MAVI_MLP_NPARAMS, "parameter count of the C ABI");

constexpr int kSplitRows = 256;  // rows per split-K partial of a weight gradient

struct mavi_mlp {
    mavi_ctx *ctx = nullptr;
    int refs = 0;          // learned propagators bound to this network
    bool closing = false;  // destroyed while referenced: freed at the last release
    bool ctx_held = false; // holds a reference to ctx (set once creation succeeded)
    float *params = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;  // FP32 master, gradient, Adam moments
    int64_t adam_t = 0;
    uint16_t *w[4] = {nullptr, nullptr, nullptr, nullptr};  // BF16 W_l^T (W3 padded to 16 rows)
    float *b3 = nullptr;                                     // output bias padded to 16
    float *ws = nullptr;                                     // split-K partials
    int64_t ws_cap = 0;
    int fused_grid = 256;
    // training work arrays, sized for work_cap rows
    int work_cap = 0;
    float *Z[3] = {nullptr, nullptr, nullptr};
    uint16_t *H[3] = {nullptr, nullptr, nullptr};
    uint16_t *dZa = nullptr, *dZb = nullptr, *dY = nullptr;
    float *Yw = nullptr, *loss_part = nullptr;
    uint16_t *Xw = nullptr;
    float *Tw = nullptr;
    // device data set (features, targets) and the minibatch gathered from it
    uint16_t *dsX = nullptr;
    float *dsT = nullptr;
    int64_t ds_count = 0, ds_cap = 0;
    uint16_t *bX = nullptr;
    float *bT = nullptr;
    int b_cap = 0;
    float *loss_hist = nullptr;
    int64_t loss_hist_cap = 0;
};

extern "C" {
MAVI_HIDDEN void mlp_retain(mavi_mlp *p);
MAVI_HIDDEN void mlp_release(mavi_mlp *p);  // frees a destroyed network at its last release
}

namespace mavi {
hipError_t ensure_work(mavi_mlp *p, int M);
hipError_t mlp_forward_fused(mavi_mlp *p, const uint16_t *X, int M, float *Y, hipStream_t st);
hipError_t mlp_forward_gemm(mavi_mlp *p, const uint16_t *X, int M, hipStream_t st);
hipError_t mlp_loss_grad_dev(mavi_mlp *p, const uint16_t *X, const float *T, int M, hipStream_t st);
hipError_t mlp_adam_step(mavi_mlp *p, float lr, hipStream_t st);
hipError_t mlp_sync_weights(mavi_mlp *p, hipStream_t st);
double mlp_loss_value(mavi_mlp *p, int M, hipStream_t st, hipError_t *err);
hipError_t mlp_dataset_reserve(mavi_mlp *p, int64_t cap);
}  // namespace mavi
