// Shapes and launch arguments of the surrogate MLP kernels (mgk_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace mgk {

constexpr int kMlpInPad = 16;    // 5 features, zero padded to one MFMA k-step
constexpr int kMlpHidden = 128;  // ReLU units per hidden layer
constexpr int kMlpLayers = 5;    // 4 hidden layers + output
constexpr int kMlpOut = 8;       // used outputs
constexpr int kMlpOutPad = 32;   // one MFMA n-tile
constexpr int kTargetStride = 8; // cls, r_n, cos(theta), sin(phi), cos(phi), spare x3

// Input and output widths of each layer (weights are K x N, row-major).
constexpr int kLayerK[kMlpLayers] = {kMlpInPad, kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden};
constexpr int kLayerN[kMlpLayers] = {kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden, kMlpOutPad};

constexpr int kFrameXStride = 8; // a frame's feature rows: the 5 features padded to 8 floats
constexpr int kBiasAll = 4 * kMlpHidden + kMlpOutPad;   // biases of all layers (416)
constexpr int kPartStride = kBiasAll + 4;               // a tile's partial sums
constexpr int kDwSplits = 4;                            // batch splits of the weight gradients

struct InferArgs {
    const float* X;     // n rows of x_stride floats: normalised features, zero padded
    int x_stride;       // kMlpInPad or kFrameXStride (the padding beyond it is implied)
    int64_t n;
    const __bf16* Wf[kMlpLayers];   // bf16 weights in the forward A-fragment order (mgk_mlp.hip)
    const float* b[kMlpLayers];
    float* Y;           // n x 8 outputs
};

// One training step's forward pass, loss and backward pass (k_train_fb).
struct TrainArgs {
    const float* X;                 // n x 16 normalised features
    const float* T;                 // n x 8 targets (the loss gives dL/dy), or null
    const float* dY;                // n x 32 a given dL/dy (when T is null)
    float* Y;                       // n x 32 forward outputs (optional)
    int n;
    int n_pad;                      // row length of HT / GT: n rounded up to a workgroup's rays
    float w_r, w_d;                 // loss weights of the disk radius and the escape direction
    const __bf16* Wf[kMlpLayers];   // forward A fragments
    const __bf16* Wg[kMlpLayers];   // backward A fragments (layers 1..4)
    const float* b[kMlpLayers];
    __bf16* HT[kMlpLayers];         // layer inputs h_l, feature-major: [K_l][n_pad]
    __bf16* GT[kMlpLayers];         // bf16 gradients g_l of the layer outputs: [N_l][n_pad]
    float* part;                    // per 32-ray tile: its sums of dz over its rays for every bias
                                    // (kBiasAll floats), then its 4 loss sums: [n_pad / 32][kPartStride]
};

// Weight gradients dW_l = h_l^T g_l over the batch, and the sums of the
// tiles' partial bias gradients and loss sums (k_train_dw).
struct DwArgs {
    const __bf16* HT[kMlpLayers];
    const __bf16* GT[kMlpLayers];
    float* gWp[kMlpLayers];         // partial weight gradients of the batch splits: [kDwSplits][K_l][N_l]
    float* gb[kMlpLayers];
    const float* part;
    float* sums;                    // [ce, r, dir, correct] sums over the batch
    int n_pad;
};

// The weight gradients summed from their partials, Adam on every parameter,
// then the bf16 fragments of the new weights (k_train_adam); with update = 0
// the weights stay and the fragments are written (refresh after weights are
// set), with the gradients summed when gWp is set.
struct AdamArgs {
    float* W[kMlpLayers];
    float* b[kMlpLayers];
    float* mW[kMlpLayers];
    float* vW[kMlpLayers];
    float* mb[kMlpLayers];
    float* vb[kMlpLayers];
    const float* gWp[kMlpLayers];   // the partial weight gradients (null: no gradient)
    float* gW[kMlpLayers];          // receives their sum
    const float* gb[kMlpLayers];
    __bf16* Wf[kMlpLayers];
    __bf16* Wg[kMlpLayers];
    float lr, b1, b2, eps, c1, c2;
    int update;                     // 0: fragments only
};

hipError_t launch_train_fb(const TrainArgs& a, hipStream_t s);
hipError_t launch_train_dw(const DwArgs& a, hipStream_t s);
hipError_t launch_train_adam(const AdamArgs& a, int t, hipStream_t s);
hipError_t launch_zero(float* p, int64_t n, hipStream_t s);
hipError_t launch_infer(const InferArgs& a, hipStream_t s);
unsigned infer_grid(int64_t n);     // workgroups of the inference kernel for n rows
int train_n_pad(int n);             // HT / GT row length for a batch of n

}  // namespace mgk
