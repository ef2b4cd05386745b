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

enum { EPI_STORE = 0, EPI_ATOMIC = 1, EPI_RELU = 2, EPI_DRELU = 3 };

struct GemmArgs {
    const void* A;
    int a_bf16;
    int64_t sam, sak;   // A(m, k) = A[m sam + k sak]
    const void* B;
    int b_bf16;
    int64_t sbk, sbn;   // B(k, n) = B[k sbk + n sbn]
    int M, N, K;
    int k_split;
    const float* bias;
    float* C;           // FP32 result (z, dz or plain)
    int64_t ldc;
    __bf16* H;          // bf16 result (h or g)
    int64_t ldh;
    const float* Z;     // saved pre-activation for EPI_DRELU
    int64_t ldz;
    int epi;
};

constexpr int kFrameXStride = 8; // a frame's feature rows: the 5 features padded to 8 floats

struct InferArgs {
    const float* X;     // n rows of x_stride floats: normalised features, zero padded
    int x_stride;       // kMlpInPad or kFrameXStride (the padding beyond it is implied)
    int64_t n;
    const __bf16* Wf[kMlpLayers];   // bf16 weights in the A-fragment order of k_repack
    const float* b[kMlpLayers];
    float* Y;           // n x 8 outputs
};

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s);
hipError_t launch_colsum(const float* X, int M, int N, float* out, hipStream_t s);
hipError_t launch_zero(float* p, int64_t n, hipStream_t s);
hipError_t launch_to_bf16(const float* x, __bf16* y, int64_t n, hipStream_t s);
hipError_t launch_repack(const float* W, int K, int N, __bf16* F, int permuted, hipStream_t s);
hipError_t launch_loss(const float* Y, const float* T, int n, float w_r, float w_d, float* dY, float* sums, hipStream_t s);
hipError_t launch_adam(float* w, float* m, float* v, const float* g, int64_t n, float lr, int t, __bf16* wb, hipStream_t s);
hipError_t launch_infer(const InferArgs& a, hipStream_t s);
unsigned infer_grid(int64_t n);   // workgroups of the inference kernel for n rows

}  // namespace mgk
