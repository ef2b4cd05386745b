// The fused inference kernel of the surrogate MLP (5 -> 4 x 128 ReLU -> 8),
// a template over where its input rows come from:
//   InferFromRows    : n rows of normalised features in HBM (mgk_mlp.hip)
//   (camera source)  : the features of a frame's paths computed in the kernel
//                      from the camera and the path's Philox samples
//                      (mgk_surrogate_frame.hip)
//
// The network is computed transposed (Z^T = W^T H^T): weights are the MFMA A
// operand, rays the columns, so a layer's FP32 accumulators, after bias and
// ReLU, are already the next layer's bf16 B operand in registers.
//
// One workgroup of 8 waves per CU holds every layer's A fragments in LDS for
// its lifetime (108 fragments of 1 KB, 110.6 KB, plus the biases) and its
// waves loop over 64-ray chunks of the input, two 32-ray tiles per wave, so
// each fragment read from LDS serves two MFMAs.  The weights are loaded once
// per workgroup and the main loop has no barrier: the waves drift apart, and
// one wave's ReLU epilogue runs while another's MFMAs do.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "mgk_mlp.hpp"

namespace mgk {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

constexpr int kStepFrags = kMlpHidden / 32;                  // A fragments (row tiles) per k-step
constexpr int kHiddenSteps = kMlpHidden / 16;                // k-steps of a hidden or the output layer
constexpr int kFragL0 = kStepFrags;                          // input layer: 1 k-step x 4 row tiles
constexpr int kFragHidden = kHiddenSteps * kStepFrags;       // 8 k-steps x 4 row tiles
constexpr int kFragOut = kHiddenSteps;                       // 8 k-steps x 1 row tile
constexpr int kFragAll = kFragL0 + 3 * kFragHidden + kFragOut;   // 108
constexpr int kInferBlock = 512;                             // 8 waves: two per SIMD
constexpr int kInferTiles = 2;                               // 32-ray tiles per wave
constexpr int kInferChunk = 32 * kInferTiles;                // rays per wave and iteration

// k-slot (g, j) of k-step ks of a B operand formed from a layer's
// accumulators: with the C layout of v_mfma_f32_32x32x16_bf16 (lane l holds
// column l % 32, rows 8 (i / 4) + 4 (l / 32) + i % 4 of element i), k-step ks
// takes elements i = 8 (ks % 2) + j of row tile ks / 2, which are the rows
// perm_k(ks, l / 32, j).  The A fragments of the next product are packed in
// this order, so accumulators become B operands without moving between lanes.
__host__ __device__ constexpr int perm_k(int ks, int g, int j) {
    return 32 * (ks >> 1) + 16 * (ks & 1) + 8 * (j >> 2) + 4 * g + (j & 3);
}

// Row of accumulator element i in lane group g (the C layout above).
__host__ __device__ constexpr int acc_row(int g, int i) { return 8 * (i >> 2) + 4 * g + (i & 3); }

// Offset (in fragments) of layer l's A fragments in the LDS image.
__host__ __device__ constexpr int frag_base(int l) { return l == 0 ? 0 : (l < kMlpLayers - 1 ? kFragL0 + (l - 1) * kFragHidden : kFragL0 + 3 * kFragHidden); }

// ReLU of two FP32 values into a packed bf16 pair: the conversion rounds to
// nearest even, then a packed signed 16-bit maximum with 0 clears each half
// whose sign bit is set (a negative value or -0 becomes +0), which is
// bf16(max(x, 0)) for every finite x.
__device__ __forceinline__ bf16x2 relu_bf16x2(float x0, float x1) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    const bf16x2 v = __builtin_convertvector(f32x2{x0, x1}, bf16x2);
    s16x2 s = __builtin_bit_cast(s16x2, v);
    s = __builtin_elementwise_max(s, s16x2{0, 0});
    return __builtin_bit_cast(bf16x2, s);
}

// ReLU of one layer's accumulators into the next layer's B operand:
// H[c][2 nt + h] holds elements i = 8 h .. 8 h + 7 of accumulator tile nt.
__device__ __forceinline__ void infer_relu(const f32x16 (&acc)[kInferTiles][kStepFrags], bf16x8 (&H)[kInferTiles][2 * kStepFrags]) {
#pragma unroll
    for (int nt = 0; nt < kStepFrags; ++nt)
#pragma unroll
        for (int c = 0; c < kInferTiles; ++c)
#pragma unroll
            for (int i = 0; i < 16; i += 2) {
                const bf16x2 v = relu_bf16x2(acc[c][nt][i], acc[c][nt][i + 1]);
                H[c][2 * nt + (i >> 3)][i & 7] = v[0];
                H[c][2 * nt + (i >> 3)][(i & 7) + 1] = v[1];
            }
}

// The bias of layer l's output rows in the accumulator layout of row tile nt
// (element i: row 32 nt + 8 (i / 4) + 4 g + i % 4): every layer's
// accumulators start from it.
__device__ __forceinline__ f32x16 bias_tile(const float* bias, int l, int nt, int g) {
    f32x16 b0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 b4 = *reinterpret_cast<const float4*>(bias + l * kMlpHidden + 32 * nt + 8 * q + 4 * g);
        b0[4 * q] = b4.x; b0[4 * q + 1] = b4.y; b0[4 * q + 2] = b4.z; b0[4 * q + 3] = b4.w;
    }
    return b0;
}

// Input rows of normalised features in HBM (x_stride floats per row, the
// features beyond it zero).  Lane (g, r) of tile c loads features 8 g .. 8 g + 7
// of row ray0 + 32 c + r.
struct InferFromRows {
    const float* X;
    int x_stride;
    static constexpr bool kExtra = false;
    __device__ __forceinline__ void load(int64_t ray0, int64_t n, int lane, bf16x8 (&xb)[kInferTiles],
                                         float (&extra)[kInferTiles][2]) const {
        const int g = lane >> 5, r = lane & 31;
#pragma unroll
        for (int c = 0; c < kInferTiles; ++c) {
            const int64_t row = ray0 + 32 * c + r;
            float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
            if (row < n && 8 * g < x_stride) {
                const float4* src = reinterpret_cast<const float4*>(X + row * x_stride + 8 * g);
                lo = src[0];
                hi = src[1];
            }
            xb[c][0] = __bf16(lo.x); xb[c][1] = __bf16(lo.y); xb[c][2] = __bf16(lo.z); xb[c][3] = __bf16(lo.w);
            xb[c][4] = __bf16(hi.x); xb[c][5] = __bf16(hi.y); xb[c][6] = __bf16(hi.z); xb[c][7] = __bf16(hi.w);
            extra[c][0] = extra[c][1] = 0.0f;
        }
    }
};

// The kernel.  Grid: one workgroup per CU (more are fine: each loads the
// weights once); every wave loops over 64-ray chunks, chunk k going to wave
// k mod (total waves).  Src::kExtra: outputs 6 and 7 of each row (padding of
// the network) carry the source's two extra values instead.
template <class Src>
__global__ void __launch_bounds__(kInferBlock, 1) k_mlp_infer_res(InferArgs a, Src src) {
    __shared__ bf16x8 wl[kFragAll * 64];
    __shared__ float bias[(kMlpLayers - 1) * kMlpHidden + kMlpOutPad];
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 5, r = lane & 31;
    {
        constexpr int kCount[kMlpLayers] = {kFragL0, kFragHidden, kFragHidden, kFragHidden, kFragOut};
#pragma unroll
        for (int l = 0; l < kMlpLayers; ++l) {
            const bf16x8* F = reinterpret_cast<const bf16x8*>(a.Wf[l]);
            for (int k = tid; k < kCount[l] * 64; k += kInferBlock) wl[frag_base(l) * 64 + k] = F[k];
        }
        for (int k = tid; k < (kMlpLayers - 1) * kMlpHidden; k += kInferBlock) bias[k] = a.b[k / kMlpHidden][k % kMlpHidden];
        if (tid < kMlpOutPad) bias[(kMlpLayers - 1) * kMlpHidden + tid] = a.b[kMlpLayers - 1][tid];
    }
    __syncthreads();
    const int64_t n_chunks = (a.n + kInferChunk - 1) / kInferChunk;
    const int64_t waves = int64_t(gridDim.x) * (kInferBlock / 64);
#pragma unroll 1
    for (int64_t ch = int64_t(blockIdx.x) * (kInferBlock / 64) + (tid >> 6); ch < n_chunks; ch += waves) {
        const int64_t ray0 = ch * kInferChunk;
        // the biases are read from LDS in every iteration: an offset the
        // compiler cannot see through keeps it from hoisting them into
        // registers across iterations (which spills)
        int zero = 0;
        asm volatile("" : "+s"(zero));
        const float* bl = bias + zero;
        bf16x8 H[kInferTiles][2 * kStepFrags];
        f32x16 acc[kInferTiles][kStepFrags];
        float extra[kInferTiles][2];
        // input layer (K = 16)
        {
            bf16x8 xb[kInferTiles];
            src.load(ray0, a.n, lane, xb, extra);
#pragma unroll
            for (int nt = 0; nt < kStepFrags; ++nt) {
                const bf16x8 A = wl[(frag_base(0) + nt) * 64 + lane];
                const f32x16 b0 = bias_tile(bl, 0, nt, g);
#pragma unroll
                for (int c = 0; c < kInferTiles; ++c) acc[c][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, xb[c], b0, 0, 0, 0);
            }
        }
        infer_relu(acc, H);
        // hidden layers
#pragma unroll 1
        for (int l = 1; l < kMlpLayers - 1; ++l) {
#pragma unroll
            for (int nt = 0; nt < kStepFrags; ++nt) {
                const f32x16 b0 = bias_tile(bl, l, nt, g);
#pragma unroll
                for (int c = 0; c < kInferTiles; ++c) acc[c][nt] = b0;
            }
#pragma unroll
            for (int ks = 0; ks < kHiddenSteps; ++ks)
#pragma unroll
                for (int nt = 0; nt < kStepFrags; ++nt) {
                    const bf16x8 A = wl[(frag_base(l) + ks * kStepFrags + nt) * 64 + lane];
#pragma unroll
                    for (int c = 0; c < kInferTiles; ++c)
                        acc[c][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, H[c][ks], acc[c][nt], 0, 0, 0);
                }
            infer_relu(acc, H);
        }
        // output layer: one 32-row tile (8 outputs used); lane group g holds
        // outputs 4 g .. 4 g + 3 (accumulator elements 0..3) of ray 32 c + r
        f32x16 o[kInferTiles];
#pragma unroll
        for (int c = 0; c < kInferTiles; ++c)
#pragma unroll
            for (int i = 0; i < 16; ++i) o[c][i] = 0.0f;
#pragma unroll
        for (int ks = 0; ks < kHiddenSteps; ++ks) {
            const bf16x8 A = wl[(frag_base(kMlpLayers - 1) + ks) * 64 + lane];
#pragma unroll
            for (int c = 0; c < kInferTiles; ++c) o[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, H[c][ks], o[c], 0, 0, 0);
        }
        const float* bo = bl + (kMlpLayers - 1) * kMlpHidden + 4 * g;
#pragma unroll
        for (int c = 0; c < kInferTiles; ++c) {
            const int64_t row = ray0 + 32 * c + r;
            float4 v = make_float4(o[c][0] + bo[0], o[c][1] + bo[1], o[c][2] + bo[2], o[c][3] + bo[3]);
            if (Src::kExtra && g == 1) { v.z = extra[c][0]; v.w = extra[c][1]; }
            if (row < a.n) *reinterpret_cast<float4*>(a.Y + row * kMlpOut + 4 * g) = v;
        }
    }
}

}  // namespace mgk
