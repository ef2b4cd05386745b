// Surrogate MLP on gfx950 bf16 matrix cores: 16 (5 used) -> 4 x 128 ReLU -> 32
// (8 used), bf16 weights and activations, FP32 accumulation.
//
//   k_gemm         : C = A B with arbitrary strides (so A^T / B^T are free),
//                    v_mfma_f32_32x32x16_bf16 tiles, 64x64 output per
//                    256-thread workgroup, optional split-K with FP32 atomics,
//                    fused epilogues (bias + ReLU -> z, h; ReLU' -> dz, g).
//                    Used for the training forward, backward-data and
//                    weight-gradient products.
//   k_mlp_infer    : the whole network in one kernel, computed transposed
//                    (Z^T = W^T H^T: weights are the MFMA A operand, rays the
//                    columns) so that a layer's FP32 accumulators, after
//                    bias + ReLU, are already the next layer's bf16 B operand
//                    in registers.  The hidden weights are staged through LDS
//                    a few k-steps at a time and serve every ray tile of the
//                    workgroup.
//   k_loss         : class cross-entropy + masked regressions -> dL/dy.
//   k_adam         : Adam on the FP32 master weights, refreshes the bf16 copy.
#include "mgk_internal.hpp"
#include "mgk_mlp.hpp"

namespace mgk {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// Fragment layout of v_mfma_f32_32x32x16_bf16 (wave64):
//   A (32 x 16): lane l holds row l % 32, k = 8 (l / 32) + j, j = 0..7
//   B (16 x 32): lane l holds col l % 32, k = 8 (l / 32) + j
//   C (32 x 32): lane l holds col l % 32, rows 8 (i / 4) + 4 (l / 32) + i % 4, i = 0..15
__device__ __forceinline__ int frag_rc(int l) { return l & 31; }
__device__ __forceinline__ int frag_k(int l, int j) { return ((l >> 5) << 3) + j; }
__device__ __forceinline__ int acc_row(int l, int i) { return ((i >> 2) << 3) + ((l >> 5) << 2) + (i & 3); }

__device__ __forceinline__ float relu(float z) { return fmaxf(z, 0.0f); }
__device__ __forceinline__ float drelu(float z) { return z > 0.0f ? 1.0f : 0.0f; }

template <class T> __device__ __forceinline__ float ld(const void* p, int64_t i) {
    return float(reinterpret_cast<const T*>(p)[i]);
}

__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (w >> 1) * 32;
    const int n0 = blockIdx.x * 64 + (w & 1) * 32;
    const int kz = g.k_split > 1 ? int(blockIdx.z) : 0;
    const int kchunk = (g.K + g.k_split - 1) / g.k_split;
    const int kb = kz * kchunk, ke = min(g.K, kb + kchunk);
    f32x16 acc;
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    const int am = m0 + frag_rc(lane), bn = n0 + frag_rc(lane);
    for (int k0 = kb; k0 < ke; k0 += 16) {
        bf16x8 a, b;
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + frag_k(lane, j);
            float av = 0.0f, bv = 0.0f;
            if (am < g.M && k < ke) {
                int64_t ia = int64_t(am) * g.sam + int64_t(k) * g.sak;
                av = g.a_bf16 ? ld<__bf16>(g.A, ia) : ld<float>(g.A, ia);
            }
            if (bn < g.N && k < ke) {
                int64_t ib = int64_t(k) * g.sbk + int64_t(bn) * g.sbn;
                bv = g.b_bf16 ? ld<__bf16>(g.B, ib) : ld<float>(g.B, ib);
            }
            a[j] = __bf16(av);
            b[j] = __bf16(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int n = n0 + frag_rc(lane);
    if (n >= g.N) return;
    for (int i = 0; i < 16; ++i) {
        const int m = m0 + acc_row(lane, i);
        if (m >= g.M) continue;
        float v = acc[i];
        const int64_t ic = int64_t(m) * g.ldc + n;
        switch (g.epi) {
            case EPI_STORE:
                if (g.bias) v += g.bias[n];
                g.C[ic] = v;
                break;
            case EPI_ATOMIC:
                atomicAdd(&g.C[ic], v);
                break;
            case EPI_RELU: {
                if (g.bias) v += g.bias[n];
                g.C[ic] = v;
                g.H[int64_t(m) * g.ldh + n] = __bf16(relu(v));
                break;
            }
            case EPI_DRELU: {
                float dz = v * drelu(g.Z[int64_t(m) * g.ldz + n]);
                g.C[ic] = dz;
                g.H[int64_t(m) * g.ldh + n] = __bf16(dz);
                break;
            }
        }
    }
}

// Column sums (bias gradients): out[n] = sum_m X[m, n]; each block sums
// kColsumRows rows (coalesced along n), one FP32 atomic per column and block.
constexpr int kColsumRows = 64;
__global__ void __launch_bounds__(256) k_colsum(const float* X, int M, int N, float* out) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int r0 = blockIdx.y * kColsumRows; r0 < M; r0 += gridDim.y * kColsumRows) {
        const int r1 = min(M, r0 + kColsumRows);
        for (int m = r0; m < r1; ++m) s += X[int64_t(m) * N + n];
    }
    atomicAdd(&out[n], s);
}

__global__ void k_zero_f32(float* p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0.0f;
}

__global__ void k_to_bf16(const float* x, __bf16* y, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) y[i] = __bf16(x[i]);
}

// Input feature of k-slot (g, j) of k-step ks in the transposed inference:
// the accumulator of output tile nt = ks / 2 holds, in lane group g, the
// features 32 nt + 8 (i / 4) + 4 g + i % 4 (i = 0..15); k-step ks takes
// i = 8 (ks % 2) + j of them as its k-slots 8 g + j.
__device__ __forceinline__ int perm_k(int ks, int g, int j) {
    return 32 * (ks >> 1) + 16 * (ks & 1) + 8 * (j >> 2) + 4 * g + (j & 3);
}

// Repack a row-major FP32 [K][N] master weight into MFMA A-fragment order of
// the transposed product (A = W^T), rounded once to bf16.  Fragment (ks, nt) holds, for lane l and j = 0..7, W[k][32 nt + l % 32] with
// k = 16 ks + 8 (l / 32) + j (natural, the input layer) or
// k = perm_k(ks, l / 32, j) (layers fed by a previous layer's accumulators);
// k-step major, so a few k-steps of all output tiles are one contiguous block:
// F[((ks * (N / 32) + nt) * 64 + lane) * 8 + j].
__global__ void k_repack(const float* W, int K, int N, __bf16* F, int permuted) {
    const int64_t total = int64_t(K) * N;
    for (int64_t o = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; o < total; o += int64_t(gridDim.x) * blockDim.x) {
        const int j = int(o & 7), lane = int((o >> 3) & 63);
        const int64_t t = o >> 9;
        const int nt = int(t % (N / 32)), ks = int(t / (N / 32));
        const int k = permuted ? perm_k(ks, lane >> 5, j) : 16 * ks + 8 * (lane >> 5) + j;
        F[o] = __bf16(W[int64_t(k) * N + 32 * nt + (lane & 31)]);
    }
}

// Loss: logits (y0 horizon, y1 disk, 0 escape) -> cross-entropy; disk rays
// regress y2 -> normalised log r; escape rays regress y3..y5 -> (cos theta,
// sin phi, cos phi).  Writes dL/dy (mean over the batch) and accumulates
// [ce, r, dir, correct] sums.
__global__ void __launch_bounds__(256) k_loss(const float* Y, const float* T, int n, float w_r, float w_d, float* dY, float* sums) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    float ce = 0.0f, lr = 0.0f, ldir = 0.0f, ok = 0.0f;
    if (i < n) {
        const float* y = Y + int64_t(i) * kMlpOutPad;
        const float* t = T + int64_t(i) * kTargetStride;
        float* d = dY + int64_t(i) * kMlpOutPad;
        for (int c = 0; c < kMlpOutPad; ++c) d[c] = 0.0f;
        const int cls = int(t[0]);
        const float inv_n = 1.0f / float(n);
        float l0 = y[0], l1 = y[1];
        float mx = fmaxf(fmaxf(l0, l1), 0.0f);
        float e0 = __expf(l0 - mx), e1 = __expf(l1 - mx), e2 = __expf(-mx);
        float s = e0 + e1 + e2;
        float p0 = e0 / s, p1 = e1 / s, p2 = e2 / s;
        float pc = cls == 0 ? p0 : (cls == 1 ? p1 : p2);
        ce = -__logf(fmaxf(pc, 1e-30f));
        int pred = (p0 >= p1 && p0 >= p2) ? 0 : (p1 >= p2 ? 1 : 2);
        ok = pred == cls ? 1.0f : 0.0f;
        d[0] = (p0 - (cls == 0 ? 1.0f : 0.0f)) * inv_n;
        d[1] = (p1 - (cls == 1 ? 1.0f : 0.0f)) * inv_n;
        if (cls == 1) {
            float e = y[2] - t[1];
            lr = w_r * e * e;
            d[2] = 2.0f * w_r * e * inv_n;
        } else if (cls == 2) {
            for (int c = 0; c < 3; ++c) {
                float e = y[3 + c] - t[2 + c];
                ldir += w_d * e * e;
                d[3 + c] = 2.0f * w_d * e * inv_n;
            }
        }
    }
    __shared__ float red[4][256];
    red[0][threadIdx.x] = ce; red[1][threadIdx.x] = lr; red[2][threadIdx.x] = ldir; red[3][threadIdx.x] = ok;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) atomicAdd(&sums[threadIdx.x], red[threadIdx.x][0]);
}

__global__ void k_adam(float* w, float* m, float* v, const float* g, int64_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2, __bf16* wb) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        float gi = g[i];
        float mi = b1 * m[i] + (1.0f - b1) * gi;
        float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = __bf16(wi);
    }
}

// ---------------------------------------------------------------------------
// Fused inference: 4 waves x MGK_INFER_TILES ray tiles x 32 rays per workgroup.
// ---------------------------------------------------------------------------
// Two 32-ray tiles per wave (each staged weight fragment read from LDS serves
// two MFMAs), two workgroups per CU, hidden weights staged four k-steps at a
// time: measured fastest of 1-2 tiles x 1-2 workgroups x 2-4 k-steps per
// stage (7.7 ms for the 66.4M feature rows of a 1080p x 32 frame).
#ifndef MGK_INFER_TILES
#define MGK_INFER_TILES 2
#define MGK_INFER_OCC 2
#endif
#ifndef MGK_INFER_STAGE
#define MGK_INFER_STAGE 4
#endif
constexpr int kInferWaves = 4;
constexpr int kStepFrags = kMlpHidden / 32;               // A fragments per hidden k-step

// (bias +) ReLU of one layer's accumulators into the next layer's B operand:
// H[c][2 nt + h] holds elements i = 8 h .. 8 h + 7 of accumulator tile nt,
// element i being row 32 nt + 8 (i / 4) + 4 g + i % 4.  The hidden layers'
// accumulators start from their bias (kAddBias false).
template <bool kAddBias, int kInferTiles>
__device__ __forceinline__ void infer_epilogue(const f32x16 (&acc)[kInferTiles][kStepFrags], const float* bias, int g,
                                               bf16x8 (&H)[kInferTiles][2 * kStepFrags]) {
#pragma unroll
    for (int nt = 0; nt < kStepFrags; ++nt) {
        float b[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) b[i] = kAddBias ? bias[32 * nt + 8 * (i >> 2) + 4 * g + (i & 3)] : 0.0f;
#pragma unroll
        for (int c = 0; c < kInferTiles; ++c)
#pragma unroll
            for (int i = 0; i < 16; ++i) H[c][2 * nt + (i >> 3)][i & 7] = __bf16(fmaxf(kAddBias ? acc[c][nt][i] + b[i] : acc[c][nt][i], 0.0f));
    }
}

// TILES 32-ray tiles per wave (each staged weight fragment serves TILES
// MFMAs); WG_PER_CU workgroups per CU (one workgroup's barrier waits and
// epilogue overlap the other's matrix work).
template <int kInferTiles, int WG_PER_CU>
__global__ void __launch_bounds__(256, WG_PER_CU) k_mlp_infer(InferArgs a) {
    constexpr int kInferRays = kInferWaves * kInferTiles * 32;
    constexpr int kStage = MGK_INFER_STAGE;               // hidden k-steps per LDS stage
    constexpr int kStageFrags = kStage * kStepFrags * 64;  // bf16x8 per stage
    constexpr int kStageLoads = kStageFrags / 256;         // of them per thread
    static_assert(kStageFrags % 256 == 0 && (2 * kStepFrags) % kStage == 0, "stage shape");
    __shared__ bf16x8 wbuf[2][kStageFrags];               // double buffered (kStage x 4 KB each)
    __shared__ float bias[4 * kMlpHidden + kMlpOutPad];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 5, r = lane & 31;
    for (int k = tid; k < 4 * kMlpHidden; k += 256) bias[k] = a.b[k / kMlpHidden][k % kMlpHidden];
    if (tid < kMlpOutPad) bias[4 * kMlpHidden + tid] = a.b[kMlpLayers - 1][tid];
    {
        const bf16x8* F1 = reinterpret_cast<const bf16x8*>(a.Wf[1]);
#pragma unroll
        for (int k = 0; k < kStageLoads; ++k) wbuf[0][tid + 256 * k] = F1[tid + 256 * k];
    }
    const int64_t ray0 = int64_t(blockIdx.x) * kInferRays + w * (kInferTiles * 32);
    bf16x8 H[kInferTiles][2 * kStepFrags];
    f32x16 acc[kInferTiles][kStepFrags];
    // layer 1 (K = 16): B = X^T, lane l holds ray l % 32, features 8 (l / 32) .. + 7
    {
        bf16x8 xb[kInferTiles];
#pragma unroll
        for (int c = 0; c < kInferTiles; ++c) {
            const int64_t row = ray0 + 32 * c + r;
            float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
            if (row < a.n && 8 * g < a.x_stride) {
                const float4* src = reinterpret_cast<const float4*>(a.X + row * a.x_stride + 8 * g);
                lo = src[0];
                hi = src[1];
            }
            xb[c][0] = __bf16(lo.x); xb[c][1] = __bf16(lo.y); xb[c][2] = __bf16(lo.z); xb[c][3] = __bf16(lo.w);
            xb[c][4] = __bf16(hi.x); xb[c][5] = __bf16(hi.y); xb[c][6] = __bf16(hi.z); xb[c][7] = __bf16(hi.w);
        }
        const bf16x8* F0 = reinterpret_cast<const bf16x8*>(a.Wf[0]);
#pragma unroll
        for (int nt = 0; nt < kStepFrags; ++nt) {
            const bf16x8 A = F0[nt * 64 + lane];
#pragma unroll
            for (int c = 0; c < kInferTiles; ++c) {
                f32x16 z;
#pragma unroll
                for (int i = 0; i < 16; ++i) z[i] = 0.0f;
                acc[c][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, xb[c], z, 0, 0, 0);
            }
        }
    }
    __syncthreads();                                      // bias and the first weight stage
    infer_epilogue<true>(acc, bias, g, H);
    // hidden layers 2..4: 16 k-steps each; the stage of step s + 1 is loaded
    // while step s runs, one barrier per step
    for (int l = 1; l < kMlpLayers - 1; ++l) {
        // the accumulators start from the layer's bias (rows of the C layout)
#pragma unroll
        for (int nt = 0; nt < kStepFrags; ++nt) {
            f32x16 b0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 b4 = *reinterpret_cast<const float4*>(bias + l * kMlpHidden + 32 * nt + 8 * q + 4 * g);
                b0[4 * q] = b4.x; b0[4 * q + 1] = b4.y; b0[4 * q + 2] = b4.z; b0[4 * q + 3] = b4.w;
            }
#pragma unroll
            for (int c = 0; c < kInferTiles; ++c) acc[c][nt] = b0;
        }
#pragma unroll
        for (int st = 0; st < 2 * kStepFrags / kStage; ++st) {
            const bool last = st == 2 * kStepFrags / kStage - 1;
            const bool has_next = !last || l + 1 < kMlpLayers - 1;
            bf16x8 nx[kStageLoads];
            if (has_next) {
                const bf16x8* Fn = last ? reinterpret_cast<const bf16x8*>(a.Wf[l + 1])
                                        : reinterpret_cast<const bf16x8*>(a.Wf[l]) + (st + 1) * kStageFrags;
#pragma unroll
                for (int k = 0; k < kStageLoads; ++k) nx[k] = Fn[tid + 256 * k];
            }
#pragma unroll
            for (int kk = 0; kk < kStage; ++kk) {
                const int ks = st * kStage + kk;
#pragma unroll
                for (int nt = 0; nt < kStepFrags; ++nt) {
                    const bf16x8 A = wbuf[st & 1][(kk * kStepFrags + nt) * 64 + lane];
#pragma unroll
                    for (int c = 0; c < kInferTiles; ++c)
                        acc[c][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, H[c][ks], acc[c][nt], 0, 0, 0);
                }
            }
            if (has_next) {
#pragma unroll
                for (int k = 0; k < kStageLoads; ++k) wbuf[(st + 1) & 1][tid + 256 * k] = nx[k];
            }
            __syncthreads();
        }
        infer_epilogue<false>(acc, bias, g, H);
    }
    // output layer: one 32-row tile (8 outputs used)
    {
        const bf16x8* F4 = reinterpret_cast<const bf16x8*>(a.Wf[kMlpLayers - 1]);
        f32x16 o[kInferTiles];
#pragma unroll
        for (int c = 0; c < kInferTiles; ++c)
#pragma unroll
            for (int i = 0; i < 16; ++i) o[c][i] = 0.0f;
#pragma unroll
        for (int ks = 0; ks < 2 * kStepFrags; ++ks) {
            const bf16x8 A = F4[ks * 64 + lane];
#pragma unroll
            for (int c = 0; c < kInferTiles; ++c) o[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, H[c][ks], o[c], 0, 0, 0);
        }
        // lane group g holds outputs 4 g .. 4 g + 3 (accumulator rows i = 0..3) of ray l % 32
        const float* bo = bias + 4 * kMlpHidden + 4 * g;
#pragma unroll
        for (int c = 0; c < kInferTiles; ++c) {
            const int64_t row = ray0 + 32 * c + r;
            if (row < a.n)
                *reinterpret_cast<float4*>(a.Y + row * kMlpOut + 4 * g) =
                    make_float4(o[c][0] + bo[0], o[c][1] + bo[1], o[c][2] + bo[2], o[c][3] + bo[3]);
        }
    }
}

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, g.k_split > 1 ? g.k_split : 1);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, g);
    return hipGetLastError();
}
hipError_t launch_colsum(const float* X, int M, int N, float* out, hipStream_t s) {
    const int blocks_y = std::min((M + kColsumRows - 1) / kColsumRows, 4096);
    hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256, blocks_y), dim3(256), 0, s, X, M, N, out);
    return hipGetLastError();
}
hipError_t launch_zero(float* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero_f32, dim3(unsigned(blocks)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
hipError_t launch_to_bf16(const float* x, __bf16* y, int64_t n, hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_to_bf16, dim3(unsigned(blocks)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}
hipError_t launch_repack(const float* W, int K, int N, __bf16* F, int permuted, hipStream_t s) {
    int64_t blocks = (int64_t(K) * N + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_repack, dim3(unsigned(blocks)), dim3(256), 0, s, W, K, N, F, permuted);
    return hipGetLastError();
}
hipError_t launch_loss(const float* Y, const float* T, int n, float w_r, float w_d, float* dY, float* sums, hipStream_t s) {
    hipLaunchKernelGGL(k_loss, dim3((n + 255) / 256), dim3(256), 0, s, Y, T, n, w_r, w_d, dY, sums);
    return hipGetLastError();
}
hipError_t launch_adam(float* w, float* m, float* v, const float* g, int64_t n, float lr, int t, __bf16* wb, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f / (1.0f - powf(b1, float(t))), c2 = 1.0f / (1.0f - powf(b2, float(t)));
    int64_t blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_adam, dim3(unsigned(blocks)), dim3(256), 0, s, w, m, v, g, n, lr, b1, b2, eps, c1, c2, wb);
    return hipGetLastError();
}
hipError_t launch_infer(const InferArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    constexpr int kTiles = MGK_INFER_TILES, kOcc = MGK_INFER_OCC;
    constexpr int kRays = kInferWaves * kTiles * 32;
    hipLaunchKernelGGL((k_mlp_infer<kTiles, kOcc>), dim3(unsigned((a.n + kRays - 1) / kRays)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace mgk
