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
#include "mgk_mlp_infer.hpp"

namespace mgk {

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
// Workgroups of the inference kernel: one per CU, fewer for short inputs.
unsigned infer_grid(int64_t n) {
    static int n_cu = 0;
    if (n_cu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n_cu <= 0)
            n_cu = 256;
    }
    const int64_t chunks = (n + kInferChunk - 1) / kInferChunk;
    const int64_t per_wg = kInferBlock / 64;
    return unsigned(std::max<int64_t>(1, std::min<int64_t>(n_cu, (chunks + per_wg - 1) / per_wg)));
}

hipError_t launch_infer(const InferArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    InferFromRows src{a.X, a.x_stride};
    hipLaunchKernelGGL(k_mlp_infer_res<InferFromRows>, dim3(infer_grid(a.n)), dim3(kInferBlock), 0, s, a, src);
    return hipGetLastError();
}

}  // namespace mgk
