// Surrogate MLP kernels (5 -> 4 x 128 ReLU -> 8, bf16 weights and
// activations, FP32 accumulation), all on v_mfma_f32_32x32x16_bf16:
//
//   k_train_fb    : one training step's forward pass, loss and backward pass
//                   through the network, fused: a workgroup of 4 waves takes
//                   128 rays of the batch (one 32-ray tile per wave),
//                   computed transposed as in inference (Z^T = W^T H^T, rays
//                   the columns) so that every layer's accumulators become the
//                   next product's B operand in registers, forward and
//                   backward (dH^T = W dZ^T).  The forward fragments fill LDS
//                   for the forward pass, the backward fragments replace them
//                   for the backward pass.  Writes the layer inputs h_l and
//                   the bf16 output gradients g_l, feature-major, for the
//                   weight gradients, and each tile's sums of the bias
//                   gradients and of the loss over its rays.
//   k_train_dw    : dW_l = h_l^T g_l: one workgroup per 32 x 32 tile of every
//                   layer's weights, its 4 waves splitting the batch and
//                   summing in LDS; two more sum the tiles' bias and loss
//                   sums (no atomics anywhere: deterministic).
//   k_train_adam  : Adam on the FP32 master weights and biases, then the
//                   bf16 forward and backward fragments of the new weights.
//   k_mlp_infer_res (mgk_mlp_infer.hpp): inference.
//
// A step is these 3 launches in place of the ~40 of separate GEMM,
// epilogue, loss and Adam kernels.
#include <algorithm>

#include "mgk_internal.hpp"
#include "mgk_mlp.hpp"
#include "mgk_mlp_infer.hpp"

namespace mgk {

__host__ __device__ constexpr int layer_k(int l) { return l == 0 ? kMlpInPad : kMlpHidden; }
__host__ __device__ constexpr int layer_n(int l) { return l == kMlpLayers - 1 ? kMlpOutPad : kMlpHidden; }

// Element (k, n) of W_l in the forward A fragments (A = W_l^T: rows are the
// outputs n, the k-dimension the inputs; layer 0 in natural input order, the
// others in perm_k order): fragment (ks, n / 32), lane 32 g + n % 32, slot j.
__device__ __forceinline__ int64_t fwd_index(int l, int k, int n) {
    int ks, g, j;
    if (l == 0) {
        ks = k >> 4; g = (k >> 3) & 1; j = k & 7;
    } else {
        ks = 2 * (k >> 5) + ((k >> 4) & 1); g = (k >> 2) & 1; j = 4 * ((k >> 3) & 1) + (k & 3);
    }
    return ((int64_t(ks) * (layer_n(l) / 32) + (n >> 5)) * 64 + 32 * g + (n & 31)) * 8 + j;
}

// Element (k, n) of W_l (l >= 1) in the backward A fragments (A = W_l: rows
// are the inputs k, the k-dimension the outputs n in perm_k order).
__device__ __forceinline__ int64_t bwd_index(int l, int k, int n) {
    const int ks = 2 * (n >> 5) + ((n >> 4) & 1), g = (n >> 2) & 1, j = 4 * ((n >> 3) & 1) + (n & 3);
    return ((int64_t(ks) * (layer_k(l) / 32) + (k >> 5)) * 64 + 32 * g + (k & 31)) * 8 + j;
}

// Backward fragments of layer l in the LDS image of k_train_fb (the output
// layer's 2 k-steps first, then the hidden layers' 8 each, top down).
constexpr int kBwdFragOut = 2 * kStepFrags;                         // 8
__host__ __device__ constexpr int bwd_base(int l) { return l == kMlpLayers - 1 ? 0 : kBwdFragOut + (kMlpLayers - 2 - l) * kFragHidden; }

constexpr int kTrainBlock = 256;        // 4 waves, one 32-ray tile each
constexpr int kTrainRays = 128;         // rays per workgroup

int train_n_pad(int n) { return (n + kTrainRays - 1) / kTrainRays * kTrainRays; }

// Copies a network's A fragments (1 KB each; layer l's from src[l] to
// LDS dst + base[l] fragments, count[l] of them) with the whole workgroup:
// every thread issues all of its 16-byte loads before its LDS writes, so the
// copy costs one memory latency, not one per layer.
template <int kThreads, int kTotal>
__device__ __forceinline__ void frags_to_lds(const __bf16* const (&src)[kMlpLayers], bf16x8* dst, const int (&base)[kMlpLayers],
                                             const int (&count)[kMlpLayers], int tid) {
    constexpr int kPer = (kTotal * 64 + kThreads - 1) / kThreads;
    bf16x8 v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        int k = u * kThreads + tid;   // fragment element across the layers, in LDS order
        int l = 0;
#pragma unroll
        for (int q = 0; q < kMlpLayers; ++q)
            if (count[q] && k >= base[q] * 64 && k < (base[q] + count[q]) * 64) l = q;
        if (k < kTotal * 64) v[u] = reinterpret_cast<const bf16x8*>(src[l])[k - base[l] * 64];
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int k = u * kThreads + tid;
        if (k < kTotal * 64) dst[k] = v[u];
    }
}

// Sum over the 32 lanes of each lane group (the rays of a tile): every lane
// of the group ends with the group's sum.
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int m = 1; m < 32; m <<= 1) v += __shfl_xor(v, m);
    return v;
}

// A wave's tile of a feature-major array (kRows x 32 rays of bf16) to HBM:
// the lanes have placed their elements in the wave's LDS buffer T [row][ray]
// (ds_write_b16 at immediate offsets from one base); every lane then stores
// 16-byte pieces of whole rows, coalesced, into A (rows of np rays) at the
// tile's first ray ray0.  The barriers order the buffer's writes and reads
// (every wave of the workgroup runs the same sequence of tiles).
template <int kRows>
__device__ __forceinline__ void tile_out(const __bf16* T, __bf16* A, int64_t np, int ray0, int lane) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kRows * 4 / 64; ++q) {
        const int c = lane + 64 * q, f = c >> 2, part = c & 3;
        *reinterpret_cast<bf16x8*>(A + int64_t(f) * np + ray0 + 8 * part) = *reinterpret_cast<const bf16x8*>(T + f * 32 + 8 * part);
    }
    __syncthreads();
}

// Accumulators (row tiles mt, element i: row 32 mt + acc_row(g, i)) of dz as
// bf16 B operands, and as a feature-major tile through the wave's buffer.
template <int kTiles>
__device__ __forceinline__ void pack_store(const f32x16 (&dz)[kTiles], int g, int r, int lane, __bf16* T, __bf16* GT, int64_t np,
                                           int ray0, bf16x8 (&G)[2 * kTiles]) {
#pragma unroll
    for (int mt = 0; mt < kTiles; ++mt)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const __bf16 v = __bf16(dz[mt][i]);
            G[2 * mt + (i >> 3)][i & 7] = v;
            T[(32 * mt + acc_row(g, i)) * 32 + r] = v;
        }
    tile_out<32 * kTiles>(T, GT, np, ray0, lane);
}

// Bias gradient of the tile: the accumulators' sums over its rays (the 32
// lanes of each lane group), stored (no atomics) among the tile's partial
// sums at the layer's offset.  A butterfly that halves the values each lane
// carries at every exchange: 16 cross-lane reads per row tile instead of 80
// for 16 separate 5-step sums.  After it, lanes r and r ^ 1 both hold the sum
// of element r / 2.
// One butterfly step: lanes exchange n of their 2 n values with the lane
// 2 n apart; a lane with that bit set keeps (and sums) the upper n.
template <int n>
__device__ __forceinline__ void halve(float (&v)[16], int r) {
    const bool up = (r & (2 * n)) != 0;
#pragma unroll
    for (int k = 0; k < n; ++k) {
        const float send = up ? v[k] : v[n + k];
        const float keep = up ? v[n + k] : v[k];
        v[k] = keep + __shfl_xor(send, 2 * n);
    }
}

template <int kTiles>
__device__ __forceinline__ void bias_grad(const f32x16 (&dz)[kTiles], int g, int r, float* part) {
#pragma unroll
    for (int mt = 0; mt < kTiles; ++mt) {
        float v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = dz[mt][i];
        halve<8>(v, r);
        halve<4>(v, r);
        halve<2>(v, r);
        halve<1>(v, r);
        const float s = v[0] + __shfl_xor(v[0], 1);
        if ((r & 1) == 0) part[32 * mt + acc_row(g, r >> 1)] = s;
    }
}

// Offset of layer l's biases among all biases.
__host__ __device__ constexpr int bias_off(int l) { return l * kMlpHidden; }

__global__ void __launch_bounds__(kTrainBlock, 1) k_train_fb(TrainArgs a) {
    __shared__ bf16x8 wl[kFragAll * 64];
    __shared__ float bias[(kMlpLayers - 1) * kMlpHidden + kMlpOutPad];
    __shared__ __bf16 tr[kTrainBlock / 64][kMlpHidden * 32];   // each wave's transpose buffer
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 5, r = lane & 31;
    __bf16* T = tr[w];
    {
        constexpr int kCount[kMlpLayers] = {kFragL0, kFragHidden, kFragHidden, kFragHidden, kFragOut};
        constexpr int kBase[kMlpLayers] = {frag_base(0), frag_base(1), frag_base(2), frag_base(3), frag_base(4)};
        frags_to_lds<kTrainBlock, kFragAll>(a.Wf, wl, kBase, kCount, tid);
        for (int k = tid; k < (kMlpLayers - 1) * kMlpHidden; k += kTrainBlock) bias[k] = a.b[k / kMlpHidden][k % kMlpHidden];
        if (tid < kMlpOutPad) bias[(kMlpLayers - 1) * kMlpHidden + tid] = a.b[kMlpLayers - 1][tid];
    }
    __syncthreads();
    const int ray0 = blockIdx.x * kTrainRays + w * 32;     // the wave's tile
    const int ray = ray0 + r;                                // this lane's column of it
    float* part = a.part + int64_t(ray0 / 32) * kPartStride; // the tile's partial sums
    const bool valid = ray < a.n;
    const int64_t np = a.n_pad;

    // ---- forward: H[l] is the input of layer l + 1 in B layout ----
    bf16x8 H0;
    {
        float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
        if (valid) {
            const float4* src = reinterpret_cast<const float4*>(a.X + int64_t(ray) * kMlpInPad + 8 * g);
            lo = src[0];
            hi = src[1];
        }
        H0[0] = __bf16(lo.x); H0[1] = __bf16(lo.y); H0[2] = __bf16(lo.z); H0[3] = __bf16(lo.w);
        H0[4] = __bf16(hi.x); H0[5] = __bf16(hi.y); H0[6] = __bf16(hi.z); H0[7] = __bf16(hi.w);
#pragma unroll
        for (int j = 0; j < 8; ++j) T[(8 * g + j) * 32 + r] = H0[j];
        tile_out<kMlpInPad>(T, a.HT[0], np, ray0, lane);
    }
    bf16x8 H[kMlpLayers - 1][2 * kStepFrags];
    f32x16 acc[kStepFrags];
#pragma unroll
    for (int nt = 0; nt < kStepFrags; ++nt)
        acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl[(frag_base(0) + nt) * 64 + lane], H0, bias_tile(bias, 0, nt, g), 0, 0, 0);
#pragma unroll
    for (int l = 1; l < kMlpLayers; ++l) {
        // ReLU into the next B operand, and h_l to HBM for the weight gradients
#pragma unroll
        for (int nt = 0; nt < kStepFrags; ++nt)
#pragma unroll
            for (int i = 0; i < 16; i += 2) {
                const bf16x2 v = relu_bf16x2(acc[nt][i], acc[nt][i + 1]);
                H[l - 1][2 * nt + (i >> 3)][i & 7] = v[0];
                H[l - 1][2 * nt + (i >> 3)][(i & 7) + 1] = v[1];
                T[(32 * nt + acc_row(g, i)) * 32 + r] = v[0];
                T[(32 * nt + acc_row(g, i + 1)) * 32 + r] = v[1];
            }
        tile_out<kMlpHidden>(T, a.HT[l], np, ray0, lane);
        if (l == kMlpLayers - 1) break;
#pragma unroll
        for (int nt = 0; nt < kStepFrags; ++nt) acc[nt] = bias_tile(bias, l, nt, g);
#pragma unroll
        for (int ks = 0; ks < kHiddenSteps; ++ks)
#pragma unroll
            for (int nt = 0; nt < kStepFrags; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl[(frag_base(l) + ks * kStepFrags + nt) * 64 + lane],
                                                                  H[l - 1][ks], acc[nt], 0, 0, 0);
    }
    // output layer: element i of o is output acc_row(g, i) of this lane's ray
    f32x16 o;
#pragma unroll
    for (int i = 0; i < 16; ++i) o[i] = 0.0f;
#pragma unroll
    for (int ks = 0; ks < kHiddenSteps; ++ks)
        o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl[(frag_base(kMlpLayers - 1) + ks) * 64 + lane], H[kMlpLayers - 2][ks], o, 0, 0, 0);
    const float* bo = bias + (kMlpLayers - 1) * kMlpHidden;
#pragma unroll
    for (int i = 0; i < 16; ++i) o[i] += bo[acc_row(g, i)];
    if (a.Y && valid)
#pragma unroll
        for (int i = 0; i < 16; ++i) a.Y[int64_t(ray) * kMlpOutPad + acc_row(g, i)] = o[i];

    // ---- dL/dy ----
    f32x16 dz4[1];
#pragma unroll
    for (int i = 0; i < 16; ++i) dz4[0][i] = 0.0f;
    if (a.T) {
        // outputs 0..3 live in lane group 0 (elements 0..3), 4..7 in group 1
        float y[8];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float other = __shfl_xor(o[i], 32);
            y[i] = g == 0 ? o[i] : other;
            y[4 + i] = g == 0 ? other : o[i];
        }
        float ce = 0.0f, lr = 0.0f, ld = 0.0f, ok = 0.0f, d[6] = {0, 0, 0, 0, 0, 0};
        if (valid) {
            const float* t = a.T + int64_t(ray) * kTargetStride;
            const int cls = int(t[0]);
            const float inv_n = 1.0f / float(a.n);
            const float mx = fmaxf(fmaxf(y[0], y[1]), 0.0f);
            const float e0 = __expf(y[0] - mx), e1 = __expf(y[1] - mx), e2 = __expf(-mx);
            const float s = e0 + e1 + e2;
            const float p0 = e0 / s, p1 = e1 / s, p2 = e2 / s;
            const float pc = cls == 0 ? p0 : (cls == 1 ? p1 : p2);
            ce = -__logf(fmaxf(pc, 1e-30f));
            const int pred = (p0 >= p1 && p0 >= p2) ? 0 : (p1 >= p2 ? 1 : 2);
            ok = pred == cls ? 1.0f : 0.0f;
            d[0] = (p0 - (cls == 0 ? 1.0f : 0.0f)) * inv_n;
            d[1] = (p1 - (cls == 1 ? 1.0f : 0.0f)) * inv_n;
            if (cls == 1) {
                const float e = y[2] - t[1];
                lr = a.w_r * e * e;
                d[2] = 2.0f * a.w_r * e * inv_n;
            } else if (cls == 2) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float e = y[3 + c] - t[2 + c];
                    ld += a.w_d * e * e;
                    d[3 + c] = 2.0f * a.w_d * e * inv_n;
                }
            }
        }
        if (g == 0) {
            dz4[0][0] = d[0]; dz4[0][1] = d[1]; dz4[0][2] = d[2]; dz4[0][3] = d[3];
        } else {
            dz4[0][0] = d[4]; dz4[0][1] = d[5];
        }
        // loss sums of the tile (both lane groups computed the same ray's loss: group 0 counts it)
        const float gate = g == 0 ? 1.0f : 0.0f;
        ce = group_sum(ce * gate); lr = group_sum(lr * gate); ld = group_sum(ld * gate); ok = group_sum(ok * gate);
        if (lane == 0) {
            part[kBiasAll + 0] = ce;
            part[kBiasAll + 1] = lr;
            part[kBiasAll + 2] = ld;
            part[kBiasAll + 3] = ok;
        }
    } else {
        if (valid)
#pragma unroll
            for (int i = 0; i < 16; ++i) dz4[0][i] = a.dY[int64_t(ray) * kMlpOutPad + acc_row(g, i)];
        if (lane < 4) part[kBiasAll + lane] = 0.0f;
    }
    bias_grad<1>(dz4, g, r, part + bias_off(kMlpLayers - 1));
    bf16x8 G[2 * kStepFrags];
    {
        bf16x8 G4[2];
        pack_store<1>(dz4, g, r, lane, T, a.GT[kMlpLayers - 1], np, ray0, G4);
        G[0] = G4[0];
        G[1] = G4[1];
    }

    // ---- backward: the backward fragments replace the forward ones ----
    __syncthreads();
    {
        constexpr int kCount[kMlpLayers] = {0, kFragHidden, kFragHidden, kFragHidden, kBwdFragOut};
        constexpr int kBase[kMlpLayers] = {0, bwd_base(1), bwd_base(2), bwd_base(3), bwd_base(4)};
        frags_to_lds<kTrainBlock, kBwdFragOut + 3 * kFragHidden>(a.Wg, wl, kBase, kCount, tid);
    }
    __syncthreads();
#pragma unroll
    for (int l = kMlpLayers - 1; l >= 1; --l) {
        // dh_l^T = W_l g_l^T (rows: the inputs of layer l), then the ReLU mask of h_l
        const int n_ks = l == kMlpLayers - 1 ? 2 : kHiddenSteps;
#pragma unroll
        for (int mt = 0; mt < kStepFrags; ++mt)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[mt][i] = 0.0f;
#pragma unroll
        for (int ks = 0; ks < n_ks; ++ks)
#pragma unroll
            for (int mt = 0; mt < kStepFrags; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl[(bwd_base(l) + ks * kStepFrags + mt) * 64 + lane], G[ks],
                                                                  acc[mt], 0, 0, 0);
#pragma unroll
        for (int mt = 0; mt < kStepFrags; ++mt)
#pragma unroll
            for (int i = 0; i < 16; ++i)
                if (!(float(H[l - 1][2 * mt + (i >> 3)][i & 7]) > 0.0f)) acc[mt][i] = 0.0f;
        bias_grad<kStepFrags>(acc, g, r, part + bias_off(l - 1));
        pack_store<kStepFrags>(acc, g, r, lane, T, a.GT[l - 1], np, ray0, G);
    }
}

// ---------------------------------------------------------------------------
// Weight gradients.
// ---------------------------------------------------------------------------
constexpr int kDwBlock = 256;   // kDwSplits workgroups per weight tile, each over a part of the batch
__host__ __device__ constexpr int dw_tiles(int l) { return ((layer_k(l) + 31) / 32) * (layer_n(l) / 32); }
constexpr int kDwTileSum = dw_tiles(0) + dw_tiles(1) + dw_tiles(2) + dw_tiles(3) + dw_tiles(4);

// Workgroups that sum the tiles' partials: 16 partial-sum slots per
// workgroup, each summed by 16 threads over interleaved tiles, then in LDS.
constexpr int kSumSlots = 16;
constexpr int kDwExtra = (kPartStride + kSumSlots - 1) / kSumSlots;

__global__ void __launch_bounds__(kDwBlock) k_train_dw(DwArgs a) {
    __shared__ float red[3][16 * 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 5, r = lane & 31;
    if (int(blockIdx.x) >= kDwTileSum * kDwSplits) {
        // bias gradients and loss sums: the tiles' partial sums (a fixed order: deterministic)
        float* sred = &red[0][0];
        const int slot = tid % kSumSlots, grp = tid / kSumSlots;   // 16 groups of tiles
        const int p = (int(blockIdx.x) - kDwTileSum * kDwSplits) * kSumSlots + slot;
        const int n_tiles = a.n_pad / 32;
        float s = 0.0f;
        if (p < kPartStride) {
#pragma unroll 8
            for (int t = grp; t < n_tiles; t += kDwBlock / kSumSlots) s += a.part[int64_t(t) * kPartStride + p];
        }
        sred[grp * kSumSlots + slot] = s;
        __syncthreads();
        if (grp == 0 && p < kPartStride) {
            for (int q = 1; q < kDwBlock / kSumSlots; ++q) s += sred[q * kSumSlots + slot];
            if (p < kBiasAll) {
                const int l = p < 4 * kMlpHidden ? p / kMlpHidden : kMlpLayers - 1;
                a.gb[l][p - bias_off(l)] = s;
            } else {
                a.sums[p - kBiasAll] = s;
            }
        }
        return;
    }
    // workgroup: tile t of layer l, over split `split` of the batch's k-steps
    const int split = int(blockIdx.x) % kDwSplits;
    int t = int(blockIdx.x) / kDwSplits, l = 0;
    while (l < kMlpLayers - 1 && t >= dw_tiles(l)) t -= dw_tiles(l++);
    const int K = layer_k(l), N = layer_n(l);
    const int nts = N / 32;
    const int mt = t / nts, nt = t % nts;
    const int64_t np = a.n_pad;
    const bool row_ok = 32 * mt + r < K;   // layer 0 has 16 input rows
    const __bf16* HA = a.HT[l] + int64_t(row_ok ? 32 * mt + r : 0) * np + 8 * g;
    const __bf16* GB = a.GT[l] + int64_t(32 * nt + r) * np + 8 * g;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    const int all_steps = int(np / 16);
    const int s_begin = split * (all_steps / kDwSplits);
    const int steps = split == kDwSplits - 1 ? all_steps : s_begin + all_steps / kDwSplits;
    bf16x8 zero;
#pragma unroll
    for (int j = 0; j < 8; ++j) zero[j] = __bf16(0.0f);
    // 16 k-steps' operands are loaded before their products (the loads
    // come from L2 or HBM; fewer in flight left the kernel waiting on them)
    constexpr int kDepth = 16;
    int ks = s_begin + w;
    for (; ks + 4 * (kDepth - 1) < steps; ks += 4 * kDepth) {
        bf16x8 A[kDepth], B[kDepth];
#pragma unroll
        for (int u = 0; u < kDepth; ++u) {
            A[u] = row_ok ? *reinterpret_cast<const bf16x8*>(HA + 16 * (ks + 4 * u)) : zero;
            B[u] = *reinterpret_cast<const bf16x8*>(GB + 16 * (ks + 4 * u));
        }
#pragma unroll
        for (int u = 0; u < kDepth; ++u) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[u], B[u], acc, 0, 0, 0);
    }
    for (; ks < steps; ks += 4) {
        const bf16x8 A = row_ok ? *reinterpret_cast<const bf16x8*>(HA + 16 * ks) : zero;
        const bf16x8 B = *reinterpret_cast<const bf16x8*>(GB + 16 * ks);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, acc, 0, 0, 0);
    }
    if (w > 0)
#pragma unroll
        for (int i = 0; i < 16; ++i) red[w - 1][i * 64 + lane] = acc[i];
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const float v = ((acc[i] + red[0][i * 64 + lane]) + red[1][i * 64 + lane]) + red[2][i * 64 + lane];
            const int row = 32 * mt + acc_row(g, i);
            if (row < K) a.gWp[l][(int64_t(split) * K + row) * N + 32 * nt + r] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Adam, fragments of the new weights.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int n_weights(int l) { return layer_k(l) * layer_n(l); }
constexpr int kParamsAll = n_weights(0) + n_weights(1) + n_weights(2) + n_weights(3) + n_weights(4) + layer_n(0) + layer_n(1) +
                           layer_n(2) + layer_n(3) + layer_n(4);

__global__ void __launch_bounds__(256) k_train_adam(AdamArgs a) {
    int idx = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (idx >= kParamsAll) return;
    int l = 0;
    bool is_b = false;
    for (; l < kMlpLayers; ++l) {
        const int nw = n_weights(l), nb = layer_n(l);
        if (idx < nw) break;
        if (idx < nw + nb) {
            idx -= nw;
            is_b = true;
            break;
        }
        idx -= nw + nb;
    }
    float *p, *m, *v;
    float gi;
    if (is_b) {
        p = a.b[l] + idx; m = a.mb[l] + idx; v = a.vb[l] + idx;
        gi = a.update ? a.gb[l][idx] : 0.0f;
    } else {
        p = a.W[l] + idx; m = a.mW[l] + idx; v = a.vW[l] + idx;
        gi = 0.0f;
        if (a.gWp[l]) {
            // the weight gradient: the sum of the batch splits' partials, in split order
            const int64_t kn = int64_t(layer_k(l)) * layer_n(l);
#pragma unroll
            for (int q = 0; q < kDwSplits; ++q) gi += a.gWp[l][q * kn + idx];
            a.gW[l][idx] = gi;
        }
    }
    float wi = *p;
    if (a.update) {
        const float mi = a.b1 * *m + (1.0f - a.b1) * gi;
        const float vi = a.b2 * *v + (1.0f - a.b2) * gi * gi;
        *m = mi;
        *v = vi;
        wi = wi - a.lr * (mi * a.c1) / (sqrtf(vi * a.c2) + a.eps);
        *p = wi;
    }
    if (!is_b) {
        const int N = layer_n(l);
        const int k = idx / N, n = idx % N;
        const __bf16 q = __bf16(wi);
        a.Wf[l][fwd_index(l, k, n)] = q;
        if (l > 0) a.Wg[l][bwd_index(l, k, n)] = q;
    }
}

__global__ void k_zero_f32(float* p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
hipError_t launch_train_fb(const TrainArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_train_fb, dim3(unsigned(a.n_pad / kTrainRays)), dim3(kTrainBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_train_dw(const DwArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_train_dw, dim3(unsigned(kDwTileSum * kDwSplits + kDwExtra)), dim3(kDwBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_train_adam(const AdamArgs& a, int t, hipStream_t s) {
    AdamArgs p = a;
    p.b1 = 0.9f; p.b2 = 0.999f; p.eps = 1e-8f;
    p.c1 = a.update ? 1.0f / (1.0f - powf(p.b1, float(t))) : 0.0f;
    p.c2 = a.update ? 1.0f / (1.0f - powf(p.b2, float(t))) : 0.0f;
    hipLaunchKernelGGL(k_train_adam, dim3(unsigned((kParamsAll + 255) / 256)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_zero(float* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero_f32, dim3(unsigned(blocks)), dim3(256), 0, s, p, n);
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
