// Display side of the splat: one thread per pixel, XYZ -> display-encoded
// colour (mgk_display.hpp).  12 bytes read and 12 written per pixel; in place
// (out == xyz) is allowed, so the pointers are not restrict-qualified.
#include "mgk_display.hpp"
#include "mgk_internal.hpp"

namespace mgk {

template <uint32_t VIEW>
__device__ __forceinline__ void display_one(const disp::Params& P, const float* xyz, float* out, uint64_t i) {
    const disp::C3 c = disp::display_pixel(P, disp::C3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
    out[3 * i] = c.x;
    out[3 * i + 1] = c.y;
    out[3 * i + 2] = c.z;
}

// VEC: both buffers 16-byte aligned.  A block takes 1024 pixels (768 float4)
// at a time: coalesced float4 loads into LDS, 4 pixels per thread out of LDS,
// the results back through LDS to coalesced float4 stores.  The last n % 1024
// pixels go one per thread.
constexpr uint32_t kDispBlock = 256, kDispChunk = 4 * kDispBlock;

template <uint32_t VIEW, bool VEC>
__global__ void __launch_bounds__(kDispBlock) k_display(disp::Params P, const float* xyz, uint64_t n, float* out) {
    P.view = VIEW;   // a compile-time view transform: the switch folds away
    if constexpr (VEC) {
        __shared__ float4 sh[3 * kDispBlock];
        const uint32_t t = threadIdx.x;
        const uint64_t n_chunk = n / kDispChunk;
        const float4* in4 = reinterpret_cast<const float4*>(xyz);
        float4* out4 = reinterpret_cast<float4*>(out);
        for (uint64_t ch = blockIdx.x; ch < n_chunk; ch += gridDim.x) {
            const uint64_t base = ch * (3 * kDispBlock);
            for (uint32_t k = 0; k < 3; ++k) sh[t + k * kDispBlock] = in4[base + t + k * kDispBlock];
            __syncthreads();
            const float4 a = sh[3 * t], b = sh[3 * t + 1], c = sh[3 * t + 2];
            const disp::C3 p0 = disp::display_pixel(P, disp::C3{a.x, a.y, a.z});
            const disp::C3 p1 = disp::display_pixel(P, disp::C3{a.w, b.x, b.y});
            const disp::C3 p2 = disp::display_pixel(P, disp::C3{b.z, b.w, c.x});
            const disp::C3 p3 = disp::display_pixel(P, disp::C3{c.y, c.z, c.w});
            sh[3 * t] = float4{p0.x, p0.y, p0.z, p1.x};
            sh[3 * t + 1] = float4{p1.y, p1.z, p2.x, p2.y};
            sh[3 * t + 2] = float4{p2.z, p3.x, p3.y, p3.z};
            __syncthreads();
            for (uint32_t k = 0; k < 3; ++k) out4[base + t + k * kDispBlock] = sh[t + k * kDispBlock];
            __syncthreads();   // before the next chunk's loads overwrite sh
        }
        const uint64_t rest = n - n_chunk * kDispChunk;
        for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + t; i < rest; i += uint64_t(gridDim.x) * blockDim.x)
            display_one<VIEW>(P, xyz, out, n_chunk * kDispChunk + i);
    } else {
        for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
            display_one<VIEW>(P, xyz, out, i);
    }
}

template <uint32_t VIEW>
void launch_view(const disp::Params& p, const float* xyz, uint64_t n, float* out, hipStream_t s) {
    const bool vec = ((reinterpret_cast<uintptr_t>(xyz) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    const uint64_t per_block = vec ? kDispChunk : kDispBlock;
    uint64_t blocks = (n + per_block - 1) / per_block;
    if (blocks > 65536ull * 4) blocks = 65536ull * 4;
    const dim3 g{unsigned(blocks)}, b{kDispBlock};
    if (vec)
        hipLaunchKernelGGL((k_display<VIEW, true>), g, b, 0, s, p, xyz, n, out);
    else
        hipLaunchKernelGGL((k_display<VIEW, false>), g, b, 0, s, p, xyz, n, out);
}

hipError_t launch_display(const disp::Params& p, const float* xyz, uint64_t n, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    switch (p.view) {
        case MAGIK_VIEW_TRANSFORM_NONE: launch_view<MAGIK_VIEW_TRANSFORM_NONE>(p, xyz, n, out, s); break;
        case MAGIK_VIEW_TRANSFORM_REINHARD_SIMPLE: launch_view<MAGIK_VIEW_TRANSFORM_REINHARD_SIMPLE>(p, xyz, n, out, s); break;
        case MAGIK_VIEW_TRANSFORM_REINHARD_EXTENDED: launch_view<MAGIK_VIEW_TRANSFORM_REINHARD_EXTENDED>(p, xyz, n, out, s); break;
        case MAGIK_VIEW_TRANSFORM_GRAN_TURISMO_7: launch_view<MAGIK_VIEW_TRANSFORM_GRAN_TURISMO_7>(p, xyz, n, out, s); break;
        case MAGIK_VIEW_TRANSFORM_AGX: launch_view<MAGIK_VIEW_TRANSFORM_AGX>(p, xyz, n, out, s); break;
        case MAGIK_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL: launch_view<MAGIK_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL>(p, xyz, n, out, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mgk
