// The per-ray part of FP64 ray generation, shared by the ray-generation kernel
// (mgk_kernels.hip) and the surrogate frame's decode (mgk_surrogate_frame.hip),
// which forms a path's initial conditions in registers instead of reading
// them back from HBM.  Both call this one function, so a path's initial
// conditions are the same bits on either route.
#pragma once

#include "mgk_device.hpp"
#include "mgk_internal.hpp"

namespace mgk {

// Sensor point and direction (device_sensor_utilities.cuh:35-80,
// thin_lens.cuh:60-76), then p = M l and u_obs . p = w . l from the camera's
// constants: y = (x, y, z, t, p_x, p_y, p_z, p_t, h0, u.p) of path i.
__device__ __forceinline__ void ray_ic_fp64(const RayGenParams& P, const IcCamera& C, uint64_t i, double (&y)[10]) {
#pragma clang fp contract(off)
    using S = double;
    const uint32_t npix = P.W * P.H;
    const uint32_t pixel = uint32_t(i % npix);
    const uint32_t sample = P.sample_begin + uint32_t(i / npix);
    const uint32_t px = pixel % P.W, py = pixel / P.W;
    float u[4];
    path_uniforms(P.seed, P.camera_id, pixel, sample, 0, u);
    S sx, sy;
    sensor_point<S>(P.W, P.H, P.sensor_width_mm, px, py, u[0], u[1], sx, sy);
    const S f_m = S(P.focal_mm) / S(1000);
    S dx = S(0) - sx, dy = S(0) - sy, dz = -f_m - S(0);
    const S dl = sqrt_<false>(dx * dx + dy * dy + dz * dz);
    dx = dx / dl; dy = dy / dl; dz = dz / dl;
    const S L = S(P.length_scale);
    const S lx = dx / L, ly = dy / L, lz = dz / L;
    const S lt = sqrt_<false>((lx * lx + ly * ly + lz * lz) - S(0));
    y[0] = C.q[0]; y[1] = C.q[1]; y[2] = C.q[2]; y[3] = C.q[3];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) y[4 + mu] = ((C.M[mu][0] * lx + C.M[mu][1] * ly) + C.M[mu][2] * lz) + C.M[mu][3] * lt;
    y[8] = P.metric == MINKOWSKI ? S(1000) : S(1);
    y[9] = ((C.w[0] * lx + C.w[1] * ly) + C.w[2] * lz) + C.w[3] * lt;
}

}  // namespace mgk
