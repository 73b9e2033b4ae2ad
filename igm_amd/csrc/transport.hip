// Mass transport of IGM's thickness equation on gfx950, fused into one pass:
// face velocities, slope-limited upwind states, fluxes, flux divergence,
//   thk <- max(thk + dt (smb - div(q)), 0),  lsurf = max(topg, -r thk + sl),
//   usurf = lsurf + thk
// (igm/utils/gradient/compute_divflux_slope_limiter.py:9-65,
//  igm/processes/thk/thk.py:27-54).  One thread per cell; the 5-point cross of
// thk and the 3 neighbouring velocities come through L1/L2 (the stencil reads
// each value ~5 times from cache, once from HBM).
//
// Plus: the CFL velocity maximum (time.py:34-49) as a deterministic two-stage
// max reduction, and the ELA-gradient surface mass balance
// (smb_simple.py:40-42).
#include "common.hpp"
#include "../../include/igm_amd.h"

namespace igm {

__device__ __forceinline__ float minmod(float a, float b) {
  // compute_divflux_slope_limiter.py:3-4
  const bool same = a * b > 0.f;
  const float fa = fabsf(a), fb = fabsf(b);
  return (fa < fb && same) ? a : ((fa > fb && same) ? b : 0.f);
}
__device__ __forceinline__ float maxmod(float a, float b) {
  // compute_divflux_slope_limiter.py:6-7
  const bool same = a * b > 0.f;
  const float fa = fabsf(a), fb = fabsf(b);
  return (fa < fb && same) ? b : ((fa > fb && same) ? a : 0.f);
}

// limited slope from the backward (sm) and forward (sp) differences
template <int SLOPE>
__device__ __forceinline__ float limited(float sm, float sp) {
  if constexpr (SLOPE == 0) return 0.f;                 // godunov
  else if constexpr (SLOPE == 1) return minmod(sm, sp); // minmod
  else return maxmod(minmod(sp, 2.f * sm), minmod(sm, 2.f * sp));  // superbee
}

// flux through the face between cells (i-1) and i along one axis, given
// h[i-2..i+1] (zero outside the grid) and the face velocity uf
template <int SLOPE>
__device__ __forceinline__ float face_flux(float hm2, float hm1, float h0, float hp1, float uf,
                                           float dx, float dt) {
  const float s_left = limited<SLOPE>((hm1 - hm2) / dx, (h0 - hm1) / dx);   // slope of cell i-1
  const float s_right = limited<SLOPE>((h0 - hm1) / dx, (hp1 - h0) / dx);   // slope of cell i
  const float w = hm1 + 0.5f * dx * (1.0f - uf * dt / dx) * s_left;
  const float e = h0 - 0.5f * dx * (1.0f + uf * dt / dx) * s_right;
  return uf * (uf > 0.f ? w : e);
}

template <int SLOPE>
__global__ void __launch_bounds__(256)
transport_kernel(int ny, int nx, const float* __restrict__ u, const float* __restrict__ v,
                 const float* __restrict__ thk, const float* __restrict__ smb,
                 const float* __restrict__ topg, float dx, float dy, float dt, float ratio_density,
                 float sealevel, float* __restrict__ thk_out, float* __restrict__ lsurf,
                 float* __restrict__ usurf, float* __restrict__ divflux) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  auto H = [&](int jj, int ii) -> float {
    return (jj >= 0 && jj < ny && ii >= 0 && ii < nx) ? thk[(int64_t)jj * nx + ii] : 0.f;
  };
  auto U = [&](int ii) { return u[(int64_t)j * nx + min(max(ii, 0), nx - 1)]; };
  auto Vv = [&](int jj) { return v[(int64_t)min(max(jj, 0), ny - 1) * nx + i]; };
  // face velocities (compute_divflux_slope_limiter.py:24-25): interior faces
  // average the two cells, the outer faces copy the edge cell
  const float uw = (i == 0) ? U(0) : 0.5f * (U(i - 1) + U(i));
  const float ue = (i == nx - 1) ? U(nx - 1) : 0.5f * (U(i) + U(i + 1));
  const float vs = (j == 0) ? Vv(0) : 0.5f * (Vv(j - 1) + Vv(j));
  const float vn = (j == ny - 1) ? Vv(ny - 1) : 0.5f * (Vv(j) + Vv(j + 1));
  const float hxm2 = H(j, i - 2), hxm1 = H(j, i - 1), h0 = H(j, i), hxp1 = H(j, i + 1), hxp2 = H(j, i + 2);
  const float hym2 = H(j - 2, i), hym1 = H(j - 1, i), hyp1 = H(j + 1, i), hyp2 = H(j + 2, i);
  const float qw = face_flux<SLOPE>(hxm2, hxm1, h0, hxp1, uw, dx, dt);
  const float qe = face_flux<SLOPE>(hxm1, h0, hxp1, hxp2, ue, dx, dt);
  const float qs = face_flux<SLOPE>(hym2, hym1, h0, hyp1, vs, dy, dt);
  const float qn = face_flux<SLOPE>(hym1, h0, hyp1, hyp2, vn, dy, dt);
  const float div = (qe - qw) / dx + (qn - qs) / dy;
  const int64_t o = (int64_t)j * nx + i;
  const float sm = smb ? smb[o] : 0.f;
  const float t = fmaxf(h0 + dt * (sm - div), 0.f);
  const float ls = fmaxf(topg[o], -ratio_density * t + sealevel);
  thk_out[o] = t;
  if (lsurf) lsurf[o] = ls;
  usurf[o] = ls + t;
  if (divflux) divflux[o] = div;
}

// max(|a|, |b|) over n values: per-block maxima, then one block
__global__ void __launch_bounds__(256)
absmax2_kernel(const float* __restrict__ a, const float* __restrict__ b, int64_t n,
               float* __restrict__ partial) {
  float m = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fmaxf(fabsf(a[i]), fabsf(b[i])));
  m = wave_max(m);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__global__ void __launch_bounds__(256)
max_final_kernel(const float* __restrict__ partial, int n, float* __restrict__ out) {
  float m = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) m = fmaxf(m, partial[i]);
  m = wave_max(m);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__global__ void __launch_bounds__(256)
smb_simple_kernel(int64_t n, const float* __restrict__ usurf, float gradabl, float gradacc,
                  float ela, float maxacc, float* __restrict__ smb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float s = usurf[i] - ela;
  const float r = s * (s < 0.f ? gradabl : gradacc);
  smb[i] = fminf(fmaxf(r, -100.f), maxacc);
}

constexpr int kMaxBlocks = 1024;

}  // namespace igm

using namespace igm;

extern "C" {

int igm_transport_f32(int32_t ny, int32_t nx, const float* ubar, const float* vbar,
                      const float* thk, const float* smb, const float* topg, double dx, double dy,
                      double dt, int32_t slope_type, double ratio_density, double sealevel,
                      float* thk_out, float* lsurf, float* usurf, float* divflux, void* stream) {
  IGM_CHECK_ARG(ubar && vbar && thk && topg && thk_out && usurf, "null pointer");
  IGM_CHECK_ARG(thk_out != thk, "thk_out must not alias thk (neighbours are read)");
  IGM_CHECK_ARG(ny >= 1 && nx >= 1 && ny <= 65535, "bad sizes");
  IGM_CHECK_ARG(slope_type >= 0 && slope_type <= 2, "slope_type: 0 godunov, 1 minmod, 2 superbee");
  dim3 grid((nx + 255) / 256, ny);
  hipStream_t st = (hipStream_t)stream;
#define IGM_TR(S)                                                                                   \
  hipLaunchKernelGGL(transport_kernel<S>, grid, dim3(256), 0, st, ny, nx, ubar, vbar, thk, smb,   \
                     topg, (float)dx, (float)dy, (float)dt, (float)ratio_density, (float)sealevel, \
                     thk_out, lsurf, usurf, divflux)
  if (slope_type == 0) IGM_TR(0);
  else if (slope_type == 1) IGM_TR(1);
  else IGM_TR(2);
#undef IGM_TR
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_absmax2_f32(const float* a, const float* b, int64_t n, float* workspace, float* out,
                    void* stream) {
  IGM_CHECK_ARG(a && b && workspace && out && n >= 1, "bad arguments");
  const int64_t nb = (n + 255) / 256;
  const int blocks = (int)(nb < kMaxBlocks ? nb : kMaxBlocks);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(absmax2_kernel, dim3(blocks), dim3(256), 0, st, a, b, n, workspace);
  IGM_CHECK_LAUNCH();
  hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(256), 0, st, workspace, blocks, out);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_smb_simple_f32(int64_t n, const float* usurf, double gradabl, double gradacc, double ela,
                       double maxacc, float* smb, void* stream) {
  IGM_CHECK_ARG(usurf && smb && n >= 1, "bad arguments");
  hipLaunchKernelGGL(smb_simple_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n, usurf, (float)gradabl, (float)gradacc, (float)ela,
                     (float)maxacc, smb);
  IGM_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
