// Vertical velocity W from the horizontal velocities, the default of IGM's
// vert_flow process (Lagrange basis, version 2, "kinematic";
// igm/processes/vert_flow/vert_flow_v2.py:17-37, 61-87, 125-147):
//
//   z_k  = sum_{l<k} thk (levels[l+1] - levels[l])     elevation of level k above the bed
//   ub_k = sum_{l<=k} w_l U_l / sum_{l<=k} w_l         (w = vert_weight), vb_k likewise
//   W_k  = -div(ub_k z_k, vb_k z_k) + U_k d(topg + z_k)/dx + V_k d(topg + z_k)/dy
//
// with the reference's stencils: the slope is centred inside and one-sided
// at the edges (edge value extrapolated as 1.5 s_0 - 0.5 s_1), the flux
// divergence centred inside and half one-sided at the edges.  One thread per
// node marches up the levels, keeping the running sums of its own column and
// of its four neighbours; the neighbours' values come through L1/L2, so HBM
// sees U, V, thk, topg about once (88 B/cell in, 40 B/cell of W out at Nz 10).
#include "common.hpp"
#include "../../include/igm_amd.h"

namespace igm {

template <int NZ>
struct VfLevels {
  float ddz[NZ];  // levels[k] - levels[k-1] (0 for k = 0)
  float w[NZ];    // vert_weight
  float icw[NZ];  // 1 / sum_{l<=k} w_l
};

template <int NZ>
__global__ void __launch_bounds__(256)
vert_flow_kernel(int ny, int nx, const float* __restrict__ U, const float* __restrict__ V,
                 const float* __restrict__ topg, const float* __restrict__ thk, VfLevels<NZ> lv,
                 float idx, float* __restrict__ W, float* __restrict__ wbase, float* __restrict__ wsurf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const int64_t plane = (int64_t)ny * nx;
  const int64_t o = (int64_t)j * nx + i;
  // columns: 0 centre, 1 west (i-1), 2 east (i+1), 3 south (j-1), 4 north (j+1)
  const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
  int64_t off[5] = {o, hw ? o - 1 : o, he ? o + 1 : o, hs ? o - nx : o, hn ? o + nx : o};
  float h[5], t[5], z[5], su[5], sv[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    h[c] = thk[off[c]];
    t[c] = topg[off[c]];
    z[c] = 0.f;
    su[c] = 0.f;
    sv[c] = 0.f;
  }
#pragma unroll
  for (int k = 0; k < NZ; ++k) {
    float u[5], v[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      u[c] = U[k * plane + off[c]];
      v[c] = V[k * plane + off[c]];
    }
    float s[5], qx[5], qy[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      if (k > 0) z[c] += h[c] * lv.ddz[k];        // cumsum of compute_dz
      su[c] += lv.w[k] * u[c];                     // cumsum(w U)
      sv[c] += lv.w[k] * v[c];
      s[c] = t[c] + z[c];                          // topg + cumdz
      qx[c] = (su[c] * lv.icw[k]) * z[c];          // ub * h
      qy[c] = (sv[c] * lv.icw[k]) * z[c];
    }
    // slopes of the level surface (vert_flow_v2.py:67-85)
    const float exl = hw ? 0.5f * s[1] + 0.5f * s[0] : 1.5f * s[0] - 0.5f * s[2];
    const float exr = he ? 0.5f * s[0] + 0.5f * s[2] : 1.5f * s[0] - 0.5f * s[1];
    const float eyl = hs ? 0.5f * s[3] + 0.5f * s[0] : 1.5f * s[0] - 0.5f * s[4];
    const float eyr = hn ? 0.5f * s[0] + 0.5f * s[4] : 1.5f * s[0] - 0.5f * s[3];
    const float sx = (exr - exl) * idx, sy = (eyr - eyl) * idx;
    // flux divergence (vert_flow_v2.py:133-147)
    const float dxl = hw ? 0.5f * (qx[1] + qx[0]) : qx[0];
    const float dxr = he ? 0.5f * (qx[0] + qx[2]) : qx[0];
    const float dyl = hs ? 0.5f * (qy[3] + qy[0]) : qy[0];
    const float dyr = hn ? 0.5f * (qy[0] + qy[4]) : qy[0];
    const float div = (dxr - dxl) * idx + (dyr - dyl) * idx;
    const float wk = (-div + u[0] * sx) + v[0] * sy;
    W[k * plane + o] = wk;
    if (k == 0 && wbase) wbase[o] = wk;
    if (k == NZ - 1 && wsurf) wsurf[o] = wk;
  }
}

template <int NZ>
void vert_flow_launch(int ny, int nx, const float* U, const float* V, const float* topg,
                      const float* thk, const double* levels, const double* vert_weight, double dx,
                      float* W, float* wbase, float* wsurf, hipStream_t st) {
  VfLevels<NZ> lv;
  float cw = 0.f;
  for (int k = 0; k < NZ; ++k) {
    lv.ddz[k] = k > 0 ? (float)levels[k] - (float)levels[k - 1] : 0.f;
    lv.w[k] = (float)vert_weight[k];
    cw += lv.w[k];
    lv.icw[k] = 1.0f / cw;
  }
  hipLaunchKernelGGL((vert_flow_kernel<NZ>), dim3((nx + 255) / 256, ny), dim3(256), 0, st, ny, nx, U,
                     V, topg, thk, lv, (float)(1.0 / dx), W, wbase, wsurf);
}

}  // namespace igm

extern "C" int igm_vert_flow_kinematic_f32(int32_t nz, int32_t ny, int32_t nx, const float* U,
                                           const float* V, const float* topg, const float* thk,
                                           const double* levels, const double* vert_weight,
                                           double dx, float* W, float* wvelbase, float* wvelsurf,
                                           void* stream) {
  using namespace igm;
  IGM_CHECK_ARG(U && V && topg && thk && levels && vert_weight && W, "null pointer");
  IGM_CHECK_ARG(ny >= 2 && nx >= 2 && ny <= 65535, "need a grid of at least 2 x 2 (the edge stencils use two cells)");
  IGM_CHECK_ARG(dx > 0.0, "dx must be positive");
  IGM_CHECK_ARG(W != U && W != V, "W must not alias U or V");
  hipStream_t st = (hipStream_t)stream;
  switch (nz) {
#define IGM_VF(N) \
  case N: vert_flow_launch<N>(ny, nx, U, V, topg, thk, levels, vert_weight, dx, W, wvelbase, wvelsurf, st); break;
    IGM_VF(2) IGM_VF(3) IGM_VF(4) IGM_VF(5) IGM_VF(6) IGM_VF(8) IGM_VF(10) IGM_VF(12) IGM_VF(16) IGM_VF(20)
#undef IGM_VF
    default:
      set_error("igm_vert_flow_kinematic_f32: nz not compiled (supported: 2,3,4,5,6,8,10,12,16,20)");
      return -1;
  }
  IGM_CHECK_LAUNCH();
  return 0;
}
