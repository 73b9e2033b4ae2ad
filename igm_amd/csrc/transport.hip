// Mass transport of IGM's thickness equation on gfx950, fused into one pass:
// face velocities, slope-limited upwind states, fluxes, flux divergence,
//   thk <- max(thk + dt (smb - div(q)), 0),  lsurf = max(topg, -r thk + sl),
//   usurf = lsurf + thk
// (igm/utils/gradient/compute_divflux_slope_limiter.py:9-65,
//  igm/processes/thk/thk.py:27-54), column-marching with DPP neighbours.
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
// h[i-2..i+1] (zero outside the grid), the face velocity uf and c = dt/dx.
// The limiters are positively homogeneous, so dx * slope = limited(cell
// differences): the reference's slope / dx and dx * slope cancel exactly in
// real arithmetic and no division is left.
template <int SLOPE>
__device__ __forceinline__ float face_flux(float hm2, float hm1, float h0, float hp1, float uf, float c) {
  const float d0 = hm1 - hm2, d1 = h0 - hm1, d2 = hp1 - h0;
  const float s_left = limited<SLOPE>(d0, d1);   // dx * slope of cell i-1
  const float s_right = limited<SLOPE>(d1, d2);  // dx * slope of cell i
  const float w = hm1 + 0.5f * (1.0f - uf * c) * s_left;
  const float e = h0 - 0.5f * (1.0f + uf * c) * s_right;
  return uf * (uf > 0.f ? w : e);
}

// Column blocks: a wave owns a strip of 64 columns -- lanes 2..61 own theirs,
// lanes 0, 1, 62, 63 are a two-column halo -- and a block of kTrRows rows.
// It issues every load of its block at once (thk of rows j0-2 .. j0+R+1,
// vbar of rows j0-1 .. j0+R, ubar, smb, topg of rows j0 .. j0+R-1), so it
// waits for memory once, then computes the R rows from registers; the column
// neighbours (i +- 1, i +- 2 of thk, i +- 1 of ubar) come from the
// neighbouring lanes by DPP moves.  Blocks of 4 rows: at 1000 x 1000 the
// launch (4250 waves) took 9.0-9.6 us against 10.3-10.9 us with 8 rows and
// 14.1-14.5 us with 16 -- more waves in flight outweigh the wider halo.
constexpr int kTrOwned = 60;
constexpr int kTrRows = 4;

template <int SLOPE>
__global__ void __launch_bounds__(256)
transport_kernel(int ny, int nx, int nstrips, int nrowblk, const float* __restrict__ u,
                 const float* __restrict__ v, const float* __restrict__ thk,
                 const float* __restrict__ smb, const float* __restrict__ topg, float idx, float idy,
                 float cx, float cy, float dt, float ratio_density, float sealevel,
                 float* __restrict__ thk_out,
                 float* __restrict__ lsurf, float* __restrict__ usurf, float* __restrict__ divflux) {
  constexpr int R = kTrRows;
  const int lane = threadIdx.x & 63;
  const int wid = xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (wid >= nstrips * nrowblk) return;  // no block-level synchronisation below
  const int strip = wid % nstrips, rblk = wid / nstrips;
  const int ic = strip * kTrOwned - 2 + lane;
  const bool col_ok = ic >= 0 && ic < nx;
  const bool own = lane >= 2 && lane < 2 + kTrOwned && col_ok;
  const int j0 = rblk * R;
  const uint32_t bytes = (uint32_t)((int64_t)ny * nx * 4);
  const rsrc_t rH = make_rsrc(thk, bytes), rU = make_rsrc(u, bytes), rV = make_rsrc(v, bytes);
  const rsrc_t rS = make_rsrc(smb, smb ? bytes : 0u), rT = make_rsrc(topg, bytes);
  // byte offset of (row, ic), kOOB outside the grid (reads 0: the zero
  // thickness the reference pads with); clamped rows for vbar
  auto off = [&](int row) -> uint32_t {
    return (col_ok && row >= 0 && row < ny) ? (uint32_t)(row * nx + ic) * 4u : kOOB;
  };
  auto offc = [&](int row) -> uint32_t { return off(row < 0 ? 0 : (row >= ny ? ny - 1 : row)); };
  auto ld = [](rsrc_t r, uint32_t o) { return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, o, 0, 0)); };
  auto st = [](float x, rsrc_t r, uint32_t o) { __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), r, o, 0, 0); };

  float h[R + 4], vv[R + 2], uu[R], ss[R], tt[R];
#pragma unroll
  for (int k = 0; k < R + 4; ++k) h[k] = ld(rH, off(j0 - 2 + k));
#pragma unroll
  for (int k = 0; k < R + 2; ++k) vv[k] = ld(rV, offc(j0 - 1 + k));
#pragma unroll
  for (int k = 0; k < R; ++k) {
    uu[k] = ld(rU, off(j0 + k));
    ss[k] = smb ? ld(rS, off(j0 + k)) : 0.f;
    tt[k] = ld(rT, off(j0 + k));
  }

  const rsrc_t rO = make_rsrc(thk_out, bytes), rUs = make_rsrc(usurf, bytes);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int j = j0 + k;
    // face velocities (compute_divflux_slope_limiter.py:24-25): interior
    // faces average the two cells, the outer faces copy the edge cell
    const float uc = uu[k];
    const float ul = lane_from_left(uc), ur = lane_from_right(uc);
    const float uw = (ic == 0) ? uc : 0.5f * (ul + uc);
    const float ue = (ic == nx - 1) ? uc : 0.5f * (uc + ur);
    const float v0 = vv[k + 1];
    const float vs = (j == 0) ? v0 : 0.5f * (vv[k] + v0);
    const float vn = (j == ny - 1) ? v0 : 0.5f * (v0 + vv[k + 2]);
    const float hc = h[k + 2];
    const float hxm1 = lane_from_left(hc), hxp1 = lane_from_right(hc);
    const float hxm2 = lane_from_left(hxm1), hxp2 = lane_from_right(hxp1);
    const float qw = face_flux<SLOPE>(hxm2, hxm1, hc, hxp1, uw, cx);
    const float qe = face_flux<SLOPE>(hxm1, hc, hxp1, hxp2, ue, cx);
    const float qs = face_flux<SLOPE>(h[k], h[k + 1], hc, h[k + 3], vs, cy);
    const float qn = face_flux<SLOPE>(h[k + 1], hc, h[k + 3], h[k + 4], vn, cy);
    const float div = (qe - qw) * idx + (qn - qs) * idy;
    const float t = fmaxf(hc + dt * (ss[k] - div), 0.f);
    const float ls = fmaxf(tt[k], -ratio_density * t + sealevel);
    const uint32_t o = (own && j < ny) ? (uint32_t)(j * nx + ic) * 4u : kOOB;
    st(t, rO, o);
    st(ls + t, rUs, o);
    if (lsurf) st(ls, make_rsrc(lsurf, bytes), o);
    if (divflux) st(div, make_rsrc(divflux, bytes), o);
  }
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
  IGM_CHECK_ARG((int64_t)ny * nx * 4 < (int64_t)kOOB, "grid above 2 GiB per field (32-bit buffer offsets)");
  // strips of 60 owned columns, blocks of kTrRows rows
  const int nstrips = (nx + kTrOwned - 1) / kTrOwned;
  const int nrowblk = (ny + kTrRows - 1) / kTrRows;
  const int waves = nstrips * nrowblk;
  dim3 grid((waves + 3) / 4);
  hipStream_t st = (hipStream_t)stream;
#define IGM_TR(S)                                                                                    \
  hipLaunchKernelGGL(transport_kernel<S>, grid, dim3(256), 0, st, ny, nx, nstrips, nrowblk, ubar, vbar, \
                     thk, smb, topg, (float)(1.0 / dx), (float)(1.0 / dy), (float)(dt / dx),           \
                     (float)(dt / dy), (float)dt, (float)ratio_density,                               \
                     (float)sealevel, thk_out, lsurf, usurf, divflux)
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
