// Blatter-Pattyn energy of IGM's ice-flow solver on gfx950: one fused pass
// that evaluates the staggered-grid energy, its adjoint (dJ/dU, dJ/dV) and,
// optionally, the legacy-Adam update, with a float64 energy reduction.
//
// Reference: igm/processes/iceflow/energy/cost_shear.py:103-179,
// cost_gravity.py:24-69, cost_sliding_weertman.py:23-43, energy.py:12-22,
// utils/gradient/compute_gradient.py:3-16, solve/solve.py:45-84.
//
// Work decomposition ("column marching"): a wavefront owns a strip of 64
// node columns and a block of consecutive node rows (the rows of a strip
// are split evenly into as many blocks as give ~1 wave per SIMD in all).
// Lane l holds node column ic = 62*strip - 1 + l; lanes 1..62 own their column, lanes 0 and 63 are a
// one-node halo.  The wave walks up the rows: at step jc it holds node rows
// jc and jc+1 of U and V (all Nz levels) in registers, computes the cell
// (jc, ic) -- the cell whose lower-left node is (jc, ic) -- taking the right
// neighbour's nodes from lane l+1 with a DPP move, and scatters the cell's
// adjoint to its 4 corner nodes: the two left corners stay in the lane, the
// two right ones move to lane l+1 with a DPP move.  Node row jc is complete
// once cell rows jc-1 and jc are done; the upper-corner contributions of cell
// row jc are carried in registers to the next step.  No LDS, no atomics: the
// gradient is a deterministic gather.
//
// Per-cell coefficients that do not depend on (U, V) are computed once per
// solve by energy_coeffs_kernel (7 planes, 28 B per cell).
#include "common.hpp"
#include "../../include/igm_amd.h"

#include <cmath>
#include <mutex>

namespace igm {

namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

constexpr int kPlanes = 7;     // q, w, gx, gy, cs, bx, by
constexpr int kStripOut = 62;  // owned columns per wave
constexpr int kPassBlock = 256;
constexpr int kWavesPerBlock = kPassBlock / kWave;

// --------------------------------------------------------- coefficients --
//
// Plane meaning for the cell with lower-left node (j, i) (S = number of cells):
//   0  q  = dx * damp / max(H, thr)        damp = 1, or 0.01 where stag4h(C) <= 0
//   1  w  = H * Bs / (p * S)               Bs = stag4h(2 A^(-1/n)), p = 1 + 1/n
//   2  gx = 0.25 * rho*g*1e-6 * H * sx / S (sx, sy: staggered slope of usurf)
//   3  gy = 0.25 * rho*g*1e-6 * H * sy / S
//   4  cs = stag4h(C) / (s * S)            s = 1 + 1/m
//   5  bx, 6 by: staggered slope of the bed usurf - thk
// The 0.25 factors absorb the 1/4 of the horizontal node average, which the
// pass keeps unscaled (sum of the 4 corner nodes); dx in q puts the vertical
// derivative in the pass's units of 1/(4 dx) (dx = dX[b, 0, 0], the scalar
// spacing of the velocity derivatives, cost_shear.py:113).
template <typename T>
__global__ void __launch_bounds__(256)
energy_coeffs_kernel(const T* __restrict__ thk, const T* __restrict__ usurf,
                     const T* __restrict__ arrh, const T* __restrict__ slid,
                     const T* __restrict__ dX, int64_t fb, int64_t fr, int64_t fc,
                     const int32_t* __restrict__ extents, int Ny, int Nx, T inv_n, T inv_p,
                     T inv_s, T thr, T rho_g, T* __restrict__ coef, T* __restrict__ hscale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int b = blockIdx.z;
  if (i >= Nx) return;
  const int ny = extents ? min(extents[2 * b], Ny) : Ny;
  const int nx = extents ? min(extents[2 * b + 1], Nx) : Nx;
  const int64_t plane = (int64_t)Ny * Nx;
  T* cb = coef + (int64_t)b * kPlanes * plane + (int64_t)j * Nx + i;
  const int64_t base = (int64_t)b * fb;
  if (i == 0 && j == 0) hscale[b] = T(1) / (T(4) * dX[base]);
  if (!(j < ny - 1 && i < nx - 1)) {
#pragma unroll
    for (int p = 0; p < kPlanes; ++p) cb[p * plane] = T(0);
    return;
  }
  const int64_t o00 = base + (int64_t)j * fr + (int64_t)i * fc;
  const int64_t o01 = o00 + fc;       // (j, i+1)
  const int64_t o10 = o00 + fr;       // (j+1, i)
  const int64_t o11 = o00 + fr + fc;  // (j+1, i+1)
  auto st4 = [](T a11, T a10, T a01, T a00) { return (a11 + a10 + a01 + a00) / T(4); };
  const T t00 = thk[o00], t01 = thk[o01], t10 = thk[o10], t11 = thk[o11];
  const T s00 = usurf[o00], s01 = usurf[o01], s10 = usurf[o10], s11 = usurf[o11];
  const T d00 = dX[o00], d01 = dX[o01], d10 = dX[o10], d11 = dX[o11];
  const T H = st4(t11, t10, t01, t00);
  const T Bs = st4(T(2) * pow(arrh[o11], -inv_n), T(2) * pow(arrh[o10], -inv_n),
                   T(2) * pow(arrh[o01], -inv_n), T(2) * pow(arrh[o00], -inv_n));
  const T Cs = st4(slid[o11], slid[o10], slid[o01], slid[o00]);
  // compute_gradient.py:10-14 on s = usurf (surface) and s = usurf - thk (bed)
  auto gx = [&](T a00, T a01, T a10, T a11) {
    const T e0 = T(2) * (a01 - a00) / (d01 + d00);
    const T e1 = T(2) * (a11 - a10) / (d11 + d10);
    return T(0.5) * (e1 + e0);
  };
  auto gy = [&](T a00, T a01, T a10, T a11) {
    const T e0 = T(2) * (a10 - a00) / (d10 + d00);
    const T e1 = T(2) * (a11 - a01) / (d11 + d01);
    return T(0.5) * (e1 + e0);
  };
  const T sx = gx(s00, s01, s10, s11), sy = gy(s00, s01, s10, s11);
  const T b00 = s00 - t00, b01 = s01 - t01, b10 = s10 - t10, b11 = s11 - t11;
  const T bx = gx(b00, b01, b10, b11), by = gy(b00, b01, b10, b11);
  const T inv_S = T(1) / (T(ny - 1) * T(nx - 1));
  const T damp = Cs > T(0) ? T(1) : T(0.01);
  const T Hc = H > thr ? H : thr;
  cb[0 * plane] = dX[base] * damp / Hc;
  cb[1 * plane] = H * Bs * inv_p * inv_S;
  cb[2 * plane] = T(0.25) * rho_g * H * sx * inv_S;
  cb[3 * plane] = T(0.25) * rho_g * H * sy * inv_S;
  cb[4 * plane] = Cs * inv_s * inv_S;
  cb[5 * plane] = bx;
  cb[6 * plane] = by;
}

// ----------------------------------------------------------------- pass --
//
// All global accesses of the pass go through buffer resources with 32-bit
// byte offsets: an offset at or beyond the resource size reads 0 and drops
// a store, in hardware.  The kernel gives every out-of-grid node (and every
// lane that does not own its node) the offset kOOB, so the zero boundary of
// the stencil and the masked stores need no branches.

// voff: per-lane byte offset (kOOB masks the lane); soff: wave-uniform byte
// offset carried in an SGPR (the level offset), so it costs no VALU add.
// Either way the range check sees an offset >= kOOB for a masked lane.
__device__ __forceinline__ float bload(rsrc_t r, uint32_t voff, uint32_t soff, float) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ double bload(rsrc_t r, uint32_t voff, uint32_t soff, double) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return __hiloint2double((int)v[1], (int)v[0]);
}
__device__ __forceinline__ void bstore(float x, rsrc_t r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), r, voff, soff, 0);
}
__device__ __forceinline__ void bstore(double x, rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  u2 v;
  v[0] = (unsigned)__double2loint(x);
  v[1] = (unsigned)__double2hiint(x);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, 0);
}

// The four Adam moments of one node and level, (m_U, v_U, m_V, v_V): one
// 16-B access in float32, two in float64.
__device__ __forceinline__ void mload(rsrc_t r, uint32_t voff, uint32_t soff, float (&m)[4]) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) m[i] = __uint_as_float(v[i]);
}
__device__ __forceinline__ void mload(rsrc_t r, uint32_t voff, uint32_t soff, double (&m)[4]) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16 * h, soff, 0);
    m[2 * h] = __hiloint2double((int)v[1], (int)v[0]);
    m[2 * h + 1] = __hiloint2double((int)v[3], (int)v[2]);
  }
}
__device__ __forceinline__ void mstore(const float (&m)[4], rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  u4 v;
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = __float_as_uint(m[i]);
  __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, 0);
}
__device__ __forceinline__ void mstore(const double (&m)[4], rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    u4 v;
    v[0] = (unsigned)__double2loint(m[2 * h]);
    v[1] = (unsigned)__double2hiint(m[2 * h]);
    v[2] = (unsigned)__double2loint(m[2 * h + 1]);
    v[3] = (unsigned)__double2hiint(m[2 * h + 1]);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, voff + 16 * h, soff, 0);
  }
}

template <typename T>
struct PassArgs {
  const T* U;
  const T* V;
  T* gU;
  T* gV;
  T* Uo;
  T* Vo;
  T* mom;         // (nbatch, Nz, Ny, Nx, 4) Adam moments, planar layout only
  const T* coef;
  const T* hscale;
  double* partials;
  const int32_t* stop;
  // finalize of the previous iteration, done by every block at its start
  // (prev_partials != nullptr; nbatch 1): the sums of prev_partials give the
  // previous COST, the EarlyStopping rule is applied to a copy of *es_in, and
  // the block decides from it whether this iteration runs; block 0 writes the
  // new state to *es_out (never *es_in, which other blocks are reading), the
  // energies to fin_out and the COST to cost_hist[prev_iter]
  const double* prev_partials;
  const IgmEarlyStop* es_in;
  IgmEarlyStop* es_out;
  double* fin_out;
  double* cost_hist;
  int prev_iter;
  uint32_t vel_bytes;   // size of every velocity-shaped resource
  uint32_t mom_bytes;
  uint32_t coef_bytes;
  uint32_t bs_bytes;    // batch stride of the velocity layout, bytes
  int Ny, Nx, plane, nstrips, nrowblk, blocks_per_batch;
  T dz[IGM_MAX_NZ];   // layer thickness dzeta_l
  T l2dz[IGM_MAX_NZ]; // log2(dzeta_l)
  T idz[IGM_MAX_NZ];  // 1/dzeta_l
  T wz[IGM_MAX_NZ];   // level weight of the gravity term: (dzeta_{k-1} + dzeta_k)/2
  T e_glen;           // (p-2)/2
  T min_sr2, max_sr2, regu2;
  T e_slid;           // s/2 - 1
  T s_slid;           // s
  T regu_w2;
  T alpha, omb1, omb2, eps;  // Adam: alpha, 1-beta1, 1-beta2, epsilon
};

// ------------------------------------------------------------- finalize --
//
// Sum of one batch entry's per-block partials in a fixed order (strided
// per-thread sums, a wave butterfly, then the 4 waves in order) into
// out[4] = {mean shear, mean sliding, mean gravity, COST}, then the
// EarlyStopping rule (iceflow/utils.py:92-111) when es is given.  Run by one
// 256-thread block.  The solver folds it into the next iteration's pass
// instead (every block redoes the sum of the previous pass's partials, which
// the kernel boundary has made visible): a "last block sums" fold inside the
// same pass would need a device-scope release per block, which on the
// multi-XCD MI355X writes back the whole L2 (measured 18 us per iteration).
__device__ __forceinline__ void es_apply(IgmEarlyStop& s, double cost, int iter) {
  s.iters_done = iter + 1;
  if (!s.has_best) {
    s.best_loss = cost;
    s.has_best = 1;
  } else {
    const double rel = (s.best_loss - cost) / fabs(s.best_loss);
    if (rel > s.relative_min_delta) {
      s.best_loss = cost;
      s.wait = 0;
    } else {
      s.wait += 1;
      if (s.wait >= s.patience) {
        s.stopped = 1;
        s.stop_iter = iter;
      }
    }
  }
}

// (shear, sliding, gravity) sums of one entry's partials, in every thread
__device__ void block_sum_partials(const double* p, int nblk, double (&r)[3]) {
  __shared__ double fin[4][3];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int i = threadIdx.x; i < nblk; i += 256) {
    s0 += p[3 * i];
    s1 += p[3 * i + 1];
    s2 += p[3 * i + 2];
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if ((threadIdx.x & 63) == 0) {
    fin[threadIdx.x >> 6][0] = s0;
    fin[threadIdx.x >> 6][1] = s1;
    fin[threadIdx.x >> 6][2] = s2;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = ((fin[0][c] + fin[1][c]) + fin[2][c]) + fin[3][c];
}

__device__ void finalize_entry(const double* p, int nblk, double* out, IgmEarlyStop* es,
                               double* cost_hist, int iter) {
  double r[3];
  block_sum_partials(p, nblk, r);
  if (threadIdx.x == 0) {
    const double cost = (r[0] + r[1]) + r[2];  // tf.add_n in energy_components order
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
    out[3] = cost;
    if (es != nullptr) {
      if (cost_hist != nullptr) cost_hist[iter] = cost;
      IgmEarlyStop st = *es;
      es_apply(st, cost, iter);
      *es = st;
    }
  }
}

// INTER = 0: planar (Nz, Ny, Nx) with separate U, V buffers;
// INTER = 1: node-interleaved Y (Ny, Nx, 2Nz), V = Y + Nz.
template <typename T, int NZ, int MODE, int INTER>
__global__ void __launch_bounds__(kPassBlock)
energy_pass_kernel(const PassArgs<T> a) {
  const int lane = threadIdx.x & (kWave - 1);
  // logical block: XCD-contiguous bands of (strip, row block) tiles
  const int lblk = xcd_remap(blockIdx.x, gridDim.x);
  const int wid = lblk * kWavesPerBlock + (threadIdx.x >> 6);
  const int b = blockIdx.y;
  double e_sh = 0.0, e_sl = 0.0, e_gr = 0.0;
  bool stopped = (a.stop != nullptr) && (*a.stop != 0);
  if (a.prev_partials != nullptr) {
    double r[3];
    block_sum_partials(a.prev_partials, a.blocks_per_batch, r);
    const double cost = (r[0] + r[1]) + r[2];
    IgmEarlyStop st = *a.es_in;
    const bool was = st.stopped != 0;
    if (!was) es_apply(st, cost, a.prev_iter);
    if (lblk == 0 && threadIdx.x == 0) {
      *a.es_out = st;  // (when already stopped: carried over unchanged)
      if (!was) {
        a.fin_out[0] = r[0];
        a.fin_out[1] = r[1];
        a.fin_out[2] = r[2];
        a.fin_out[3] = cost;
        if (a.cost_hist != nullptr) a.cost_hist[a.prev_iter] = cost;
      }
    }
    stopped = stopped || (st.stopped != 0);
  }
  constexpr uint32_t ES = sizeof(T);

  if (!stopped && wid < a.nstrips * a.nrowblk) {
    const int strip = wid % a.nstrips;
    const int rblk = wid / a.nstrips;
    const int ic = strip * kStripOut - 1 + lane;
    const bool col_ok = (ic >= 0) && (ic < a.Nx);
    const bool own = (lane >= 1) && (lane <= kStripOut) && col_ok;
    // node rows [j0, j1) of this wave: the Ny rows split evenly over nrowblk
    const int j0 = (int)(((int64_t)rblk * a.Ny) / a.nrowblk);
    const int j1 = (int)(((int64_t)(rblk + 1) * a.Ny) / a.nrowblk);
    const rsrc_t rU = make_rsrc(a.U, a.vel_bytes), rV = make_rsrc(a.V, a.vel_bytes);
    const rsrc_t rC = make_rsrc(a.coef, a.coef_bytes);
    const uint32_t bbase = (uint32_t)b * a.bs_bytes;
    const uint32_t cbase = (uint32_t)b * (uint32_t)(kPlanes * a.plane) * ES;
    const uint32_t kstep = INTER ? ES : (uint32_t)a.plane * ES;   // bytes between levels
    const T hh = a.hscale[b];
    const T hh2 = hh * hh;
    const T h05 = T(0.5) * hh2, h025 = T(0.25) * hh2;

    // byte offset of node (row, ic), level 0 -- or kOOB outside the grid
    auto node_off = [&](int row, bool extra) -> uint32_t {
      const bool ok = extra && col_ok && row >= 0 && row < a.Ny;
      const uint32_t e = (uint32_t)(row * a.Nx + ic) * (INTER ? 2u * NZ : 1u);
      return ok ? bbase + e * ES : kOOB;
    };
    // node row `row` of U and V (rows this wave never uses -- past j1 --
    // read nothing: the last step's prefetch would be wasted bytes)
    auto load_row = [&](T (&u)[NZ], T (&v)[NZ], int row) {
      const uint32_t o = node_off(row, row <= j1);
#pragma unroll
      for (int k = 0; k < NZ; ++k) {
        u[k] = bload(rU, o, k * kstep, T());
        v[k] = bload(rV, o, k * kstep, T());
      }
    };

    // upper-corner adjoint contributions of the previous cell row
    T cu[NZ], cv[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      cu[k] = T(0);
      cv[k] = T(0);
    }

    // coefficient planes of cell row jc (kOOB rows read 0)
    const uint32_t ps = (uint32_t)a.plane * ES;
    auto load_coef = [&](T (&c)[kPlanes], int jc) {
      const uint32_t co =
          (col_ok && jc >= 0 && jc < j1) ? cbase + (uint32_t)(jc * a.Nx + ic) * ES : kOOB;
#pragma unroll
      for (int p = 0; p < kPlanes; ++p) c[p] = bload(rC, co, p * ps, T());
    };

    // float32 fused Adam: the moments of node row jc are loaded at the start
    // of step jc and the update runs at its end, after the level sweep.
    // (The float64 pass keeps its loads at the point of use, where its
    // register budget stays within the VGPR file.)
    constexpr bool kDeferAdam = (MODE == 2) && (sizeof(T) == 4);

    // One step: node rows jc (up, vp) and jc+1 (uc, vc) and the coefficients
    // of cell row jc (cc) are in registers, loaded one step earlier.  The
    // step issues the loads of the next step first (vector memory returns in
    // order, so the sweep below never waits on them), computes cell row jc
    // and emits node row jc.
    auto step = [&](int jc, const T (&up)[NZ], const T (&vp)[NZ], const T (&uc)[NZ],
                    const T (&vc)[NZ], T (&un)[NZ], T (&vn)[NZ], const T (&cc)[kPlanes],
                    T (&cn)[kPlanes]) {
      const bool emit_row = own && (jc >= j0);
      const uint32_t on = node_off(jc, emit_row);  // output node (jc, ic), masked
      // moments of node (jc, ic): 4 x the velocity node offset, plus the batch
      const uint32_t om = emit_row ? (uint32_t)b * (uint32_t)(NZ * a.plane) * 4u * ES +
                                         (uint32_t)(jc * a.Nx + ic) * 4u * ES
                                   : kOOB;
      const uint32_t mstep = 4u * kstep;
      const rsrc_t rM = make_rsrc(a.mom, a.mom_bytes);
      T mo[NZ][4], gu_[NZ], gv_[NZ];
      if constexpr (kDeferAdam) {
#pragma unroll
        for (int k = 0; k < NZ; ++k) mload(rM, om, k * mstep, mo[k]);
      }
      load_coef(cn, jc + 1);     // next step's coefficients
      load_row(un, vn, jc + 2);  // next step's upper node row
      // keep the loads above ahead of the sweep (the scheduler would sink them)
      __builtin_amdgcn_sched_barrier(0);
      const T q = cc[0], w = cc[1], gx = cc[2], gy = cc[3], csl = cc[4], bx = cc[5], by = cc[6];
      const T whh = w * hh2;

      // legacy-Adam update of node (jc, ic), level k, from its gradient
      auto adam = [&](int k, T g_u, T g_v, T (&m)[4]) {
        m[0] += (g_u - m[0]) * a.omb1;
        m[1] += (g_u * g_u - m[1]) * a.omb2;
        m[2] += (g_v - m[2]) * a.omb1;
        m[3] += (g_v * g_v - m[3]) * a.omb2;
        bstore(up[k] - (m[0] * a.alpha) * rcp(sqrt_(m[1]) + a.eps), make_rsrc(a.Uo, a.vel_bytes), on,
               k * kstep);
        bstore(vp[k] - (m[2] * a.alpha) * rcp(sqrt_(m[3]) + a.eps), make_rsrc(a.Vo, a.vel_bytes), on,
               k * kstep);
        mstore(m, rM, om, k * mstep);
      };

      // Adjoint scatter of level k to the 4 corner nodes, then the output of
      // node (jc, ic), level k (masked stores).  With A = gs - gx and
      // B = gs + gx, the corners get A -+ gy (left) and B -+ gy (right); the
      // right ones arrive from lane l-1: node row jc receives
      // X - Z, node row jc+1 (carried) X + Z, X = A + B(l-1), Z = gy + gy(l-1).
      auto scatter = [&](T& carry, T gx_, T gy_, T gs_) -> T {
        const T A = gs_ - gx_, B = gs_ + gx_;
        const T X = A + lane_from_left(B);
        const T Z = gy_ + lane_from_left(gy_);
        const T g = (carry - Z) + X;
        carry = X + Z;
        return g;
      };
      auto emit = [&](int k, T gux, T guy, T gsu, T gvx, T gvy, T gsv) {
        const uint32_t sk = k * kstep;
        const T g_u = scatter(cu[k], gux, guy, gsu);
        const T g_v = scatter(cv[k], gvx, gvy, gsv);
        if constexpr (MODE == 1) {
          bstore(g_u, make_rsrc(a.gU, a.vel_bytes), on, sk);
          bstore(g_v, make_rsrc(a.gV, a.vel_bytes), on, sk);
        } else if constexpr (kDeferAdam) {
          gu_[k] = g_u;
          gv_[k] = g_v;
        } else if constexpr (MODE == 2) {
          T m[4];
          mload(rM, om, k * mstep, m);
          adam(k, g_u, g_v, m);
        }
      };

      // Level sweep; layer l = k-1 is handled once level k is known.  The
      // horizontal derivatives are kept in units of 1/(4 dx) (ux = 4 dx *
      // dU/dx at the level), the vertical ones likewise (q holds damp*dx/Hc);
      // hh2 = 1/(4 dx)^2 restores the strain rate where it is needed.
      T p_ux = 0, p_uy = 0, p_vx = 0, p_vy = 0, p_su = 0, p_sv = 0;  // level k-1
      T g_ux = 0, g_uy = 0, g_vx = 0, g_vy = 0, g_su = 0, g_sv = 0;  // its adjoint
      T c_sh = 0, c_sl = 0, wsu = 0, wsv = 0;
#pragma unroll
      for (int k = 0; k < NZ; ++k) {
        const T su = up[k] + uc[k], du = uc[k] - up[k];
        const T sv = vp[k] + vc[k], dv = vc[k] - vp[k];
        const T su_r = lane_from_right(su), du_r = lane_from_right(du);
        const T sv_r = lane_from_right(sv), dv_r = lane_from_right(dv);
        const T ux = su_r - su, uy = du + du_r;
        const T vx = sv_r - sv, vy = dv + dv_r;
        const T sU = su + su_r, sV = sv + sv_r;  // 4 x node average
        T n_ux = 0, n_uy = 0, n_vx = 0, n_vy = 0;
        T n_su = a.wz[k] * gx, n_sv = a.wz[k] * gy;  // gravity (linear)
        wsu += a.wz[k] * sU;
        wsv += a.wz[k] * sV;
        if (k == 0) {  // Weertman sliding on the basal level
          const T ub = T(0.25) * sU, vb = T(0.25) * sV;
          const T c2 = ub * bx + vb * by;
          const T N = (ub * ub + vb * vb) + a.regu_w2 + c2 * c2;
          const T t = pow_pos(N, a.e_slid);  // N^(s/2 - 1)
          c_sl = csl * N * t;
          const T gs = T(0.25) * csl * a.s_slid * t;
          n_su += gs * (ub + c2 * bx);
          n_sv += gs * (vb + c2 * by);
        }
        if (k >= 1) {
          const int l = k - 1;
          const T dUdx = p_ux + ux, dUdy = p_uy + uy;
          const T dVdx = p_vx + vx, dVdy = p_vy + vy;
          const T rz = q * a.idz[l];
          const T dUdz = (sU - p_su) * rz, dVdz = (sV - p_sv) * rz;
          const T Dux = dUdx - dUdz * bx, Duy = dUdy - dUdz * by;
          const T Dvx = dVdx - dVdz * bx, Dvy = dVdy - dVdz * by;
          const T P = Dvx + Duy;  // 2 Exy
          const T Ezz = -(Dux + Dvy);
          // sr2 = hh2 (0.5 (Dux^2 + Dvy^2 + Ezz^2) + Exy^2 + 0.25 (dUdz^2 + dVdz^2))
          const T t1 = (Dux * Dux + Dvy * Dvy) + Ezz * Ezz;
          const T t2 = (P * P + dUdz * dUdz) + dVdz * dVdz;
          const T sr2 = h05 * t1 + h025 * t2;
          const T sclip = clamp_(sr2, a.min_sr2, a.max_sr2);
          const bool in_rng = (sclip == sr2);  // the clip passes its gradient inside the range
          const T sc = sclip + a.regu2;
          const T wl = scaled_pow(sc, a.e_glen, a.dz[l], a.l2dz[l]);  // dzeta_l sc^e
          c_sh += wl * sr2;
          // adjoint of the layer (in the scaled variables):
          // G = w hh2 wl (1 + e sr2 / sc), zero ratio outside the clip range
          const T WL = whh * wl;
          const T er = in_rng ? a.e_glen * sr2 : T(0);
          const T G = WL + WL * (er * rcp(sc));
          const T Gh = T(0.5) * G;
          const T gDux = G * (Dux - Ezz), gDvy = G * (Dvy - Ezz), gDxy = Gh * P;
          const T gdUdz = Gh * dUdz - gDux * bx - gDxy * by;
          const T gdVdz = Gh * dVdz - gDxy * bx - gDvy * by;
          g_ux += gDux;
          g_uy += gDxy;
          g_vx += gDxy;
          g_vy += gDvy;
          n_ux = gDux;
          n_uy = gDxy;
          n_vx = gDxy;
          n_vy = gDvy;
          const T gzu = gdUdz * rz, gzv = gdVdz * rz;
          g_su -= gzu;
          n_su += gzu;
          g_sv -= gzv;
          n_sv += gzv;
          emit(k - 1, g_ux, g_uy, g_su, g_vx, g_vy, g_sv);
        }
        p_ux = ux; p_uy = uy; p_vx = vx; p_vy = vy; p_su = sU; p_sv = sV;
        g_ux = n_ux; g_uy = n_uy; g_vx = n_vx; g_vy = n_vy; g_su = n_su; g_sv = n_sv;
      }
      emit(NZ - 1, g_ux, g_uy, g_su, g_vx, g_vy, g_sv);

      if constexpr (kDeferAdam) {
        // keep the updates after the sweep (the scheduler would hoist them
        // into it, right after each gradient, and wait there for the moments)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < NZ; ++k) adam(k, gu_[k], gv_[k], mo[k]);
      }

      if (emit_row) {
        e_sh += (double)(w * c_sh);
        e_sl += (double)c_sl;
        e_gr += (double)(gx * wsu + gy * wsv);
      }
    };

    // three row and coefficient buffers in rotation, unrolled so that no
    // register moves
    T bu0[NZ], bv0[NZ], bu1[NZ], bv1[NZ], bu2[NZ], bv2[NZ];
    T cf0[kPlanes], cf1[kPlanes], cf2[kPlanes];
    load_row(bu0, bv0, j0 - 1);
    load_row(bu1, bv1, j0);
    load_coef(cf0, j0 - 1);
    for (int jc = j0 - 1; jc < j1; jc += 3) {
      step(jc, bu0, bv0, bu1, bv1, bu2, bv2, cf0, cf1);
      if (jc + 1 >= j1) break;
      step(jc + 1, bu1, bv1, bu2, bv2, bu0, bv0, cf1, cf2);
      if (jc + 2 >= j1) break;
      step(jc + 2, bu2, bv2, bu0, bv0, bu1, bv1, cf2, cf0);
    }
  }

  // deterministic block reduction of the three energy terms
  __shared__ double red[kWavesPerBlock][3];
  e_sh = wave_sum(e_sh);
  e_sl = wave_sum(e_sl);
  e_gr = wave_sum(e_gr);
  if (lane == 0) {
    red[threadIdx.x >> 6][0] = e_sh;
    red[threadIdx.x >> 6][1] = e_sl;
    red[threadIdx.x >> 6][2] = e_gr;
  }
  __syncthreads();
  if (threadIdx.x == 0 && !stopped) {
    double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int wv = 0; wv < kWavesPerBlock; ++wv) {
      s0 += red[wv][0];
      s1 += red[wv][1];
      s2 += red[wv][2];
    }
    double* o = a.partials + 3 * ((int64_t)b * a.blocks_per_batch + lblk);
    o[0] = s0;
    o[1] = s1;
    o[2] = s2;
  }
}

// Waves per launch the row split aims at: 4 per CU of MI355X's 256, one per
// SIMD, all resident at once, so that every SIMD gets one equally long wave
// (measured at 1000 x 1000: 1024 waves beat 2048-4096, whose extra halo rows
// cost more than the latency a second wave would hide).  IGM_ENERGY_WAVES
// overrides it (for tuning sweeps).
constexpr int kTargetWaves = 1024;
constexpr int kMinRows = 4;  // fewer rows per wave would be mostly halo

int target_waves() {
  static int n = [] {
    const char* e = getenv("IGM_ENERGY_WAVES");
    const int r = e ? atoi(e) : kTargetWaves;
    return (r >= 64 && r <= (1 << 20)) ? r : kTargetWaves;
  }();
  return n;
}

struct PassGeom {
  int nstrips, nrowblk, blocks;
};
inline PassGeom pass_geom(int nbatch, int Ny, int Nx) {
  PassGeom g;
  g.nstrips = (Nx + kStripOut - 1) / kStripOut;
  const int64_t per = (int64_t)g.nstrips * nbatch;
  int64_t want = (target_waves() + per / 2) / per;
  const int64_t most = Ny / kMinRows > 1 ? Ny / kMinRows : 1;
  g.nrowblk = (int)(want < 1 ? 1 : (want > most ? most : want));
  const int waves = g.nstrips * g.nrowblk;
  g.blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  return g;
}

// ------------------------------------------------------------- finalize --

constexpr int kFinBlock = 256;

__global__ void __launch_bounds__(kFinBlock)
energy_finalize_kernel(const double* __restrict__ partials, int blocks_per_batch,
                       double* __restrict__ out, IgmEarlyStop* es, double* cost_hist, int iter) {
  const int b = blockIdx.x;
  if (es != nullptr && es->stopped) return;
  finalize_entry(partials + 3 * (int64_t)b * blocks_per_batch, blocks_per_batch, out + 4 * b, es,
                 cost_hist, iter);
}

// ------------------------------------------------------ velocity outputs --

template <typename T, int NZ>
struct VwArr {
  T w[NZ];
};

// ------------------------------------------------------------- dispatch --

template <typename T>
int coeffs_impl(const T* thk, const T* usurf, const T* arrh, const T* slid, const T* dX,
                int64_t fb, int64_t fr, int64_t fc, const int32_t* extents, int nbatch, int Ny,
                int Nx, const IgmPhysics* phys, T* coef, T* hscale, void* stream) {
  IGM_CHECK_ARG(thk && usurf && arrh && slid && dX && coef && hscale && phys, "null pointer");
  IGM_CHECK_ARG(nbatch >= 1 && Ny >= 2 && Nx >= 2, "need nbatch >= 1 and a grid of at least 2 x 2");
  IGM_CHECK_ARG(Ny <= 65535, "Ny must be <= 65535");
  IGM_CHECK_ARG(phys->force_negative_gravitational_energy == 0,
                "force_negative_gravitational_energy is not supported");
  const T inv_n = T(1.0 / phys->exp_glen);
  const T p = T(1.0 + 1.0 / phys->exp_glen);
  const T s = T(1.0 + 1.0 / phys->exp_weertman);
  const T rho_g = T(phys->ice_density * phys->gravity_cst * 1e-6);
  dim3 grid((Nx + 255) / 256, Ny, nbatch);
  hipLaunchKernelGGL(energy_coeffs_kernel<T>, grid, dim3(256), 0, (hipStream_t)stream, thk,
                     usurf, arrh, slid, dX, fb, fr, fc, extents, Ny, Nx, inv_n, T(1) / p, T(1) / s,
                     T(phys->thr_ice_thk), rho_g, coef, hscale);
  IGM_CHECK_LAUNCH();
  return 0;
}

template <typename T, int MODE, int INTER>
void launch_pass_nz(int nz, dim3 grid, hipStream_t st, const PassArgs<T>& a, bool* ok) {
  *ok = true;
  switch (nz) {
#define IGM_NZ_CASE(N) \
  case N: hipLaunchKernelGGL((energy_pass_kernel<T, N, MODE, INTER>), grid, dim3(kPassBlock), 0, st, a); break;
    IGM_NZ_CASE(2)
    IGM_NZ_CASE(3)
    IGM_NZ_CASE(4)
    IGM_NZ_CASE(5)
    IGM_NZ_CASE(6)
    IGM_NZ_CASE(8)
    IGM_NZ_CASE(10)
    IGM_NZ_CASE(12)
    IGM_NZ_CASE(16)
    IGM_NZ_CASE(20)
#undef IGM_NZ_CASE
    default: *ok = false;
  }
}

template <typename T>
int pass_impl(int mode, int nz, int nbatch, int Ny, int Nx, const T* U, const T* V, int64_t bs,
              int64_t ks, int64_t rs, int64_t cs, const T* coef, const T* hscale,
              const double* dzeta, const IgmPhysics* phys, T* gU, T* gV, T* Uo, T* Vo, T* mom,
              const IgmAdam* adam, double* partials, const int32_t* stop,
              const double* prev_partials, const IgmEarlyStop* es_in, IgmEarlyStop* es_out,
              double* fin_out, double* cost_hist, int prev_iter, void* stream) {
  IGM_CHECK_ARG(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
  IGM_CHECK_ARG(U && V && coef && hscale && dzeta && phys && partials, "null pointer");
  IGM_CHECK_ARG(nbatch >= 1 && Ny >= 2 && Nx >= 2, "need nbatch >= 1 and a grid of at least 2 x 2");
  IGM_CHECK_ARG(nz >= 2 && nz <= IGM_MAX_NZ, "nz out of range");
  IGM_CHECK_ARG(nbatch <= 65535, "nbatch must be <= 65535");
  if (mode == 1) IGM_CHECK_ARG(gU && gV, "mode 1 needs gU and gV");
  if (mode == 2) {
    IGM_CHECK_ARG(Uo && Vo && mom && adam, "mode 2 needs Uout, Vout, moments and adam");
    IGM_CHECK_ARG(Uo != U && Vo != V, "Uout/Vout must not alias U/V");
  }
  // supported layouts: planar (Nz, Ny, Nx) or node-interleaved (Ny, Nx, 2 Nz)
  const int64_t plane = (int64_t)Ny * Nx;
  int inter = -1;
  if (cs == 1 && rs == Nx && ks == plane) inter = 0;
  if (ks == 1 && cs == 2 * nz && rs == (int64_t)Nx * 2 * nz && V == U + nz) inter = 1;
  IGM_CHECK_ARG(inter >= 0, "unsupported velocity layout (need planar (Nz,Ny,Nx) or interleaved (Ny,Nx,2Nz))");
  IGM_CHECK_ARG(mode != 2 || inter == 0, "mode 2 needs the planar velocity layout");
  const int64_t per = inter ? plane * 2 * nz : plane * nz;  // elements of one batch entry
  if (nbatch > 1) IGM_CHECK_ARG(bs >= per, "batch stride smaller than one entry");
  const int64_t span = (int64_t)(nbatch - 1) * (nbatch > 1 ? bs : 0) + per;
  const int64_t vel_bytes = span * (int64_t)sizeof(T);
  const int64_t coef_bytes = (int64_t)nbatch * kPlanes * plane * (int64_t)sizeof(T);
  const int64_t mom_bytes = (int64_t)nbatch * nz * plane * 4 * (int64_t)sizeof(T);
  IGM_CHECK_ARG(vel_bytes < (int64_t)kOOB && coef_bytes < (int64_t)kOOB && mom_bytes < (int64_t)kOOB,
                "problem too large for 32-bit buffer offsets (2 GiB per buffer)");
  PassArgs<T> a{};
  a.U = U; a.V = V; a.gU = gU; a.gV = gV; a.Uo = Uo; a.Vo = Vo;
  a.mom = mom;
  a.coef = coef; a.hscale = hscale; a.partials = partials; a.stop = stop;
  if (prev_partials) {
    IGM_CHECK_ARG(nbatch == 1, "the fused finalize needs nbatch == 1");
    IGM_CHECK_ARG(es_in && es_out && fin_out && es_in != es_out, "fused finalize needs es_in != es_out and out");
    IGM_CHECK_ARG(prev_partials != partials, "prev_partials must not alias partials");
  }
  a.prev_partials = prev_partials; a.es_in = es_in; a.es_out = es_out; a.fin_out = fin_out;
  a.cost_hist = cost_hist; a.prev_iter = prev_iter;
  a.vel_bytes = (uint32_t)vel_bytes;
  a.coef_bytes = (uint32_t)coef_bytes;
  a.mom_bytes = (uint32_t)mom_bytes;
  a.bs_bytes = (uint32_t)((nbatch > 1 ? bs : 0) * (int64_t)sizeof(T));
  a.plane = (int)plane;
  a.Ny = Ny; a.Nx = Nx;
  const PassGeom g = pass_geom(nbatch, Ny, Nx);
  a.nstrips = g.nstrips; a.nrowblk = g.nrowblk; a.blocks_per_batch = g.blocks;
  for (int k = 0; k < IGM_MAX_NZ; ++k) { a.dz[k] = a.idz[k] = a.wz[k] = a.l2dz[k] = T(0); }
  for (int l = 0; l < nz - 1; ++l) {
    IGM_CHECK_ARG(dzeta[l] > 0.0, "dzeta must be positive");
    a.dz[l] = T(dzeta[l]);
    a.idz[l] = T(1.0 / dzeta[l]);
    a.l2dz[l] = T(std::log2(dzeta[l]));
  }
  for (int k = 0; k < nz; ++k) {
    const double lo = k > 0 ? dzeta[k - 1] : 0.0;
    const double hi = k < nz - 1 ? dzeta[k] : 0.0;
    a.wz[k] = T(0.5 * (lo + hi));
  }
  const double p = 1.0 + 1.0 / phys->exp_glen;
  const double s = 1.0 + 1.0 / phys->exp_weertman;
  a.e_glen = T((p - 2.0) / 2.0);
  // The clip bounds min_sr^2, max_sr^2 live in the working precision
  // (cost_shear.py:172): in float32, 1e40 becomes +inf.
  a.min_sr2 = T(phys->min_sr * phys->min_sr);
  a.max_sr2 = T(phys->max_sr * phys->max_sr);
  a.regu2 = T(phys->regu_glen * phys->regu_glen);
  a.e_slid = T(s / 2.0 - 1.0);
  a.s_slid = T(s);
  a.regu_w2 = T(phys->regu_weertman * phys->regu_weertman);
  if (adam) {
    a.alpha = T(adam->alpha);
    a.omb1 = T(1) - T(adam->beta1);
    a.omb2 = T(1) - T(adam->beta2);
    a.eps = T(adam->epsilon);
  }
  dim3 grid(g.blocks, nbatch);
  bool ok = false;
  hipStream_t st = (hipStream_t)stream;
#define IGM_MODE_LAYOUT(M, I) \
  if (mode == M && inter == I) launch_pass_nz<T, M, I>(nz, grid, st, a, &ok);
  IGM_MODE_LAYOUT(0, 0) IGM_MODE_LAYOUT(1, 0) IGM_MODE_LAYOUT(2, 0)
  IGM_MODE_LAYOUT(0, 1) IGM_MODE_LAYOUT(1, 1)
#undef IGM_MODE_LAYOUT
  IGM_CHECK_ARG(ok, "nz not compiled (supported: 2,3,4,5,6,8,10,12,16,20)");
  IGM_CHECK_LAUNCH();
  return 0;
}

}  // namespace igm

// The velocity-outputs kernel takes the vertical weights by value.
namespace igm {

template <typename T, int NZ>
__global__ void __launch_bounds__(256)
velout_kernel(int Ny, int Nx, const T* __restrict__ U, const T* __restrict__ V, int64_t ks,
              int64_t rs, int64_t cs, const T* __restrict__ thk, VwArr<T, NZ> vw, T* Uo, T* Vo,
              T* ub, T* vb, T* us, T* vs, T* ubar, T* vbar) {
  const int64_t n = (int64_t)Ny * Nx;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int j = (int)(idx / Nx), i = (int)(idx % Nx);
  const int64_t o = (int64_t)j * rs + (int64_t)i * cs;
  const bool ice = thk[idx] > T(0);
  T sbu = 0, sbv = 0, u0 = 0, v0 = 0, u1 = 0, v1 = 0;
#pragma unroll
  for (int k = 0; k < NZ; ++k) {
    const T u = ice ? U[o + k * ks] : T(0);
    const T v = ice ? V[o + k * ks] : T(0);
    if (Uo) Uo[k * n + idx] = u;
    if (Vo) Vo[k * n + idx] = v;
    sbu += u * vw.w[k];
    sbv += v * vw.w[k];
    if (k == 0) { u0 = u; v0 = v; }
    if (k == NZ - 1) { u1 = u; v1 = v; }
  }
  if (ub) ub[idx] = u0;
  if (vb) vb[idx] = v0;
  if (us) us[idx] = u1;
  if (vs) vs[idx] = v1;
  if (ubar) ubar[idx] = sbu;
  if (vbar) vbar[idx] = sbv;
}

template <typename T, int NZ>
void velout_launch(int Ny, int Nx, const T* U, const T* V, int64_t ks, int64_t rs, int64_t cs,
                   const T* thk, const double* vert_weight, T* Uo, T* Vo, T* ub, T* vb, T* us,
                   T* vs, T* ubar, T* vbar, hipStream_t st) {
  VwArr<T, NZ> vw;
  for (int k = 0; k < NZ; ++k) vw.w[k] = T(vert_weight[k]);
  const int64_t n = (int64_t)Ny * Nx;
  hipLaunchKernelGGL((velout_kernel<T, NZ>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     Ny, Nx, U, V, ks, rs, cs, thk, vw, Uo, Vo, ub, vb, us, vs, ubar, vbar);
}

template <typename T>
int velout_impl2(int nz, int Ny, int Nx, const T* U, const T* V, int64_t ks, int64_t rs,
                 int64_t cs, const T* thk, const double* vert_weight, T* Uo, T* Vo, T* ub,
                 T* vb, T* us, T* vs, T* ubar, T* vbar, void* stream) {
  IGM_CHECK_ARG(U && V && thk && vert_weight, "null pointer");
  IGM_CHECK_ARG(Ny >= 1 && Nx >= 1, "empty grid");
  hipStream_t st = (hipStream_t)stream;
  switch (nz) {
#define IGM_VZ(N) \
  case N: velout_launch<T, N>(Ny, Nx, U, V, ks, rs, cs, thk, vert_weight, Uo, Vo, ub, vb, us, vs, ubar, vbar, st); break;
    IGM_VZ(2) IGM_VZ(3) IGM_VZ(4) IGM_VZ(5) IGM_VZ(6) IGM_VZ(8) IGM_VZ(10) IGM_VZ(12) IGM_VZ(16) IGM_VZ(20)
#undef IGM_VZ
    default:
      set_error("igm_velocity_outputs: nz not compiled (supported: 2,3,4,5,6,8,10,12,16,20)");
      return -1;
  }
  IGM_CHECK_LAUNCH();
  return 0;
}
}  // namespace igm

extern "C" {

const char* igm_last_error(void) { return igm::last_error(); }

int64_t igm_energy_workspace_doubles(int32_t nbatch, int32_t Ny, int32_t Nx) {
  if (nbatch < 1 || Ny < 2 || Nx < 2) return 0;
  const igm::PassGeom g = igm::pass_geom(nbatch, Ny, Nx);
  return 3 * (int64_t)nbatch * g.blocks;
}

int igm_energy_coeffs_f32(const float* thk, const float* usurf, const float* arrhenius,
                          const float* slidingco, const float* dX, int64_t fb, int64_t fr,
                          int64_t fc, const int32_t* extents, int32_t nbatch, int32_t Ny,
                          int32_t Nx, const IgmPhysics* phys, float* coef, float* hscale,
                          void* stream) {
  return igm::coeffs_impl<float>(thk, usurf, arrhenius, slidingco, dX, fb, fr, fc, extents, nbatch,
                                 Ny, Nx, phys, coef, hscale, stream);
}
int igm_energy_coeffs_f64(const double* thk, const double* usurf, const double* arrhenius,
                          const double* slidingco, const double* dX, int64_t fb, int64_t fr,
                          int64_t fc, const int32_t* extents, int32_t nbatch, int32_t Ny,
                          int32_t Nx, const IgmPhysics* phys, double* coef, double* hscale,
                          void* stream) {
  return igm::coeffs_impl<double>(thk, usurf, arrhenius, slidingco, dX, fb, fr, fc, extents,
                                  nbatch, Ny, Nx, phys, coef, hscale, stream);
}

int igm_energy_pass_f32(int32_t mode, int32_t nz, int32_t nbatch, int32_t Ny, int32_t Nx,
                        const float* U, const float* V, int64_t bstride, int64_t kstride,
                        int64_t rstride, int64_t cstride, const float* coef, const float* hscale,
                        const double* dzeta, const IgmPhysics* phys, float* gU, float* gV,
                        float* Uout, float* Vout, float* moments, const IgmAdam* adam,
                        double* partials, const int32_t* stop_flag,
                        const double* prev_partials, const IgmEarlyStop* es_in,
                        IgmEarlyStop* es_out, double* out, double* cost_hist, int32_t prev_iter,
                        void* stream) {
  return igm::pass_impl<float>(mode, nz, nbatch, Ny, Nx, U, V, bstride, kstride, rstride, cstride,
                               coef, hscale, dzeta, phys, gU, gV, Uout, Vout, moments, adam,
                               partials, stop_flag, prev_partials, es_in, es_out, out, cost_hist,
                               prev_iter, stream);
}
int igm_energy_pass_f64(int32_t mode, int32_t nz, int32_t nbatch, int32_t Ny, int32_t Nx,
                        const double* U, const double* V, int64_t bstride, int64_t kstride,
                        int64_t rstride, int64_t cstride, const double* coef,
                        const double* hscale, const double* dzeta, const IgmPhysics* phys,
                        double* gU, double* gV, double* Uout, double* Vout, double* moments,
                        const IgmAdam* adam, double* partials, const int32_t* stop_flag,
                        const double* prev_partials, const IgmEarlyStop* es_in,
                        IgmEarlyStop* es_out, double* out, double* cost_hist, int32_t prev_iter,
                        void* stream) {
  return igm::pass_impl<double>(mode, nz, nbatch, Ny, Nx, U, V, bstride, kstride, rstride,
                                cstride, coef, hscale, dzeta, phys, gU, gV, Uout, Vout, moments,
                                adam, partials, stop_flag, prev_partials, es_in, es_out, out,
                                cost_hist, prev_iter, stream);
}

int igm_energy_finalize(const double* partials, int32_t nbatch, int32_t Ny, int32_t Nx,
                        double* out, IgmEarlyStop* es, double* cost_hist, int32_t iter,
                        void* stream) {
  IGM_CHECK_ARG(partials && out, "null pointer");
  IGM_CHECK_ARG(nbatch >= 1 && Ny >= 2 && Nx >= 2, "bad sizes");
  IGM_CHECK_ARG(es == nullptr || nbatch == 1, "early stopping needs nbatch == 1");
  const igm::PassGeom g = igm::pass_geom(nbatch, Ny, Nx);
  hipLaunchKernelGGL(igm::energy_finalize_kernel, dim3(nbatch), dim3(igm::kFinBlock), 0,
                     (hipStream_t)stream, partials, g.blocks, out, es, cost_hist, iter);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_velocity_outputs_f32(int32_t nz, int32_t Ny, int32_t Nx, const float* U, const float* V,
                             int64_t kstride, int64_t rstride, int64_t cstride, const float* thk,
                             const double* vert_weight, float* Uo, float* Vo, float* uvelbase,
                             float* vvelbase, float* uvelsurf, float* vvelsurf, float* ubar,
                             float* vbar, void* stream) {
  return igm::velout_impl2<float>(nz, Ny, Nx, U, V, kstride, rstride, cstride, thk, vert_weight, Uo,
                                  Vo, uvelbase, vvelbase, uvelsurf, vvelsurf, ubar, vbar, stream);
}
int igm_velocity_outputs_f64(int32_t nz, int32_t Ny, int32_t Nx, const double* U,
                             const double* V, int64_t kstride, int64_t rstride, int64_t cstride,
                             const double* thk, const double* vert_weight, double* Uo, double* Vo,
                             double* uvelbase, double* vvelbase, double* uvelsurf,
                             double* vvelsurf, double* ubar, double* vbar, void* stream) {
  return igm::velout_impl2<double>(nz, Ny, Nx, U, V, kstride, rstride, cstride, thk, vert_weight,
                                   Uo, Vo, uvelbase, vvelbase, uvelsurf, vvelsurf, ubar, vbar,
                                   stream);
}

}  // extern "C"
