// Pieces shared by the energy passes of energy.hip (one solver iteration
// per launch) and energy_pair.hip (two iterations per launch): the argument
// block, buffer-resource access, the float64 partial sums with the
// EarlyStopping rule, the level sweep of one cell row and the row split.
//
// Reference: igm/processes/iceflow/energy/cost_shear.py:103-179,
// cost_gravity.py:24-69, cost_sliding_weertman.py:23-43,
// solve/solve.py:45-84, iceflow/utils.py:92-111.
#pragma once

#include "common.hpp"
#include "../../include/igm_amd.h"

#include <cmath>
#include <cstdlib>

namespace igm {

constexpr int kPlanes = 7;     // q, w, gx, gy, cs, bx, by
constexpr int kStripOut = 62;  // owned columns per wave, one iteration per pass
constexpr int kPairOut = 60;   // owned columns per wave, two iterations per pass
constexpr int kPassBlock = 256;
constexpr int kWavesPerBlock = kPassBlock / kWave;

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
  T* mom_out;     // two-iteration pass: moments written here (read from mom)
  const T* coef;
  const T* hscale;
  double* partials;
  const int32_t* stop;
  // finalize of the previous pass, done by every block at its start
  // (prev_partials != nullptr; nbatch 1): see finalize_previous
  const double* prev_partials;
  int prev_count;  // solver iterations the previous pass ran (1 or 2)
  int prev_nblk;   // blocks of the previous pass (its partials)
  // device-driven solve pass (alpha_tab != nullptr): the pass decides from
  // the EarlyStopping state how many iterations it runs; alpha_tab[i] is the
  // Adam step of iteration i < nbitmax
  const float* alpha_tab;
  int nbitmax;
  int finalize_only;  // account the previous pass, run nothing
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
  T alpha2;                  // two-iteration pass: alpha of the second iteration
};


// ------------------------------------------------------ partial sums --
//
// EarlyStopping (iceflow/utils.py:92-111): one step of the rule for the COST
// of iteration `iter`.
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

// (shear, sliding, gravity) sums of nblk per-block records p[stride * i ..
// stride * i + 2], in every thread of a 256-thread block, in a fixed order
// (strided per-thread sums, a wave butterfly, then the 4 waves in order).
// Ends with a barrier, so that a second call may reuse the shared buffer.
__device__ inline void block_sum_partials(const double* p, int nblk, int stride, double (&r)[3]) {
  __shared__ double fin[4][3];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int i = threadIdx.x; i < nblk; i += 256) {
    s0 += p[stride * i];
    s1 += p[stride * i + 1];
    s2 += p[stride * i + 2];
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
  __syncthreads();
}

// Finalize of the previous pass, done by every block of a pass at its start
// (the kernel boundary has made the previous pass's partials visible).  For
// each of the a.prev_count iterations that pass ran, in order: the sums of
// its partials give the COST (tf.add_n in energy_components order,
// solve.py:57-59) and the EarlyStopping rule is applied to a private copy of
// *es_in -- every block reaches the same state.  Block 0 writes that state
// to *es_out (never *es_in, which other blocks are reading), the energies of
// the last accounted iteration to fin_out[4] and each COST to cost_hist.
// Returns whether the state says stop.
template <typename T>
__device__ bool finalize_previous(const PassArgs<T>& a, bool block0) {
  IgmEarlyStop st = *a.es_in;
  const int stride = 3 * a.prev_count;
  for (int t = 0; t < a.prev_count; ++t) {
    double r[3];
    block_sum_partials(a.prev_partials + 3 * t, a.prev_nblk, stride, r);
    if (st.stopped) continue;  // (carried over unchanged)
    const double cost = (r[0] + r[1]) + r[2];
    es_apply(st, cost, a.prev_iter + t);
    if (block0 && threadIdx.x == 0) {
      a.fin_out[0] = r[0];
      a.fin_out[1] = r[1];
      a.fin_out[2] = r[2];
      a.fin_out[3] = cost;
      if (a.cost_hist != nullptr) a.cost_hist[a.prev_iter + t] = cost;
    }
  }
  if (block0 && threadIdx.x == 0) *a.es_out = st;
  return st.stopped != 0;
}

// Iterations a device-driven solve pass runs when the EarlyStopping state
// before it is s and its first iteration is f: none once stopped or done;
// two when the first cannot end the solve (the rule has no best COST yet,
// or one more non-improving COST leaves wait below patience) and two
// remain; else one.
__device__ __forceinline__ int pass_iterations(const IgmEarlyStop& s, int f, int nbitmax) {
  if (s.stopped || f >= nbitmax) return 0;
  const bool first_cannot_stop = !s.has_best || (s.wait + 1 < s.patience);
  return (first_cannot_stop && f + 1 < nbitmax) ? 2 : 1;
}

// ------------------------------------------------------- level sweep --
//
// Energy terms of one cell, in the working precision.
template <typename T>
struct CellEnergy {
  T sh, sl, gr;
};

// Level sweep of the cell whose lower-left node is (jc, ic): node rows jc
// (up, vp) and jc+1 (uc, vc) of all levels and the cell's coefficients cc,
// all in registers.  The right neighbour's nodes come from lane l+1 (DPP).
// For every level k, once its adjoint is complete, emit(k, g_u, g_v) gets
// dJ/dU and dJ/dV of node (jc, ic): this cell's lower-left corner, the
// lower-right corner of the cell of lane l-1 (DPP), and the carry (cu, cv)
// of cell row jc-1, which the call replaces by this row's upper-corner
// contributions.  Layer l = k-1 is processed once level k is known, so only
// two levels of derivatives are live.  Horizontal derivatives are kept in
// units of 1/(4 dx) (ux = 4 dx dU/dx at the level), the vertical ones
// likewise (q holds damp*dx/Hc); hh2 = 1/(4 dx)^2 restores the strain rate
// where it is needed, h05 = hh2/2, h025 = hh2/4.
//
// LevelSweep runs the sweep one level at a time: level(k, ...) for k = 0 ..
// NZ-1 in order, then finish().  level(k) emits level k-1 (k >= 1) and
// finish() emits level NZ-1, so a caller can interleave two sweeps whose
// inputs depend on each other's outputs (the two-iteration pass).
template <typename T, int NZ>
struct LevelSweep {
  const PassArgs<T>& a;
  const T hh2, h05, h025;
  const T q, w, gx, gy, csl, bx, by, whh;
  T p_ux = 0, p_uy = 0, p_vx = 0, p_vy = 0, p_su = 0, p_sv = 0;  // level k-1
  T g_ux = 0, g_uy = 0, g_vx = 0, g_vy = 0, g_su = 0, g_sv = 0;  // its adjoint
  T c_sh = 0, c_sl = 0, wsu = 0, wsv = 0;

  __device__ __forceinline__ LevelSweep(const PassArgs<T>& a_, T hh2_, T h05_, T h025_,
                                        const T (&cc)[kPlanes])
      : a(a_), hh2(hh2_), h05(h05_), h025(h025_), q(cc[0]), w(cc[1]), gx(cc[2]), gy(cc[3]),
        csl(cc[4]), bx(cc[5]), by(cc[6]), whh(cc[1] * hh2_) {}

  // Adjoint scatter of one level to the 4 corner nodes.  With A = gs - gx
  // and B = gs + gx, the corners get A -+ gy (left) and B -+ gy (right); the
  // right ones arrive from lane l-1: node row jc receives X - Z, node row
  // jc+1 (carried) X + Z, X = A + B(l-1), Z = gy + gy(l-1).
  __device__ __forceinline__ static T scatter(T& carry, T gx_, T gy_, T gs_) {
#pragma clang fp contract(on)
    const T A = gs_ - gx_, B = gs_ + gx_;
    const T X = A + lane_from_left(B);
    const T Z = gy_ + lane_from_left(gy_);
    const T g = (carry - Z) + X;
    carry = X + Z;
    return g;
  }

  template <typename Emit>
  __device__ __forceinline__ void out(int k, T (&cu)[NZ], T (&cv)[NZ], Emit&& emit) {
    const T g_u = scatter(cu[k], g_ux, g_uy, g_su);
    const T g_v = scatter(cv[k], g_vx, g_vy, g_sv);
    emit(k, g_u, g_v);
  }

  // Level k from the nodes (upk, vpk) of row jc and (uck, vck) of row jc+1.
  // Multiply-adds are fused within one expression only, never across
  // statements: the fusion then follows the source, not the surrounding
  // kernel, and every kernel that runs this sweep rounds the same way.
  template <typename Emit>
  __device__ __forceinline__ void level(int k, T upk, T vpk, T uck, T vck, T (&cu)[NZ],
                                        T (&cv)[NZ], Emit&& emit) {
#pragma clang fp contract(on)
    const T su = upk + uck, du = uck - upk;
    const T sv = vpk + vck, dv = vck - vpk;
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
      out(k - 1, cu, cv, emit);
    }
    p_ux = ux; p_uy = uy; p_vx = vx; p_vy = vy; p_su = sU; p_sv = sV;
    g_ux = n_ux; g_uy = n_uy; g_vx = n_vx; g_vy = n_vy; g_su = n_su; g_sv = n_sv;
  }

  // Emits level NZ-1 and returns the cell's (shear, sliding, gravity) energy.
  template <typename Emit>
  __device__ __forceinline__ CellEnergy<T> finish(T (&cu)[NZ], T (&cv)[NZ], Emit&& emit) {
#pragma clang fp contract(on)
    out(NZ - 1, cu, cv, emit);
    return {w * c_sh, c_sl, gx * wsu + gy * wsv};
  }
};

template <typename T, int NZ, typename Emit>
__device__ __forceinline__ CellEnergy<T> level_sweep(const PassArgs<T>& a, T hh2, T h05, T h025,
                                                    const T (&up)[NZ], const T (&vp)[NZ],
                                                    const T (&uc)[NZ], const T (&vc)[NZ],
                                                    const T (&cc)[kPlanes], T (&cu)[NZ],
                                                    T (&cv)[NZ], Emit&& emit) {
  LevelSweep<T, NZ> sw(a, hh2, h05, h025, cc);
#pragma unroll
  for (int k = 0; k < NZ; ++k) sw.level(k, up[k], vp[k], uc[k], vc[k], cu, cv, emit);
  return sw.finish(cu, cv, emit);
}

// One legacy-Adam step (solve.py:84, TF's legacy Adam with alpha carrying
// the bias correction) of the slots m = (m_U, v_U, m_V, v_V) from the
// gradient (g_u, g_v): updates m and returns the stepped velocities of the
// node, u - alpha m_U / (sqrt(v_U) + eps) and likewise for v.  Fused
// multiply-adds follow the source, as in level_sweep.
template <typename T>
__device__ __forceinline__ void adam_step(T (&m)[4], T g_u, T g_v, T omb1, T omb2, T alpha, T eps,
                                          T u, T v, T& u_new, T& v_new) {
#pragma clang fp contract(on)
  m[0] += (g_u - m[0]) * omb1;
  m[1] += (g_u * g_u - m[1]) * omb2;
  m[2] += (g_v - m[2]) * omb1;
  m[3] += (g_v * g_v - m[3]) * omb2;
  u_new = u - (m[0] * alpha) * rcp(sqrt_(m[1]) + eps);
  v_new = v - (m[2] * alpha) * rcp(sqrt_(m[3]) + eps);
}

// ------------------------------------------------------------ row split --
//
// Waves per launch the row split aims at: one full round of MI355X's 256
// CUs.  The two-iteration pass holds 256 registers per lane, one wave per
// SIMD: 1024 waves (at 1000 x 1000, 119.5 us per launch against 124.0 at
// 896 and 193.2 at 1088, where a second round starts).  The one-iteration
// passes get 2048: two waves per SIMD for the float energy and energy +
// adjoint passes (164 and 179 registers; energy + adjoint 42.3 against
// 45.4 us at 1024 and 50.5 at 3072; the fused Adam pass 101.4 against
// 103.8 us).  One geometry serves every one-iteration pass, whatever its
// dtype and mode, because the finalize and the workspace size are derived
// from it.  IGM_ENERGY_WAVES and IGM_PAIR_WAVES override them (for tuning
// sweeps).
constexpr int kTargetWaves = 2048;
constexpr int kPairTargetWaves = 1024;
constexpr int kMinRows = 4;  // fewer rows per wave would be mostly halo

inline int env_waves(const char* name, int dflt) {
  const char* e = getenv(name);
  const int r = e ? atoi(e) : dflt;
  return (r >= 64 && r <= (1 << 20)) ? r : dflt;
}
inline int target_waves() {
  static int n = env_waves("IGM_ENERGY_WAVES", kTargetWaves);
  return n;
}
inline int pair_target_waves() {
  static int n = env_waves("IGM_PAIR_WAVES", kPairTargetWaves);
  return n;
}

struct PassGeom {
  int nstrips, nrowblk, blocks;
};
// strips of `strip_out` owned columns; the rows of each strip split evenly
inline PassGeom pass_geom(int nbatch, int Ny, int Nx, int strip_out, int target) {
  PassGeom g;
  g.nstrips = (Nx + strip_out - 1) / strip_out;
  const int64_t per = (int64_t)g.nstrips * nbatch;
  int64_t want = (target + per / 2) / per;
  const int64_t most = Ny / kMinRows > 1 ? Ny / kMinRows : 1;
  g.nrowblk = (int)(want < 1 ? 1 : (want > most ? most : want));
  const int waves = g.nstrips * g.nrowblk;
  g.blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  return g;
}
// the geometry of the one-iteration passes, and of the two-iteration pass
inline PassGeom one_pass_geom(int nbatch, int Ny, int Nx) {
  return pass_geom(nbatch, Ny, Nx, kStripOut, target_waves());
}
inline PassGeom pair_pass_geom(int Ny, int Nx) { return pass_geom(1, Ny, Nx, kPairOut, pair_target_waves()); }

// Fills the physics and vertical-grid fields of the argument block (host).
template <typename T>
int fill_physics(PassArgs<T>& a, int nz, const double* dzeta, const IgmPhysics* phys) {
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
  return 0;
}

}  // namespace igm
