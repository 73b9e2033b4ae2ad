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
#include "energy_pass.hpp"

#include <mutex>

namespace igm {

// four dwords, the operand of the 16-byte buffer loads and stores
typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));

namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

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

// ------------------------------------------------------------- finalize --
//
// Sum of one batch entry's per-block partials (block_sum_partials) into
// out[4] = {mean shear, mean sliding, mean gravity, COST} for each of the
// `count` iterations of the pass that wrote them (records of 3 * count
// doubles per block), then the EarlyStopping rule (iceflow/utils.py:92-111)
// when es is given.  Run by one 256-thread block.  The solver folds it into
// the next iteration's pass instead (finalize_previous): a "last block sums"
// fold inside the same pass would need a device-scope release per block,
// which on the multi-XCD MI355X writes back the whole L2 (measured 18 us per
// iteration).
__device__ void finalize_entry(const double* p, int nblk, int count, double* out,
                               IgmEarlyStop* es, double* cost_hist, int iter) {
  IgmEarlyStop st{};
  if (es != nullptr) st = *es;
  for (int t = 0; t < count; ++t) {
    double r[3];
    block_sum_partials(p + 3 * t, nblk, 3 * count, r);
    if (threadIdx.x == 0) {
      const double cost = (r[0] + r[1]) + r[2];  // tf.add_n in energy_components order
      out[0] = r[0];
      out[1] = r[1];
      out[2] = r[2];
      out[3] = cost;
      if (es != nullptr && !st.stopped) {
        if (cost_hist != nullptr) cost_hist[iter + t] = cost;
        es_apply(st, cost, iter + t);
      }
    }
  }
  if (threadIdx.x == 0 && es != nullptr) *es = st;
}

// INTER = 0: planar (Nz, Ny, Nx) with separate U, V buffers;
// INTER = 1: node-interleaved Y (Ny, Nx, 2Nz), V = Y + Nz, one level per access;
// INTER = 2: the same layout, float32, even Nz, 16-B aligned: a node's 2 Nz
//   values (80 B at Nz 10) move as Nz/2 16-B accesses per lane, and mode 1
//   stores the gradient of a node after its level sweep the same way (the
//   one-level accesses of a 64-lane wave touch 64 cache lines 80 B apart).
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
    const bool st = finalize_previous(a, lblk == 0);
    stopped = stopped || st;
  }
  constexpr uint32_t ES = sizeof(T);
  constexpr bool kInter = INTER != 0;
  constexpr bool kVecNode = INTER == 2 && NZ % 2 == 0 && sizeof(T) == 4;

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
    const uint32_t kstep = kInter ? ES : (uint32_t)a.plane * ES;   // bytes between levels
    const T hh = a.hscale[b];
    const T hh2 = hh * hh;
    const T h05 = T(0.5) * hh2, h025 = T(0.25) * hh2;

    // byte offset of node (row, ic), level 0 -- or kOOB outside the grid
    auto node_off = [&](int row, bool extra) -> uint32_t {
      const bool ok = extra && col_ok && row >= 0 && row < a.Ny;
      const uint32_t e = (uint32_t)(row * a.Nx + ic) * (kInter ? 2u * NZ : 1u);
      return ok ? bbase + e * ES : kOOB;
    };
    // node row `row` of U and V (rows this wave never uses -- past j1 --
    // read nothing: the last step's prefetch would be wasted bytes)
    auto load_row = [&](T (&u)[NZ], T (&v)[NZ], int row) {
      const uint32_t o = node_off(row, row <= j1);
      if constexpr (kVecNode) {
        // U levels then V levels of the node: V = U + Nz, all through rU
        float f[2 * NZ];
#pragma unroll
        for (int i = 0; i < NZ / 2; ++i) {
          const auto x = __builtin_amdgcn_raw_buffer_load_b128(rU, o, 16 * i, 0);
#pragma unroll
          for (int e = 0; e < 4; ++e) f[4 * i + e] = __uint_as_float(x[e]);
        }
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
          u[k] = f[k];
          v[k] = f[NZ + k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
          u[k] = bload(rU, o, k * kstep, T());
          v[k] = bload(rV, o, k * kstep, T());
        }
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

      // legacy-Adam update of node (jc, ic), level k, from its gradient
      auto adam = [&](int k, T g_u, T g_v, T (&m)[4]) {
        T un_, vn_;
        adam_step(m, g_u, g_v, a.omb1, a.omb2, a.alpha, a.eps, up[k], vp[k], un_, vn_);
        bstore(un_, make_rsrc(a.Uo, a.vel_bytes), on, k * kstep);
        bstore(vn_, make_rsrc(a.Vo, a.vel_bytes), on, k * kstep);
        mstore(m, rM, om, k * mstep);
      };
      // output of node (jc, ic), level k (masked stores)
      auto emit = [&](int k, T g_u, T g_v) {
        if constexpr (MODE == 1 && kVecNode) {
          gu_[k] = g_u;  // stored as whole nodes after the sweep
          gv_[k] = g_v;
        } else if constexpr (MODE == 1) {
          bstore(g_u, make_rsrc(a.gU, a.vel_bytes), on, k * kstep);
          bstore(g_v, make_rsrc(a.gV, a.vel_bytes), on, k * kstep);
        } else if constexpr (kDeferAdam) {
          gu_[k] = g_u;
          gv_[k] = g_v;
        } else if constexpr (MODE == 2) {
          T m[4];
          mload(rM, om, k * mstep, m);
          adam(k, g_u, g_v, m);
        }
      };
      const CellEnergy<T> e = level_sweep<T, NZ>(a, hh2, h05, h025, up, vp, uc, vc, cc, cu, cv, emit);
      if constexpr (MODE == 1 && kVecNode) {
        // the node's gradient: gU levels then gV levels (gV = gU + Nz)
        const rsrc_t rG = make_rsrc(a.gU, a.vel_bytes);
        float f[2 * NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
          f[k] = gu_[k];
          f[NZ + k] = gv_[k];
        }
#pragma unroll
        for (int i = 0; i < NZ / 2; ++i) {
          const v4u32 x = {__float_as_uint(f[4 * i]), __float_as_uint(f[4 * i + 1]),
                           __float_as_uint(f[4 * i + 2]), __float_as_uint(f[4 * i + 3])};
          __builtin_amdgcn_raw_buffer_store_b128(x, rG, on, 16 * i, 0);
        }
      }

      if constexpr (kDeferAdam) {
        // keep the updates after the sweep (the scheduler would hoist them
        // into it, right after each gradient, and wait there for the moments)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < NZ; ++k) adam(k, gu_[k], gv_[k], mo[k]);
      }

      if (emit_row) {
        e_sh += (double)e.sh;
        e_sl += (double)e.sl;
        e_gr += (double)e.gr;
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

// ------------------------------------------------------------- finalize --

constexpr int kFinBlock = 256;

__global__ void __launch_bounds__(kFinBlock)
energy_finalize_kernel(const double* __restrict__ partials, int blocks_per_batch, int count,
                       double* __restrict__ out, IgmEarlyStop* es, double* cost_hist, int iter) {
  const int b = blockIdx.x;
  if (es != nullptr && es->stopped) return;
  finalize_entry(partials + 3 * count * (int64_t)b * blocks_per_batch, blocks_per_batch, count,
                 out + 4 * b, es, cost_hist, iter);
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
  // interleaved nodes as 16-B accesses: float32, even Nz (a node is a whole
  // number of 16-B chunks), Y (and gY, with gV = gU + Nz) 16-B aligned and a
  // batch stride of whole 16-B chunks
  if (inter == 1 && sizeof(T) == 4 && nz % 2 == 0 && (reinterpret_cast<uintptr_t>(U) & 15) == 0 &&
      (nbatch == 1 || (bs * (int64_t)sizeof(T)) % 16 == 0) &&
      (mode != 1 || ((reinterpret_cast<uintptr_t>(gU) & 15) == 0 && gV == gU + nz)))
    inter = 2;
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
  const PassGeom g = one_pass_geom(nbatch, Ny, Nx);
  a.nstrips = g.nstrips; a.nrowblk = g.nrowblk; a.blocks_per_batch = g.blocks;
  a.prev_count = 1;
  a.prev_nblk = g.blocks;
  if (const int rc = fill_physics(a, nz, dzeta, phys)) return rc;
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
  if constexpr (sizeof(T) == 4) {
    IGM_MODE_LAYOUT(0, 2) IGM_MODE_LAYOUT(1, 2)
  }
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
// ------------------------------------------- allocation-free evaluation --
//
// The per-block partial sums of igm_energy_eval_f32_noalloc live in this
// static device array instead of a caller workspace: 65,536 doubles
// (512 KiB), held by every GPU the library runs on until the process ends.
// Host code never passes the array's name to a kernel: it looks up the
// array's device address once per device (hipGetSymbolAddress, with that
// device current) and passes that.  The pass overwrites every partial it
// reads back, so nothing has to start at zero; calls that overlap on one
// GPU race for the array.
constexpr int64_t kScratchDoubles = 65536;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device (nullptr + error set on failure).
double* scratch_on_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("igm_energy_eval_f32_noalloc: no current HIP device (or more than 64)");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error(std::string("igm_energy_eval_f32_noalloc: hipGetSymbolAddress failed: ") +
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace igm

extern "C" {

const char* igm_last_error(void) { return igm::last_error(); }

int igm_energy_eval_f32_noalloc(int32_t mode, int32_t nz, int32_t nbatch, int32_t Ny, int32_t Nx,
                                const float* U, const float* V, int64_t bstride, int64_t kstride,
                                int64_t rstride, int64_t cstride, const float* coef,
                                const float* hscale, const double* dzeta, const IgmPhysics* phys,
                                float* gU, float* gV, double* out, void* stream) {
  IGM_CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 (energy) or 1 (energy and gradient)");
  IGM_CHECK_ARG(out != nullptr, "null pointer");
  IGM_CHECK_ARG(nbatch >= 1 && nbatch <= 65535 && Ny >= 2 && Nx >= 2, "bad sizes");
  const igm::PassGeom g = igm::one_pass_geom(nbatch, Ny, Nx);
  const int64_t need = 3 * (int64_t)nbatch * g.blocks;
  if (need > igm::kScratchDoubles) {
    igm::set_error("igm_energy_eval_f32_noalloc: this batch needs " + std::to_string(need) +
                   " partial sums, more than the " + std::to_string(igm::kScratchDoubles) +
                   " of the static scratch (use igm_energy_pass_f32 with a workspace)");
    return -1;
  }
  double* part = igm::scratch_on_current_device();
  if (part == nullptr) return -1;
  const int rc = igm::pass_impl<float>(mode, nz, nbatch, Ny, Nx, U, V, bstride, kstride, rstride,
                                       cstride, coef, hscale, dzeta, phys, gU, gV, nullptr, nullptr,
                                       nullptr, nullptr, part, nullptr, nullptr, nullptr, nullptr,
                                       nullptr, nullptr, 0, stream);
  if (rc != 0) return rc;
  hipLaunchKernelGGL(igm::energy_finalize_kernel, dim3(nbatch), dim3(igm::kFinBlock), 0,
                     (hipStream_t)stream, part, g.blocks, 1, out, nullptr, nullptr, 0);
  IGM_CHECK_LAUNCH();
  return 0;
}

int64_t igm_energy_workspace_doubles(int32_t nbatch, int32_t Ny, int32_t Nx) {
  if (nbatch < 1 || Ny < 2 || Nx < 2) return 0;
  // room for either pass: 3 doubles per block of the one-iteration pass,
  // 6 per block of the two-iteration pass (one field)
  const igm::PassGeom g = igm::one_pass_geom(nbatch, Ny, Nx);
  const igm::PassGeom g2 = igm::pair_pass_geom(Ny, Nx);
  const int64_t one = 3 * (int64_t)nbatch * g.blocks, two = 6 * (int64_t)g2.blocks;
  return one > two ? one : two;
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
  const igm::PassGeom g = igm::one_pass_geom(nbatch, Ny, Nx);
  hipLaunchKernelGGL(igm::energy_finalize_kernel, dim3(nbatch), dim3(igm::kFinBlock), 0,
                     (hipStream_t)stream, partials, g.blocks, 1, out, es, cost_hist, iter);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_energy_finalize_pair(const double* partials, int32_t Ny, int32_t Nx, double* out,
                             IgmEarlyStop* es, double* cost_hist, int32_t iter, void* stream) {
  IGM_CHECK_ARG(partials && out, "null pointer");
  IGM_CHECK_ARG(Ny >= 2 && Nx >= 2, "bad sizes");
  const igm::PassGeom g = igm::pair_pass_geom(Ny, Nx);
  hipLaunchKernelGGL(igm::energy_finalize_kernel, dim3(1), dim3(igm::kFinBlock), 0,
                     (hipStream_t)stream, partials, g.blocks, 2, out, es, cost_hist, iter);
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
