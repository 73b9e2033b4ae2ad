// Two iterations of IGM's energy solver in one pass on gfx950: energy,
// adjoint and legacy-Adam update, twice, with temporal blocking of the
// column march (float32, planar layout, one field).
//
// Reference: igm/processes/iceflow/solve/solve.py:45-84 -- two consecutive
// iterations of the loop, with no EarlyStopping check between them -- and
// energy/cost_shear.py:103-179, cost_gravity.py:24-69,
// cost_sliding_weertman.py:23-43.
//
// The one-iteration pass reads U, V, the Adam moments and the coefficients
// and writes U, V and the moments: about 500 B per horizontal cell at Nz 10,
// which makes an iteration HBM-bound.  This pass moves those bytes once for
// two iterations.  A wave marches up its rows with two stages one row apart.
// At step c:
//   stage A computes cell row c at the input velocity U0, which completes
//     the gradient of node row c; the first Adam update turns node row c
//     into U1 and its moments into m1, kept in registers and in LDS;
//   stage B computes cell row c-1 at U1 (node row c-1 from LDS, row c in
//     registers), which completes node row c-1 for the second iteration; the
//     second Adam update writes U2 and the moments m2 of that row.
// U1 and m1 never leave the CU.  The two stages run level by level, B one
// level behind A, as two independent instruction streams.
//
// The price is a wider halo.  U1 is only exact where stage A's gradient is
// complete, lanes 1..62 of the 64, so stage B's gradient is complete on lanes
// 2..61: a wave owns 60 columns.  Vertically, stage A starts two rows below
// the wave's first node row and ends one above its last, so a wave reads node
// rows j0-2 .. j1+1 of U0 and the moments of rows j0-1 .. j1.  Because
// neighbouring waves read moment rows that this pass writes, the moments are
// read from one buffer and written to another.
//
// Both stages run the level sweep of the one-iteration pass (LevelSweep)
// with the same operands, so a pair pass gives the same results as two
// one-iteration passes, bit for bit.
#include "energy_pass.hpp"

#include <type_traits>

namespace igm {

// step modes of energy_pair_kernel
constexpr int kOneIter = 0;   // one iteration
constexpr int kStageA = 1;    // two iterations, stage A only
constexpr int kPairStep = 2;  // two iterations, stages A and B

// Prologue of a device-driven solve pass (a.alpha_tab set), done by every
// block.  The state *es_in is the one the previous pass started from, so the
// previous pass's iteration count is recomputed from it exactly as that pass
// decided it; those iterations are then accounted in order (their COSTs from
// the previous partials, the EarlyStopping rule of iceflow/utils.py:92-111
// on a private copy).  Block 0 writes the new state to *es_out and the COSTs
// to cost_hist.  Returns the iterations this pass runs (0, 1 or 2) and its
// first iteration in *it0.
__device__ int solve_prologue(const PassArgs<float>& a, bool block0, int* it0) {
  IgmEarlyStop st = *a.es_in;
  const int f_prev = st.iters_done;
  const int c_prev = a.prev_partials != nullptr ? pass_iterations(st, f_prev, a.nbitmax) : 0;
  // The previous pass's COSTs can raise `wait` by at most one each (the
  // first only sets the best when there is none).  If even then wait + 1
  // stays below patience, the rule stops nowhere in that pass and cannot stop
  // at this pass's first iteration: the count below is pass_iterations' for
  // any COSTs.  Only block 0, which writes the state and the COST history,
  // then needs the sums; every other block starts its work at once.
  const int max_wait = st.wait + (st.has_best ? c_prev : (c_prev > 0 ? c_prev - 1 : 0));
  if (!block0 && !st.stopped && max_wait + 1 < st.patience) {
    const int f = f_prev + c_prev;
    *it0 = f;
    if (a.finalize_only || f >= a.nbitmax) return 0;
    return f + 1 < a.nbitmax ? 2 : 1;
  }
  for (int t = 0; t < c_prev; ++t) {
    double r[3];
    block_sum_partials(a.prev_partials + 3 * t, a.prev_nblk, 6, r);
    const double cost = (r[0] + r[1]) + r[2];
    es_apply(st, cost, f_prev + t);
    if (block0 && threadIdx.x == 0) {
      a.fin_out[0] = r[0];
      a.fin_out[1] = r[1];
      a.fin_out[2] = r[2];
      a.fin_out[3] = cost;
      if (a.cost_hist != nullptr) a.cost_hist[f_prev + t] = cost;
    }
  }
  if (c_prev > 0) st.passes += 1;
  if (block0 && threadIdx.x == 0) *a.es_out = st;
  *it0 = st.iters_done;
  return a.finalize_only ? 0 : pass_iterations(st, st.iters_done, a.nbitmax);
}

// Velocity layouts of a pass's input and output (template bits): planar, U
// and V in two (Nz, Ny, Nx) buffers, or interleaved, one (Nz, Ny, Nx, 2)
// buffer holding (U, V) per node and level -- one 8-B access where planar
// takes two 4-B ones.  A solve keeps its intermediate fields interleaved.
constexpr int kInInterleaved = 1;
constexpr int kOutInterleaved = 2;

template <int NZ, int LAYOUT>
__global__ void __launch_bounds__(kPassBlock)
energy_pair_kernel(const PassArgs<float> a) {
  constexpr bool kInIL = (LAYOUT & kInInterleaved) != 0;
  constexpr bool kOutIL = (LAYOUT & kOutInterleaved) != 0;
  using T = float;
  constexpr uint32_t ES = sizeof(T);
  const int lane = threadIdx.x & (kWave - 1);
  // logical block: XCD-contiguous bands of (strip, row block) tiles
  const int lblk = xcd_remap(blockIdx.x, gridDim.x);
  const int wid = lblk * kWavesPerBlock + (threadIdx.x >> 6);
  double e1_sh = 0.0, e1_sl = 0.0, e1_gr = 0.0;  // energy at U0 (first iteration)
  double e2_sh = 0.0, e2_sl = 0.0, e2_gr = 0.0;  // energy at U1 (second iteration)
  // per wave: two slots of one node row's m1 (16 B) and U1 (8 B) per level
  // and lane -- 120 KB per block at Nz 10, 144 KB at Nz 12, within gfx950's
  // 160 KB of LDS
  __shared__ float4 lds_m[kWavesPerBlock][2][NZ][kWave];
  __shared__ float2 lds_q[kWavesPerBlock][2][NZ][kWave];
  bool stopped = (a.stop != nullptr) && (*a.stop != 0);
  bool pair = true;  // two iterations (else one)
  float alpha1 = a.alpha, alpha2 = a.alpha2;
  if (a.alpha_tab != nullptr) {
    int it0;
    const int n = solve_prologue(a, lblk == 0, &it0);
    stopped = stopped || n == 0;
    pair = n == 2;
    if (n > 0) {
      alpha1 = a.alpha_tab[it0];
      alpha2 = pair ? a.alpha_tab[it0 + 1] : 0.0f;
    }
  } else if (a.prev_partials != nullptr) {
    const bool st = finalize_previous(a, lblk == 0);
    stopped = stopped || st;
  }

  if (!stopped && wid < a.nstrips * a.nrowblk) {
    const int strip = wid % a.nstrips;
    const int rblk = wid / a.nstrips;
    const int ic = strip * kPairOut - 2 + lane;
    const bool col_ok = (ic >= 0) && (ic < a.Nx);
    const bool own = (lane >= 2) && (lane <= kPairOut + 1) && col_ok;
    // node rows [j0, j1) of this wave: the Ny rows split evenly over nrowblk
    const int j0 = (int)(((int64_t)rblk * a.Ny) / a.nrowblk);
    const int j1 = (int)(((int64_t)(rblk + 1) * a.Ny) / a.nrowblk);
    // (an interleaved buffer holds twice the bytes of a planar one; its V
    // resource is unused)
    const rsrc_t rU = make_rsrc(a.U, kInIL ? 2u * a.vel_bytes : a.vel_bytes);
    const rsrc_t rV = make_rsrc(kInIL ? a.U : a.V, a.vel_bytes);
    const rsrc_t rUo = make_rsrc(a.Uo, kOutIL ? 2u * a.vel_bytes : a.vel_bytes);
    const rsrc_t rVo = make_rsrc(kOutIL ? a.Uo : a.Vo, a.vel_bytes);
    const rsrc_t rC = make_rsrc(a.coef, a.coef_bytes);
    // no input moments (a fresh optimizer): a resource of size 0, whose every
    // read is out of range and returns 0 -- the zero moments, with no bytes
    // moved; no output moments (the last pass of a solve): every store is
    // out of range and dropped
    const rsrc_t rMi = make_rsrc(a.mom, a.mom != nullptr ? a.mom_bytes : 0u);
    const rsrc_t rMo = make_rsrc(a.mom_out, a.mom_out != nullptr ? a.mom_bytes : 0u);
    const uint32_t kstep = (uint32_t)a.plane * ES;  // bytes between levels
    const uint32_t mstep = 4u * kstep;              // moments: 4 slots per node and level
    const T hh = a.hscale[0];
    const T hh2 = hh * hh;
    const T h05 = T(0.5) * hh2, h025 = T(0.25) * hh2;

    // byte offset of node (row, ic), level 0 -- or kOOB outside the grid.
    // The conditions here and below combine with & rather than &&, and both
    // offsets are computed, so that the compiler selects instead of
    // branching: a branch would split a step into blocks, and at their joins
    // the memory-counter waits become conservative (they then also wait for
    // the stores stage B issued last).
    auto node_off = [&](int row, bool extra) -> uint32_t {
      const bool ok = extra & col_ok & (row >= 0) & (row < a.Ny);
      const uint32_t o = (uint32_t)(row * a.Nx + ic) * ES;
      return ok ? o : kOOB;
    };
    // node row `row` of U0 and V0 (stage A reads rows j0-2 .. j1+1)
    auto load_row = [&](T (&u)[NZ], T (&v)[NZ], int row) {
      const uint32_t o = node_off(row, row <= j1 + 1);
#pragma unroll
      for (int k = 0; k < NZ; ++k) {
        if constexpr (kInIL) {
          const auto w = __builtin_amdgcn_raw_buffer_load_b64(rU, o == kOOB ? kOOB : 2u * o,
                                                              2u * k * kstep, 0);
          u[k] = __uint_as_float(w[0]);
          v[k] = __uint_as_float(w[1]);
        } else {
          u[k] = bload(rU, o, k * kstep, T());
          v[k] = bload(rV, o, k * kstep, T());
        }
      }
    };
    // (u, v) of node offset o (kOOB: dropped), level k, into the output
    auto store_uv = [&](T u, T v, uint32_t o, int k) {
      if constexpr (kOutIL) {
        typedef unsigned int u2 __attribute__((ext_vector_type(2)));
        u2 w;
        w[0] = __float_as_uint(u);
        w[1] = __float_as_uint(v);
        __builtin_amdgcn_raw_buffer_store_b64(w, rUo, o == kOOB ? kOOB : 2u * o, 2u * k * kstep, 0);
      } else {
        bstore(u, rUo, o, k * kstep);
        bstore(v, rVo, o, k * kstep);
      }
    };
    // coefficient planes of cell row jc (stage A reads rows j0-2 .. j1)
    const uint32_t ps = (uint32_t)a.plane * ES;
    auto load_coef = [&](T (&c)[kPlanes], int jc) {
      const bool ok = col_ok & (jc >= 0) & (jc < a.Ny) & (jc <= j1);
      const uint32_t o = (uint32_t)(jc * a.Nx + ic) * ES;
      const uint32_t co = ok ? o : kOOB;
#pragma unroll
      for (int p = 0; p < kPlanes; ++p) c[p] = bload(rC, co, p * ps, T());
    };
    // byte offset of the moments of node row `row`, level 0 (rows j0-1 .. j1
    // give U1; the others read 0)
    auto mom_off = [&](int row) -> uint32_t {
      const bool ok = col_ok & (row >= 0) & (row < a.Ny) & (row >= j0 - 1) & (row <= j1);
      const uint32_t o = (uint32_t)(row * a.Nx + ic) * (4u * ES);
      return ok ? o : kOOB;
    };

    // upper-corner adjoint contributions of the previous cell row, per stage
    T cuA[NZ], cvA[NZ], cuB[NZ], cvB[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      cuA[k] = cvA[k] = cuB[k] = cvB[k] = T(0);
    }

    // Stage B's inputs from the previous step -- U1 and the moments m1 of
    // node row c-1 -- wait in LDS.  Each wave has a private pair of slots
    // that alternate with the row parity: step c writes row c to slot c & 1
    // in its first update, and its stage B reads row c-1 from the other slot.
    // LDS accesses of one wave complete in order, so no barrier is needed.
    float4 (&sm)[2][NZ][kWave] = lds_m[threadIdx.x >> 6];
    float2 (&sq)[2][NZ][kWave] = lds_q[threadIdx.x >> 6];

    // Moments of the node row of the step's first update, in registers.  The
    // first update of level k consumes m[k] and reloads it with the next
    // row's level k, so every level's moments arrive a step ahead.
    T m[NZ][4];
    {
      const uint32_t om = mom_off(j0 - 2);
#pragma unroll
      for (int k = 0; k < NZ; ++k) mload(rMi, om, k * mstep, m[k]);
    }

    // One step c; the mode is kOneIter (one iteration: the first update is
    // the output), kStageA (two iterations, no stage B: the wave's first two
    // steps, cell rows j0-2 and j0-1, would only feed node rows below its
    // own) or kPairStep (stages A and B).  In registers on entry: U0 node rows
    // c (a_lo) and c+1 (a_hi) and the coefficients of cell rows c (k_cur) and
    // c-1 (k_prv).  The step first issues the next step's loads (vector memory
    // returns in order, so the sweeps never wait on them): the coefficients
    // of cell row c+1 and U0 of node row c+2.
    //
    // Stage A (cell row c at U0) and stage B (cell row c-1 at U1) run level
    // by level, B one level behind: A's level k emits the gradient of node
    // row c at level k-1, whose first update gives U1 of that level, which B's
    // level k-1 reads.  The two sweeps are independent streams of
    // instructions in one basic block, which the scheduler interleaves.
    auto step = [&](auto mode, int c, const T (&a_lo_u)[NZ], const T (&a_lo_v)[NZ],
                    const T (&a_hi_u)[NZ], const T (&a_hi_v)[NZ], T (&a_nx_u)[NZ],
                    T (&a_nx_v)[NZ], const T (&k_cur)[kPlanes], T (&k_nx)[kPlanes],
                    const T (&k_prv)[kPlanes]) {
      constexpr int M = decltype(mode)::value;
      load_coef(k_nx, c + 1);
      load_row(a_nx_u, a_nx_v, c + 2);
      // keep the loads above ahead of the sweeps (the scheduler would sink them)
      __builtin_amdgcn_sched_barrier(0);

      const int sn = c & 1, sp = sn ^ 1;
      const uint32_t om_next = mom_off(c + 1);
      // one iteration: U1 and m1 of node row c are the output (owned rows)
      const bool own_c = own & (c >= j0) & (c < j1);  // node and cell row c are the wave's
      const uint32_t o1 = M == kOneIter ? node_off(c, own_c) : kOOB;
      const uint32_t om1 = o1 == kOOB ? kOOB : 4u * o1;
      // stage B's output: U2 and m2 of node row c-1 (owned rows)
      const bool emit2 = own & (c - 1 >= j0);
      const uint32_t o2 = M == kPairStep ? node_off(c - 1, emit2) : kOOB;
      const uint32_t om2 = o2 == kOOB ? kOOB : 4u * o2;

      T q_u[NZ], q_v[NZ];  // U1 of node row c
      T p_u[NZ], p_v[NZ];  // U1 of node row c-1 (stage B)
      // first update of node row c at level k, from stage A's gradient
      auto first = [&](int k, T g_u, T g_v) {
        adam_step(m[k], g_u, g_v, a.omb1, a.omb2, alpha1, a.eps, a_lo_u[k], a_lo_v[k], q_u[k],
                  q_v[k]);
        if constexpr (M == kOneIter) {
          store_uv(q_u[k], q_v[k], o1, k);
          mstore(m[k], rMo, om1, k * mstep);
        } else {
          sm[sn][k][lane] = make_float4(m[k][0], m[k][1], m[k][2], m[k][3]);
          sq[sn][k][lane] = make_float2(q_u[k], q_v[k]);
        }
        mload(rMi, om_next, k * mstep, m[k]);
      };
      // second update of node row c-1 at level k, from stage B's gradient
      auto second = [&](int k, T g_u, T g_v) {
        const float4 t = sm[sp][k][lane];
        T mk[4] = {t.x, t.y, t.z, t.w};
        T un_, vn_;
        adam_step(mk, g_u, g_v, a.omb1, a.omb2, alpha2, a.eps, p_u[k], p_v[k], un_, vn_);
        store_uv(un_, vn_, o2, k);
        mstore(mk, rMo, om2, k * mstep);
      };
      auto stage_b_level = [&](LevelSweep<T, NZ>& sb, int k) {
        const float2 t = sq[sp][k][lane];
        p_u[k] = t.x;
        p_v[k] = t.y;
        sb.level(k, p_u[k], p_v[k], q_u[k], q_v[k], cuB, cvB, second);
      };

      LevelSweep<T, NZ> sa(a, hh2, h05, h025, k_cur);
      LevelSweep<T, NZ> sb(a, hh2, h05, h025, k_prv);
#pragma unroll
      for (int k = 0; k < NZ; ++k) {
        sa.level(k, a_lo_u[k], a_lo_v[k], a_hi_u[k], a_hi_v[k], cuA, cvA, first);
        if constexpr (M == kPairStep) {
          if (k >= 1) stage_b_level(sb, k - 1);
        }
      }
      const CellEnergy<T> eA = sa.finish(cuA, cvA, first);
      e1_sh += own_c ? (double)eA.sh : 0.0;
      e1_sl += own_c ? (double)eA.sl : 0.0;
      e1_gr += own_c ? (double)eA.gr : 0.0;
      if constexpr (M == kPairStep) {
        stage_b_level(sb, NZ - 1);
        const CellEnergy<T> eB = sb.finish(cuB, cvB, second);
        e2_sh += emit2 ? (double)eB.sh : 0.0;
        e2_sl += emit2 ? (double)eB.sl : 0.0;
        e2_gr += emit2 ? (double)eB.gr : 0.0;
      }
    };
    using OneIter = std::integral_constant<int, kOneIter>;
    using StageA = std::integral_constant<int, kStageA>;
    using PairStep = std::integral_constant<int, kPairStep>;

    // U0 rows and coefficient rows in rotation with period 3 (the loops are
    // unrolled three times, so the rotation costs no moves).  A wave has at
    // least two node rows (j1 >= j0 + 1), so its first two steps, c = j0-2 and
    // j0-1, always run; in a two-iteration pass they have no stage B.
    T A0u[NZ], A0v[NZ], A1u[NZ], A1v[NZ], A2u[NZ], A2v[NZ];
    T K0[kPlanes], K1[kPlanes], K2[kPlanes];
    load_row(A0u, A0v, j0 - 2);
    load_row(A1u, A1v, j0 - 1);
    load_coef(K0, j0 - 2);
    if (!pair) {
      for (int c = j0 - 2; c <= j1; c += 3) {
        step(OneIter(), c, A0u, A0v, A1u, A1v, A2u, A2v, K0, K1, K2);
        if (c + 1 > j1) break;
        step(OneIter(), c + 1, A1u, A1v, A2u, A2v, A0u, A0v, K1, K2, K0);
        if (c + 2 > j1) break;
        step(OneIter(), c + 2, A2u, A2v, A0u, A0v, A1u, A1v, K2, K0, K1);
      }
    } else {
      step(StageA(), j0 - 2, A0u, A0v, A1u, A1v, A2u, A2v, K0, K1, K2);
      step(StageA(), j0 - 1, A1u, A1v, A2u, A2v, A0u, A0v, K1, K2, K0);
      for (int c = j0; c <= j1; c += 3) {
        step(PairStep(), c, A2u, A2v, A0u, A0v, A1u, A1v, K2, K0, K1);
        if (c + 1 > j1) break;
        step(PairStep(), c + 1, A0u, A0v, A1u, A1v, A2u, A2v, K0, K1, K2);
        if (c + 2 > j1) break;
        step(PairStep(), c + 2, A1u, A1v, A2u, A2v, A0u, A0v, K1, K2, K0);
      }
    }
  }

  // deterministic block reduction: 6 doubles per block, the three energy
  // terms of each iteration
  __shared__ double red[kWavesPerBlock][6];
  double e[6] = {e1_sh, e1_sl, e1_gr, e2_sh, e2_sl, e2_gr};
#pragma unroll
  for (int t = 0; t < 6; ++t) e[t] = wave_sum(e[t]);
  if (lane == 0) {
#pragma unroll
    for (int t = 0; t < 6; ++t) red[threadIdx.x >> 6][t] = e[t];
  }
  __syncthreads();
  if (threadIdx.x == 0 && !stopped) {
    double* o = a.partials + 6 * (int64_t)lblk;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      double s = 0;
#pragma unroll
      for (int wv = 0; wv < kWavesPerBlock; ++wv) s += red[wv][t];
      o[t] = s;
    }
  }
}

int pair_impl(int nz, int Ny, int Nx, const float* U, const float* V, const float* coef,
              const float* hscale, const double* dzeta, const IgmPhysics* phys, float* Uo,
              float* Vo, const float* mom_in, float* mom_out, const IgmAdam* adam1,
              const IgmAdam* adam2, double* partials, const int32_t* stop,
              const double* prev_partials, int prev_count, const IgmEarlyStop* es_in,
              IgmEarlyStop* es_out, double* fin_out, double* cost_hist, int prev_iter,
              const float* alpha_tab, int nbitmax, int finalize_only, void* stream) {
  IGM_CHECK_ARG(U && coef && hscale && dzeta && phys && partials, "null pointer");
  IGM_CHECK_ARG(Uo && adam1 && adam2, "needs Uout and both Adam steps");
  // V (Vout) NULL: U (Uout) holds (U, V) interleaved, (Nz, Ny, Nx, 2)
  const int layout = (V == nullptr ? kInInterleaved : 0) | (Vo == nullptr ? kOutInterleaved : 0);
  IGM_CHECK_ARG(Uo != U && (V == nullptr || Uo != V) && (Vo == nullptr || (Vo != U && Vo != V)),
                "Uout/Vout must not alias U/V");
  IGM_CHECK_ARG(mom_out == nullptr || mom_out != mom_in,
                "the moments are read from one buffer and written to another");
  IGM_CHECK_ARG(Ny >= 2 && Nx >= 2, "need a grid of at least 2 x 2");
  IGM_CHECK_ARG(nz >= 2 && nz <= IGM_MAX_NZ, "nz out of range");
  IGM_CHECK_ARG(adam1->beta1 == adam2->beta1 && adam1->beta2 == adam2->beta2 &&
                    adam1->epsilon == adam2->epsilon,
                "both iterations must share beta1, beta2 and epsilon");
  const int64_t plane = (int64_t)Ny * Nx;
  const int64_t vel_bytes = plane * nz * (int64_t)sizeof(float);
  const int64_t coef_bytes = kPlanes * plane * (int64_t)sizeof(float);
  const int64_t mom_bytes = 4 * vel_bytes;
  IGM_CHECK_ARG(mom_bytes < (int64_t)kOOB && coef_bytes < (int64_t)kOOB,
                "problem too large for 32-bit buffer offsets (2 GiB per buffer)");
  PassArgs<float> a{};
  a.U = U; a.V = V; a.Uo = Uo; a.Vo = Vo;
  a.mom = const_cast<float*>(mom_in);
  a.mom_out = mom_out;
  a.coef = coef; a.hscale = hscale; a.partials = partials; a.stop = stop;
  const PassGeom g = pair_pass_geom(Ny, Nx);
  if (alpha_tab) {
    IGM_CHECK_ARG(nbitmax >= 1, "nbitmax must be >= 1");
    IGM_CHECK_ARG(es_in && es_out && fin_out && es_in != es_out,
                  "a solve pass needs es_in != es_out and out");
    IGM_CHECK_ARG(prev_partials != partials, "prev_partials must not alias partials");
    IGM_CHECK_ARG(prev_partials || !finalize_only, "finalize_only needs the previous partials");
    a.prev_nblk = g.blocks;
  } else if (prev_partials) {
    IGM_CHECK_ARG(prev_count == 1 || prev_count == 2,
                  "prev_count must be 1 (one-iteration pass) or 2 (two-iteration pass)");
    IGM_CHECK_ARG(es_in && es_out && fin_out && es_in != es_out,
                  "fused finalize needs es_in != es_out and out");
    IGM_CHECK_ARG(prev_partials != partials, "prev_partials must not alias partials");
    a.prev_count = prev_count;
    a.prev_nblk = prev_count == 1 ? one_pass_geom(1, Ny, Nx).blocks : g.blocks;
  }
  a.prev_partials = prev_partials; a.es_in = es_in; a.es_out = es_out; a.fin_out = fin_out;
  a.cost_hist = cost_hist; a.prev_iter = prev_iter;
  a.alpha_tab = alpha_tab; a.nbitmax = nbitmax; a.finalize_only = finalize_only ? 1 : 0;
  a.vel_bytes = (uint32_t)vel_bytes;
  a.coef_bytes = (uint32_t)coef_bytes;
  a.mom_bytes = (uint32_t)mom_bytes;
  a.plane = (int)plane;
  a.Ny = Ny; a.Nx = Nx;
  a.nstrips = g.nstrips; a.nrowblk = g.nrowblk; a.blocks_per_batch = g.blocks;
  if (const int rc = fill_physics(a, nz, dzeta, phys)) return rc;
  a.alpha = float(adam1->alpha);
  a.alpha2 = float(adam2->alpha);
  a.omb1 = 1.0f - float(adam1->beta1);
  a.omb2 = 1.0f - float(adam1->beta2);
  a.eps = float(adam1->epsilon);
  hipStream_t st = (hipStream_t)stream;
  switch (nz) {
#define IGM_PAIR_CASE(N)                                                                                \
  case N:                                                                                               \
    switch (layout) {                                                                                   \
      case 0: hipLaunchKernelGGL((energy_pair_kernel<N, 0>), dim3(g.blocks), dim3(kPassBlock), 0, st, a); break; \
      case 1: hipLaunchKernelGGL((energy_pair_kernel<N, 1>), dim3(g.blocks), dim3(kPassBlock), 0, st, a); break; \
      case 2: hipLaunchKernelGGL((energy_pair_kernel<N, 2>), dim3(g.blocks), dim3(kPassBlock), 0, st, a); break; \
      default: hipLaunchKernelGGL((energy_pair_kernel<N, 3>), dim3(g.blocks), dim3(kPassBlock), 0, st, a); break; \
    }                                                                                                   \
    break;
    IGM_PAIR_CASE(2)
    IGM_PAIR_CASE(3)
    IGM_PAIR_CASE(4)
    IGM_PAIR_CASE(5)
    IGM_PAIR_CASE(6)
    IGM_PAIR_CASE(8)
    IGM_PAIR_CASE(10)
    IGM_PAIR_CASE(12)
#undef IGM_PAIR_CASE
    default:
      set_error("igm_energy_pass2_f32: nz not compiled (supported: 2,3,4,5,6,8,10,12)");
      return -1;
  }
  IGM_CHECK_LAUNCH();
  return 0;
}

}  // namespace igm

extern "C" {

int igm_energy_pass2_f32(int32_t nz, int32_t Ny, int32_t Nx, const float* U, const float* V,
                         const float* coef, const float* hscale, const double* dzeta,
                         const IgmPhysics* phys, float* Uout, float* Vout,
                         const float* moments_in, float* moments_out, const IgmAdam* adam1,
                         const IgmAdam* adam2, double* partials, const int32_t* stop_flag,
                         const double* prev_partials, int32_t prev_count,
                         const IgmEarlyStop* es_in, IgmEarlyStop* es_out, double* out,
                         double* cost_hist, int32_t prev_iter, void* stream) {
  return igm::pair_impl(nz, Ny, Nx, U, V, coef, hscale, dzeta, phys, Uout, Vout, moments_in,
                        moments_out, adam1, adam2, partials, stop_flag, prev_partials, prev_count,
                        es_in, es_out, out, cost_hist, prev_iter, nullptr, 0, 0, stream);
}

int igm_energy_solve_pass_f32(int32_t nz, int32_t Ny, int32_t Nx, const float* U, const float* V,
                              const float* coef, const float* hscale, const double* dzeta,
                              const IgmPhysics* phys, float* Uout, float* Vout,
                              const float* moments_in, float* moments_out, const IgmAdam* adam,
                              const float* alpha_tab, int32_t nbitmax, double* partials,
                              const double* prev_partials, const IgmEarlyStop* es_in,
                              IgmEarlyStop* es_out, double* out, double* cost_hist,
                              int32_t finalize_only, void* stream) {
  return igm::pair_impl(nz, Ny, Nx, U, V, coef, hscale, dzeta, phys, Uout, Vout, moments_in,
                        moments_out, adam, adam, partials, nullptr, prev_partials, 0, es_in,
                        es_out, out, cost_hist, 0, alpha_tab, nbitmax, finalize_only, stream);
}

}  // extern "C"
