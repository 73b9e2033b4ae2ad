// Convolutional emulator of IGM's ice flow on gfx950: forward and backward
// kernels of the "cnn" network (emulate/neural_network.py:9-66):
//   16 x [Conv2D 3x3 "same" + bias -> LeakyReLU(0.01)] -> Conv2D 1x1 + bias.
//
// Data layout: activations are NHWC bf16 (B, H, W, 32); the network output Y
// is NHWC fp32 (B, H, W, 2Nz).  Per-batch (ny, nx) extents mask padded
// batches: every kernel writes zeros outside a batch entry's extent, which
// makes a padded batch behave exactly like per-entry "same" zero padding.
//
// Kernels:
//  conv_in_kernel       layer 1 (5 -> 32), fp32 VALU (raw inputs such as the
//                       surface elevation need fp32: bf16 spacing is 16 m at
//                       2500-3000 m), bias + LeakyReLU, bf16 out.
//  conv3x3_mfma_kernel  32 -> 32 3x3 layers on v_mfma_f32_16x16x32_bf16:
//                       forward (bias + LeakyReLU) and dgrad (same kernel,
//                       flipped/transposed weights, output gated by the
//                       LeakyReLU derivative of the layer below).  Input
//                       rows are staged in LDS (pixel stride 80 B: the
//                       16-B fragment reads of 8 lanes hit distinct banks).
//  head_mfma_kernel     1x1 32 -> 2Nz (<= 32) output layer, fp32 out.
//  wgrad_mfma_kernel    weight gradient: K = pixels; tiles are transposed
//                       into LDS as [channel][pixel] (three x-shifted copies
//                       of the activation rows keep every fragment read
//                       16-B aligned); per-block partial sums, then a
//                       deterministic two-stage reduction
//                       (partials_group_sum_kernel, wgrad_reduce_kernel).
//  head_bwd_kernel      dY -> dA16 * LeakyReLU'(A16) (fp32 VALU) and a padded
//                       bf16 copy of dY for the head's wgrad.
//  conv_in_wgrad_kernel layer-1 weight gradient, fp32 VALU.
//  Adam: per-tensor clip_by_norm + legacy Adam on the fp32 master weights,
//  then bf16 packing of the MFMA operands.
#include "common.hpp"

#include <cstdlib>
#include "../../include/igm_amd.h"

namespace igm {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kC = 32;        // hidden channels
constexpr int kCin = 5;       // input fields
constexpr int kTileW = 64;    // pixels per tile row
constexpr int kTileH = 4;     // output rows per forward tile (one per wave)
constexpr int kPxStride = 40; // LDS pixel stride in bf16 (80 B)
constexpr int kChStride = 72; // LDS channel-row stride in bf16 (144 B) for wgrad

__device__ __forceinline__ float leaky(float x) { return x > 0.f ? x : 0.01f * x; }
__device__ __forceinline__ float bf2f(bf16 h) { return (float)h; }
__device__ __forceinline__ bf16 f2bf(float f) { return (bf16)f; }

__device__ __forceinline__ void extent_of(const int32_t* ext, int b, int H, int W, int& ny, int& nx) {
  ny = ext ? min(ext[2 * b], H) : H;
  nx = ext ? min(ext[2 * b + 1], W) : W;
}

// ------------------------------------------------------------ layer 1 --
// One thread per output pixel.  The 45 inputs of its 3x3 window are loaded
// up front; the weights (HWIO [3][3][5][32] fp32) are wave-uniform, so they
// come through scalar loads (read-only, no LDS) as the FMA operand.  Each
// output channel accumulates bias, then taps 0..8, then fields 0..4.
__global__ void __launch_bounds__(256)
conv_in_kernel(const float* __restrict__ f0, const float* __restrict__ f1,
               const float* __restrict__ f2, const float* __restrict__ f3,
               const float* __restrict__ f4, int64_t fb, int64_t fr, int64_t fc,
               const int32_t* __restrict__ ext, int H, int W, const float* __restrict__ w,
               const float* __restrict__ bias, bf16* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const int b = blockIdx.z;
  if (x >= W) return;
  int ny, nx;
  extent_of(ext, b, H, W, ny, nx);
  bf16* o = out + (((int64_t)b * H + y) * W + x) * kC;
  if (y >= ny || x >= nx) {
    uint4 z = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) reinterpret_cast<uint4*>(o)[q] = z;
    return;
  }
  const float* fs[kCin] = {f0, f1, f2, f3, f4};
  float v[9][kCin];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    const bool in = yy >= 0 && yy < ny && xx >= 0 && xx < nx;
    const int64_t off = (int64_t)b * fb + (int64_t)yy * fr + (int64_t)xx * fc;
#pragma unroll
    for (int ci = 0; ci < kCin; ++ci) v[t][ci] = in ? fs[ci][off] : 0.f;
  }
  float acc[kC];
#pragma unroll
  for (int c = 0; c < kC; ++c) acc[c] = bias[c];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int ci = 0; ci < kCin; ++ci) {
      const float* wr = w + (t * kCin + ci) * kC;
#pragma unroll
      for (int c = 0; c < kC; ++c) acc[c] += v[t][ci] * wr[c];
    }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    bf16x8 h;
#pragma unroll
    for (int e = 0; e < 8; ++e) h[e] = f2bf(leaky(acc[q * 8 + e]));
    reinterpret_cast<bf16x8*>(o)[q] = h;
  }
}

// Gradient of the network with respect to its 5 input fields, through
// layer 1: the transpose of conv_in_kernel's cross-correlation,
//   dX[b, y, x, ci] = sum_{r, s, co} W[r][s][ci][co] * dz1[b, y - r + 1, x - s + 1, co],
// zero padding, fp32 accumulation in a fixed order; dz1 (bf16 NHWC, 32
// channels) is the gradient at layer 1's pre-activation output.  Zeros
// outside an entry's extent.
__global__ void __launch_bounds__(256)
conv_in_dgrad_kernel(const bf16* __restrict__ dz1, const int32_t* __restrict__ ext, int H, int W,
                     const float* __restrict__ w, float* __restrict__ dX) {
  __shared__ float ws[9 * kCin * kC];
  for (int i = threadIdx.x; i < 9 * kCin * kC; i += blockDim.x) ws[i] = w[i];
  __syncthreads();
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const int b = blockIdx.z;
  if (x >= W) return;
  int ny, nx;
  extent_of(ext, b, H, W, ny, nx);
  float* o = dX + (((int64_t)b * H + y) * W + x) * kCin;
  float acc[kCin];
#pragma unroll
  for (int ci = 0; ci < kCin; ++ci) acc[ci] = 0.f;
  if (y < ny && x < nx) {
#pragma unroll 1
    for (int r = 0; r < 3; ++r) {
      const int yy = y - r + 1;
#pragma unroll 1
      for (int s = 0; s < 3; ++s) {
        const int xx = x - s + 1;
        if (yy < 0 || yy >= ny || xx < 0 || xx >= nx) continue;
        const uint4* src = reinterpret_cast<const uint4*>(dz1 + (((int64_t)b * H + yy) * W + xx) * kC);
        const float* wr = ws + (r * 3 + s) * kCin * kC;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint4 u = src[q];
          const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int co = 8 * q + e;
            const uint32_t bits = (e & 1) ? (wv[e >> 1] & 0xFFFF0000u) : (wv[e >> 1] << 16);
            const float d = __uint_as_float(bits);
#pragma unroll
            for (int ci = 0; ci < kCin; ++ci) acc[ci] += d * wr[ci * kC + co];
          }
        }
      }
    }
  }
#pragma unroll
  for (int ci = 0; ci < kCin; ++ci) o[ci] = acc[ci];
}

// ------------------------------------------------ 3x3 MFMA conv (fwd/dgrad) --
// wt: [9][32 out][32 in] bf16.  MODE 0: out = LeakyReLU(conv + bias).
// MODE 1 (dgrad): out = conv * (aux > 0 ? 1 : 0.01), aux = activation of the
// layer below (its sign is the sign of its pre-activation).
// Block = 4 waves; tile = 4 output rows x 64 pixels; wave w owns row y0 + w.
// Persistent: logical block k walks tiles [k*per, (k+1)*per) of the tile
// list ordered with the column-tile index fastest; logical blocks are laid
// out XCD by XCD (xcd_remap), so each XCD sweeps a band of tile rows and the
// halo rows and columns of neighbouring tiles meet in its L2.  The 18 weight
// fragments are loaded once per block.
template <int MODE>
__global__ void __launch_bounds__(256)
conv3x3_mfma_kernel(const bf16* __restrict__ in, const bf16* __restrict__ wt,
                    const float* __restrict__ bias, const bf16* __restrict__ aux,
                    const int32_t* __restrict__ ext, int B, int H, int W, int per,
                    bf16* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) bf16 tile[kTileH + 2][kTileW + 2][kPxStride];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nty = (H + kTileH - 1) / kTileH;
  const int ntx = (W + kTileW - 1) / kTileW;
  // (32-bit tile indices: the host keeps the activations below 2 GiB, so
  // there are fewer than 2^25 tiles)
  const int ntiles = B * ntx * nty;
  // XCD-contiguous runs of tiles: vertically adjacent tiles (ntx apart in
  // the x-fastest order) then share one L2 for their halo rows
  const int t0 = xcd_remap(blockIdx.x, gridDim.x) * per;
  if (t0 >= ntiles) return;
  const int t1 = t0 + per < ntiles ? t0 + per : ntiles;

  // weight fragments: A operand (16 out-channels x 32 in-channels per tap)
  bf16x8 wf[9][2];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
      wf[t][cb] = *reinterpret_cast<const bf16x8*>(wt + ((t * kC + cb * 16 + (lane & 15)) * kC + 8 * (lane >> 4)));
  float bias_r[2][4];
  if constexpr (MODE == 0) {
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int e = 0; e < 4; ++e) bias_r[cb][e] = bias[cb * 16 + 4 * (lane >> 4) + e];
  }

  // Staging of a tile's input rows y0-1 .. y0+4, pixels x0-1 .. x0+64 (zero
  // outside the entry's extent): kChunks 16-B chunks, kPerThread per thread,
  // all loads issued before any is used.  The next tile's loads are issued
  // before the current tile's MFMAs, so they are in flight during them.
  constexpr int kChunks = (kTileH + 2) * (kTileW + 2) * 4;
  constexpr int kPerThread = (kChunks + 255) / 256;
  uint4 stage[kPerThread];
  struct Tile {
    int b, y0, x0, ny, nx;
    bool valid;
  };
  auto tile_of = [&](int ti) {
    Tile t;
    t.x0 = (ti % ntx) * kTileW;
    t.y0 = ((ti / ntx) % nty) * kTileH;
    t.b = ti / (nty * ntx);
    extent_of(ext, t.b, H, W, t.ny, t.nx);
    t.valid = t.y0 < t.ny && t.x0 < t.nx;
    return t;
  };
  // branch-free staging loads of tile ti (all masked when ti >= t1 or the
  // tile is padding): the destination registers stay fixed across the loop
  const rsrc_t rin = make_rsrc(in, (uint32_t)((int64_t)B * H * W * kC * 2));
  auto issue_loads = [&](const Tile& t, bool live) {
    const int64_t img = (int64_t)t.b * H * W;
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int c = tid + 256 * k;
      const int chunk = c & 3;
      const int px = (c >> 2) % (kTileW + 2);
      const int row = (c >> 2) / (kTileW + 2);
      const int yy = t.y0 - 1 + row, xx = t.x0 - 1 + px;
      const bool ok = live && t.valid && c < kChunks && yy >= 0 && yy < t.ny && xx >= 0 && xx < t.nx;
      const uint32_t off = ok ? (uint32_t)(((img + (int64_t)yy * W + xx) * kC + chunk * 8) * 2) : kOOB;
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(rin, off, 0, 0);
      stage[k] = uint4{v[0], v[1], v[2], v[3]};
    }
  };

  Tile cur = tile_of(t0);
  issue_loads(cur, true);
  for (int ti = t0; ti < t1; ++ti) {
    const int64_t img = (int64_t)cur.b * H * W;
    const int y = cur.y0 + wave;
    const int x0 = cur.x0;
    const int ny = cur.ny, nx = cur.nx;
    if (!cur.valid) {  // tile outside the valid region: zeros (block-uniform)
      cur = tile_of(ti + 1 < t1 ? ti + 1 : ti);
      issue_loads(cur, ti + 1 < t1);
      if (y < H) {
        const int x = x0 + lane;
        if (x < W) {
          uint4 z = {0, 0, 0, 0};
          uint4* o = reinterpret_cast<uint4*>(out + ((img + (int64_t)y * W + x) * kC));
#pragma unroll
          for (int q = 0; q < 4; ++q) o[q] = z;
        }
      }
      continue;
    }
    // (the compute below is reached only through this path, after the wait
    // for the staged loads, so the compiler places no memory waits in it)
    // MODE 1: the epilogue's activation signs are loaded here, before the
    // barrier and so before the next tile's staging loads.  Loads complete
    // in issue order, so the epilogue's wait for them leaves the staging
    // loads in flight (issued after them, it would wait for those too).
    uint2 auxv[4][2];
    if constexpr (MODE == 1) {
      const rsrc_t raux = make_rsrc(aux, (uint32_t)((int64_t)B * H * W * kC * 2));
#pragma unroll
      for (int pb = 0; pb < 4; ++pb) {
        const int x = x0 + pb * 16 + (lane & 15);
        const bool ok = y < H && x < W;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          const int co = cb * 16 + 4 * (lane >> 4);
          const uint32_t off = ok ? (uint32_t)(((img + (int64_t)y * W + x) * kC + co) * 2) : kOOB;
          const auto v = __builtin_amdgcn_raw_buffer_load_b64(raux, off, 0, 0);
          auxv[pb][cb] = uint2{v[0], v[1]};
        }
      }
    }
    __syncthreads();  // the previous tile's fragment reads are done
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int c = tid + 256 * k;
      if (c < kChunks) {
        const int chunk = c & 3;
        const int px = (c >> 2) % (kTileW + 2);
        const int row = (c >> 2) / (kTileW + 2);
        *reinterpret_cast<uint4*>(&tile[row][px][chunk * 8]) = stage[k];
      }
    }
    __syncthreads();
    cur = tile_of(ti + 1 < t1 ? ti + 1 : ti);
    issue_loads(cur, ti + 1 < t1);

    f32x4 acc[4][2];
#pragma unroll
    for (int pb = 0; pb < 4; ++pb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) acc[pb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};

#pragma unroll
    for (int pb = 0; pb < 4; ++pb) {
      bf16x8 bfr[9];
#pragma unroll
      for (int t = 0; t < 9; ++t)
        bfr[t] = *reinterpret_cast<const bf16x8*>(
            &tile[wave + t / 3][pb * 16 + (lane & 15) + t % 3][8 * (lane >> 4)]);
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        acc[pb][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[t][0], bfr[t], acc[pb][0], 0, 0, 0);
        acc[pb][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[t][1], bfr[t], acc[pb][1], 0, 0, 0);
      }
    }

    if (y < H) {
#pragma unroll
      for (int pb = 0; pb < 4; ++pb) {
        const int x = x0 + pb * 16 + (lane & 15);
        if (x >= W) continue;
        const bool valid = (y < ny) && (x < nx);
        const int64_t po = (img + (int64_t)y * W + x) * kC;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          const int co = cb * 16 + 4 * (lane >> 4);
          float v[4];
          if constexpr (MODE == 0) {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = leaky(acc[pb][cb][e] + bias_r[cb][e]);
          } else {
            const bf16* g = reinterpret_cast<const bf16*>(&auxv[pb][cb]);
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = acc[pb][cb][e] * (bf2f(g[e]) > 0.f ? 1.f : 0.01f);
          }
          bf16 h[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) h[e] = valid ? f2bf(v[e]) : f2bf(0.f);
          *reinterpret_cast<uint2*>(out + po + co) = *reinterpret_cast<const uint2*>(h);
        }
      }
    }
  }
}

// persistent grid for the 3x3 kernels: enough blocks to fill the chip a few
// times over, each walking a contiguous run of tiles
// 512 blocks (2 per CU, each pipelining a run of ~8 tiles): measured at
// 1000 x 1000, 35.3 us per 32->32 layer against 36.6 / 39.5 / 46.2 us for
// 1024 / 2048 / 4096 blocks and 48.6 us for 256.  (IGM_CONV_BLOCKS
// overrides the target block count, for tuning sweeps.)
inline int64_t conv_target_blocks() {
  static int64_t n = [] {
    const char* e = getenv("IGM_CONV_BLOCKS");
    const long long r = e ? atoll(e) : 512;
    return (int64_t)((r >= 64 && r <= (1 << 22)) ? r : 512);
  }();
  return n;
}

inline void conv_grid(int B, int H, int W, int* blocks, int* per) {
  const int64_t nty = (H + kTileH - 1) / kTileH, ntx = (W + kTileW - 1) / kTileW;
  const int64_t ntiles = (int64_t)B * nty * ntx;
  const int64_t kTarget = conv_target_blocks();
  const int64_t p = (ntiles + kTarget - 1) / kTarget;
  *per = (int)(p < 1 ? 1 : p);
  *blocks = (int)((ntiles + *per - 1) / *per);
}

// ------------------------------------------------------- 1x1 head (MFMA) --
// wt: [32 out (zero-padded beyond cout)][32 in] bf16; Y fp32 (B,H,W,cout).
__global__ void __launch_bounds__(256)
head_mfma_kernel(const bf16* __restrict__ in, const bf16* __restrict__ wt,
                 const float* __restrict__ bias, const int32_t* __restrict__ ext, int H, int W,
                 int cout, float* __restrict__ Y) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t npx = (int64_t)H * W;
  const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * 64;  // 64 pixels per wave
  const int b = blockIdx.y;
  if (p0 >= npx) return;
  int ny, nx;
  extent_of(ext, b, H, W, ny, nx);
  bf16x8 wf[2];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb)
    wf[cb] = *reinterpret_cast<const bf16x8*>(wt + ((cb * 16 + (lane & 15)) * kC + 8 * (lane >> 4)));
  const bf16* ib = in + (int64_t)b * npx * kC;
  float* Yb = Y + (int64_t)b * npx * cout;
  // this lane's output channels co = cb*16 + 4*(lane/16) + e: their biases,
  // and whether all 4 are real (cout % 4 == 0 makes a pixel's row of cout
  // floats 16-B aligned: one 16-B store per lane and channel block)
  float bl[2][4];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int co = cb * 16 + 4 * (lane >> 4) + e;
      bl[cb][e] = co < cout ? bias[co] : 0.f;
    }
  const bool vec = (cout % 4) == 0 && (reinterpret_cast<uintptr_t>(Y) & 15) == 0;
#pragma unroll
  for (int pb = 0; pb < 4; ++pb) {
    const int64_t p = p0 + pb * 16 + (lane & 15);
    f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
    for (int kc = 0; kc < 1; ++kc) {
      bf16x8 bfr;
      if (p < npx) bfr = *reinterpret_cast<const bf16x8*>(ib + p * kC + 8 * (lane >> 4));
      else {
#pragma unroll
        for (int e = 0; e < 8; ++e) bfr[e] = f2bf(0.f);
      }
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[0], bfr, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[1], bfr, acc[1], 0, 0, 0);
    }
    if (p >= npx) continue;
    const int yy = (int)(p / W), xx = (int)(p % W);
    const bool valid = yy < ny && xx < nx;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int co = cb * 16 + 4 * (lane >> 4);
      float o[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) o[e] = valid ? acc[cb][e] + bl[cb][e] : 0.f;
      if (vec) {
        if (co < cout) *reinterpret_cast<float4*>(Yb + p * cout + co) = make_float4(o[0], o[1], o[2], o[3]);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (co + e < cout) Yb[p * cout + co + e] = o[e];
      }
    }
  }
}

// ------------------------------------------------------- head backward --
// dZ16 = (dY W17^T) * LeakyReLU'(A16) (bf16); dYb = bf16(dY) zero-padded to 32.
// w17: HWIO [32 in][cout] fp32.
__global__ void __launch_bounds__(256)
head_bwd_kernel(const float* __restrict__ dY, const bf16* __restrict__ a16,
                const float* __restrict__ w17, const int32_t* __restrict__ ext, int H, int W,
                int cout, bf16* __restrict__ dz16, bf16* __restrict__ dYb) {
  // ws[ci][co] = w17[ci][co], zero for co >= cout: the loops below run over
  // whole groups of 4 output channels with static register indices (the
  // padded terms add exact zeros)
  __shared__ __attribute__((aligned(16))) float ws[kC][kC];
  for (int i = threadIdx.x; i < kC * kC; i += blockDim.x) {
    const int ci = i / kC, co = i % kC;
    ws[ci][co] = co < cout ? w17[ci * cout + co] : 0.f;
  }
  __syncthreads();
  const int64_t npx = (int64_t)H * W;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (p >= npx) return;
  int ny, nx;
  extent_of(ext, b, H, W, ny, nx);
  const int yy = (int)(p / W), xx = (int)(p % W);
  const bool valid = yy < ny && xx < nx;
  const int64_t q = (int64_t)b * npx + p;
  if (!valid) {  // outside the extent: zero gradients, no arithmetic
    const uint4 z = {0, 0, 0, 0};
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) {
      reinterpret_cast<uint4*>(dz16 + q * kC)[c4] = z;
      reinterpret_cast<uint4*>(dYb + q * kC)[c4] = z;
    }
    return;
  }
  // dY's row of cout floats: 16-B loads when cout % 4 == 0 (the row is then
  // 16-B aligned in a 16-B aligned dY), 4-B loads otherwise
  float d[kC];
  if ((cout % 4) == 0 && (reinterpret_cast<uintptr_t>(dY) & 15) == 0) {
#pragma unroll
    for (int g = 0; g < kC / 4; ++g) {
      const float4 t = 4 * g < cout ? *reinterpret_cast<const float4*>(dY + q * cout + 4 * g)
                                    : make_float4(0.f, 0.f, 0.f, 0.f);
      d[4 * g] = t.x;
      d[4 * g + 1] = t.y;
      d[4 * g + 2] = t.z;
      d[4 * g + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int c = 0; c < kC; ++c) d[c] = c < cout ? dY[q * cout + c] : 0.f;
  }
  // s[ci] = sum over co = 0 .. cout-1, in order, of d[co] * w[ci][co]
  float s[kC];
#pragma unroll
  for (int ci = 0; ci < kC; ++ci) s[ci] = 0.f;
#pragma unroll
  for (int g = 0; g < kC / 4; ++g) {
    if (4 * g < cout) {  // wave-uniform
#pragma unroll
      for (int ci = 0; ci < kC; ++ci) {
        const float4 wv = *reinterpret_cast<const float4*>(&ws[ci][4 * g]);
        s[ci] += d[4 * g] * wv.x;
        s[ci] += d[4 * g + 1] * wv.y;
        s[ci] += d[4 * g + 2] * wv.z;
        s[ci] += d[4 * g + 3] * wv.w;
      }
    }
  }
  bf16x8 outv[4], outy[4];
#pragma unroll
  for (int c4 = 0; c4 < 4; ++c4) {
    const bf16x8 av = reinterpret_cast<const bf16x8*>(a16 + q * kC)[c4];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int ci = c4 * 8 + e;
      outv[c4][e] = f2bf(s[ci] * (bf2f(av[e]) > 0.f ? 1.f : 0.01f));
      outy[c4][e] = f2bf(d[ci]);
    }
  }
#pragma unroll
  for (int c4 = 0; c4 < 4; ++c4) {
    reinterpret_cast<bf16x8*>(dz16 + q * kC)[c4] = outv[c4];
    reinterpret_cast<bf16x8*>(dYb + q * kC)[c4] = outy[c4];
  }
}

// ------------------------------------------------------------ wgrad MFMA --
// dW[tap][co][ci] = sum_p A[p + delta_tap][ci] * dZ[p][co],  db[co] = sum_p dZ[p][co].
// Persistent blocks over tiles of kWgRows rows x 64 pixels; TAPS = 9 (3x3)
// or 1 (1x1).  Wave w owns the output blocks (co-block w&1, ci-block w>>1)
// of all taps.  LDS holds the tile's dZ rows and its activation rows (with a
// one-row halo for 3x3) transposed ([channel][pixel]), written as packed
// 2-pixel dwords; the halo rows are loaded once for the kWgRows rows.  The
// x-shift s of a tap is made in registers from two aligned 16-byte reads:
// s=0 is the first read, s=2 a dword select, s=1 a 16-bit funnel shift
// (v_alignbit).  A block loads its next tile into registers before the
// current tile's MFMAs.  partials: [gridDim.x][TAPS*1024 + 32] fp32.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

constexpr int kWgRows = 4;  // dZ rows per tile (8 measured slower on c4 batches)

template <int TAPS>
__global__ void __launch_bounds__(256)
wgrad_mfma_kernel(const bf16* __restrict__ act, const bf16* __restrict__ dz,
                  const int32_t* __restrict__ ext, int B, int H, int W, int accumulate,
                  float* __restrict__ partials) {
  constexpr int AROWS = TAPS == 9 ? kWgRows + 2 : kWgRows;  // activation rows in LDS
  constexpr int NT = TAPS;  // accumulators per wave
  __shared__ __attribute__((aligned(16))) uint32_t lds_dz[kWgRows][kC][kChStride / 2];
  __shared__ __attribute__((aligned(16))) uint32_t lds_act[AROWS][kC][kChStride / 2];
  __shared__ float dbred[256][8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cb = wave & 1, cib = wave >> 1;
  const int ntx = (W + kTileW - 1) / kTileW;
  const int nty = (H + kWgRows - 1) / kWgRows;
  const int ntiles = B * nty * ntx;  // < 2^25: activations below 2 GiB

  f32x4 acc[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  // accumulate: this block's partial sums of the previous sub-batches, loaded
  // now so that their latency hides behind the tiles (added at the end)
  float* P = partials + (int64_t)blockIdx.x * (TAPS * kC * kC + kC);
  const int ci = cib * 16 + (lane & 15);
  f32x4 prev[NT];
#pragma unroll
  for (int tap = 0; tap < NT; ++tap)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int co = cb * 16 + 4 * (lane >> 4) + e;
      prev[tap][e] = accumulate ? P[(tap * kC + co) * kC + ci] : 0.f;
    }
  float dba[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) dba[e] = 0.f;

  // Staging of a tile: dZ rows y0 .. y0 + kWgRows - 1 and activation rows
  // (y0 - 1 .. y0 + kWgRows for 3x3), each as 16-B channel chunks of pixel
  // pairs, in registers.  The next tile's loads are issued before the
  // current tile's MFMAs, so they are in flight during them; masked lanes
  // (outside the entry's extent, or no next tile) read zeros.
  constexpr int kPairs = (kTileW + 2) / 2;
  constexpr int kDzItems = kWgRows * 32 * 4;
  constexpr int kDzPer = (kDzItems + 255) / 256;
  constexpr int kActItems = AROWS * kPairs * 4;
  constexpr int kActPer = (kActItems + 255) / 256;
  const uint32_t vbytes = (uint32_t)((int64_t)B * H * W * kC * 2);
  const rsrc_t rdz = make_rsrc(dz, vbytes), ract = make_rsrc(act, vbytes);
  uint4 sdz[kDzPer][2], sact[kActPer][2];
  struct Tile {
    int b, y0, x0, ny, nx;
    bool valid;
  };
  auto tile_of = [&](int t) {
    Tile r;
    r.b = t / (nty * ntx);
    r.y0 = ((t / ntx) % nty) * kWgRows;
    r.x0 = (t % ntx) * kTileW;
    extent_of(ext, r.b, H, W, r.ny, r.nx);
    r.valid = r.y0 < r.ny && r.x0 < r.nx;
    return r;
  };
  auto ld16 = [&](rsrc_t r, bool ok, int64_t pix, int chunk) {
    const uint32_t off = ok ? (uint32_t)((pix * kC + chunk * 8) * 2) : kOOB;
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return uint4{v[0], v[1], v[2], v[3]};
  };
  auto issue_loads = [&](const Tile& t, bool live) {
    const int64_t img = (int64_t)t.b * H * W;
    const bool on = live && t.valid;
#pragma unroll
    for (int k = 0; k < kDzPer; ++k) {
      const int c = tid + 256 * k;
      const int chunk = c & 3, q = (c >> 2) & 31, row = c >> 7;
      const int yd = t.y0 + row, xa = t.x0 + 2 * q;
      const bool rok = on && c < kDzItems && yd < t.ny;
      const int64_t pix = img + (int64_t)yd * W + xa;
      sdz[k][0] = ld16(rdz, rok && xa < t.nx, pix, chunk);
      sdz[k][1] = ld16(rdz, rok && xa + 1 < t.nx, pix + 1, chunk);
    }
#pragma unroll
    for (int k = 0; k < kActPer; ++k) {
      const int c = tid + 256 * k;
      const int chunk = c & 3, q = (c >> 2) % kPairs, row = (c >> 2) / kPairs;
      const int yy = (TAPS == 9) ? t.y0 - 1 + row : t.y0 + row;
      const int xa = t.x0 - 1 + 2 * q;
      const bool rok = on && c < kActItems && yy >= 0 && yy < t.ny;
      const int64_t pix = img + (int64_t)yy * W + xa;
      sact[k][0] = ld16(ract, rok && xa >= 0 && xa < t.nx, pix, chunk);
      sact[k][1] = ld16(ract, rok && xa + 1 >= 0 && xa + 1 < t.nx, pix + 1, chunk);
    }
  };
  // staged registers -> LDS, channel-major with pixel pairs per dword:
  // lds_dz[row][ch][q] holds pixels (x0 + 2q, x0 + 2q + 1), lds_act[row][ch][q]
  // pixels (x0 - 1 + 2q, x0 + 2q); thread tid always takes channel chunk
  // tid & 3, and sums its dZ channels into the bias gradient
  auto store_tile = [&]() {
#pragma unroll
    for (int k = 0; k < kDzPer; ++k) {
      const int c = tid + 256 * k;
      if (c < kDzItems) {
        const int chunk = c & 3, q = (c >> 2) & 31, row = c >> 7;
        const uint32_t a[4] = {sdz[k][0].x, sdz[k][0].y, sdz[k][0].z, sdz[k][0].w};
        const uint32_t c2[4] = {sdz[k][1].x, sdz[k][1].y, sdz[k][1].z, sdz[k][1].w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          // channel 2i: low halves; channel 2i+1: high halves (bf16 bit patterns)
          lds_dz[row][chunk * 8 + 2 * i][q] = (a[i] & 0xFFFFu) | (c2[i] << 16);
          lds_dz[row][chunk * 8 + 2 * i + 1][q] = (a[i] >> 16) | (c2[i] & 0xFFFF0000u);
          dba[2 * i] += __uint_as_float(a[i] << 16) + __uint_as_float(c2[i] << 16);
          dba[2 * i + 1] += __uint_as_float(a[i] & 0xFFFF0000u) + __uint_as_float(c2[i] & 0xFFFF0000u);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kActPer; ++k) {
      const int c = tid + 256 * k;
      if (c < kActItems) {
        const int chunk = c & 3, q = (c >> 2) % kPairs, row = (c >> 2) / kPairs;
        const uint32_t a[4] = {sact[k][0].x, sact[k][0].y, sact[k][0].z, sact[k][0].w};
        const uint32_t c2[4] = {sact[k][1].x, sact[k][1].y, sact[k][1].z, sact[k][1].w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          lds_act[row][chunk * 8 + 2 * i][q] = (a[i] & 0xFFFFu) | (c2[i] << 16);
          lds_act[row][chunk * 8 + 2 * i + 1][q] = (a[i] >> 16) | (c2[i] & 0xFFFF0000u);
        }
      }
    }
  };

  int tile = blockIdx.x;
  Tile cur = tile_of(tile < ntiles ? tile : 0);
  issue_loads(cur, tile < ntiles);
  for (; tile < ntiles; tile += gridDim.x) {
    const bool valid = cur.valid;  // block-uniform
    if (valid) {
      __syncthreads();  // the previous tile's fragment reads are done
      store_tile();
      __syncthreads();
    }
    const int nt = tile + gridDim.x;
    cur = tile_of(nt < ntiles ? nt : tile);
    issue_loads(cur, nt < ntiles);
    if (!valid) continue;
#pragma unroll
    for (int rr = 0; rr < kWgRows; ++rr) {
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        const int base = (ks * 32 + 8 * (lane >> 4)) / 2;  // dword index of the 8-pixel group
        const u32x4 a4 = *reinterpret_cast<const u32x4*>(&lds_dz[rr][cb * 16 + (lane & 15)][base]);
        const bf16x8 afr = as_bf16x8(a4);
#pragma unroll
        for (int r = 0; r < (TAPS == 9 ? 3 : 1); ++r) {
          // activation pixels x0 - 1 + t with t = 2*base .. 2*base + 15, of
          // activation row rr + r (dZ row rr, tap row r)
          const uint32_t* src = &lds_act[rr + r][cib * 16 + (lane & 15)][base];
          const u32x4 c0 = *reinterpret_cast<const u32x4*>(src);
          const u32x4 c1 = *reinterpret_cast<const u32x4*>(src + 4);
          u32x4 s1;
          s1[0] = __builtin_amdgcn_alignbit(c0[1], c0[0], 16);
          s1[1] = __builtin_amdgcn_alignbit(c0[2], c0[1], 16);
          s1[2] = __builtin_amdgcn_alignbit(c0[3], c0[2], 16);
          s1[3] = __builtin_amdgcn_alignbit(c1[0], c0[3], 16);
          if constexpr (TAPS == 9) {
            // s = 0: pixels x0 - 1 + p  -> t = p          (c0)
            // s = 1: pixels x0 + p      -> t = p + 1      (16-bit funnel shift)
            // s = 2: pixels x0 + 1 + p  -> t = p + 2      (dwords 1..4)
            u32x4 s2;
            s2[0] = c0[1]; s2[1] = c0[2]; s2[2] = c0[3]; s2[3] = c1[0];
            acc[r * 3 + 0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, as_bf16x8(c0), acc[r * 3 + 0], 0, 0, 0);
            acc[r * 3 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, as_bf16x8(s1), acc[r * 3 + 1], 0, 0, 0);
            acc[r * 3 + 2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, as_bf16x8(s2), acc[r * 3 + 2], 0, 0, 0);
          } else {
            // 1x1: pixel x0 + p -> t = p + 1
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, as_bf16x8(s1), acc[0], 0, 0, 0);
          }
        }
      }
    }
  }
  // partial dW: D[co][ci], lane: ci = cib*16 + lane%16, co = cb*16 + 4*(lane/16) + e
#pragma unroll
  for (int tap = 0; tap < NT; ++tap)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int co = cb * 16 + 4 * (lane >> 4) + e;
      P[(tap * kC + co) * kC + ci] = accumulate ? prev[tap][e] + acc[tap][e] : acc[tap][e];
    }
  // partial db: thread t holds channels (t&3)*8 .. +8
#pragma unroll
  for (int e = 0; e < 8; ++e) dbred[tid][e] = dba[e];
  __syncthreads();
  if (tid < kC) {
    const int chunk = tid >> 3, e = tid & 7;
    float s = 0.f;
    for (int t = chunk; t < 256; t += 4) s += dbred[t][e];
    P[TAPS * kC * kC + tid] = accumulate ? P[TAPS * kC * kC + tid] + s : s;
  }
}

// out_hwio[tap][ci][co] = sum_blk P[blk][(tap*32 + co)*32 + ci] (co < cout);
// db[co] = sum_blk P[blk][TAPS*1024 + co].  A block reduces 64 consecutive
// partial entries: thread (g, j) sums blocks g, g+4, ... of entry j, then the
// four group sums are added in a fixed order (deterministic).
// Deterministic reduction of per-block partials P[nblk][stride], in two
// stages.  Stage 1 (partials_group_sum_kernel): group g of R rows,
// rows g*R .. min(nblk, (g+1)*R) - 1, is summed in row order into its first
// row, in place; each (group, column) belongs to one thread, so the write
// races with nothing, and the loads are issued 8 at a time.  Stage 2 sums
// the group rows in group order.  The fixed order makes the result
// independent of scheduling.
__global__ void __launch_bounds__(256)
partials_group_sum_kernel(float* __restrict__ P, int nblk, int stride, int ncols, int R) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int r0 = blockIdx.y * R;
  if (c >= ncols || r0 >= nblk) return;
  const int n = min(nblk, r0 + R) - r0;
  float* p = P + (int64_t)r0 * stride + c;
  float s = 0.f;
  int r = 0;
  for (; r + 8 <= n; r += 8) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = p[(int64_t)(r + i) * stride];
#pragma unroll
    for (int i = 0; i < 8; ++i) s += v[i];
  }
  for (; r < n; ++r) s += p[(int64_t)r * stride];
  p[0] = s;
}

// Stage 2 for a 3x3 or 1x1 layer: the group sums of entry src = ((tap * 32
// + co) * 32 + ci) (then the 32 bias entries) scattered to dW (HWIO, cout
// output channels) and db.
__global__ void __launch_bounds__(256)
wgrad_reduce_kernel(const float* __restrict__ P, int ngroups, int R, int taps, int cout,
                    int accumulate, float* __restrict__ dW, float* __restrict__ db) {
  const int stride = taps * kC * kC + kC;
  const int src = blockIdx.x * 256 + threadIdx.x;
  if (src >= stride) return;
  float t = 0.f;
  for (int g = 0; g < ngroups; ++g) t += P[(int64_t)g * R * stride + src];
  if (src < taps * kC * kC) {
    const int ci = src % kC, co = (src / kC) % kC, tap = src / (kC * kC);
    if (co < cout) {
      float* d = dW + (tap * kC + ci) * cout + co;
      *d = accumulate ? *d + t : t;
    }
  } else {
    const int co = src - taps * kC * kC;
    if (co < cout) db[co] = accumulate ? db[co] + t : t;
  }
}

// Stage 1 + 2 on the host side: G groups of R rows, G <= kMaxGroups.
constexpr int kMaxGroups = 16;
inline void group_rows(int nblk, int* R, int* G) {
  *G = nblk < kMaxGroups ? nblk : kMaxGroups;
  *R = (nblk + *G - 1) / *G;
  *G = (nblk + *R - 1) / *R;
}

// ------------------------------------------------------ layer-1 wgrad --
// dW1[r][s][ci][co] (HWIO, fp32) = sum_p X[p + delta][ci] * dZ1[p][co]; db1.
// Persistent blocks over tiles of 4 rows x 64 pixels; wave w takes row w of
// a tile.  Lane l accumulates a 6 x 4 register tile of the output: the
// (tap, ci) pairs m = 6 (l / 8) .. +5 (m = tap * 5 + ci; 48 slots, the last
// three empty) by the output channels co = 4 (l % 8) .. +3.  Per pixel, six
// LDS reads of X and one 16-B read of dZ feed 24 FMAs.  At the end the four
// waves' tiles are summed in wave order into partials[blk][1440 + 32]
// (P[m * 32 + co], then the 32 bias sums).
constexpr int kWgTileW = 64, kWgTileH = 4;
constexpr int kWgOut = 9 * kCin * kC + kC;  // 1472 partial entries per block
__global__ void __launch_bounds__(256)
conv_in_wgrad_kernel(const float* __restrict__ f0, const float* __restrict__ f1,
                     const float* __restrict__ f2, const float* __restrict__ f3,
                     const float* __restrict__ f4, int64_t fb, int64_t fr, int64_t fc,
                     const int32_t* __restrict__ ext, int B, int H, int W,
                     const bf16* __restrict__ dz, int accumulate, float* __restrict__ partials) {
  __shared__ float xs[kWgTileH + 2][kWgTileW + 2][kCin];
  __shared__ float4 dzs[kWgTileH][kWgTileW][kC / 4];
  __shared__ float red[4][kWgOut];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int mb = lane >> 3, nb = lane & 7;
  const float* fs[kCin] = {f0, f1, f2, f3, f4};
  float acc[6][4], dbacc[4] = {0.f, 0.f, 0.f, 0.f};
  int xoff[6];  // LDS offset of the lane's (tap, ci) pairs from xs[w][px][0]
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int m = min(6 * mb + j, 9 * kCin - 1), tap = m / kCin, ci = m % kCin;
    xoff[j] = ((tap / 3) * (kWgTileW + 2) + tap % 3) * kCin + ci;
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[j][q] = 0.f;
  }
  const int ntx = (W + kWgTileW - 1) / kWgTileW, nty = (H + kWgTileH - 1) / kWgTileH;
  const int ntiles = B * nty * ntx;  // < 2^25: layer-1 gradient below 2 GiB
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int b = tile / (nty * ntx);
    const int y0 = ((tile / ntx) % nty) * kWgTileH;
    const int x0 = (tile % ntx) * kWgTileW;
    int ny, nx;
    extent_of(ext, b, H, W, ny, nx);
    if (y0 >= ny || x0 >= nx) continue;
    __syncthreads();
    // X with a one-pixel halo, zero outside the entry's extent
    for (int c = tid; c < (kWgTileH + 2) * (kWgTileW + 2) * kCin; c += 256) {
      const int ci = c % kCin, t = (c / kCin) % (kWgTileW + 2), row = c / (kCin * (kWgTileW + 2));
      const int yy = y0 - 1 + row, xx = x0 - 1 + t;
      float v = 0.f;
      if (yy >= 0 && yy < ny && xx >= 0 && xx < nx)
        v = fs[ci][(int64_t)b * fb + (int64_t)yy * fr + (int64_t)xx * fc];
      xs[row][t][ci] = v;
    }
    // dZ1 (bf16, 32 channels), zero outside the extent
    for (int c = tid; c < kWgTileH * kWgTileW * 4; c += 256) {
      const int chunk = c & 3, px = (c >> 2) % kWgTileW, row = (c >> 2) / kWgTileW;
      const int y = y0 + row, x = x0 + px;
      bf16x8 v;
      if (y < ny && x < nx) {
        v = *reinterpret_cast<const bf16x8*>(dz + (((int64_t)b * H + y) * W + x) * kC + chunk * 8);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = f2bf(0.f);
      }
      dzs[row][px][2 * chunk] = make_float4(bf2f(v[0]), bf2f(v[1]), bf2f(v[2]), bf2f(v[3]));
      dzs[row][px][2 * chunk + 1] = make_float4(bf2f(v[4]), bf2f(v[5]), bf2f(v[6]), bf2f(v[7]));
    }
    __syncthreads();
    const float* xw = &xs[w][0][0];
    for (int px = 0; px < kWgTileW; ++px) {
      const float4 d = dzs[w][px][nb];
      dbacc[0] += d.x;
      dbacc[1] += d.y;
      dbacc[2] += d.z;
      dbacc[3] += d.w;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const float xv = xw[px * kCin + xoff[j]];
        acc[j][0] += xv * d.x;
        acc[j][1] += xv * d.y;
        acc[j][2] += xv * d.z;
        acc[j][3] += xv * d.w;
      }
    }
  }
  // the four waves' tiles, summed in wave order
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int m = 6 * mb + j;
    if (m < 9 * kCin) {
#pragma unroll
      for (int q = 0; q < 4; ++q) red[w][m * kC + 4 * nb + q] = acc[j][q];
    }
  }
  if (mb == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[w][9 * kCin * kC + 4 * nb + q] = dbacc[q];
  }
  __syncthreads();
  float* P = partials + (int64_t)blockIdx.x * kWgOut;
  constexpr int kPer = (kWgOut + 255) / 256;
  float prev[kPer];  // (all loads before the first store)
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + 256 * k;
    prev[k] = (accumulate && i < kWgOut) ? P[i] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + 256 * k;
    if (i < kWgOut) {
      const float v = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
      P[i] = accumulate ? prev[k] + v : v;
    }
  }
}

// Stage 2 as a plain column sum: out[i] = sum over the group rows of P[.][i],
// i < count
__global__ void __launch_bounds__(256)
sum_partials_kernel(const float* __restrict__ P, int ngroups, int R, int stride, int count,
                    int accumulate, float* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  float t = 0.f;
  for (int g = 0; g < ngroups; ++g) t += P[(int64_t)g * R * stride + i];
  out[i] = accumulate ? out[i] + t : t;
}

// ------------------------------------------------------------ optimizer --
// Per-tensor sum of squares (one block per tensor), deterministic.
__global__ void __launch_bounds__(256)
sqnorm_kernel(const float* __restrict__ g, const int64_t* __restrict__ offs, float* __restrict__ sq) {
  const int t = blockIdx.x;
  const int64_t a = offs[t], e = offs[t + 1];
  __shared__ float red[256];
  float s = 0.f;
  for (int64_t i = a + threadIdx.x; i < e; i += 256) s += g[i] * g[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) sq[t] = red[0];
}

// tf.clip_by_norm per tensor, then legacy Adam on the fp32 master weights.
__global__ void __launch_bounds__(256)
adam_clip_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                 float* __restrict__ v, int64_t n, const int64_t* __restrict__ offs, int ntens,
                 const float* __restrict__ sq, float clipnorm, float alpha, float omb1, float omb2,
                 float eps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = ntens;  // offs[lo] <= i < offs[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offs[mid] <= i) lo = mid; else hi = mid;
  }
  float gi = g[i];
  if (clipnorm > 0.f) {
    const float l2sum = sq[lo];
    const float l2norm = l2sum > 0.f ? sqrtf(l2sum) : l2sum;
    gi = gi * clipnorm / fmaxf(l2norm, clipnorm);
  }
  float mi = m[i], vi = v[i];
  mi += (gi - mi) * omb1;
  vi += (gi * gi - vi) * omb2;
  m[i] = mi;
  v[i] = vi;
  w[i] -= (mi * alpha) / (sqrtf(vi) + eps);
}

// bf16 MFMA operands from the fp32 HWIO master weights of the 3x3 layers:
// fwd[L][tap][co][ci] = W[tap][ci][co];  dgrad[L][tap][ci][co] = W[8 - tap][ci][co].
__global__ void __launch_bounds__(256)
pack3x3_kernel(const float* __restrict__ w, const int64_t* __restrict__ offs, int nlayers,
               bf16* __restrict__ fwd, bf16* __restrict__ dgr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int per = 9 * kC * kC;
  if (i >= (int64_t)nlayers * per) return;
  const int L = (int)(i / per), r = (int)(i % per);
  const int tap = r / (kC * kC), a = (r / kC) % kC, c = r % kC;
  const float* W = w + offs[L];
  // fwd element [tap][co=a][ci=c]
  fwd[i] = f2bf(W[(tap * kC + c) * kC + a]);
  // dgrad element [tap][out=ci=a][in=co=c] = W[8-tap][ci=a][co=c]
  dgr[i] = f2bf(W[((8 - tap) * kC + a) * kC + c]);
}

// head operand [32 co (zero-padded)][32 ci] from HWIO [ci][cout].
__global__ void pack_head_kernel(const float* __restrict__ w, int cout, bf16* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kC * kC) return;
  const int co = i / kC, ci = i % kC;
  out[i] = f2bf(co < cout ? w[ci * cout + co] : 0.f);
}

inline int grid_for(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace igm

// =================================================================== C-ABI ==
using namespace igm;

extern "C" {

int igm_conv_in_fwd(const float* f0, const float* f1, const float* f2, const float* f3,
                    const float* f4, int64_t fb, int64_t fr, int64_t fc, const int32_t* extents,
                    int32_t B, int32_t H, int32_t W, const float* w_hwio, const float* bias,
                    void* out_bf16, void* stream) {
  IGM_CHECK_ARG(f0 && f1 && f2 && f3 && f4 && w_hwio && bias && out_bf16, "null pointer");
  IGM_CHECK_ARG(B >= 1 && H >= 1 && W >= 1 && H <= 65535 && B <= 65535, "bad sizes");
  hipLaunchKernelGGL(conv_in_kernel, dim3(grid_for(W), H, B), dim3(256), 0, (hipStream_t)stream,
                     f0, f1, f2, f3, f4, fb, fr, fc, extents, H, W, w_hwio, bias,
                     (bf16*)out_bf16);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_conv3x3_fwd_bf16(const void* in, const void* wpacked, const float* bias,
                         const int32_t* extents, int32_t B, int32_t H, int32_t W, void* out,
                         void* stream) {
  IGM_CHECK_ARG(in && wpacked && bias && out, "null pointer");
  IGM_CHECK_ARG(in != out, "in and out must differ");
  IGM_CHECK_ARG(B >= 1 && H >= 1 && W >= 1 && B <= 65535, "bad sizes");
  IGM_CHECK_ARG((int64_t)B * H * W * kC * 2 < (int64_t)kOOB, "activations above 2 GiB (32-bit buffer offsets)");
  int blocks, per;
  conv_grid(B, H, W, &blocks, &per);
  hipLaunchKernelGGL(conv3x3_mfma_kernel<0>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                     (const bf16*)in, (const bf16*)wpacked, bias, (const bf16*)nullptr, extents, B, H, W,
                     per, (bf16*)out);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_conv3x3_dgrad_bf16(const void* dz, const void* wpacked_dgrad, const void* act_below,
                           const int32_t* extents, int32_t B, int32_t H, int32_t W, void* dz_below,
                           void* stream) {
  IGM_CHECK_ARG(dz && wpacked_dgrad && act_below && dz_below, "null pointer");
  IGM_CHECK_ARG(dz != dz_below, "in and out must differ");
  IGM_CHECK_ARG(B >= 1 && H >= 1 && W >= 1 && B <= 65535, "bad sizes");
  IGM_CHECK_ARG((int64_t)B * H * W * kC * 2 < (int64_t)kOOB, "activations above 2 GiB (32-bit buffer offsets)");
  int blocks, per;
  conv_grid(B, H, W, &blocks, &per);
  hipLaunchKernelGGL(conv3x3_mfma_kernel<1>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                     (const bf16*)dz, (const bf16*)wpacked_dgrad, (const float*)nullptr,
                     (const bf16*)act_below, extents, B, H, W, per, (bf16*)dz_below);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_conv1x1_head_fwd(const void* in, const void* wpacked, const float* bias,
                         const int32_t* extents, int32_t B, int32_t H, int32_t W, int32_t cout,
                         float* Y, void* stream) {
  IGM_CHECK_ARG(in && wpacked && bias && Y, "null pointer");
  IGM_CHECK_ARG(cout >= 1 && cout <= kC, "cout must be in [1, 32]");
  IGM_CHECK_ARG(B >= 1 && H >= 1 && W >= 1 && B <= 65535, "bad sizes");
  const int64_t npx = (int64_t)H * W;
  dim3 grid((unsigned)((npx + 255) / 256), B);
  hipLaunchKernelGGL(head_mfma_kernel, grid, dim3(256), 0, (hipStream_t)stream, (const bf16*)in,
                     (const bf16*)wpacked, bias, extents, H, W, cout, Y);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_conv1x1_head_bwd(const float* dY, const void* a16, const float* w_hwio,
                         const int32_t* extents, int32_t B, int32_t H, int32_t W, int32_t cout,
                         void* dz16, void* dY_bf16, void* stream) {
  IGM_CHECK_ARG(dY && a16 && w_hwio && dz16 && dY_bf16, "null pointer");
  IGM_CHECK_ARG(cout >= 1 && cout <= kC, "cout must be in [1, 32]");
  IGM_CHECK_ARG(B >= 1 && H >= 1 && W >= 1 && B <= 65535, "bad sizes");
  const int64_t npx = (int64_t)H * W;
  dim3 grid((unsigned)((npx + 255) / 256), B);
  hipLaunchKernelGGL(head_bwd_kernel, grid, dim3(256), 0, (hipStream_t)stream, dY,
                     (const bf16*)a16, w_hwio, extents, H, W, cout, (bf16*)dz16, (bf16*)dY_bf16);
  IGM_CHECK_LAUNCH();
  return 0;
}

int64_t igm_wgrad_workspace_floats(int32_t taps, int32_t nblocks) {
  if (taps == 9) return (int64_t)nblocks * (9 * 1024 + 32);
  if (taps == 1) return (int64_t)nblocks * (1024 + 32);
  if (taps == 0) return (int64_t)nblocks * (9 * 5 * 32 + 32);  // layer 1
  return 0;
}

int igm_conv_wgrad_partials_bf16(int32_t taps, const void* act, const void* dz,
                                 const int32_t* extents, int32_t B, int32_t H, int32_t W,
                                 int32_t nblocks, float* partials, int32_t accumulate,
                                 void* stream) {
  IGM_CHECK_ARG(act && dz && partials, "null pointer");
  IGM_CHECK_ARG(taps == 9 || taps == 1, "taps must be 9 or 1");
  IGM_CHECK_ARG(nblocks >= 1 && B >= 1 && H >= 1 && W >= 1, "bad sizes");
  IGM_CHECK_ARG((int64_t)B * H * W * kC * 2 < (int64_t)kOOB, "activations above 2 GiB (32-bit buffer offsets)");
  hipStream_t st = (hipStream_t)stream;
  const int acc = accumulate ? 1 : 0;
  if (taps == 9)
    hipLaunchKernelGGL(wgrad_mfma_kernel<9>, dim3(nblocks), dim3(256), 0, st, (const bf16*)act,
                       (const bf16*)dz, extents, B, H, W, acc, partials);
  else
    hipLaunchKernelGGL(wgrad_mfma_kernel<1>, dim3(nblocks), dim3(256), 0, st, (const bf16*)act,
                       (const bf16*)dz, extents, B, H, W, acc, partials);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_conv_in_wgrad_partials(const float* f0, const float* f1, const float* f2, const float* f3,
                               const float* f4, int64_t fb, int64_t fr, int64_t fc,
                               const int32_t* extents, int32_t B, int32_t H, int32_t W,
                               const void* dz1, int32_t nblocks, float* partials,
                               int32_t accumulate, void* stream) {
  IGM_CHECK_ARG(f0 && f1 && f2 && f3 && f4 && dz1 && partials, "null pointer");
  IGM_CHECK_ARG(nblocks >= 1 && B >= 1 && H >= 1 && W >= 1, "bad sizes");
  IGM_CHECK_ARG((int64_t)B * H * W * kC * 2 < (int64_t)kOOB, "activations above 2 GiB (32-bit buffer offsets)");
  hipLaunchKernelGGL(conv_in_wgrad_kernel, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, f0,
                     f1, f2, f3, f4, fb, fr, fc, extents, B, H, W, (const bf16*)dz1,
                     accumulate ? 1 : 0, partials);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_conv_wgrad_reduce_f32(int32_t taps, int32_t cout, int32_t nblocks, float* partials,
                              int32_t accumulate, float* dW_hwio, float* db, void* stream) {
  IGM_CHECK_ARG(partials && dW_hwio && db, "null pointer");
  IGM_CHECK_ARG(taps == 9 || taps == 1 || taps == 0, "taps must be 9, 1 or 0 (layer 1)");
  IGM_CHECK_ARG(cout >= 1 && cout <= kC, "cout must be in [1, 32]");
  IGM_CHECK_ARG(taps != 0 || cout == kC, "layer 1 has 32 outputs");
  IGM_CHECK_ARG(nblocks >= 1, "bad sizes");
  hipStream_t st = (hipStream_t)stream;
  int R, G;
  group_rows(nblocks, &R, &G);
  if (taps == 0) {
    // the layer-1 partial layout is [blk][1440 + 32] = [blk][n]: a plain column sum
    const int n = kWgOut;
    hipLaunchKernelGGL(partials_group_sum_kernel, dim3((n + 255) / 256, G), dim3(256), 0, st,
                       partials, nblocks, n, n, R);
    IGM_CHECK_LAUNCH();
    hipLaunchKernelGGL(sum_partials_kernel, dim3((9 * kCin * kC + 255) / 256), dim3(256), 0, st,
                       partials, G, R, n, 9 * kCin * kC, accumulate, dW_hwio);
    IGM_CHECK_LAUNCH();
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partials + 9 * kCin * kC,
                       G, R, n, kC, accumulate, db);
    IGM_CHECK_LAUNCH();
    return 0;
  }
  const int entries = taps * kC * kC + kC;
  hipLaunchKernelGGL(partials_group_sum_kernel, dim3((entries + 255) / 256, G), dim3(256), 0, st,
                     partials, nblocks, entries, entries, R);
  IGM_CHECK_LAUNCH();
  hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((entries + 255) / 256), dim3(256), 0, st, partials,
                     G, R, taps, cout, accumulate, dW_hwio, db);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_conv_wgrad_bf16(int32_t taps, const void* act, const void* dz, const int32_t* extents,
                        int32_t B, int32_t H, int32_t W, int32_t cout, int32_t nblocks,
                        float* workspace, int32_t accumulate, float* dW_hwio, float* db,
                        void* stream) {
  IGM_CHECK_ARG(act && dz && workspace && dW_hwio && db, "null pointer");
  IGM_CHECK_ARG(taps == 9 || taps == 1, "taps must be 9 or 1");
  IGM_CHECK_ARG(cout >= 1 && cout <= kC, "cout must be in [1, 32]");
  if (const int rc = igm_conv_wgrad_partials_bf16(taps, act, dz, extents, B, H, W, nblocks, workspace,
                                                  0, stream))
    return rc;
  return igm_conv_wgrad_reduce_f32(taps, cout, nblocks, workspace, accumulate, dW_hwio, db, stream);
}

int igm_conv_in_wgrad(const float* f0, const float* f1, const float* f2, const float* f3,
                      const float* f4, int64_t fb, int64_t fr, int64_t fc, const int32_t* extents,
                      int32_t B, int32_t H, int32_t W, const void* dz1, int32_t nblocks,
                      float* workspace, int32_t accumulate, float* dW_hwio, float* db,
                      void* stream) {
  IGM_CHECK_ARG(f0 && f1 && f2 && f3 && f4 && dz1 && workspace && dW_hwio && db, "null pointer");
  if (const int rc = igm_conv_in_wgrad_partials(f0, f1, f2, f3, f4, fb, fr, fc, extents, B, H, W,
                                                dz1, nblocks, workspace, 0, stream))
    return rc;
  return igm_conv_wgrad_reduce_f32(0, kC, nblocks, workspace, accumulate, dW_hwio, db, stream);
}

int igm_conv_in_dgrad(const void* dz1, const float* w_hwio, const int32_t* extents, int32_t B,
                      int32_t H, int32_t W, float* dX, void* stream) {
  IGM_CHECK_ARG(dz1 && w_hwio && dX, "null pointer");
  IGM_CHECK_ARG(B >= 1 && H >= 1 && W >= 1 && B <= 65535 && H <= 65535, "bad sizes");
  hipLaunchKernelGGL(conv_in_dgrad_kernel, dim3((W + 255) / 256, H, B), dim3(256), 0,
                     (hipStream_t)stream, (const bf16*)dz1, extents, H, W, w_hwio, dX);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_adam_clip_f32(float* w, const float* g, float* m, float* v, int64_t n, const int64_t* offs,
                      int32_t ntens, float* sq_workspace, double clipnorm, const IgmAdam* adam,
                      void* stream) {
  IGM_CHECK_ARG(w && g && m && v && offs && sq_workspace && adam, "null pointer");
  IGM_CHECK_ARG(ntens >= 1 && n >= 1, "bad sizes");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(sqnorm_kernel, dim3(ntens), dim3(256), 0, st, g, offs, sq_workspace);
  IGM_CHECK_LAUNCH();
  hipLaunchKernelGGL(adam_clip_kernel, dim3(grid_for(n)), dim3(256), 0, st, w, g, m, v, n, offs,
                     ntens, sq_workspace, (float)clipnorm, (float)adam->alpha,
                     (float)(1.0 - adam->beta1), (float)(1.0 - adam->beta2), (float)adam->epsilon);
  IGM_CHECK_LAUNCH();
  return 0;
}

int igm_pack_conv_weights(const float* w, const int64_t* offs3x3, int32_t nlayers3x3,
                          const float* w_head, int32_t cout, void* fwd, void* dgrad, void* head,
                          void* stream) {
  IGM_CHECK_ARG(w && w_head && fwd && dgrad && head, "null pointer");
  IGM_CHECK_ARG(nlayers3x3 == 0 || offs3x3, "null offs3x3");
  IGM_CHECK_ARG(cout >= 1 && cout <= kC && nlayers3x3 >= 0, "bad sizes");
  hipStream_t st = (hipStream_t)stream;
  if (nlayers3x3 > 0) {
    hipLaunchKernelGGL(pack3x3_kernel, dim3(grid_for((int64_t)nlayers3x3 * 9 * kC * kC)), dim3(256),
                       0, st, w, offs3x3, nlayers3x3, (bf16*)fwd, (bf16*)dgrad);
    IGM_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(pack_head_kernel, dim3(grid_for(kC * kC)), dim3(256), 0, st, w_head, cout,
                     (bf16*)head);
  IGM_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
