// Entry-parallel fused training step over packed wavefront tiles (K2 + K3 + K4).
//
// The rows of a batch are packed (on the host, once per zoom level) into tiles
// of whole rows holding at most 64 entries and 64 rows; a row with more than 64
// entries forms a tile of its own.  A wavefront takes one tile at a time and
// gives each lane one entry:
//   1. the lane evaluates the phase network of its entry at fl32(p + o*h) and
//      keeps the hidden activations in registers;
//   2. a segmented inclusive scan over the lanes (bounded by the first lane of
//      each row, fixed summation tree) forms the row sums, so the last lane of
//      a row holds r = c + sum_e w_e NN(pos_e) and adds 0.5 r^2 to the loss;
//   3. every lane back-propagates r * w_e through its kept activations into
//      per-lane gradient registers of both phase networks.
// Tiles of one long row are processed one chunk of 64 entries at a time (the
// forward pass is recomputed for the gradient).  The grid is persistent (at
// most one workgroup of 8 wavefronts per CU, wavefronts stride over the tiles),
// so the gradient registers are reduced once per wavefront: DPP wave sums, LDS
// across the 8 wavefronts, one partial row per workgroup.  In the fused mode
// the last workgroup to arrive reduces the partial rows (parameter-major, each
// wavefront a set of parameters, lanes over workgroups with coalesced loads,
// float64 in a fixed order) and applies the optimizer, so one launch is one
// optimizer step and results are bitwise reproducible.
//
// Reference: loss = mean(optax.l2_loss(lhs, rhs)) over the batch
// (jax_dips/solvers/poisson/trainer.py:892-912), value_and_grad and the optax
// update (trainer.py:783-789), DoubleMLP (nn/mlp/MLP.py:93-139).
#include <limits.h>

#include "dips_common.h"
#include "dips_hip.h"
#include "dips_mlp.h"
#include "dips_step.h"

namespace dips {

constexpr int kTileCap = 64;                          // entries and rows of a packed tile
constexpr int kMaxTileBlocks = 256;                   // one workgroup per CU
constexpr int kBlocksPerLane = kMaxTileBlocks / DIPS_WAVE;
constexpr int kReduceChunk = 8;                       // parameters per wavefront per load round
constexpr int kLdsStride = 68;                        // floats per staged operand row: 64 entries + 4
constexpr int kStageFloats = 2 * 16 * kLdsStride;     // A and B operand stacks of one wavefront

// Network shapes with a compiled tile kernel (Lm, Hm, Lp, Hp).
#define DIPS_TILE_SHAPES(X) \
  X(1, 1, 2, 10)            \
  X(1, 3, 2, 10)            \
  X(1, 1, 1, 1)             \
  X(2, 10, 2, 10)

struct TileArgs {
  const float* xs;
  const float* ys;
  const float* zs;
  int ny, nz, npts;
  float hx, hy, hz;
  const int32_t* row_point;
  const int32_t* row_ptr;
  const float* row_c;
  int64_t nrows;
  const uint32_t* ent_code;
  const float* ent_w;
  int64_t nnz;
  const int32_t* tile_row;
  const int32_t* ent_row;
  int64_t tile_begin;  // first boundary of the batch's tiles in tile_row
  int ntiles;
};

__device__ __forceinline__ int64_t clamp64(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The last workgroup of a fused launch: reduce the parameter-major partial rows
// (partial[p * nblk + b], nblk <= kMaxTileBlocks) in a fixed order and apply
// the optimizer to params / m / v.
template <int kStride, int kTileWaves>
__device__ void tile_reduce_apply(const float* partial, int nblk, const FusedOpt& fo, float* params) {
  constexpr int kTileThreads = kTileWaves * DIPS_WAVE;
  constexpr int PT = kStride - 1;
  constexpr int kPer = (kStride + kTileWaves - 1) / kTileWaves;
  __shared__ float gsh[kStride];
  __shared__ double dred[kTileWaves];
  __shared__ float gnorm_s;
  __shared__ int count_s;
  const int lane = threadIdx.x & (DIPS_WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / DIPS_WAVE);
  for (int q0 = 0; q0 < kPer; q0 += kReduceChunk) {
    float vals[kReduceChunk][kBlocksPerLane];
#pragma unroll
    for (int qq = 0; qq < kReduceChunk; ++qq) {
      const int p = wave + (q0 + qq) * kTileWaves;
#pragma unroll
      for (int i = 0; i < kBlocksPerLane; ++i) {
        const int b = lane + i * DIPS_WAVE;
        // plain, unconditional (index-clamped) loads after the acquire fence of
        // the caller, so that all loads of a round are in flight together
        const int pc = p < kStride ? p : kStride - 1;
        const int bc = b < nblk ? b : nblk - 1;
        const float val = partial[(int64_t)pc * nblk + bc];
        vals[qq][i] = b < nblk ? val : 0.0f;  // p >= kStride: a duplicate, never stored
      }
    }
#pragma unroll
    for (int qq = 0; qq < kReduceChunk; ++qq) {
      const int p = wave + (q0 + qq) * kTileWaves;
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < kBlocksPerLane; ++i) acc += (double)vals[qq][i];
      acc = dips_wave_sum_d(acc);
      if (lane == 0 && p < kStride) gsh[p] = (float)(acc * (double)fo.inv_b);
    }
  }
  __syncthreads();
  double sq = 0.0;
  for (int p = threadIdx.x; p < PT; p += kTileThreads) sq += (double)gsh[p] * (double)gsh[p];
  sq = dips_wave_sum_d(sq);
  if (lane == 0) dred[wave] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kTileWaves; ++w) t += dred[w];
    gnorm_s = (float)sqrt(t);
    count_s = *fo.count;
  }
  __syncthreads();
  const int c = count_s;
  const float gn = gnorm_s;
  const bool clip = (fo.opt.kind == 0) && !(gn < fo.opt.max_norm);
  const float lr_t = schedule_lr(fo.opt, c);
  const float t1 = (float)(c + 1);
  const float bc1 = 1.0f - powf(0.9f, t1);
  const float bc2 = 1.0f - powf(0.999f, t1);
  for (int p = threadIdx.x; p < PT; p += kTileThreads)
    apply_update(fo.opt, c, lr_t, bc1, bc2, clip, gn, p, gsh[p], params, fo.m, fo.v);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (c >= 0 && c < fo.loss_cap) fo.loss_hist[c] = gsh[PT];
    *fo.count = c + 1;
    *fo.sync = 0u;
  }
}

typedef float f4 __attribute__((ext_vector_type(4)));

// Weight gradients of a phase network with 8 <= H <= 12 accumulate on the
// matrix cores: per tile, dW = sum over the 64 entries of (input (x) delta),
// a K = 64 contraction of two 16-wide operand stacks staged through LDS:
//   block 0:  rows [x, y, z, 1, a_{L-1}]   cols [delta_0, seed]
//             -> W0 (rows 0-2), b0 (row 3), Wout (col H), bout (row 3, col H)
//   block l:  rows [a_{l-1}, 1]            cols [delta_l]
//             -> W_l (rows 0..H-1), b_l (row H)
// Smaller networks accumulate per lane in registers.
template <int H>
struct MfmaGrad {
  static constexpr bool value = H >= 8 && H <= 12;
};


// D += A^T B over the 64 staged entries: 16 MFMA 16x16x4 steps, step s taking
// entries 16*(lane/16) + s; operand rows are read as 4 float4 per lane.
__device__ __forceinline__ void stage_mfma(const float* AT, const float* BT, int lane, f4& acc) {
  const int m = lane & 15, kq = lane >> 4;
  const f4* ap = reinterpret_cast<const f4*>(AT + m * kLdsStride + 16 * kq);
  const f4* bp = reinterpret_cast<const f4*>(BT + m * kLdsStride + 16 * kq);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const f4 av = ap[q], bv = bp[q];
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bv.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bv.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bv.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bv.w, acc, 0, 0, 0);
  }
}

// Accumulate one chunk of 64 entries (lanes with on == false contribute zero)
// into the L accumulator blocks of a network.  Called in wave-uniform control flow.
template <int L, int H>
__device__ __forceinline__ void mfma_grad_chunk(float* AT, float* BT, int lane, bool on, float x, float y, float z,
                                                const float (&act)[L][H], const float (&dl)[L][H], float seed,
                                                f4 (&acc)[L]) {
  const float one = on ? 1.0f : 0.0f;
  AT[0 * kLdsStride + lane] = on ? x : 0.0f;
  AT[1 * kLdsStride + lane] = on ? y : 0.0f;
  AT[2 * kLdsStride + lane] = on ? z : 0.0f;
  AT[3 * kLdsStride + lane] = one;
#pragma unroll
  for (int i = 0; i < H; ++i) AT[(4 + i) * kLdsStride + lane] = on ? act[L - 1][i] : 0.0f;
#pragma unroll
  for (int j = 0; j < H; ++j) BT[j * kLdsStride + lane] = on ? dl[0][j] : 0.0f;
  BT[H * kLdsStride + lane] = on ? seed : 0.0f;
  __builtin_amdgcn_wave_barrier();
  stage_mfma(AT, BT, lane, acc[0]);
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int l = 1; l < L; ++l) {
#pragma unroll
    for (int i = 0; i < H; ++i) AT[i * kLdsStride + lane] = on ? act[l - 1][i] : 0.0f;
    AT[H * kLdsStride + lane] = one;
#pragma unroll
    for (int j = 0; j < H; ++j) BT[j * kLdsStride + lane] = on ? dl[l][j] : 0.0f;
    __builtin_amdgcn_wave_barrier();
    stage_mfma(AT, BT, lane, acc[l]);
    __builtin_amdgcn_wave_barrier();
  }
}

// Parameter index of element (i, j) of accumulator block blk, or -1 if unused.
template <int L, int H>
__device__ __forceinline__ int mfma_grad_param(int blk, int i, int j) {
  using NL = NetLayout<L, H>;
  if (blk == 0) {
    if (i < 3 && j < H) return NL::w(0) + i * H + j;
    if (i == 3 && j < H) return NL::b(0) + j;
    if (i >= 4 && i < 4 + H && j == H) return NL::wout() + (i - 4);
    if (i == 3 && j == H) return NL::bout();
    return -1;
  }
  if (i < H && j < H) return NL::w(blk) + i * H + j;
  if (i == H && j < H) return NL::b(blk) + j;
  return -1;
}

// One phase network's gradient accumulators: MFMA blocks or per-lane registers.
template <int L, int H, bool kGrad>
struct NetGrad {
  static constexpr bool kMfma = MfmaGrad<H>::value;
  static constexpr int P = NetLayout<L, H>::size();
  float g[kGrad && !kMfma ? P : 1];
  f4 acc[kGrad && kMfma ? L : 1];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < (kGrad && !kMfma ? P : 1); ++i) g[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < (kGrad && kMfma ? L : 1); ++i) acc[i] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  }

  // Gradient contribution of one chunk of entries; mine = this lane's entry
  // belongs to this network (act holds its activations).  Wave-uniform call.
  __device__ __forceinline__ void add(const float* P_, float* AT, float* BT, int lane, bool mine, float x, float y,
                                      float z, float seed, const float (&act)[L][H]) {
    // a fresh (opaque) weight pointer: the backward pass reloads the weights
    // instead of keeping the forward pass's loads live in vector registers
    const float* Q = P_;
    asm volatile("" : "+s"(Q));
    if constexpr (kGrad && kMfma) {
      float dl[L][H] = {};
      if (mine) mlp_deltas<L, H>(Q, seed, act, dl);
      mfma_grad_chunk<L, H>(AT, BT, lane, mine, x, y, z, act, dl, seed, acc);
    } else if constexpr (kGrad) {
      if (mine) mlp_backward_kept<L, H>(Q, x, y, z, seed, act, g);
    }
  }

  // Writes this wavefront's sums into red[off .. off + P).
  __device__ __forceinline__ void flush(float* red, int off, int lane) {
    if constexpr (kGrad && kMfma) {
#pragma unroll
      for (int blk = 0; blk < L; ++blk)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int p = mfma_grad_param<L, H>(blk, 4 * (lane >> 4) + r, lane & 15);
          if (p >= 0) red[off + p] = acc[blk][r];
        }
    } else if constexpr (kGrad) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const float s = dips_wave_sum(g[p]);
        if (lane == 0) red[off + p] = s;
      }
    }
  }
};

// Workgroup size per shape: 16 wavefronts (four per SIMD, 128 registers each)
// unless both networks have two or more hidden layers on the matrix cores,
// which need more registers per lane: then 8 wavefronts.
template <int Lm, int Hm, int Lp, int Hp>
struct TileCfg {
  static constexpr bool kWide = Lm >= 2 && Lp >= 2 && MfmaGrad<Hm>::value && MfmaGrad<Hp>::value;
  static constexpr int kThreads = kWide ? 512 : 1024;
  static constexpr int kWaves = kThreads / DIPS_WAVE;
  static constexpr size_t kLdsBytes = sizeof(float) * kStageFloats * kWaves;
};

template <int Lm, int Hm, int Lp, int Hp, bool kGrad, bool kFused>
__global__ __launch_bounds__((TileCfg<Lm, Hm, Lp, Hp>::kThreads)) void tile_step_kernel(
    TileArgs a, float* params, float* __restrict__ partial, FusedOpt fo) {
  constexpr int kTileThreads = TileCfg<Lm, Hm, Lp, Hp>::kThreads;
  constexpr int kTileWaves = TileCfg<Lm, Hm, Lp, Hp>::kWaves;
  using NP = NetLayout<Lp, Hp>;
  using NM = NetLayout<Lm, Hm>;
  constexpr int PP = NP::size();
  constexpr int PM = NM::size();
  constexpr int PT = PP + PM;
  extern __shared__ f4 tile_lds[];
  float* const lds = reinterpret_cast<float*>(tile_lds);

  const int lane = threadIdx.x & (DIPS_WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / DIPS_WAVE);
  float* const AT = lds + wave * kStageFloats;
  float* const BT = AT + 16 * kLdsStride;
  for (int i = lane; i < kStageFloats; i += DIPS_WAVE) AT[i] = 0.0f;

  NetGrad<Lp, Hp, kGrad> gp;
  NetGrad<Lm, Hm, kGrad> gm;
  gp.zero();
  gm.zero();
  float lsum = 0.0f;

  const int nw = (int)gridDim.x * kTileWaves;
  for (int t = (int)blockIdx.x * kTileWaves + wave; t < a.ntiles; t += nw) {
    // weights are re-read per tile (scalar loads that hit the scalar cache)
    // rather than held in vector registers across the loop
    const float* Pp = params;
    asm volatile("" : "+s"(Pp));
    const float* Pm = Pp + PP;
    const int64_t tb = a.tile_begin + t;
    const int64_t r0 = clamp64(a.tile_row[tb], 0, a.nrows);
    const int64_t r1 = clamp64(a.tile_row[tb + 1], r0, a.nrows);
    const int64_t e0 = clamp64(a.row_ptr[r0], 0, a.nnz);
    const int64_t e1 = clamp64(a.row_ptr[r1], e0, a.nnz);
    if (e1 - e0 <= kTileCap && r1 - r0 <= kTileCap) {
      const int n = (int)(e1 - e0);
      const int nr = (int)(r1 - r0);
      // rows without entries contribute their constant only
      if (lane < nr) {
        const int64_t row = r0 + lane;
        const int pt = a.row_point[row];
        if (a.row_ptr[row + 1] <= a.row_ptr[row] && pt >= 0 && pt < a.npts) {
          const float c = a.row_c[row];
          lsum = fmaf(0.5f * c, c, lsum);
        }
      }
      const bool active = lane < n;
      float v = 0.0f, w = 0.0f, c = 0.0f, x = 0.0f, y = 0.0f, z = 0.0f;
      int head = lane, last = lane;
      bool valid = false, plus = false;
      float actp[Lp][Hp] = {};
      float actm[Lm][Hm] = {};
      if (active) {
        const int64_t e = e0 + lane;
        const int64_t row = clamp64(a.ent_row[e], r0, r1 - 1);
        head = (int)clamp64((int64_t)a.row_ptr[row] - e0, 0, lane);
        last = (int)clamp64((int64_t)a.row_ptr[row + 1] - 1 - e0, lane, n - 1);
        const int pt = a.row_point[row];
        valid = pt >= 0 && pt < a.npts;
        if (valid) {
          const int k = pt % a.nz;
          const int ij = pt / a.nz;
          const int j = ij % a.ny;
          const int i = ij / a.ny;
          const uint32_t code = a.ent_code[e];
          w = a.ent_w[e];
          c = a.row_c[row];
          int ox, oy, oz;
          dips_decode_offset(code, ox, oy, oz);
          x = a.xs[i] + (float)ox * a.hx;
          y = a.ys[j] + (float)oy * a.hy;
          z = a.zs[k] + (float)oz * a.hz;
          plus = (code & 128u) != 0;
          const float u = plus ? mlp_forward_keep<Lp, Hp>(Pp, x, y, z, actp)
                               : mlp_forward_keep<Lm, Hm>(Pm, x, y, z, actm);
          v = w * u;
        }
      }
      // segmented inclusive scan: lane l ends with the sum of lanes head(l)..l
#pragma unroll
      for (int d = 1; d < DIPS_WAVE; d <<= 1) {
        const float o = __shfl_up(v, d, DIPS_WAVE);
        if (lane - d >= head) v += o;
      }
      const float tot = __shfl(v, last, DIPS_WAVE);
      const bool on = active && valid;
      const float r = tot + c;
      if (on && lane == last) lsum = fmaf(0.5f * r, r, lsum);
      const float seed = r * w;
      gp.add(Pp, AT, BT, lane, on && plus, x, y, z, seed, actp);
      gm.add(Pm, AT, BT, lane, on && !plus, x, y, z, seed, actm);
    }
  }
  // long rows (tiles skipped above): one wavefront per row, 64 entries per
  // chunk; the gradient pass recomputes the forward activations.  A loop of
  // its own so that its register demand does not add to the packed path's.
#pragma unroll 1
  for (int t = (int)blockIdx.x * kTileWaves + wave; t < a.ntiles; t += nw) {
    const int64_t tb = a.tile_begin + t;
    const int64_t r0 = clamp64(a.tile_row[tb], 0, a.nrows);
    const int64_t r1 = clamp64(a.tile_row[tb + 1], r0, a.nrows);
    const int64_t e0 = clamp64(a.row_ptr[r0], 0, a.nnz);
    const int64_t e1 = clamp64(a.row_ptr[r1], e0, a.nnz);
    if (e1 - e0 <= kTileCap && r1 - r0 <= kTileCap) continue;
    {
#pragma unroll 1
      for (int64_t row = r0; row < r1; ++row) {
        const int pt = a.row_point[row];
        if (pt < 0 || pt >= a.npts) continue;
        const int k = pt % a.nz;
        const int ij = pt / a.nz;
        const int j = ij % a.ny;
        const int i = ij / a.ny;
        const float px = a.xs[i], py = a.ys[j], pz = a.zs[k];
        const int64_t beg = clamp64(a.row_ptr[row], 0, a.nnz);
        const int64_t end = clamp64(a.row_ptr[row + 1], beg, a.nnz);
        float acc = 0.0f;
#pragma unroll 1
        for (int64_t e = beg + lane; e < end; e += DIPS_WAVE) {
          const float* Wp = params;  // opaque per chunk: no weights hoisted into registers
          asm volatile("" : "+s"(Wp));
          const uint32_t code = a.ent_code[e];
          int ox, oy, oz;
          dips_decode_offset(code, ox, oy, oz);
          const float x = px + (float)ox * a.hx;
          const float y = py + (float)oy * a.hy;
          const float z = pz + (float)oz * a.hz;
          const float u = (code & 128u) ? mlp_forward<Lp, Hp>(Wp, x, y, z) : mlp_forward<Lm, Hm>(Wp + PP, x, y, z);
          acc = fmaf(a.ent_w[e], u, acc);
        }
        const float r = dips_wave_sum(acc) + a.row_c[row];
        if (lane == 0) lsum = fmaf(0.5f * r, r, lsum);
        if (!kGrad) continue;
#pragma unroll 1
        for (int64_t base = beg; base < end; base += DIPS_WAVE) {
          const float* Wp = params;
          asm volatile("" : "+s"(Wp));
          const int64_t e = base + lane;
          const bool on = e < end;
          bool plus = false;
          float x = 0.0f, y = 0.0f, z = 0.0f, seed = 0.0f;
          float actp[Lp][Hp] = {};
          float actm[Lm][Hm] = {};
          if (on) {
            const uint32_t code = a.ent_code[e];
            int ox, oy, oz;
            dips_decode_offset(code, ox, oy, oz);
            x = px + (float)ox * a.hx;
            y = py + (float)oy * a.hy;
            z = pz + (float)oz * a.hz;
            seed = r * a.ent_w[e];
            plus = (code & 128u) != 0;
            if (plus)
              (void)mlp_forward_keep<Lp, Hp>(Wp, x, y, z, actp);
            else
              (void)mlp_forward_keep<Lm, Hm>(Wp + PP, x, y, z, actm);
          }
          gp.add(Wp, AT, BT, lane, on && plus, x, y, z, seed, actp);
          gm.add(Wp + PP, AT, BT, lane, on && !plus, x, y, z, seed, actm);
        }
      }
    }
  }

  // workgroup reduction: per-wavefront sums into LDS (the staging area is
  // reused), then the wavefronts are added in order
  constexpr int stride = PT + 1;
  __syncthreads();
  float* const red = lds;  // [kTileWaves][stride]
  gp.flush(red + wave * stride, 0, lane);
  gm.flush(red + wave * stride, PP, lane);
  {
    const float s = dips_wave_sum(lsum);
    if (lane == 0) red[wave * stride + PT] = s;
  }
  __syncthreads();
  const int nblk = (int)gridDim.x;
  for (int p = threadIdx.x; p < stride; p += kTileThreads) {
    float s = 0.0f;
    if (kGrad || p == PT) {
#pragma unroll
      for (int w = 0; w < kTileWaves; ++w) s += red[w * stride + p];
    }
    if (kFused)
      partial[(int64_t)p * nblk + blockIdx.x] = s;
    else
      partial[(int64_t)blockIdx.x * stride + p] = s;
  }
  if (!kFused) return;

  // One thread per workgroup fences and counts: the barrier completes the
  // workgroup's partial stores, the agent-scope release makes them visible to
  // the other XCDs (one L2 write-back per workgroup), and the last workgroup
  // to arrive acquires (invalidates its caches) before reading every row.
  __shared__ int am_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int last = (atomicAdd(fo.sync, 1u) == (uint32_t)(nblk - 1)) ? 1 : 0;
    if (last) __threadfence();
    am_last = last;
  }
  __syncthreads();
  if (!am_last) return;
  tile_reduce_apply<stride, kTileWaves>(partial, nblk, fo, params);
}

// workgroups of a launch over ntiles tiles with waves wavefronts per workgroup
static int tile_blocks(int64_t ntiles, int waves) {
  int64_t b = (ntiles + waves - 1) / waves;
  return (int)(b < 1 ? 1 : (b > kMaxTileBlocks ? kMaxTileBlocks : b));
}

int tile_supported(const dips_net_t* n) {
#define X(A, B, C, D) \
  if (n->Lm == A && n->Hm == B && n->Lp == C && n->Hp == D) return 1;
  DIPS_TILE_SHAPES(X)
#undef X
  return 0;
}

int64_t tile_partial_size(const dips_net_t* net) { return (int64_t)kMaxTileBlocks * partial_stride(net); }

template <int Lm, int Hm, int Lp, int Hp>
static void tile_launch_shape(const TileArgs& a, float* params, float* partial, int mode, const FusedOpt& fo,
                              int nblk, hipStream_t st) {
  dim3 grid(nblk), block(TileCfg<Lm, Hm, Lp, Hp>::kThreads);
  const size_t lds = TileCfg<Lm, Hm, Lp, Hp>::kLdsBytes;
  // the dynamic LDS limit is raised once per kernel instantiation
#define DIPS_TILE_LAUNCH(G, F)                                                                                   \
  static const bool lds_set_##G##F = hipFuncSetAttribute((const void*)tile_step_kernel<Lm, Hm, Lp, Hp, G, F>,  \
                                                         hipFuncAttributeMaxDynamicSharedMemorySize,            \
                                                         (int)lds) == hipSuccess;                               \
  (void)lds_set_##G##F;                                                                                          \
  hipLaunchKernelGGL((tile_step_kernel<Lm, Hm, Lp, Hp, G, F>), grid, block, lds, st, a, params, partial, fo)
  if (mode == 2) {
    DIPS_TILE_LAUNCH(true, true);
  } else if (mode == 1) {
    DIPS_TILE_LAUNCH(true, false);
  } else {
    DIPS_TILE_LAUNCH(false, false);
  }
#undef DIPS_TILE_LAUNCH
}

int tile_step_launch(const dips_grid_t* g, const dips_records_t* rec, const dips_tiles_t* tiles, int64_t tile_begin,
                     int64_t ntiles, const dips_net_t* net, const float* params, float* partial, int64_t partial_cap,
                     int want_grad, int* nblk_out, hipStream_t stream, const FusedOpt* fused) {
  if (!tiles || !tiles->tile_row || (rec->nnz > 0 && !tiles->ent_row)) {
    set_error("tile step: invalid tiles");
    return DIPS_EINVAL;
  }
  if (ntiles < 1 || ntiles > INT_MAX || tile_begin < 0 || tile_begin + ntiles + 1 > tiles->nbounds) {
    set_error("tile step: tile range outside the tile table");
    return DIPS_EINVAL;
  }
  if (rec->nrows >= INT_MAX || (int64_t)g->nx * g->ny * g->nz >= INT_MAX) {
    set_error("tile step: grid or records too large for 32-bit point indices");
    return DIPS_EINVAL;
  }
  TileArgs a{g->xs,         g->ys,          g->zs,       g->ny,          g->nz,       g->nx * g->ny * g->nz,
             rec->hx,       rec->hy,        rec->hz,     rec->row_point, rec->row_ptr, rec->row_c,
             rec->nrows,    rec->ent_code,  rec->ent_w,  rec->nnz,       tiles->tile_row, tiles->ent_row,
             tile_begin,    (int)ntiles};
  FusedOpt fo{};
  if (fused) fo = *fused;
  const int mode = fused ? 2 : (want_grad ? 1 : 0);
#define X(A, B, C, D)                                                                                   \
  if (net->Lm == A && net->Hm == B && net->Lp == C && net->Hp == D) {                                  \
    const int nblk = tile_blocks(ntiles, TileCfg<A, B, C, D>::kWaves);                                 \
    if (partial_cap < (int64_t)nblk * partial_stride(net)) {                                           \
      set_error("tile step: partial buffer too small");                                                \
      return DIPS_EINVAL;                                                                              \
    }                                                                                                  \
    *nblk_out = nblk;                                                                                  \
    tile_launch_shape<A, B, C, D>(a, const_cast<float*>(params), partial, mode, fo, nblk, stream);    \
    return hip_status(hipGetLastError(), "tile_step_kernel launch");                                   \
  }
  DIPS_TILE_SHAPES(X)
#undef X
  set_error("tile step: network shape has no compiled tile kernel");
  return DIPS_EUNSUPPORTED;
}

int tile_train_step(const dips_grid_t* g, const dips_records_t* rec, const dips_tiles_t* tiles, int64_t tile_begin,
                    int64_t ntiles, const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params,
                    float* m, float* v, int32_t* count, float* loss_hist, int loss_cap, float* partial,
                    int64_t partial_cap, uint32_t* sync, hipStream_t stream) {
  FusedOpt fo{*opt, inv_batch, m, v, count, loss_hist, loss_cap, sync};
  int nb = 0;
  return tile_step_launch(g, rec, tiles, tile_begin, ntiles, net, params, partial, partial_cap, 1, &nb, stream, &fo);
}

// ent_row[e] = row of entry e (one thread per row)
__global__ void entry_rows_kernel(const int32_t* __restrict__ row_ptr, int64_t nrows, int64_t nnz,
                                  int32_t* __restrict__ ent_row) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  const int64_t beg = clamp64(row_ptr[row], 0, nnz);
  const int64_t end = clamp64(row_ptr[row + 1], beg, nnz);
  for (int64_t e = beg; e < end; ++e) ent_row[e] = (int32_t)row;
}

int entry_rows_launch(const dips_records_t* rec, int32_t* ent_row, hipStream_t st) {
  if (rec->nrows >= INT_MAX) {
    set_error("dips_entry_rows: too many rows for 32-bit row indices");
    return DIPS_EINVAL;
  }
  const int64_t nb = (rec->nrows + 255) / 256;
  hipLaunchKernelGGL(entry_rows_kernel, dim3((unsigned)nb), dim3(256), 0, st, rec->row_ptr, rec->nrows, rec->nnz,
                     ent_row);
  return hip_status(hipGetLastError(), "entry_rows_kernel launch");
}

// Host: pack the rows of each batch into tiles (see the file comment).
int64_t pack_tiles(const int32_t* row_ptr, int64_t nrows, const int64_t* batch_row_begin, const int64_t* batch_rows,
                   int32_t nbatches, int32_t* tile_row, int64_t tile_cap, int64_t* batch_tile_begin) {
  if (!row_ptr || !batch_row_begin || !batch_rows || nbatches < 1 || nrows < 1 || nrows >= INT_MAX) {
    set_error("dips_pack_tiles: invalid arguments");
    return -1;
  }
  int64_t k = 0;
  bool overflow = false;
  auto emit = [&](int64_t r) {
    if (tile_row) {
      if (k < tile_cap)
        tile_row[k] = (int32_t)r;
      else
        overflow = true;
    }
    ++k;
  };
  for (int32_t b = 0; b < nbatches; ++b) {
    const int64_t begin = batch_row_begin[b], end = begin + batch_rows[b];
    if (batch_rows[b] < 1 || begin < 0 || end > nrows) {
      set_error("dips_pack_tiles: batch outside the records");
      return -1;
    }
    if (batch_tile_begin) batch_tile_begin[b] = k;
    int64_t r = begin;
    emit(r);
    while (r < end) {
      int64_t len = (int64_t)row_ptr[r + 1] - row_ptr[r];
      if (len < 0) {
        set_error("dips_pack_tiles: row offsets decrease");
        return -1;
      }
      if (len > kTileCap) {
        ++r;
        emit(r);
        continue;
      }
      int64_t n = 0;
      int rows = 0;
      while (r < end && rows < kTileCap) {
        len = (int64_t)row_ptr[r + 1] - row_ptr[r];
        if (len < 0) {
          set_error("dips_pack_tiles: row offsets decrease");
          return -1;
        }
        if (len > kTileCap || n + len > kTileCap) break;
        n += len;
        ++r;
        ++rows;
      }
      emit(r);
    }
  }
  if (batch_tile_begin) batch_tile_begin[nbatches] = k;
  if (overflow) {
    set_error("dips_pack_tiles: tile_row capacity too small");
    return -1;
  }
  return k;
}

}  // namespace dips
