// Entry-parallel fused training step over packed wavefront tiles (loss, gradient, optimizer).
//
// The rows of a batch are packed (on the host, once per zoom level) into tiles
// of whole rows holding at most 64 slots, one slot per entry and one for a row
// without entries; a row with more than 64 entries forms a tile of its own.
// Each tile is laid out once per level as 64 slots (SoA: position, weight, row
// constant at the row's last slot, head/last lane and phase), so a wavefront
// reads its tile with six coalesced loads and no dependent lookups:
//   1. every lane evaluates the phase network of its slot at the stored
//      position fl32(p + o*h) and keeps the hidden activations in registers;
//   2. a segmented inclusive scan over the lanes (bounded by the row's first
//      lane, fixed summation tree) forms r = c + sum_e w_e NN(pos_e) at the
//      row's last lane, which adds 0.5 r^2 to the loss and broadcasts r;
//   3. every lane back-propagates r * w_e through its kept activations.
// Weight gradients of networks of width 8..12 accumulate on the matrix cores
// (K = 64 contraction over the tile's slots), smaller networks per lane.
// Tiles of one long row are processed one chunk of 64 entries at a time from
// the CSR records (the forward pass is recomputed for the gradient).  The grid
// is persistent (one workgroup per CU, wavefronts stride over the tiles), so
// the gradients are reduced once per wavefront, then across the workgroup in
// LDS into one partial row.  In the fused mode the last workgroup to arrive
// reduces the partial rows in a fixed order (float64) and applies the
// optimizer: one launch is one optimizer step, bitwise reproducible.
//
// Reference: loss = mean(optax.l2_loss(lhs, rhs)) over the batch
// (jax_dips/solvers/poisson/trainer.py:892-912), value_and_grad and the optax
// update (trainer.py:783-789), DoubleMLP (nn/mlp/MLP.py:93-139).
#include <limits.h>

#include <algorithm>
#include <vector>

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

// Network shapes with a compiled tile kernel (Lm, Hm, Lp, Hp), preconditioner
// hidden widths (PC1, PC2; 0 = no preconditioner, PC2 = 0 = one hidden layer)
// and hidden activation (0 = tanh, 1 = CELU: the hash network's coordinate
// MLPs with the identity encoder, 3 -> [16] -> 1 in the reference's LPBE
// configuration, examples/benchmark_LPBE/conf/lpbe.yaml:48-60).  [8, 4] is the
// preconditioner of the reference's LPBE configuration
// (examples/benchmark_LPBE/conf/lpbe.yaml:63-67); its hidden activation is
// tanh whatever the networks' activation, so (1, 16, 1, 16, 8, 4, 1) is that
// configuration in full.
#define DIPS_TILE_SHAPES(X)  \
  X(1, 1, 2, 10, 0, 0, 0)    \
  X(1, 3, 2, 10, 0, 0, 0)    \
  X(1, 1, 1, 1, 0, 0, 0)     \
  X(2, 10, 2, 10, 0, 0, 0)   \
  X(2, 16, 2, 16, 0, 0, 0)   \
  X(1, 1, 2, 10, 8, 4, 0)    \
  X(1, 1, 2, 10, 8, 0, 0)    \
  X(2, 10, 2, 10, 8, 4, 0)   \
  X(1, 16, 1, 16, 0, 0, 1)   \
  X(2, 16, 2, 16, 0, 0, 1)   \
  X(1, 16, 1, 16, 8, 4, 1)

struct TileStepArgs {
  PointSet ps;
  float hx, hy, hz;
  const int32_t* row_point;
  const int32_t* row_ptr;
  const float* row_c;
  int64_t nrows;
  const uint32_t* ent_code;
  const float* ent_w;
  int64_t nnz;
  const int32_t* tile_row;
  const float* sx;  // slot arrays, 64 per tile boundary index
  const float* sy;
  const float* sz;
  const float* sw;
  const float* sc;
  const uint32_t* sinfo;
  const int32_t* srow;       // row of each used slot (preconditioner only)
  const float* row_coeffs;   // (nrows,26) cell coefficients (preconditioner only)
  float pc_scale;
  int64_t tile_begin;  // first boundary of the batch's tiles in tile_row
  int ntiles;
};

// slot_info bits: 0..5 first lane of the row, 6..11 last lane, 12 phase (+),
// 13 slot in use
constexpr uint32_t kSlotPlus = 1u << 12;
constexpr uint32_t kSlotUsed = 1u << 13;

__device__ __forceinline__ int64_t clamp64(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The last workgroup of a fused launch: reduce the parameter-major partial rows
// (partial[p * nblk + b], nblk <= kMaxTileBlocks) in a fixed order and apply
// the optimizer to params / m / v.
template <int kStride, int kTileWaves>
__device__ void tile_reduce_apply(const float* partial, int nblk, const FusedOpt& fo, float* params) {
  constexpr int kTileThreads = kTileWaves * DIPS_WAVE;
  constexpr int PT = kStride - 1;
  constexpr int kPer = (kStride + kTileWaves - 1) / kTileWaves;  // parameters per wavefront
  constexpr int kOwn = (PT + kTileThreads - 1) / kTileThreads;   // parameters updated per thread
  __shared__ float gsh[kStride];
  __shared__ double dred[kTileWaves];
  __shared__ float sched[3];
  const int lane = threadIdx.x & (DIPS_WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / DIPS_WAVE);
  // every global read is issued up front so that their latencies overlap: the
  // optimizer state of the parameters this thread updates, the step counter,
  // and the partial rows (plain, unconditional index-clamped loads after the
  // caller's acquire; each wavefront reduces a set of parameters, its lanes
  // stride over the workgroups)
  float pv[kOwn], mv[kOwn], vv[kOwn];
#pragma unroll
  for (int q = 0; q < kOwn; ++q) {
    const int p = threadIdx.x + q * kTileThreads;
    const int pc = p < PT ? p : PT - 1;
    pv[q] = gld(params, pc);
    mv[q] = gld(fo.m, pc);
    vv[q] = gld(fo.v, pc);
  }
  const int c = gld(fo.count, 0);
  if (wave == 0) {
    // step-dependent scalars, computed once (powf is long) while the partial
    // rows are loaded by the other wavefronts
    const float t1 = (float)(c + 1);
    if (lane == 0) {
      sched[0] = schedule_lr(fo.opt, c);
      sched[1] = 1.0f - powf(0.9f, t1);
      sched[2] = 1.0f - powf(0.999f, t1);
    }
  }
  // one round of loads when the wavefront's parameters fit (<= 32 per wavefront: the preconditioned
  // networks' 420 parameters load in one round instead of four), else rounds of 16 (the 707 of the 2x16
  // networks: three); the summation order is the same
  constexpr int kChunk = kPer <= 4 * kReduceChunk ? kPer : 2 * kReduceChunk;
  for (int q0 = 0; q0 < kPer; q0 += kChunk) {
    float vals[kChunk][kBlocksPerLane];
#pragma unroll
    for (int qq = 0; qq < kChunk; ++qq) {
      const int p = wave + (q0 + qq) * kTileWaves;
#pragma unroll
      for (int i = 0; i < kBlocksPerLane; ++i) {
        const int b = lane + i * DIPS_WAVE;
        const int pc = p < kStride ? p : kStride - 1;
        const int bc = b < nblk ? b : nblk - 1;
        const float val = gld(partial, (int64_t)pc * nblk + bc);
        vals[qq][i] = b < nblk ? val : 0.0f;  // p >= kStride: a duplicate, never stored
      }
    }
#pragma unroll
    for (int qq = 0; qq < kChunk; ++qq) {
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
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kTileWaves; ++w) t += dred[w];
  const float gn = (float)sqrt(t);
  const bool clip = (fo.opt.kind == 0) && !(gn < fo.opt.max_norm);
  const float lr_t = sched[0], bc1 = sched[1], bc2 = sched[2];
#pragma unroll
  for (int q = 0; q < kOwn; ++q) {
    const int p = threadIdx.x + q * kTileThreads;
    if (p < PT) {
      apply_update_value(fo.opt, lr_t, bc1, bc2, clip, gn, gsh[p], pv[q], mv[q], vv[q]);
      gst(params, p, pv[q]);
      gst(fo.m, p, mv[q]);
      gst(fo.v, p, vv[q]);
    }
  }
  if (threadIdx.x == 0) {
    if (c >= 0 && c < fo.loss_cap) gst(fo.loss_hist, c, gsh[PT]);
    gst(fo.count, 0, c + 1);
    if (fo.sync) *fo.sync = 0u;
  }
}

// Reduction of the parameter-major partial rows of a tile_step_kernel launch
// (kOut = 1) and the optimizer, as a kernel of its own: the kernel boundary
// orders the partial rows before their reads.
template <int kStride>
__global__ __launch_bounds__(1024) void tile_apply_kernel(const float* partial, int nblk, FusedOpt fo,
                                                         float* params) {
  tile_reduce_apply<kStride, 16>(partial, nblk, fo, params);
}

typedef float f4 __attribute__((ext_vector_type(4)));

// Weight gradients of a phase network with 8 <= H <= 12 accumulate on the
// matrix cores: per tile, dW = sum over the 64 entries of (input (x) delta),
// a K = 64 contraction of two 16-wide operand stacks staged through LDS:
//   block 0:  rows [x, y, z, 1, a_{L-1}]   cols [delta_0, seed]
//             -> W0 (rows 0-2), b0 (row 3), Wout (col H), bout (row 3, col H)
//   block l:  rows [a_{l-1}, 1]            cols [delta_l]
//             -> W_l (rows 0..H-1), b_l (row H)
// Smaller networks accumulate per lane in registers.  Width 13..16 leaves no
// operand row for a bias: the weight matrices go to the matrix cores
//   block 0:  rows [x, y, z, 1]   cols [delta_0]   -> W0, b0
//   block l:  rows [a_{l-1}]      cols [delta_l]   -> W_l
// and the biases b_l (l >= 1), Wout and bout accumulate per lane.
template <int H>
struct MfmaGrad {
  static constexpr bool value = H >= 8 && H <= 12;
};
template <int H>
struct MfmaGradWide {
  static constexpr bool value = H >= 13 && H <= 16;
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

// Accumulate one chunk of 64 entries into the L accumulator blocks of a
// network.  Lanes that do not contribute carry seed 0 (hence zero deltas) and
// finite activations and positions, so their products vanish without masking.
// Called in wave-uniform control flow.
template <int L, int H>
__device__ __forceinline__ void mfma_grad_chunk(float* AT, float* BT, int lane, float x, float y, float z,
                                                const float (&act)[L][H], const float (&dl)[L][H], float seed,
                                                f4 (&acc)[L]) {
  AT[0 * kLdsStride + lane] = x;
  AT[1 * kLdsStride + lane] = y;
  AT[2 * kLdsStride + lane] = z;
  AT[3 * kLdsStride + lane] = 1.0f;
#pragma unroll
  for (int i = 0; i < H; ++i) AT[(4 + i) * kLdsStride + lane] = act[L - 1][i];
#pragma unroll
  for (int j = 0; j < H; ++j) BT[j * kLdsStride + lane] = dl[0][j];
  BT[H * kLdsStride + lane] = seed;
  __builtin_amdgcn_wave_barrier();
  stage_mfma(AT, BT, lane, acc[0]);
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int l = 1; l < L; ++l) {
#pragma unroll
    for (int i = 0; i < H; ++i) AT[i * kLdsStride + lane] = act[l - 1][i];
    AT[H * kLdsStride + lane] = 1.0f;
#pragma unroll
    for (int j = 0; j < H; ++j) BT[j * kLdsStride + lane] = dl[l][j];
    __builtin_amdgcn_wave_barrier();
    stage_mfma(AT, BT, lane, acc[l]);
    __builtin_amdgcn_wave_barrier();
  }
}

// The matrix-core part of one chunk for width 13..16 (MfmaGradWide).
template <int L, int H>
__device__ __forceinline__ void mfma_grad_chunk_wide(float* AT, float* BT, int lane, float x, float y, float z,
                                                     const float (&act)[L][H], const float (&dl)[L][H],
                                                     f4 (&acc)[L]) {
  AT[0 * kLdsStride + lane] = x;
  AT[1 * kLdsStride + lane] = y;
  AT[2 * kLdsStride + lane] = z;
  AT[3 * kLdsStride + lane] = 1.0f;
#pragma unroll
  for (int j = 0; j < H; ++j) BT[j * kLdsStride + lane] = dl[0][j];
  __builtin_amdgcn_wave_barrier();
  stage_mfma(AT, BT, lane, acc[0]);
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int l = 1; l < L; ++l) {
#pragma unroll
    for (int i = 0; i < H; ++i) AT[i * kLdsStride + lane] = act[l - 1][i];
#pragma unroll
    for (int j = 0; j < H; ++j) BT[j * kLdsStride + lane] = dl[l][j];
    __builtin_amdgcn_wave_barrier();
    stage_mfma(AT, BT, lane, acc[l]);
    __builtin_amdgcn_wave_barrier();
  }
}

template <int L, int H>
__device__ __forceinline__ int mfma_grad_param_wide(int blk, int i, int j) {
  using NL = NetLayout<L, H>;
  if (j >= H) return -1;
  if (blk == 0) {
    if (i < 3) return NL::w(0) + i * H + j;
    return (i == 3) ? NL::b(0) + j : -1;
  }
  return (i < H) ? NL::w(blk) + i * H + j : -1;
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
template <int L, int H, bool kGrad, int Act = kActTanh>
struct NetGrad {
  static constexpr bool kWide = MfmaGradWide<H>::value;
  static constexpr bool kMfma = MfmaGrad<H>::value || kWide;
  static constexpr int P = NetLayout<L, H>::size();
  // per-lane sums of the wide layout: b_1 .. b_{L-1}, Wout, bout
  static constexpr int kLane = (L - 1) * H + H + 1;
  float g[kGrad && !kMfma ? P : (kGrad && kWide ? kLane : 1)];
  f4 acc[kGrad && kMfma ? L : 1];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < (kGrad && !kMfma ? P : (kGrad && kWide ? kLane : 1)); ++i) g[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < (kGrad && kMfma ? L : 1); ++i) acc[i] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  }

  // Gradient contribution of one chunk of entries; mine = this lane's entry
  // belongs to this network (act holds its activations; the activations of
  // other lanes must be finite, e.g. zero).  Wave-uniform call.
  __device__ __forceinline__ void add(dips_weight_t* P_, float* AT, float* BT, int lane, bool mine, float x, float y,
                                      float z, float seed, const float (&act)[L][H]) {
    // a fresh (opaque) weight pointer: the backward pass reloads the weights
    // instead of keeping the forward pass's loads live in registers
    dips_weight_t* Q = P_;
    asm volatile("" : "+s"(Q));
    // a chunk without entries of this network adds exact zeros: skip it (most
    // tiles hold one phase only)
    if constexpr (kGrad && kMfma) {
      if (!__any(mine)) return;
    }
    if constexpr (kGrad && kWide) {
      const float s = mine ? seed : 0.0f;
      float dl[L][H];
      mlp_deltas<L, H, Act>(Q, s, act, dl);
#pragma unroll
      for (int l = 1; l < L; ++l)
#pragma unroll
        for (int j = 0; j < H; ++j) g[(l - 1) * H + j] += dl[l][j];
#pragma unroll
      for (int i = 0; i < H; ++i) g[(L - 1) * H + i] = fmaf(s, act[L - 1][i], g[(L - 1) * H + i]);
      g[kLane - 1] += s;
      mfma_grad_chunk_wide<L, H>(AT, BT, lane, x, y, z, act, dl, acc);
    } else if constexpr (kGrad && kMfma) {
      const float s = mine ? seed : 0.0f;
      float dl[L][H];
      mlp_deltas<L, H, Act>(Q, s, act, dl);
      mfma_grad_chunk<L, H>(AT, BT, lane, x, y, z, act, dl, s, acc);
    } else if constexpr (kGrad) {
      if (mine) mlp_backward_kept<L, H, Act>(Q, x, y, z, seed, act, g);
    }
  }

  // Writes this wavefront's sums into red[off .. off + P).
  __device__ __forceinline__ void flush(float* red, int off, int lane) {
    if constexpr (kGrad && kWide) {
      using NL = NetLayout<L, H>;
#pragma unroll
      for (int blk = 0; blk < L; ++blk)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int p = mfma_grad_param_wide<L, H>(blk, 4 * (lane >> 4) + r, lane & 15);
          if (p >= 0) red[off + p] = acc[blk][r];
        }
#pragma unroll
      for (int q = 0; q < kLane; ++q) {
        const float s = dips_wave_sum(g[q]);
        const int p = q < (L - 1) * H ? NL::b(1 + q / H) + q % H : NL::wout() + (q - (L - 1) * H);
        if (lane == 0) red[off + p] = s;
      }
    } else if constexpr (kGrad && kMfma) {
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

// Learned preconditioner (nn/preconditioner.py:10-35): P = 0.5 + s*sigmoid(z),
// z = MLP(c) on the 26 cell coefficients of the row, hidden widths PC1 and
// (PC2 > 0) PC2, tanh.  Evaluated at the row's last lane; its weight gradients
// accumulate on the matrix cores in three 16x16 blocks (K = the tile's slots):
//   0: rows c[0..15]            cols d1          -> W1 rows 0..15
//   1: rows [c[16..25], 1]      cols d1          -> W1 rows 16..25, b1
//   2: rows [a1, 1, a2, 1]      cols [d2, dz]    -> W2, b2, Wout, bout  (PC2 > 0)
//      rows [a1, 1]             cols [dz]        -> Wout, bout          (PC2 = 0)
template <int PC1, int PC2>
struct Precond {
  static constexpr int kIn = 26;
  static constexpr int kLast = PC2 > 0 ? PC2 : PC1;
  static constexpr int w1() { return 0; }
  static constexpr int b1() { return kIn * PC1; }
  static constexpr int w2() { return b1() + PC1; }
  static constexpr int b2() { return w2() + PC1 * PC2; }
  static constexpr int wout() { return PC2 > 0 ? b2() + PC2 : b1() + PC1; }
  static constexpr int bout() { return wout() + kLast; }
  static constexpr int size() { return bout() + 1; }
  static_assert(PC1 == 0 || (PC1 <= 15 && PC2 <= 14 && PC1 + PC2 + 2 <= 16), "preconditioner too wide for the staging");

  template <typename WP>
  __device__ __forceinline__ static float forward(WP W, const float (&c)[kIn], float (&a1)[PC1 > 0 ? PC1 : 1],
                                                  float (&a2)[PC2 > 0 ? PC2 : 1], float& sig) {
#pragma unroll
    for (int j = 0; j < PC1; ++j) {
      float s = W[b1() + j];
#pragma unroll
      for (int i = 0; i < kIn; ++i) s = fmaf(c[i], W[w1() + i * PC1 + j], s);
      a1[j] = dips_tanh(s);
    }
    float z = W[bout()];
    if constexpr (PC2 > 0) {
#pragma unroll
      for (int j = 0; j < PC2; ++j) {
        float s = W[b2() + j];
#pragma unroll
        for (int i = 0; i < PC1; ++i) s = fmaf(a1[i], W[w2() + i * PC2 + j], s);
        a2[j] = dips_tanh(s);
      }
#pragma unroll
      for (int i = 0; i < PC2; ++i) z = fmaf(a2[i], W[wout() + i], z);
    } else {
#pragma unroll
      for (int i = 0; i < PC1; ++i) z = fmaf(a1[i], W[wout() + i], z);
    }
    sig = 1.0f / (1.0f + expf(-z));
    return sig;
  }

  template <typename WP>
  __device__ __forceinline__ static void deltas(WP W, float dz, const float (&a1)[PC1 > 0 ? PC1 : 1],
                                                const float (&a2)[PC2 > 0 ? PC2 : 1], float (&d1)[PC1 > 0 ? PC1 : 1],
                                                float (&d2)[PC2 > 0 ? PC2 : 1]) {
    if constexpr (PC2 > 0) {
#pragma unroll
      for (int j = 0; j < PC2; ++j) d2[j] = dz * W[wout() + j] * (1.0f - a2[j] * a2[j]);
#pragma unroll
      for (int i = 0; i < PC1; ++i) {
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < PC2; ++j) acc = fmaf(W[w2() + i * PC2 + j], d2[j], acc);
        d1[i] = acc * (1.0f - a1[i] * a1[i]);
      }
    } else {
#pragma unroll
      for (int i = 0; i < PC1; ++i) d1[i] = dz * W[wout() + i] * (1.0f - a1[i] * a1[i]);
    }
  }

  // parameter index (relative to the preconditioner block) of element (i, j) of block blk, or -1
  __device__ __forceinline__ static int param(int blk, int i, int j) {
    if (blk == 0) return (j < PC1) ? w1() + i * PC1 + j : -1;
    if (blk == 1) {
      if (j >= PC1) return -1;
      if (i < 10) return w1() + (16 + i) * PC1 + j;
      return (i == 10) ? b1() + j : -1;
    }
    if constexpr (PC2 > 0) {
      if (i < PC1 && j < PC2) return w2() + i * PC2 + j;
      if (i == PC1 && j < PC2) return b2() + j;
      if (i > PC1 && i <= PC1 + PC2 && j == PC2) return wout() + (i - PC1 - 1);
      if (i == PC1 + PC2 + 1 && j == PC2) return bout();
      return -1;
    } else {
      if (j != 0) return -1;
      if (i < PC1) return wout() + i;
      return (i == PC1) ? bout() : -1;
    }
  }
};

// Gradient accumulators of the preconditioner (three MFMA blocks).
template <int PC1, int PC2, bool kGrad>
struct PrecondGrad {
  static constexpr bool kOn = kGrad && PC1 > 0;
  using PN = Precond<(PC1 > 0 ? PC1 : 1), PC2>;
  f4 acc[kOn ? 3 : 1];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < (kOn ? 3 : 1); ++i) acc[i] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  }

  // One chunk of lanes: lanes with dz != 0 contribute (their c, a1, a2 finite).
  // Wave-uniform call.
  template <typename WP>
  __device__ __forceinline__ void add(WP W, float* AT, float* BT, int lane, float dz, const float (&c)[26],
                                      const float (&a1)[PC1 > 0 ? PC1 : 1], const float (&a2)[PC2 > 0 ? PC2 : 1]) {
    if constexpr (kOn) {
      float d1[PC1];
      float d2[PC2 > 0 ? PC2 : 1];
      PN::deltas(W, dz, a1, a2, d1, d2);
      // block 0
#pragma unroll
      for (int i = 0; i < 16; ++i) AT[i * kLdsStride + lane] = c[i];
#pragma unroll
      for (int j = 0; j < PC1; ++j) BT[j * kLdsStride + lane] = d1[j];
      __builtin_amdgcn_wave_barrier();
      stage_mfma(AT, BT, lane, acc[0]);
      __builtin_amdgcn_wave_barrier();
      // block 1 (same B operand)
#pragma unroll
      for (int i = 0; i < 10; ++i) AT[i * kLdsStride + lane] = c[16 + i];
      AT[10 * kLdsStride + lane] = 1.0f;
      __builtin_amdgcn_wave_barrier();
      stage_mfma(AT, BT, lane, acc[1]);
      __builtin_amdgcn_wave_barrier();
      // block 2
#pragma unroll
      for (int i = 0; i < PC1; ++i) AT[i * kLdsStride + lane] = a1[i];
      AT[PC1 * kLdsStride + lane] = 1.0f;
      if constexpr (PC2 > 0) {
#pragma unroll
        for (int i = 0; i < PC2; ++i) AT[(PC1 + 1 + i) * kLdsStride + lane] = a2[i];
        AT[(PC1 + PC2 + 1) * kLdsStride + lane] = 1.0f;
#pragma unroll
        for (int j = 0; j < PC2; ++j) BT[j * kLdsStride + lane] = d2[j];
        BT[PC2 * kLdsStride + lane] = dz;
      } else {
        BT[0 * kLdsStride + lane] = dz;
      }
      __builtin_amdgcn_wave_barrier();
      stage_mfma(AT, BT, lane, acc[2]);
      __builtin_amdgcn_wave_barrier();
    }
  }

  __device__ __forceinline__ void flush(float* red, int off, int lane) {
    if constexpr (kOn) {
#pragma unroll
      for (int blk = 0; blk < 3; ++blk)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int p = PN::param(blk, 4 * (lane >> 4) + r, lane & 15);
          if (p >= 0) red[off + p] = acc[blk][r];
        }
    }
  }
};

// Workgroup size per shape: 16 wavefronts (four per SIMD, 128 registers each)
// unless both networks have two or more hidden layers on the matrix cores,
// which need more registers per lane: then 8 wavefronts.
template <int Lm, int Hm, int Lp, int Hp, int PC1 = 0>
struct TileCfg {
  static constexpr bool kWide = (Lm >= 2 && Lp >= 2 && MfmaGrad<Hm>::value && MfmaGrad<Hp>::value) || PC1 > 0 ||
                                MfmaGradWide<Hm>::value || MfmaGradWide<Hp>::value;
  static constexpr int kThreads = kWide ? 512 : 1024;
  static constexpr int kWaves = kThreads / DIPS_WAVE;
  static constexpr size_t kLdsBytes = sizeof(float) * kStageFloats * kWaves;
};

// kOut: 0 = block-major partial rows (for dips_opt_step), 1 = parameter-major
// rows (for tile_apply_kernel), 2 = parameter-major rows and the fused
// last-workgroup reduction + optimizer.
//
// With a preconditioner (PC1 > 0) the row's residual is r = P * r0 with
// P = 0.5 + s*sigmoid(MLP(cell coefficients)) evaluated at the row's last
// lane (trainer.py:845-847, discretization.py:418-419); the entries' seeds
// become r * P * w and the preconditioner's parameters (after both networks)
// receive r * r0 * dP/dtheta.
//
// The body of one workgroup: block blk of nblk workgroups over the tiles of
// a.  Writes the workgroup's partial row (kOut 0: partial[blk * stride + p],
// else partial[p * nblk + blk]).
template <int Lm, int Hm, int Lp, int Hp, bool kGrad, int kOut, int PC1, int PC2, int Act>
__device__ __forceinline__ void tile_step_body(const TileStepArgs& a, float* params, float* __restrict__ partial,
                                               int blk, int nblk) {
  constexpr int kTileThreads = TileCfg<Lm, Hm, Lp, Hp, PC1>::kThreads;
  constexpr int kTileWaves = TileCfg<Lm, Hm, Lp, Hp, PC1>::kWaves;
  using NP = NetLayout<Lp, Hp>;
  using NM = NetLayout<Lm, Hm>;
  using PN = Precond<(PC1 > 0 ? PC1 : 1), PC2>;
  constexpr int PP = NP::size();
  constexpr int PM = NM::size();
  constexpr int PC = PC1 > 0 ? PN::size() : 0;
  constexpr int PT = PP + PM + PC;
  constexpr bool kSameShape = Lm == Lp && Hm == Hp;
  extern __shared__ f4 tile_lds[];
  float* const lds = reinterpret_cast<float*>(tile_lds);

  const int lane = threadIdx.x & (DIPS_WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / DIPS_WAVE);
  float* const AT = lds + wave * kStageFloats;
  float* const BT = AT + 16 * kLdsStride;
  // the staging rows a block does not write hold stale values; they only feed
  // accumulator entries that are never flushed (no zero fill needed)

  NetGrad<Lp, Hp, kGrad, Act> gp;
  NetGrad<Lm, Hm, kGrad, Act> gm;
  PrecondGrad<PC1, PC2, kGrad> gc;
  gp.zero();
  gm.zero();
  gc.zero();
  float lsum = 0.0f;

  const int nw = nblk * kTileWaves;
  const int t0 = blk * kTileWaves + wave;
  // slot data of the wavefront's next tile is loaded one tile ahead (the last
  // prefetch re-reads the final tile, unused)
  auto slot_of = [&](int t) { return (a.tile_begin + min(t, a.ntiles - 1)) * kTileCap + lane; };
  int64_t s0 = slot_of(t0);
  uint32_t n_info = gld(a.sinfo, s0);
  float n_x = gld(a.sx, s0), n_y = gld(a.sy, s0), n_z = gld(a.sz, s0), n_w = gld(a.sw, s0), n_c = gld(a.sc, s0);
  // tiles without a used slot (a single row of more than 64 entries) are
  // marked here, by the wavefront's iteration index, for the second loop
  uint64_t longmask = 0;
  int iter = 0;
  for (int t = t0; t < a.ntiles; t += nw, ++iter) {
    // weights are re-read per tile (scalar loads that hit the scalar cache)
    // rather than held in vector registers across the loop
    const float* Praw = params;
    asm volatile("" : "+s"(Praw));
    dips_weight_t* Pp = dips_weights(Praw);
    dips_weight_t* Pm = Pp + PP;
    const uint32_t info = n_info;
    const float x = n_x, y = n_y, z = n_z, w = n_w, c = n_c;
    {
      const int64_t sn = slot_of(t + nw);
      n_info = gld(a.sinfo, sn);
      n_x = gld(a.sx, sn);
      n_y = gld(a.sy, sn);
      n_z = gld(a.sz, sn);
      n_w = gld(a.sw, sn);
      n_c = gld(a.sc, sn);
    }
    const bool on = (info & kSlotUsed) != 0;
    if (!__any(on)) {
      if (iter < 64) longmask |= 1ull << iter;
      continue;
    }
    const bool plus = (info & kSlotPlus) != 0;
    const int head = on ? (int)(info & 63u) : lane;
    const int last = on ? (int)((info >> 6) & 63u) : lane;
    float actp[Lp][Hp] = {};
    float actm[Lm][Hm] = {};
    float v = 0.0f;
    if constexpr (kSameShape) {
      // the two phase networks in separate branches: merged into one call
      // they would read the weights through a per-lane pointer
      float up = 0.0f, um = 0.0f;
      if (on && plus) up = mlp_forward_keep<Lp, Hp, Act>(Pp, x, y, z, actp);
      asm volatile("" : "+s"(Pm));
      if (on && !plus) um = mlp_forward_keep<Lm, Hm, Act>(Pm, x, y, z, actm);
      if (on) v = fmaf(w, plus ? up : um, c);
    } else if (on) {
      const float u = plus ? mlp_forward_keep<Lp, Hp, Act>(Pp, x, y, z, actp) : mlp_forward_keep<Lm, Hm, Act>(Pm, x, y, z, actm);
      v = fmaf(w, u, c);  // c is nonzero only in the row's last slot
    }
    // segmented inclusive scan: lane l ends with the sum of lanes head(l)..l
    v = dips_segmented_scan(v, lane, head);
    float r = __shfl(v, last, DIPS_WAVE);
    float rmul = r;  // d(0.5 r^2)/d(sum w u): r, times P when preconditioned
    if constexpr (PC1 > 0) {
      // the preconditioner at each row's last lane (other lanes: zero inputs,
      // finite activations, zero seed)
      const bool is_last = on && lane == last;
      float cf[26] = {};
      if (is_last) {
        const int64_t row = clamp64(gld(a.srow, slot_of(t)), 0, a.nrows - 1);
        const float* cr = a.row_coeffs + row * 26;
#pragma unroll
        for (int i = 0; i < 26; ++i) cf[i] = gld(cr, i);
      }
      dips_weight_t* Pc = Pp + PP + PM;
      float a1[PC1], a2[PC2 > 0 ? PC2 : 1], sig;
      PN::forward(Pc, cf, a1, a2, sig);
      const float pre = __shfl(0.5f + a.pc_scale * sig, last, DIPS_WAVE);
      const float r0 = r;
      r = pre * r0;
      rmul = r * pre;
      if constexpr (kGrad) {
        const float dz = is_last ? r * r0 * a.pc_scale * sig * (1.0f - sig) : 0.0f;
        gc.add(Pc, AT, BT, lane, dz, cf, a1, a2);
      }
    }
    if (on && lane == last) lsum = fmaf(0.5f * r, r, lsum);
    const float seed = rmul * w;
    gp.add(Pp, AT, BT, lane, on && plus, x, y, z, seed, actp);
    gm.add(Pm, AT, BT, lane, on && !plus, x, y, z, seed, actm);
  }
  // long rows (tiles skipped above): one wavefront per row, 64 entries per
  // chunk; the gradient pass recomputes the forward activations.  A loop of
  // its own so that its register demand does not add to the packed path's.
  iter = 0;
#pragma unroll 1
  for (int t = t0; t < a.ntiles; t += nw, ++iter) {
    if (iter < 64 && !((longmask >> iter) & 1)) continue;  // a packed tile, done above
    const int64_t tb = a.tile_begin + t;
    const int64_t r0 = clamp64(gld(a.tile_row, tb), 0, a.nrows);
    const int64_t r1 = clamp64(gld(a.tile_row, tb + 1), r0, a.nrows);
    // the tiles without slots: more than 64 rows or slots (tile_slots_kernel)
    if (r1 - r0 <= kTileCap) {
      int len = 0;
      if (lane < r1 - r0) len = max(gld(a.row_ptr, r0 + lane + 1) - gld(a.row_ptr, r0 + lane), 1);
      if (dips_wave_sum_i(len) <= kTileCap) continue;
    }
    {
#pragma unroll 1
      for (int64_t row = r0; row < r1; ++row) {
        const int pt = gld(a.row_point, row);
        float px, py, pz;
        if (!point_xyz(a.ps, pt, px, py, pz)) continue;
        const int64_t beg = clamp64(gld(a.row_ptr, row), 0, a.nnz);
        const int64_t end = clamp64(gld(a.row_ptr, row + 1), beg, a.nnz);
        float acc = 0.0f;
#pragma unroll 1
        for (int64_t e = beg + lane; e < end; e += DIPS_WAVE) {
          const float* Wraw = params;  // opaque per chunk: no weights hoisted into registers
          asm volatile("" : "+s"(Wraw));
          dips_weight_t* Wp = dips_weights(Wraw);
          const uint32_t code = gld(a.ent_code, e);
          int ox, oy, oz;
          dips_decode_offset(code, ox, oy, oz);
          const float x = px + (float)ox * a.hx;
          const float y = py + (float)oy * a.hy;
          const float z = pz + (float)oz * a.hz;
          const bool pl = (code & 128u) != 0;
          float u;
          if constexpr (kSameShape) {
            dips_weight_t* Wm = Wp + PP;
            float up = 0.0f, um = 0.0f;
            if (pl) up = mlp_forward<Lp, Hp, Act>(Wp, x, y, z);
            asm volatile("" : "+s"(Wm));
            if (!pl) um = mlp_forward<Lm, Hm, Act>(Wm, x, y, z);
            u = pl ? up : um;
          } else {
            u = pl ? mlp_forward<Lp, Hp, Act>(Wp, x, y, z) : mlp_forward<Lm, Hm, Act>(Wp + PP, x, y, z);
          }
          acc = fmaf(gld(a.ent_w, e), u, acc);
        }
        float r = dips_wave_sum(acc) + gld(a.row_c, row);
        float rmul = r;
        if constexpr (PC1 > 0) {
          // every lane evaluates the row's preconditioner; lane 0 carries its seed
          const float* Wraw = params;
          asm volatile("" : "+s"(Wraw));
          dips_weight_t* Pc = dips_weights(Wraw) + PP + PM;
          float cf[26];
          const float* cr = a.row_coeffs + row * 26;
#pragma unroll
          for (int i = 0; i < 26; ++i) cf[i] = gld(cr, i);
          float a1[PC1], a2[PC2 > 0 ? PC2 : 1], sig;
          PN::forward(Pc, cf, a1, a2, sig);
          const float pre = 0.5f + a.pc_scale * sig;
          const float r0 = r;
          r = pre * r0;
          rmul = r * pre;
          if constexpr (kGrad) {
            const float dz = lane == 0 ? r * r0 * a.pc_scale * sig * (1.0f - sig) : 0.0f;
            gc.add(Pc, AT, BT, lane, dz, cf, a1, a2);
          }
        }
        if (lane == 0) lsum = fmaf(0.5f * r, r, lsum);
        if (!kGrad) continue;
#pragma unroll 1
        for (int64_t base = beg; base < end; base += DIPS_WAVE) {
          const float* Wraw = params;
          asm volatile("" : "+s"(Wraw));
          dips_weight_t* Wp = dips_weights(Wraw);
          const int64_t e = base + lane;
          const bool on = e < end;
          bool plus = false;
          float x = 0.0f, y = 0.0f, z = 0.0f, seed = 0.0f;
          float actp[Lp][Hp] = {};
          float actm[Lm][Hm] = {};
          if (on) {
            const uint32_t code = gld(a.ent_code, e);
            int ox, oy, oz;
            dips_decode_offset(code, ox, oy, oz);
            x = px + (float)ox * a.hx;
            y = py + (float)oy * a.hy;
            z = pz + (float)oz * a.hz;
            seed = rmul * gld(a.ent_w, e);
            plus = (code & 128u) != 0;
          }
          dips_weight_t* Wm = Wp + PP;
          if constexpr (kSameShape) {
            if (on && plus) (void)mlp_forward_keep<Lp, Hp, Act>(Wp, x, y, z, actp);
            asm volatile("" : "+s"(Wm));
            if (on && !plus) (void)mlp_forward_keep<Lm, Hm, Act>(Wm, x, y, z, actm);
          } else if (on) {
            if (plus)
              (void)mlp_forward_keep<Lp, Hp, Act>(Wp, x, y, z, actp);
            else
              (void)mlp_forward_keep<Lm, Hm, Act>(Wm, x, y, z, actm);
          }
          gp.add(Wp, AT, BT, lane, on && plus, x, y, z, seed, actp);
          gm.add(Wm, AT, BT, lane, on && !plus, x, y, z, seed, actm);
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
  gc.flush(red + wave * stride, PP + PM, lane);
  {
    const float s = dips_wave_sum(lsum);
    if (lane == 0) red[wave * stride + PT] = s;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < stride; p += kTileThreads) {
    float s = 0.0f;
    if (kGrad || p == PT) {
#pragma unroll
      for (int w = 0; w < kTileWaves; ++w) s += red[w * stride + p];
    }
    if (kOut == 0)
      gst(partial, (int64_t)blk * stride + p, s);
    else
      gst(partial, (int64_t)p * nblk + blk, s);
  }
}

template <int Lm, int Hm, int Lp, int Hp, int PC1, int PC2>
constexpr int tile_stride() {
  return NetLayout<Lp, Hp>::size() + NetLayout<Lm, Hm>::size() +
         (PC1 > 0 ? Precond<(PC1 > 0 ? PC1 : 1), PC2>::size() : 0) + 1;
}

template <int Lm, int Hm, int Lp, int Hp, bool kGrad, int kOut, int PC1 = 0, int PC2 = 0, int Act = kActTanh>
__global__ __launch_bounds__((TileCfg<Lm, Hm, Lp, Hp, PC1>::kThreads)) void tile_step_kernel(
    TileStepArgs a, float* params, float* __restrict__ partial, FusedOpt fo) {
  constexpr int kTileWaves = TileCfg<Lm, Hm, Lp, Hp, PC1>::kWaves;
  constexpr int stride = tile_stride<Lm, Hm, Lp, Hp, PC1, PC2>();
  const int nblk = (int)gridDim.x;
  tile_step_body<Lm, Hm, Lp, Hp, kGrad, kOut, PC1, PC2, Act>(a, params, partial, (int)blockIdx.x, nblk);
  if (kOut != 2) return;

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
__host__ __device__ static inline int tile_blocks(int64_t ntiles, int waves) {
  int64_t b = (ntiles + waves - 1) / waves;
  return (int)(b < 1 ? 1 : (b > kMaxTileBlocks ? kMaxTileBlocks : b));
}

static bool tile_shape_is(const dips_net_t* n, int Lm, int Hm, int Lp, int Hp, int pc1, int pc2, int act) {
  const int w1 = n->pc_w1 > 0 ? n->pc_w1 : 0, w2 = n->pc_w1 > 0 ? n->pc_w2 : 0;
  return n->Lm == Lm && n->Hm == Hm && n->Lp == Lp && n->Hp == Hp && w1 == pc1 && w2 == pc2 && n->act == act;
}

int tile_supported(const dips_net_t* n) {
#define X(A, B, C, D, E, F, G) \
  if (tile_shape_is(n, A, B, C, D, E, F, G)) return 1;
  DIPS_TILE_SHAPES(X)
#undef X
  return 0;
}

int64_t tile_partial_size(const dips_net_t* net) { return (int64_t)kMaxTileBlocks * partial_stride(net); }

template <int Lm, int Hm, int Lp, int Hp, int PC1, int PC2, int Act>
static void tile_launch_shape(const TileStepArgs& a, float* params, float* partial, int mode, const FusedOpt& fo,
                              int nblk, hipStream_t st) {
  dim3 grid(nblk), block(TileCfg<Lm, Hm, Lp, Hp, PC1>::kThreads);
  const size_t lds = TileCfg<Lm, Hm, Lp, Hp, PC1>::kLdsBytes;
  // the dynamic LDS limit is raised once per kernel instantiation
#define DIPS_TILE_LAUNCH(G, F)                                                                                   \
  static const bool lds_set_##G##F =                                                                             \
      hipFuncSetAttribute((const void*)tile_step_kernel<Lm, Hm, Lp, Hp, G, F, PC1, PC2, Act>,                    \
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;                   \
  (void)lds_set_##G##F;                                                                                          \
  hipLaunchKernelGGL((tile_step_kernel<Lm, Hm, Lp, Hp, G, F, PC1, PC2, Act>), grid, block, lds, st, a, params,      \
                     partial, fo)
  if (mode == 2) {
    DIPS_TILE_LAUNCH(true, 2);
  } else if (mode == 3) {
    DIPS_TILE_LAUNCH(true, 1);
    constexpr int kStride =
        NetLayout<Lp, Hp>::size() + NetLayout<Lm, Hm>::size() + (PC1 > 0 ? Precond<(PC1 > 0 ? PC1 : 1), PC2>::size() : 0) + 1;
    hipLaunchKernelGGL((tile_apply_kernel<kStride>), dim3(1), dim3(1024), 0, st, partial, nblk, fo, params);
  } else if (mode == 1) {
    DIPS_TILE_LAUNCH(true, 0);
  } else {
    DIPS_TILE_LAUNCH(false, 0);
  }
#undef DIPS_TILE_LAUNCH
}

int tile_step_launch(const dips_grid_t* g, const dips_records_t* rec, const dips_tiles_t* tiles, int64_t tile_begin,
                     int64_t ntiles, const dips_net_t* net, const float* params, float* partial, int64_t partial_cap,
                     int want_grad, int* nblk_out, hipStream_t stream, const FusedOpt* fused) {
  if (!tiles || !tiles->tile_row || !tiles->slot_x || !tiles->slot_y || !tiles->slot_z || !tiles->slot_w ||
      !tiles->slot_c || !tiles->slot_info) {
    set_error("tile step: invalid tiles");
    return DIPS_EINVAL;
  }
  if (ntiles < 1 || ntiles > INT_MAX || tile_begin < 0 || tile_begin + ntiles + 1 > tiles->nbounds) {
    set_error("tile step: tile range outside the tile table");
    return DIPS_EINVAL;
  }
  if (rec->nrows >= INT_MAX || point_set(g).ngrid + point_set(g).nextra >= INT_MAX) {
    set_error("tile step: grid or records too large for 32-bit point indices");
    return DIPS_EINVAL;
  }
  const bool pc = net->pc_w1 > 0;
  if (pc && (!tiles->slot_row || !rec->row_coeffs)) {
    set_error("tile step: the preconditioner needs slot rows and per-row cell coefficients");
    return DIPS_EINVAL;
  }
  TileStepArgs a{point_set(g),
                 rec->hx,
                 rec->hy,
                 rec->hz,
                 rec->row_point,
                 rec->row_ptr,
                 rec->row_c,
                 rec->nrows,
                 rec->ent_code,
                 rec->ent_w,
                 rec->nnz,
                 tiles->tile_row,
                 tiles->slot_x,
                 tiles->slot_y,
                 tiles->slot_z,
                 tiles->slot_w,
                 tiles->slot_c,
                 tiles->slot_info,
                 pc ? tiles->slot_row : nullptr,
                 pc ? rec->row_coeffs : nullptr,
                 pc ? net->pc_scale : 0.0f,
                 tile_begin,
                 (int)ntiles};
  FusedOpt fo{};
  if (fused) fo = *fused;
  // with an optimizer: one fused launch when an arrival counter is given,
  // else the step kernel followed by tile_apply_kernel
  const int mode = fused ? (fused->sync ? 2 : 3) : (want_grad ? 1 : 0);
#define X(A, B, C, D, E, F, G)                                                                          \
  if (tile_shape_is(net, A, B, C, D, E, F, G)) {                                                       \
    const int nblk = tile_blocks(ntiles, TileCfg<A, B, C, D, E>::kWaves);                              \
    if (partial_cap < (int64_t)nblk * partial_stride(net)) {                                           \
      set_error("tile step: partial buffer too small");                                                \
      return DIPS_EINVAL;                                                                              \
    }                                                                                                  \
    *nblk_out = nblk;                                                                                  \
    tile_launch_shape<A, B, C, D, E, F, G>(a, const_cast<float*>(params), partial, mode, fo, nblk, stream); \
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

// ---------------------------------------------------------------------------
// Batched training of many problems.  A global step is one persistent launch
// whose workgroups pull items from a device work queue: item (p, k) is block k
// of the nblk_p workgroup blocks that problem p's step uses on its own
// (nblk_p = min(cap, blocks for the tiles of its batch s mod nbatches_p)), and
// writes exactly that block's partial row, whichever workgroup takes it.  Then
// one reduction + optimizer workgroup per problem, in block order.  With
// cap = kMaxTileBlocks every problem follows the trajectory the
// single-problem two-launch step gives it, bit for bit.  The global step and
// the queue head live in device memory (BatchCtrl), advanced by a one-thread
// kernel after each step, so an epoch of launches can be captured in a graph
// and replayed.
struct BatchDesc {
  TileStepArgs a;                   // tile_begin / ntiles are set per step
  const int64_t* batch_tile_begin;  // device copy, nbatches + 1
  int32_t nbatches;
  int32_t steps;
  int32_t blk_begin;                // first item of the problem
  int32_t blk_count;                // items of the problem: its most blocks over the batches
  float* params;
  float* partial;                   // blk_count * stride floats
  FusedOpt fo;
};

struct BatchCtrl {
  uint32_t queue;  // next item of the current global step
  int32_t step;    // current global step
};

// the step's tile range and block count of problem d at global step s
__device__ __forceinline__ bool batch_step_range(const BatchDesc& d, int s, int waves, int cap, int64_t& tb, int& nt,
                                                 int& nblk) {
  if (s >= d.steps) return false;
  const int b = s % d.nbatches;
  tb = gld(d.batch_tile_begin, b);
  nt = (int)(gld(d.batch_tile_begin, b + 1) - tb - 1);
  nblk = min(tile_blocks(nt, waves), cap);
  return true;
}

template <int Lm, int Hm, int Lp, int Hp, int PC1, int PC2, int Act>
__global__ __launch_bounds__((TileCfg<Lm, Hm, Lp, Hp, PC1>::kThreads)) void tile_batch_step_kernel(
    const BatchDesc* __restrict__ descs, int nprob, int nitems, int cap, BatchCtrl* __restrict__ ctrl) {
  constexpr int kTileWaves = TileCfg<Lm, Hm, Lp, Hp, PC1>::kWaves;
  const int s = ctrl->step;
  __shared__ int item_s;
  for (;;) {
    __syncthreads();  // every thread has read the previous item
    if (threadIdx.x == 0) item_s = (int)atomicAdd(&ctrl->queue, 1u);
    __syncthreads();
    const int item = item_s;
    if (item >= nitems) return;
    // owner problem: the last p with blk_begin <= item
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (descs[mid].blk_begin <= item)
        lo = mid;
      else
        hi = mid - 1;
    }
    const BatchDesc& d = descs[lo];
    int64_t tb;
    int nt, nblk;
    if (!batch_step_range(d, s, kTileWaves, cap, tb, nt, nblk)) continue;
    const int k = item - d.blk_begin;
    if (k >= nblk) continue;
    TileStepArgs a = d.a;
    a.tile_begin = tb;
    a.ntiles = nt;
    tile_step_body<Lm, Hm, Lp, Hp, true, 1, PC1, PC2, Act>(a, d.params, d.partial, k, nblk);
  }
}

template <int kStride, int kWaves>
__global__ __launch_bounds__(1024) void tile_batch_apply_kernel(const BatchDesc* __restrict__ descs, int cap,
                                                                const BatchCtrl* __restrict__ ctrl) {
  const BatchDesc& d = descs[blockIdx.x];
  int64_t tb;
  int nt, nblk;
  if (!batch_step_range(d, ctrl->step, kWaves, cap, tb, nt, nblk)) return;
  tile_reduce_apply<kStride, 16>(d.partial, nblk, d.fo, d.params);
}

// after a step: empty queue, next global step
__global__ void batch_ctrl_next_kernel(BatchCtrl* ctrl) {
  if (threadIdx.x == 0) {
    ctrl->queue = 0u;
    ctrl->step = ctrl->step + 1;
  }
}

__global__ void batch_ctrl_init_kernel(BatchCtrl* ctrl, int step) {
  if (threadIdx.x == 0) {
    ctrl->queue = 0u;
    ctrl->step = step;
  }
}

namespace {
struct BatchLayout {
  int64_t desc_off, table_off, partial_off, ctrl_off, bytes;
  int32_t nblk_total;
  std::vector<int32_t> blk_count;
};

constexpr int64_t kAlign = 256;
int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

int tile_waves_of(const dips_net_t* net) {
#define X(A, B, C, D, E, F, G) \
  if (tile_shape_is(net, A, B, C, D, E, F, G)) return TileCfg<A, B, C, D, E>::kWaves;
  DIPS_TILE_SHAPES(X)
#undef X
  return 0;
}

int batch_cap(const dips_batch_cfg_t* cfg) {
  const int c = cfg ? cfg->blk_cap : 0;
  return c <= 0 ? kMaxTileBlocks : std::min(c, kMaxTileBlocks);
}

int batch_layout(const dips_batch_problem_t* probs, int32_t nprob, const dips_net_t* net, int cap, BatchLayout& L) {
  const int waves = tile_waves_of(net);
  if (!waves) {
    set_error("batch train: network shape has no compiled tile kernel");
    return DIPS_EUNSUPPORTED;
  }
  if (!probs || nprob < 1) {
    set_error("batch train: no problems");
    return DIPS_EINVAL;
  }
  const int64_t stride = partial_stride(net);
  int64_t tables = 0, partials = 0, blocks = 0;
  L.blk_count.assign(nprob, 0);
  for (int p = 0; p < nprob; ++p) {
    const dips_batch_problem_t& q = probs[p];
    if (!q.batch_tile_begin || q.nbatches < 1 || q.batch_size < 1) {
      set_error("batch train: problem " + std::to_string(p) + ": invalid batch table");
      return DIPS_EINVAL;
    }
    int cnt = 1;
    for (int b = 0; b < q.nbatches; ++b) {
      const int64_t nt = q.batch_tile_begin[b + 1] - q.batch_tile_begin[b] - 1;
      if (nt < 1 || q.batch_tile_begin[b] < 0 || q.batch_tile_begin[b + 1] > q.tiles.nbounds) {
        set_error("batch train: problem " + std::to_string(p) + ": batch without tiles or outside the tile table");
        return DIPS_EINVAL;
      }
      cnt = std::max(cnt, std::min(tile_blocks(nt, waves), cap));
    }
    L.blk_count[p] = cnt;
    blocks += cnt;
    tables += q.nbatches + 1;
    partials += (int64_t)cnt * stride;
  }
  if (blocks > INT_MAX / 2) {
    set_error("batch train: too many workgroups");
    return DIPS_EINVAL;
  }
  L.nblk_total = (int32_t)blocks;
  L.desc_off = 0;
  L.table_off = align_up((int64_t)sizeof(BatchDesc) * nprob);
  L.partial_off = align_up(L.table_off + (int64_t)sizeof(int64_t) * tables);
  L.ctrl_off = align_up(L.partial_off + (int64_t)sizeof(float) * partials);
  L.bytes = align_up(L.ctrl_off + (int64_t)sizeof(BatchCtrl));
  return DIPS_OK;
}

// workgroups of the persistent step launch: as many as are resident at once, or the configured count
template <int Lm, int Hm, int Lp, int Hp, int PC1, int PC2, int Act>
int batch_grid_shape(const dips_batch_cfg_t* cfg) {
  if (cfg && cfg->grid_blocks > 0) return cfg->grid_blocks;
  return resident_blocks((const void*)tile_batch_step_kernel<Lm, Hm, Lp, Hp, PC1, PC2, Act>,
                         TileCfg<Lm, Hm, Lp, Hp, PC1>::kThreads, TileCfg<Lm, Hm, Lp, Hp, PC1>::kLdsBytes,
                         kMaxTileBlocks);
}
}  // namespace

int64_t batch_workspace(const dips_batch_problem_t* probs, int32_t nprob, const dips_net_t* net,
                        const dips_batch_cfg_t* cfg) {
  BatchLayout L;
  if (batch_layout(probs, nprob, net, batch_cap(cfg), L)) return -1;
  return L.bytes;
}

template <int Lm, int Hm, int Lp, int Hp, int PC1, int PC2, int Act>
static void batch_launch_shape(const BatchDesc* descs, BatchCtrl* ctrl, const dips_batch_plan_t& plan, hipStream_t st) {
  constexpr int kWaves = TileCfg<Lm, Hm, Lp, Hp, PC1>::kWaves;
  const size_t lds = TileCfg<Lm, Hm, Lp, Hp, PC1>::kLdsBytes;
  static const bool lds_set =
      hipFuncSetAttribute((const void*)tile_batch_step_kernel<Lm, Hm, Lp, Hp, PC1, PC2, Act>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;
  (void)lds_set;
  hipLaunchKernelGGL((tile_batch_step_kernel<Lm, Hm, Lp, Hp, PC1, PC2, Act>), dim3(plan.grid_blocks),
                     dim3(TileCfg<Lm, Hm, Lp, Hp, PC1>::kThreads), lds, st, descs, plan.nprob, plan.nitems,
                     plan.blk_cap, ctrl);
  constexpr int kStride = tile_stride<Lm, Hm, Lp, Hp, PC1, PC2>();
  hipLaunchKernelGGL((tile_batch_apply_kernel<kStride, kWaves>), dim3(plan.nprob), dim3(1024), 0, st, descs,
                     plan.blk_cap, ctrl);
  hipLaunchKernelGGL(batch_ctrl_next_kernel, dim3(1), dim3(64), 0, st, ctrl);
}

int batch_setup(const dips_batch_problem_t* probs, int32_t nprob, int32_t nepochs, const dips_net_t* net,
                const dips_opt_t* opt, const dips_batch_cfg_t* cfg, void* work, int64_t work_bytes,
                dips_batch_plan_t* plan, hipStream_t stream) {
  BatchLayout L;
  const int cap = batch_cap(cfg);
  int st = batch_layout(probs, nprob, net, cap, L);
  if (st) return st;
  if (!work || work_bytes < L.bytes) {
    set_error("batch train: workspace too small (need " + std::to_string(L.bytes) + " bytes)");
    return DIPS_EINVAL;
  }
  if (nepochs < 0 || !opt || !plan || (cfg && cfg->grid_blocks < 0)) {
    set_error("batch train: invalid epochs, optimizer, plan or grid size");
    return DIPS_EINVAL;
  }
  const bool pc = net->pc_w1 > 0;
  char* base = static_cast<char*>(work);
  std::vector<BatchDesc> descs(nprob);
  std::vector<int64_t> tables;
  int32_t blk = 0;
  int64_t maxsteps = 0;
  int64_t poff = 0;
  const int64_t stride = partial_stride(net);
  for (int p = 0; p < nprob; ++p) {
    const dips_batch_problem_t& q = probs[p];
    if (q.rec.nrows >= INT_MAX || point_set(&q.grid).ngrid + point_set(&q.grid).nextra >= INT_MAX) {
      set_error("batch train: problem " + std::to_string(p) + ": too large for 32-bit point indices");
      return DIPS_EINVAL;
    }
    const dips_tiles_t& t = q.tiles;
    if (!t.tile_row || !t.slot_x || !t.slot_y || !t.slot_z || !t.slot_w || !t.slot_c || !t.slot_info ||
        (pc && (!t.slot_row || !q.rec.row_coeffs)) || !q.params || !q.m || !q.v || !q.count || !q.loss_hist) {
      set_error("batch train: problem " + std::to_string(p) + ": null buffer");
      return DIPS_EINVAL;
    }
    const int64_t steps = (int64_t)nepochs * q.nbatches;
    if (steps > INT_MAX) {
      set_error("batch train: too many steps");
      return DIPS_EINVAL;
    }
    maxsteps = std::max(maxsteps, steps);
    BatchDesc& d = descs[p];
    d.a = TileStepArgs{point_set(&q.grid), q.rec.hx,  q.rec.hy,      q.rec.hz,      q.rec.row_point,
                       q.rec.row_ptr,     q.rec.row_c, q.rec.nrows,  q.rec.ent_code, q.rec.ent_w,
                       q.rec.nnz,         t.tile_row, t.slot_x,      t.slot_y,      t.slot_z,
                       t.slot_w,          t.slot_c,   t.slot_info,   pc ? t.slot_row : nullptr,
                       pc ? q.rec.row_coeffs : nullptr, pc ? net->pc_scale : 0.0f, 0, 0};
    d.batch_tile_begin = reinterpret_cast<const int64_t*>(base + L.table_off) + tables.size();
    tables.insert(tables.end(), q.batch_tile_begin, q.batch_tile_begin + q.nbatches + 1);
    d.nbatches = q.nbatches;
    d.steps = (int32_t)steps;
    d.blk_begin = blk;
    d.blk_count = L.blk_count[p];
    blk += L.blk_count[p];
    d.params = q.params;
    d.partial = reinterpret_cast<float*>(base + L.partial_off) + poff;
    poff += (int64_t)L.blk_count[p] * stride;
    d.fo = FusedOpt{*opt, 1.0f / (float)q.batch_size, q.m, q.v, q.count, q.loss_hist, q.loss_cap, nullptr};
  }
  int grid = 0;
#define X(A, B, C, D, E, F, G) \
  if (!grid && tile_shape_is(net, A, B, C, D, E, F, G)) grid = batch_grid_shape<A, B, C, D, E, F, G>(cfg);
  DIPS_TILE_SHAPES(X)
#undef X
  if (grid < 1) {
    set_error("batch train: network shape has no compiled tile kernel");
    return DIPS_EUNSUPPORTED;
  }
  // the tables are read by the kernels: copy, then wait so the host vectors may go
  hipError_t e = hipMemcpyAsync(base + L.desc_off, descs.data(), sizeof(BatchDesc) * nprob, hipMemcpyHostToDevice,
                                stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(base + L.table_off, tables.data(), sizeof(int64_t) * tables.size(), hipMemcpyHostToDevice,
                       stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(batch_ctrl_init_kernel, dim3(1), dim3(64), 0, stream,
                       reinterpret_cast<BatchCtrl*>(base + L.ctrl_off), 0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_status(e, "batch train: table copy");
  *plan = dips_batch_plan_t{nprob, L.nblk_total, cap, grid, (int32_t)maxsteps, (int32_t)(L.ctrl_off / kAlign)};
  return DIPS_OK;
}

int batch_steps(const dips_net_t* net, const dips_batch_plan_t* plan, int32_t nsteps, void* work, hipStream_t stream) {
  if (!net || !plan || !work || nsteps < 0 || plan->nprob < 1 || plan->nitems < 1 || plan->grid_blocks < 1 ||
      plan->blk_cap < 1 || plan->ctrl_units < 1) {
    set_error("batch steps: invalid plan");
    return DIPS_EINVAL;
  }
  char* base = static_cast<char*>(work);
  const BatchDesc* dd = reinterpret_cast<const BatchDesc*>(base);
  BatchCtrl* ctrl = reinterpret_cast<BatchCtrl*>(base + (int64_t)plan->ctrl_units * kAlign);
  for (int32_t s = 0; s < nsteps; ++s) {
#define X(A, B, C, D, E, F, G)                                         \
  if (tile_shape_is(net, A, B, C, D, E, F, G)) {                      \
    batch_launch_shape<A, B, C, D, E, F, G>(dd, ctrl, *plan, stream); \
  } else
    DIPS_TILE_SHAPES(X)
#undef X
    {
      set_error("batch steps: network shape has no compiled tile kernel");
      return DIPS_EUNSUPPORTED;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_status(e, "batch steps: launch");
  }
  return DIPS_OK;
}

int batch_train(const dips_batch_problem_t* probs, int32_t nprob, int32_t nepochs, const dips_net_t* net,
                const dips_opt_t* opt, void* work, int64_t work_bytes, hipStream_t stream) {
  dips_batch_plan_t plan;
  int st = batch_setup(probs, nprob, nepochs, net, opt, nullptr, work, work_bytes, &plan, stream);
  if (st) return st;
  return batch_steps(net, &plan, plan.max_steps, work, stream);
}

// Slot layout of every tile (once per zoom level): one wavefront per tile
// boundary index k, tile k = rows [tile_row[k], tile_row[k+1]).  Lane s of
// the tile fills slot s: its row j (rows take max(entries, 1) consecutive
// slots), the entry's position fl32(p + o*h), weight, phase, the row's first
// and last lane, and the row constant in the row's last slot.  Tiles with more
// than 64 rows or slots (single long rows) and the empty groups between
// batches get unused slots.
constexpr int kSlotWaves = 4;

struct SlotArgs {
  PointSet ps;
  float hx, hy, hz;
  const int32_t* row_point;
  const int32_t* row_ptr;
  const float* row_c;
  int64_t nrows;
  const uint32_t* ent_code;
  const float* ent_w;
  int64_t nnz;
  const int32_t* tile_row;
  int64_t ngroups;  // nbounds (the last boundary starts no tile)
  float* sx;
  float* sy;
  float* sz;
  float* sw;
  float* sc;
  uint32_t* sinfo;
  int32_t* srow;  // may be null
};

__global__ __launch_bounds__(kSlotWaves * DIPS_WAVE) void tile_slots_kernel(SlotArgs a) {
  __shared__ int rowof[kSlotWaves][kTileCap];
  const int lane = threadIdx.x & (DIPS_WAVE - 1);
  const int wave = threadIdx.x / DIPS_WAVE;
  const int64_t k = (int64_t)blockIdx.x * kSlotWaves + wave;
  if (k >= a.ngroups) return;  // whole wavefront
  const int64_t r0 = clamp64(a.tile_row[k], 0, a.nrows);
  const int64_t r1 = (k + 1 < a.ngroups) ? clamp64(a.tile_row[k + 1], r0, a.nrows) : r0;
  const int64_t nr = r1 - r0;
  int len = 0, tl = 0;
  if (nr <= kTileCap && lane < nr) {
    tl = max(a.row_ptr[r0 + lane + 1] - a.row_ptr[r0 + lane], 0);
    len = max(tl, 1);
  }
  // inclusive scan of the slot counts over the lanes
  int incl = len;
#pragma unroll
  for (int d = 1; d < DIPS_WAVE; d <<= 1) {
    const int o = __shfl_up(incl, d, DIPS_WAVE);
    if (lane >= d) incl += o;
  }
  const int start = incl - len;
  const int total = __shfl(incl, DIPS_WAVE - 1, DIPS_WAVE);
  const bool ok = nr <= kTileCap && total <= kTileCap;
  rowof[wave][lane] = -1;
  __builtin_amdgcn_wave_barrier();
  if (ok && lane < nr)
    for (int q = 0; q < len; ++q) rowof[wave][start + q] = lane;
  __builtin_amdgcn_wave_barrier();
  const int j = ok ? rowof[wave][lane] : -1;
  const int jj = j < 0 ? 0 : j;
  const int sj = __shfl(start, jj, DIPS_WAVE);
  const int lj = __shfl(len, jj, DIPS_WAVE);
  const int tj = __shfl(tl, jj, DIPS_WAVE);
  uint32_t info = 0u;
  float x = 0.0f, y = 0.0f, z = 0.0f, w = 0.0f, c = 0.0f;
  int32_t srow = -1;
  if (j >= 0) {
    const int64_t row = r0 + j;
    srow = (int32_t)row;
    const int pt = a.row_point[row];
    if (point_xyz(a.ps, pt, x, y, z)) {
      const int head = sj, last = sj + lj - 1;
      info = kSlotUsed | (uint32_t)head | ((uint32_t)last << 6);
      if (tj > 0) {
        const int64_t e = clamp64((int64_t)a.row_ptr[row] + (lane - sj), 0, a.nnz - 1);
        const uint32_t code = a.ent_code[e];
        int ox, oy, oz;
        dips_decode_offset(code, ox, oy, oz);
        x = x + (float)ox * a.hx;
        y = y + (float)oy * a.hy;
        z = z + (float)oz * a.hz;
        w = a.ent_w[e];
        if (code & 128u) info |= kSlotPlus;
      }
      if (lane == last) c = a.row_c[row];
    }
  }
  const int64_t s = k * kTileCap + lane;
  a.sx[s] = x;
  a.sy[s] = y;
  a.sz[s] = z;
  a.sw[s] = w;
  a.sc[s] = c;
  a.sinfo[s] = info;
  if (a.srow) a.srow[s] = srow;
}

int tile_slots_launch(const dips_grid_t* g, const dips_records_t* rec, const int32_t* tile_row, int64_t nbounds,
                      float* sx, float* sy, float* sz, float* sw, float* sc, uint32_t* sinfo, int32_t* srow,
                      hipStream_t st) {
  if (rec->nrows >= INT_MAX || point_set(g).ngrid + point_set(g).nextra >= INT_MAX) {
    set_error("dips_tile_slots: grid or records too large for 32-bit indices");
    return DIPS_EINVAL;
  }
  SlotArgs a{point_set(g),
             rec->hx,    rec->hy,       rec->hz,    rec->row_point, rec->row_ptr, rec->row_c,
             rec->nrows, rec->ent_code, rec->ent_w, rec->nnz, tile_row, nbounds,
             sx,         sy,            sz,         sw,       sc,       sinfo,    srow};
  const int64_t nb = (nbounds + kSlotWaves - 1) / kSlotWaves;
  hipLaunchKernelGGL(tile_slots_kernel, dim3((unsigned)nb), dim3(kSlotWaves * DIPS_WAVE), 0, st, a);
  return hip_status(hipGetLastError(), "tile_slots_kernel launch");
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
        const int64_t slots = len > 0 ? len : 1;  // a row without entries takes one slot
        if (len > kTileCap || n + slots > kTileCap) break;
        n += slots;
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
