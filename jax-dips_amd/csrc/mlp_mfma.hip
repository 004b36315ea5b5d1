// MFMA path of the training step for wide networks (hidden width up to 64).
//
// Work unit: a tile of 16 record entries of one phase (entries are sorted by
// phase inside each batch), evaluated by one wavefront with
// v_mfma_f32_16x16x4f32.  Activations are kept feature-major: for a layer of
// width H = 16*NB, lane l holds feature 16*fb + 4*(l/16) + r of tile position
// l%16 in register (fb, r) -- exactly the accumulator layout of the MFMA.  With
// the contraction index of the next layer ordered as k = 16*fb + 4*(l/16) + r
// for step (fb, r), that accumulator is directly the B operand of the next
// layer, so layers chain without data movement.  Weight operands are packed
// once per workgroup into LDS in that order.  Weight gradients contract over
// positions, so the two operands (activation and delta of a tile) pass through
// a padded LDS transpose and the 16x16 dW blocks accumulate in AGPRs across
// all tiles of the wavefront.
//
// Per optimizer step: fwd (u per entry) -> rowsum (r per row, loss partials)
// -> bwd (parameter-gradient partials per workgroup) -> reduce -> apply.
//
// Reference: the same loss and gradient as loss.hip (trainer.py:783-789,
// :892-912; DoubleMLP, nn/mlp/MLP.py:93-139).
#include <math.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "dips_common.h"
#include "dips_hip.h"
#include "dips_mlp.h"

namespace dips {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kMT = 256;  // threads per workgroup (4 wavefronts)
constexpr int kTP = 20;   // padded row length of a transposed 16-position tile (16-byte aligned rows)
constexpr int kMfmaBwdBlocks = 512;   // workgroups of the backward kernel (both phases), two per CU
constexpr int kMfmaFwdBlocks = 1024;  // workgroups of the forward kernel (both phases), four per CU

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// ---- packed weights of one phase net in LDS ---------------------------------
// [L0: NB*64] [for hidden layer 1..L-1: NB*NB blocks of kBlk, bias NB*16] [wout NB*16] [bout 1]
// Block (ob, fb) of hidden layer l holds W_l[16fb + 4(lane/16) + r][16ob + lane%16]
// as lane's 4-float unit at unit(lane) = lane + lane/16: the forward pass reads
// one 16-byte unit per lane (the A operands of 4 MFMA steps), and the backward
// pass reads the transpose from the same copy, 4 floats per lane 4 apart; the
// unit left free after every 16 spreads both patterns over all LDS banks.
template <int L, int NB>
struct Packed {
  static constexpr int H = 16 * NB;
  static constexpr int kBlk = 4 * 68;  // floats of one 16x16 block: 64 units of 4 plus 4 free units
  static constexpr int kMat = NB * NB * kBlk;
  static constexpr int kPer = kMat + H;
  static constexpr int l0() { return 0; }
  static constexpr int fwd(int l) { return NB * 64 + (l - 1) * kPer; }
  static constexpr int bias(int l) { return fwd(l) + kMat; }
  static constexpr int wout() { return NB * 64 + (L - 1) * kPer; }
  static constexpr int bout() { return wout() + H; }
  static constexpr int size() { return bout() + 1; }
  static __device__ __forceinline__ int unit(int lane) { return lane + (lane >> 4); }
};

// flat-parameter offsets of one phase net of true width Hh
struct FlatNet {
  int Hh, L, base;
  __device__ int w(int l) const { return base + (l == 0 ? 0 : 4 * Hh + (l - 1) * (Hh * Hh + Hh)); }
  __device__ int b(int l) const { return w(l) + (l == 0 ? 3 * Hh : Hh * Hh); }
  __device__ int wout() const { return base + 4 * Hh + (L - 1) * (Hh * Hh + Hh); }
  __device__ int bout() const { return wout() + Hh; }
};

// Layer 0, the hidden-layer matrices and the hidden biases are stored times
// kTanhExp2Scale, so that the accumulator of a hidden layer is already the
// argument of dips_tanh_exp2 (the backward pass divides the factor out of its
// gradients); the output layer is stored as it is.
template <int L, int NB>
__device__ void pack_weights(const float* __restrict__ params, FlatNet fn, float* __restrict__ lds) {
  using PK = Packed<L, NB>;
  const int Hh = fn.Hh;
  for (int idx = threadIdx.x; idx < PK::size(); idx += blockDim.x) {
    float v = 0.0f;
    if (idx < NB * 64) {
      const int ob = idx / 64, lane = idx % 64;
      const int out = 16 * ob + lane % 16, k = lane / 16;
      if (out < Hh) v = kTanhExp2Scale * ((k < 3) ? params[fn.w(0) + k * Hh + out] : params[fn.b(0) + out]);
    } else if (idx >= PK::wout()) {
      const int h = idx - PK::wout();
      if (h < PK::H) {
        if (h < Hh) v = params[fn.wout() + h];
      } else {
        v = params[fn.bout()];
      }
    } else {
      const int rel = idx - NB * 64;
      const int l = rel / PK::kPer + 1;
      const int q = rel % PK::kPer;
      if (q < PK::kMat) {
        const int blk = q / PK::kBlk, t = q % PK::kBlk;
        const int up = t / 4, r = t % 4;  // unit position up = 17 a + b holds lane 16 a + b (b < 16)
        if (up % 17 < 16) {
          const int lane = 16 * (up / 17) + up % 17;
          const int ob = blk / NB, fb = blk % NB;
          const int in = 16 * fb + 4 * (lane / 16) + r, out = 16 * ob + lane % 16;
          if (in < Hh && out < Hh) v = kTanhExp2Scale * params[fn.w(l) + in * Hh + out];
        }
      } else {
        const int h = q - PK::kMat;
        if (h < Hh) v = kTanhExp2Scale * params[fn.b(l) + h];
      }
    }
    lds[idx] = v;
  }
}

// layer 0 of one tile (K = 4: x, y, z and the bias row) from its packed weights at W[0, NB*64)
template <int L, int NB>
__device__ __forceinline__ void tile_layer0(const float* __restrict__ W, float xin, float (&act)[L][4 * NB]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int ob = 0; ob < NB; ++ob) {
    f4 c = {0.f, 0.f, 0.f, 0.f};
    c = mfma4(W[ob * 64 + lane], xin, c);
#pragma unroll
    for (int r = 0; r < 4; ++r) act[0][ob * 4 + r] = dips_tanh_exp2(c[r]);
  }
}

// layer 0 from this lane's operands held in registers (w0[ob] = the packed value at ob * 64 + lane)
template <int L, int NB>
__device__ __forceinline__ void tile_layer0_regs(const float (&w0)[NB], float xin, float (&act)[L][4 * NB]) {
#pragma unroll
  for (int ob = 0; ob < NB; ++ob) {
    f4 c = {0.f, 0.f, 0.f, 0.f};
    c = mfma4(w0[ob], xin, c);
#pragma unroll
    for (int r = 0; r < 4; ++r) act[0][ob * 4 + r] = dips_tanh_exp2(c[r]);
  }
}

// forward of one tile: act[l][fb*4+r] feature-major; returns nothing (u by caller)
template <int L, int NB>
__device__ __forceinline__ void tile_forward(const float* __restrict__ W, float xin, float (&act)[L][4 * NB]) {
  using PK = Packed<L, NB>;
  const int lane = threadIdx.x & 63;
  static_assert(PK::l0() == 0, "layer 0 leads the packed weights");
  tile_layer0<L, NB>(W, xin, act);
#pragma unroll
  for (int l = 1; l < L; ++l) {
    f4 c[NB];
#pragma unroll
    for (int ob = 0; ob < NB; ++ob) {
#pragma unroll
      for (int r = 0; r < 4; ++r) c[ob][r] = W[PK::bias(l) + 16 * ob + 4 * (lane / 16) + r];
    }
#pragma unroll
    for (int fb = 0; fb < NB; ++fb) {
#pragma unroll
      for (int ob = 0; ob < NB; ++ob) {
        const f4 w4 = *reinterpret_cast<const f4*>(&W[PK::fwd(l) + (ob * NB + fb) * PK::kBlk + 4 * PK::unit(lane)]);
#pragma unroll
        for (int r = 0; r < 4; ++r) c[ob] = mfma4(w4[r], act[l - 1][fb * 4 + r], c[ob]);
      }
    }
#pragma unroll
    for (int ob = 0; ob < NB; ++ob) {
#pragma unroll
      for (int r = 0; r < 4; ++r) act[l][ob * 4 + r] = dips_tanh_exp2(c[ob][r]);
    }
  }
}

// ---- hidden layers on the bf16 matrix cores, to float32 accuracy --------------
// Each float32 operand x is split exactly into three bf16 parts by truncation:
// x0 = x with the low 16 bits cleared, x1 likewise of x - x0, x2 = x - x0 - x1
// (x has 24 significant bits, each part holds 8: x0 + x1 + x2 == x bit for bit
// for normal x, |x1| < 2^-7 |x|, |x2| < 2^-15 |x|).  A product a w is then the
// sum of the six part products with part indices summing to at most 2, each an
// exact bf16 x bf16 product accumulated in float32 by v_mfma_f32_16x16x32_bf16;
// the dropped terms a1 w2 + a2 w1 + a2 w2 are below 2^-21 |a w|.  One such MFMA
// (8,192 multiply-adds) takes ~16 cycles against 256 for the same work as eight
// v_mfma_f32_16x16x4f32 (the float32 MFMA does not overlap with VALU work on a
// SIMD; the bf16 one overlaps with the tanh's transcendentals).
//
// Operand layout of a K = 32 chunk c of a hidden layer (input blocks fb = 2c, 2c+1):
// slot t of lane group g = lane/16 is input feature 16 (2c + t/4) + 4 g + t%4,
// which is where the activation of the previous layer already sits (register
// (fb, r) of the feature-major layout), so the B operand is the 8 activations
// of blocks 2c and 2c+1, split, and layers still chain without data movement.
typedef uint32_t u4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f4 mfma_bf(u4 a, u4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

__device__ __forceinline__ float bf_trunc(float x) { return __uint_as_float(__float_as_uint(x) & 0xffff0000u); }

// bf16 bits of part p of x (the three parts sum to x exactly)
__device__ __forceinline__ uint32_t bf_part(float x, int p) {
  const float x0 = bf_trunc(x);
  const float r1 = x - x0;
  const float x1 = bf_trunc(r1);
  const float v = p == 0 ? x0 : (p == 1 ? x1 : r1 - x1);
  return __float_as_uint(v) >> 16;
}

// the three parts of (x, y) as packed bf16 pairs (x in the low half)
__device__ __forceinline__ void bf_split_pair(float x, float y, uint32_t& h, uint32_t& m, uint32_t& l) {
  const uint32_t xb = __float_as_uint(x), yb = __float_as_uint(y);
  h = __builtin_amdgcn_perm(yb, xb, 0x07060302u);
  const float x1 = x - __uint_as_float(xb & 0xffff0000u);
  const float y1 = y - __uint_as_float(yb & 0xffff0000u);
  const uint32_t x1b = __float_as_uint(x1), y1b = __float_as_uint(y1);
  m = __builtin_amdgcn_perm(y1b, x1b, 0x07060302u);
  const float x2 = x1 - __uint_as_float(x1b & 0xffff0000u);
  const float y2 = y1 - __uint_as_float(y1b & 0xffff0000u);
  l = __builtin_amdgcn_perm(__float_as_uint(y2), __float_as_uint(x2), 0x07060302u);
}

// Packed weights of one phase net for the split-bf16 forward pass:
// [L0: NB*64 floats][hidden biases: (L-1) x H][wout: H][bout: 1] (padded to 16 bytes), then
// per hidden layer l, output block ob, chunk c and part p one 64-lane set of 16-byte units
// (lane = 16 g + i holds W_l[16 (2c + t/4) + 4 g + t%4][16 ob + i], t = 0..7, part p).
// Scaled like Packed: every hidden-layer weight and bias times kTanhExp2Scale.
template <int L, int NB>
struct PackedBf {
  static constexpr int H = 16 * NB;
  static constexpr int NC = (NB + 1) / 2;
  static constexpr int l0() { return 0; }
  static constexpr int bias(int l) { return NB * 64 + (l - 1) * H; }
  static constexpr int wout() { return NB * 64 + (L - 1) * H; }
  static constexpr int bout() { return wout() + H; }
  static constexpr int nf() { return (bout() + 1 + 3) & ~3; }
  static constexpr int units() { return (L - 1) * NB * NC * 3 * 64; }
  static constexpr int size() { return nf() + 4 * units(); }  // floats
  static constexpr int set(int l, int ob, int c, int p) { return (((l - 1) * NB + ob) * NC + c) * 3 + p; }
};

template <int L, int NB>
__device__ void pack_weights_bf(const float* __restrict__ params, FlatNet fn, float* __restrict__ lds) {
  using PK = PackedBf<L, NB>;
  const int Hh = fn.Hh;
  for (int idx = threadIdx.x; idx < PK::nf(); idx += blockDim.x) {
    float v = 0.0f;
    if (idx < NB * 64) {
      const int ob = idx / 64, lane = idx % 64;
      const int out = 16 * ob + lane % 16, k = lane / 16;
      if (out < Hh) v = kTanhExp2Scale * ((k < 3) ? params[fn.w(0) + k * Hh + out] : params[fn.b(0) + out]);
    } else if (idx < PK::wout()) {
      const int l = (idx - NB * 64) / PK::H + 1, h = (idx - NB * 64) % PK::H;
      if (h < Hh) v = kTanhExp2Scale * params[fn.b(l) + h];
    } else if (idx < PK::bout()) {
      const int h = idx - PK::wout();
      if (h < Hh) v = params[fn.wout() + h];
    } else if (idx == PK::bout()) {
      v = params[fn.bout()];
    }
    lds[idx] = v;
  }
  uint32_t* wb = reinterpret_cast<uint32_t*>(lds + PK::nf());
  for (int d = threadIdx.x; d < 4 * PK::units(); d += blockDim.x) {
    const int u = d % 4, lane = (d / 4) % 64;
    int rest = d / 256;
    const int p = rest % 3;
    rest /= 3;
    const int c = rest % PK::NC;
    rest /= PK::NC;
    const int ob = rest % NB, l = rest / NB + 1;
    const int g = lane / 16, out = 16 * ob + lane % 16;
    uint32_t word = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = 2 * u + h, fb = 2 * c + t / 4;
      const int in = 16 * fb + 4 * g + t % 4;
      float w = 0.0f;
      if (fb < NB && in < Hh && out < Hh) w = kTanhExp2Scale * params[fn.w(l) + in * Hh + out];
      word |= bf_part(w, p) << (16 * h);
    }
    wb[d] = word;
  }
}

// forward of one tile with the hidden layers on the bf16 matrix cores (layer 0, K = 4, in float32)
template <int L, int NB>
__device__ __forceinline__ void tile_forward_bf(const float* __restrict__ W, float xin, float (&act)[L][4 * NB]) {
  using PK = PackedBf<L, NB>;
  const int lane = threadIdx.x & 63;
  const u4* __restrict__ wb = reinterpret_cast<const u4*>(W + PK::nf());
  static_assert(PK::l0() == 0, "layer 0 leads the packed weights");
  tile_layer0<L, NB>(W, xin, act);
#pragma unroll
  for (int l = 1; l < L; ++l) {
    f4 c[NB];
#pragma unroll
    for (int ob = 0; ob < NB; ++ob) {
#pragma unroll
      for (int r = 0; r < 4; ++r) c[ob][r] = W[PK::bias(l) + 16 * ob + 4 * (lane / 16) + r];
    }
#pragma unroll
    for (int cc = 0; cc < PK::NC; ++cc) {
      u4 ah, am, al;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int fb = 2 * cc + u / 2, r = (2 * u) % 4;
        const float x = fb < NB ? act[l - 1][fb * 4 + r] : 0.0f;
        const float y = fb < NB ? act[l - 1][fb * 4 + r + 1] : 0.0f;
        uint32_t h, m, lo;
        bf_split_pair(x, y, h, m, lo);
        ah[u] = h;
        am[u] = m;
        al[u] = lo;
      }
#pragma unroll
      for (int ob = 0; ob < NB; ++ob) {
        const u4 w0 = wb[PK::set(l, ob, cc, 0) * 64 + lane];
        const u4 w1 = wb[PK::set(l, ob, cc, 1) * 64 + lane];
        const u4 w2 = wb[PK::set(l, ob, cc, 2) * 64 + lane];
        // smallest terms first
        c[ob] = mfma_bf(w2, ah, c[ob]);
        c[ob] = mfma_bf(w1, am, c[ob]);
        c[ob] = mfma_bf(w0, al, c[ob]);
        c[ob] = mfma_bf(w1, ah, c[ob]);
        c[ob] = mfma_bf(w0, am, c[ob]);
        c[ob] = mfma_bf(w0, ah, c[ob]);
      }
    }
#pragma unroll
    for (int ob = 0; ob < NB; ++ob) {
#pragma unroll
      for (int r = 0; r < 4; ++r) act[l][ob * 4 + r] = dips_tanh_exp2(c[ob][r]);
    }
  }
}

// output u of the tile position l%16 (identical in the 4 lanes of that position)
template <int L, int NB, class PK = Packed<L, NB>>
__device__ __forceinline__ float tile_output(const float* __restrict__ W, const float (&act)[L][4 * NB]) {
  const int lane = threadIdx.x & 63;
  float s = 0.0f;
#pragma unroll
  for (int fb = 0; fb < NB; ++fb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) s = fmaf(act[L - 1][fb * 4 + r], W[PK::wout() + 16 * fb + 4 * (lane / 16) + r], s);
  }
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  return s + W[PK::bout()];
}

// ---- tile entry data ---------------------------------------------------------
struct MfmaTileArgs {
  PointSet ps;
  float hx, hy, hz;
  const int32_t* row_point;  // global row -> grid point
  const int32_t* t_row;      // phase-sorted entries: global row
  const uint32_t* t_code;    // phase-sorted entries: code (bit 7 = phase)
  const float* t_w;          // phase-sorted entries: weight
  const int32_t* t_csr;      // phase-sorted entries: CSR entry index
  const int32_t* t_group;    // items = groups of entries at one position: group g is the phase-sorted entries
                             // [t_group[g], t_group[g+1]); null: every entry is an item of its own
  const float* i_pos;        // item positions of the level, SoA: x at [item], y at [i_count + item], z at
                             // [2 i_count + item] (built once per level by mfma_items_kernel)
  int64_t i_count;
};

// input of lane group k (x, y, z or the constant 1 of the bias row) for item `item`; 0 / 1 past the end
__device__ __forceinline__ float item_input(const MfmaTileArgs& T, int64_t item, int64_t item_end, int k) {
  if (k == 3) return 1.0f;
  return item < item_end ? T.i_pos[k * T.i_count + item] : 0.0f;
}

// phase-sorted entries [b, e) of item `item`
__device__ __forceinline__ void item_members(const MfmaTileArgs& T, int64_t item, int64_t& b, int64_t& e) {
  if (T.t_group) {
    b = T.t_group[item];
    e = T.t_group[item + 1];
  } else {
    b = item;
    e = item + 1;
  }
}

// position of phase-sorted entry e; valid = false for padding lanes
__device__ __forceinline__ bool entry_pos(const MfmaTileArgs& T, int64_t e, int64_t e_end, float& x, float& y, float& z,
                                          int& row) {
  row = -1;
  x = y = z = 0.0f;
  if (e >= e_end) return false;
  row = T.t_row[e];
  const int pt = T.row_point[row];
  float px, py, pz;
  if (!point_xyz(T.ps, pt, px, py, pz)) return false;
  int ox, oy, oz;
  dips_decode_offset(T.t_code[e], ox, oy, oz);
  x = px + (float)ox * T.hx;
  y = py + (float)oy * T.hy;
  z = pz + (float)oz * T.hz;
  return true;
}

// ---- per level: item positions and the entry -> item map ------------------------
// One thread per item of a batch: its position (that of its first member entry;
// all members share it) into i_pos, and its index into e_item at the CSR slot of
// every member, through which the row sums read u per item.
__global__ __launch_bounds__(256) void mfma_items_kernel(MfmaTileArgs T, int64_t i0, int64_t n, int32_t* __restrict__ e_item,
                                                         float* __restrict__ i_pos) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t item = i0 + t;
  int64_t mb, me;
  item_members(T, item, mb, me);
  float x, y, z;
  int row;
  entry_pos(T, mb, me, x, y, z, row);  // a row naming no valid point gives the position (0, 0, 0)
  i_pos[item] = x;
  i_pos[T.i_count + item] = y;
  i_pos[2 * T.i_count + item] = z;
  for (int64_t e = mb; e < me; ++e) e_item[T.t_csr[e]] = (int32_t)item;
}

// ---- kernel 1: u for every item of the batch -----------------------------------
// Persistent: workgroups [0, nblk_p) take the + tiles (+ items [e0, e0+np)),
// the others the - tiles; each packs its network's weights into LDS once and
// its wavefronts stride over the tiles.  An item (an entry, or a group of
// entries at one position) is evaluated once and its u written to u_item[item - e0].
constexpr int kFT = 512;  // threads per forward workgroup: 8 wavefronts share one packed weight copy
constexpr int kFW = kFT / 64;

// the forward kernel's hidden layers: split bf16 from two 16-wide blocks on (where the matrix work
// outweighs the splits), float32 MFMA below
template <int L, int NB>
struct FwdPack {
  static constexpr bool kBf = NB >= 2;
  static constexpr int size() { return kBf ? PackedBf<L, NB>::size() : Packed<L, NB>::size(); }
};
template <int L, int NB>
__device__ __forceinline__ void fwd_pack(const float* __restrict__ params, FlatNet fn, float* __restrict__ lds) {
  if constexpr (FwdPack<L, NB>::kBf)
    pack_weights_bf<L, NB>(params, fn, lds);
  else
    pack_weights<L, NB>(params, fn, lds);
}
// the forward pass of one tile (split bf16 or float32 by FwdPack) into act
template <int L, int NB>
__device__ __forceinline__ void fwd_tile_act(const float* __restrict__ W, float xin, float (&act)[L][4 * NB]) {
  if constexpr (FwdPack<L, NB>::kBf)
    tile_forward_bf<L, NB>(W, xin, act);
  else
    tile_forward<L, NB>(W, xin, act);
}

// Stored activations of a tile: hidden layers 1..L-1 (layer 0, four MFMA steps, is recomputed more cheaply
// than it is stored), f4 ((l - 1) * NB + fb) of lane `lane` at act_tile[((l - 1) * NB + fb) * 64 + lane]
// (act_tile = the tile's 64 * act_f4 f4s): one 1 KB contiguous store or load per wavefront and f4.
template <int L, int NB>
__device__ __forceinline__ void store_act(f4* __restrict__ dst, const float (&act)[L][4 * NB]) {
#pragma unroll
  for (int l = 1; l < L; ++l)
#pragma unroll
    for (int fb = 0; fb < NB; ++fb)
      __builtin_nontemporal_store(f4{act[l][fb * 4], act[l][fb * 4 + 1], act[l][fb * 4 + 2], act[l][fb * 4 + 3]},
                                  dst + ((l - 1) * NB + fb) * 64);
}
template <int L, int NB>
__device__ __forceinline__ void load_act(const f4* __restrict__ src, float (&act)[L][4 * NB]) {
#pragma unroll
  for (int l = 1; l < L; ++l)
#pragma unroll
    for (int fb = 0; fb < NB; ++fb) {
      const f4 v = __builtin_nontemporal_load(src + ((l - 1) * NB + fb) * 64);
#pragma unroll
      for (int r = 0; r < 4; ++r) act[l][fb * 4 + r] = v[r];
    }
}

// u of one tile; with act_tile (this lane's f4 0 of the tile's stored activations) also stores them
template <int L, int NB>
__device__ __forceinline__ float fwd_tile_u(const float* __restrict__ W, float xin, f4* __restrict__ act_tile) {
  float act[L][4 * NB];
  fwd_tile_act<L, NB>(W, xin, act);
  if (act_tile) store_act<L, NB>(act_tile, act);
  if constexpr (FwdPack<L, NB>::kBf)
    return tile_output<L, NB, PackedBf<L, NB>>(W, act);
  else
    return tile_output<L, NB>(W, act);
}

template <int Lm, int NBm, int Lp, int NBp>
__global__ __launch_bounds__(kFT, (FwdPack<Lp, NBp>::kBf || FwdPack<Lm, NBm>::kBf) ? 6 : 8) void mfma_fwd_kernel(MfmaTileArgs T, const float* __restrict__ params, int Hm, int Hp,
                                                       int64_t e0, int64_t np_, int64_t nm_, float* __restrict__ u_item,
                                                       int nblk_p, f4* __restrict__ act_out, int act_f4) {
  extern __shared__ __align__(16) float lds[];
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  const bool plus = (int)blockIdx.x < nblk_p;
  const int64_t ntiles = plus ? (np_ + 15) / 16 : (nm_ + 15) / 16;
  const int nblk = plus ? nblk_p : (int)gridDim.x - nblk_p;
  const int bid = plus ? (int)blockIdx.x : (int)blockIdx.x - nblk_p;
  const int PPp = 4 * Hp + (Lp - 1) * (Hp * Hp + Hp) + Hp + 1;
  if (plus)
    fwd_pack<Lp, NBp>(params, FlatNet{Hp, Lp, 0}, lds);
  else
    fwd_pack<Lm, NBm>(params, FlatNet{Hm, Lm, PPp}, lds);
  __syncthreads();
  const int64_t eb = plus ? e0 : e0 + np_;
  const int64_t ee = plus ? e0 + np_ : e0 + np_ + nm_;
  const int k = lane / 16;
  int64_t tile = (int64_t)bid * kFW + wave;
  float xin_next = item_input(T, eb + tile * 16 + (lane & 15), ee, k);
  for (; tile < ntiles; tile += (int64_t)nblk * kFW) {
    const int64_t item = eb + tile * 16 + (lane & 15);
    // this tile's input was loaded one tile ahead; load the next one's now
    const float xin = xin_next;
    xin_next = item_input(T, item + (int64_t)nblk * kFW * 16, ee, k);
    // stored activations (if asked for) of tile `tile` of the + net at global tile `tile`, of the - net after
    // the (np + 15) / 16 tiles of the + net
    f4* at = act_out ? act_out + ((plus ? tile : (np_ + 15) / 16 + tile) * act_f4) * 64 + lane : nullptr;
    const float u = plus ? fwd_tile_u<Lp, NBp>(lds, xin, at) : fwd_tile_u<Lm, NBm>(lds, xin, at);
    if (item < ee && lane < 16) u_item[item - e0] = u;
  }
}

// ---- XCD-contiguous chunks -------------------------------------------------------
// Workgroups go to the 8 XCDs round-robin by index; a grid of 8 * ceil(n / 8)
// workgroups maps workgroup b to chunk (b % 8) ceil(n / 8) + b / 8, so that each
// XCD (and its L2) takes one contiguous eighth of the n chunks.  Chunks >= n exit.
constexpr int kXcds = 8;
inline unsigned xcd_grid(int64_t n) { return (unsigned)(kXcds * ((n + kXcds - 1) / kXcds)); }
__device__ __forceinline__ int64_t xcd_chunk(int64_t n) {
  const int64_t per = (n + kXcds - 1) / kXcds;
  return (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// ---- kernel 2: r per row, loss partials ---------------------------------------
// u of CSR entry e: u[e] (per-entry values, e_item null), else u[e_item[e]] (per-item values).
// Launched on xcd_grid(rowsum_blocks(nrows)) workgroups of kRST threads; loss_part[chunk] per chunk of
// kMT rows.
constexpr int kRST = 1024;
__global__ __launch_bounds__(kRST) void mfma_rowsum_kernel(const int32_t* __restrict__ row_ptr,
                                                          const float* __restrict__ row_c, int64_t row_begin,
                                                          int64_t nrows, const float* __restrict__ ent_w, int64_t nnz,
                                                          const float* __restrict__ u, const int32_t* __restrict__ e_item,
                                                          float* __restrict__ r_out, double* __restrict__ loss_part) {
  const int64_t nch = (nrows + kMT - 1) / kMT;
  const int64_t chunk = xcd_chunk(nch);
  if (chunk >= nch) return;  // the whole workgroup
  // kRowLanes lanes per row (its entries read contiguously across them, so that the long rows of
  // interface points do not stall a whole wavefront and every load instruction is coalesced),
  // kRST / kRowLanes rows per pass, kMT rows per workgroup
  constexpr int kRowLanes = 8, kPer = 2, kRowsPass = kRST / kRowLanes;
  static_assert(kMT % kRowsPass == 0, "whole passes");
  const int sub = threadIdx.x % kRowLanes, slot = threadIdx.x / kRowLanes;
  double l = 0.0;
#pragma unroll
  for (int pass = 0; pass < kMT / kRowsPass; ++pass) {
    const int64_t i = chunk * kMT + pass * kRowsPass + slot;
    if (i < nrows) {  // uniform over the kRowLanes lanes of a row
      const int64_t row = row_begin + i;
      const int beg = max(row_ptr[row], 0);
      const int end = (int)min((int64_t)row_ptr[row + 1], nnz);
      // lane `sub` sums entries beg + sub + kRowLanes j in order; every load of an iteration is issued
      // before the first use (the gathers through e_item depend on one load, not on the previous entry)
      float acc = 0.0f;
      for (int e0 = beg + sub; e0 < end; e0 += kRowLanes * kPer) {
        float w[kPer], uv[kPer];
        int it[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
          const int e = e0 + kRowLanes * j;
          w[j] = e < end ? ent_w[e] : 0.0f;
          it[j] = e < end ? (e_item ? e_item[e] : e) : -1;
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) uv[j] = it[j] >= 0 ? u[it[j]] : 0.0f;
#pragma unroll
        for (int j = 0; j < kPer; ++j)
          if (it[j] >= 0) acc = fmaf(w[j], uv[j], acc);
      }
      // the lanes' partial sums in a fixed tree (addition is commutative: all lanes of the row agree)
      acc += __shfl_xor(acc, 1, 64);
      acc += __shfl_xor(acc, 2, 64);
      acc += __shfl_xor(acc, 4, 64);
      const float r = row_c[row] + acc;
      if (sub == 0) {
        r_out[i] = r;
        l += 0.5 * (double)r * (double)r;
      }
    }
  }
  l = dips_wave_sum_d(l);
  __shared__ double red[kRST / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kRST / 64; ++w) t += red[w];
    loss_part[chunk] = t;
  }
}

// ---- kernel 2b: seed per item ----------------------------------------------------
// seed of an item = sum over its member entries (in order) of r(row) w: the
// derivative of the batch's half sum of squares with respect to the item's u.
// Written to seed[item - i0] (the buffer that held u, consumed by the row sums).
// Launched on xcd_grid(ceil(n / 256)) workgroups of 256 threads.
__global__ __launch_bounds__(256) void mfma_seed_kernel(MfmaTileArgs T, int64_t i0, int64_t n, int64_t row_begin,
                                                        const float* __restrict__ r_row, float* __restrict__ seed) {
  const int64_t t = xcd_chunk((n + 255) / 256) * 256 + threadIdx.x;
  if (t >= n) return;
  int64_t mb, me;
  item_members(T, i0 + t, mb, me);
  // members in chunks of 4, loads of a chunk issued before the first use; the sum keeps the member order
  float s = 0.0f;
  for (int64_t c = mb; c < me; c += 4) {
    float w[4], rv[4];
    int64_t rr[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool in = c + j < me;
      w[j] = in ? T.t_w[c + j] : 0.0f;
      rr[j] = in ? (int64_t)T.t_row[c + j] - row_begin : -1;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) rv[j] = rr[j] >= 0 ? r_row[rr[j]] : 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (c + j == mb)
        s = rv[j] * w[j];
      else if (c + j < me)
        s = fmaf(rv[j], w[j], s);
    }
  }
  seed[t] = s;
}

// ---- kernel 3: parameter-gradient partials ---------------------------------------
// Cooperative backward: a workgroup of kBW wavefronts takes kBW tiles of one
// phase per round.  Every wavefront recomputes its tile's forward pass and
// back-propagates its deltas (registers); per layer, each wavefront stages its
// tile's operands (activation and delta, feature-major rows of 16 positions)
// in LDS, and after a barrier each wavefront contracts only the weight-gradient
// blocks it owns over all kBW tiles (K = 16 kBW positions).  Owning 4 of the
// 16 blocks of a 64-wide layer keeps a lane's registers within 256, and the
// packed weights (one copy, read transposed for the deltas) plus the stage fit
// twice into a CU's LDS: two workgroups per CU, whose barriers are independent,
// so one workgroup's wavefronts issue while the other's wait.
// Bias, output-weight and output-bias sums stay per lane and are reduced over
// the 16 positions with row-local DPP sums at the end.  The packed weights carry
// the factor kTanhExp2Scale, so each layer's delta comes out that factor larger
// than the one after it; the flush divides it out.
constexpr int kBW = 4;              // wavefronts per backward workgroup
constexpr int kBT = 64 * kBW;       // its threads

// Transposed split-bf16 copy of the hidden matrices for the backward pass: per hidden
// layer l, input block ib, chunk c of output blocks (2c, 2c+1) and part p one 64-lane
// set of 16-byte units, lane 16 g + i holding W_l[16 ib + i][16 (2c + t/4) + 4 g + t%4]
// (t = 0..7) times kTanhExp2Scale, part p: the A operand of the split-bf16 product
// W_l d_l, whose B operand is the split delta in its feature-major registers.
template <int L, int NB>
struct TransBf {
  static constexpr int NC = (NB + 1) / 2;
  static constexpr int units() { return (L > 1 ? L - 1 : 0) * NB * NC * 3 * 64; }
  static constexpr int size() { return 4 * units(); }  // floats
  static constexpr int set(int l, int ib, int c, int p) { return (((l - 1) * NB + ib) * NC + c) * 3 + p; }
};

template <int L, int NB>
__device__ void pack_trans_bf(const float* __restrict__ params, FlatNet fn, float* __restrict__ lds) {
  using TB = TransBf<L, NB>;
  const int Hh = fn.Hh;
  uint32_t* wb = reinterpret_cast<uint32_t*>(lds);
  for (int d = threadIdx.x; d < 4 * TB::units(); d += blockDim.x) {
    const int u = d % 4, lane = (d / 4) % 64;
    int rest = d / 256;
    const int p = rest % 3;
    rest /= 3;
    const int c = rest % TB::NC;
    rest /= TB::NC;
    const int ib = rest % NB, l = rest / NB + 1;
    const int g = lane / 16, in = 16 * ib + lane % 16;
    uint32_t word = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = 2 * u + h, ob = 2 * c + t / 4;
      const int out = 16 * ob + 4 * g + t % 4;
      float w = 0.0f;
      if (ob < NB && in < Hh && out < Hh) w = kTanhExp2Scale * params[fn.w(l) + in * Hh + out];
      word |= bf_part(w, p) << (16 * h);
    }
    wb[d] = word;
  }
}

// this lane's layer-0 operands (the packed value at ob * 64 + lane) and output weights (feature
// 16 fb + 4 (lane / 16) + r at fb * 4 + r), for the backward pass without the float32 copy in LDS
template <int NB>
__device__ __forceinline__ void lane_weights(const float* __restrict__ params, FlatNet fn, float (&w0)[NB],
                                             float (&wo)[4 * NB]) {
  const int lane = threadIdx.x & 63, Hh = fn.Hh;
#pragma unroll
  for (int ob = 0; ob < NB; ++ob) {
    const int out = 16 * ob + lane % 16, k = lane / 16;
    w0[ob] = out < Hh ? kTanhExp2Scale * ((k < 3) ? params[fn.w(0) + k * Hh + out] : params[fn.b(0) + out]) : 0.0f;
  }
#pragma unroll
  for (int q = 0; q < 4 * NB; ++q) {
    const int h = 16 * (q / 4) + 4 * (lane / 16) + q % 4;
    wo[q] = h < Hh ? params[fn.wout() + h] : 0.0f;
  }
}

template <int L, int NB>
struct CoopAcc {
  static constexpr int kW = (NB * NB + kBW - 1) / kBW;  // hidden-layer blocks owned per wavefront
  f4 dW[L > 1 ? L - 1 : 1][kW];                         // block q = wave * kW + j: (q / NB, q % NB)
  f4 dW0;                                               // layer 0, out block = wave (wave < NB)
  f4 dB[L > 1 ? L - 1 : 1];                             // hidden-layer biases, out block = wave (wave < NB):
                                                        // delta rows contracted with a row of ones (row 0)
  float dwo[4 * NB];                                    // per-lane partial output-weight sums
  float dbo;
};

// Operand stage geometry: rows of TP floats; with kPerm the rows of each group of 4
// are rotated (row 4G + r sits at 4G + (r + G) % 4), which with TP = 16 spreads the
// transposing writes (lanes of one write instruction hit rows 4g + r, g = 0..3)
// over all 64 banks; TP = 20 spreads them by padding instead.
template <int TP, bool kPerm>
struct Stage {
  static constexpr int kTPs = TP;
  static __device__ __forceinline__ int row(int r) { return kPerm ? (r & ~3) | ((r + (r >> 2)) & 3) : r; }
};
using StagePad = Stage<kTP, false>;
using StageRot = Stage<16, true>;

// transposes a feature-major tile matrix (NB*16 features x 16 positions) through LDS
template <int NB, class SG = StagePad>
__device__ __forceinline__ void to_lds(float* __restrict__ buf, const float (&m)[4 * NB]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int fb = 0; fb < NB; ++fb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[SG::row(16 * fb + 4 * (lane / 16) + r) * SG::kTPs + (lane & 15)] = m[fb * 4 + r];
  }
}

// sum over the 16 lanes of each row (lane 16g + 15 receives it): row-local DPP steps
__device__ __forceinline__ float row16_sum(float v) {
  v += dips_dpp_f<0x111, 0xF, true>(v);
  v += dips_dpp_f<0x112, 0xF, true>(v);
  v += dips_dpp_f<0x114, 0xF, true>(v);
  v += dips_dpp_f<0x118, 0xF, true>(v);
  return v;
}

// D += A^T B over the kBW staged tiles of the workgroup: A rows arow.. (arow < 0: a row
// of ones and 15 rows of zeros), B rows brow..  The 4 contraction steps of lane group kq
// take positions 4 kq .. 4 kq + 3 of a tile, one 16-byte LDS read per operand (rows of
// SG::kTPs floats, 16-byte aligned).
template <class SG = StagePad>
__device__ __forceinline__ f4 coop_block(const float* __restrict__ stage, int stride_w, int arow, int brow, bool a_inputs,
                                         f4 acc) {
  const int lane = threadIdx.x & 63;
  const int m = lane & 15, kq = lane / 16;
  constexpr int TP = SG::kTPs;
#pragma unroll
  for (int s = 0; s < kBW; ++s) {
    const float* A = stage + s * stride_w;
    const float* D = A + stride_w / 2;
    const f4 b4 = *reinterpret_cast<const f4*>(&D[SG::row(brow + m) * TP + 4 * kq]);
    f4 a4;
    if (arow < 0)  // a row of ones (row 0) and zeros: the delta rows summed over positions
      a4 = m == 0 ? f4{1.f, 1.f, 1.f, 1.f} : f4{0.f, 0.f, 0.f, 0.f};
    else if (a_inputs)
      a4 = m < 4 ? *reinterpret_cast<const f4*>(&A[SG::row(m) * TP + 4 * kq]) : f4{0.f, 0.f, 0.f, 0.f};
    else
      a4 = *reinterpret_cast<const f4*>(&A[SG::row(arow + m) * TP + 4 * kq]);
#pragma unroll
    for (int st = 0; st < 4; ++st) acc = mfma4(a4[st], b4[st], acc);
  }
  return acc;
}

// writes the workgroup's gradient of one phase net into out[0 .. flat size)
template <int L, int NB>
__device__ void flush_coop(CoopAcc<L, NB>& G, FlatNet fn, float* __restrict__ red, float* __restrict__ out) {
  using CA = CoopAcc<L, NB>;
  const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
  const int Hh = fn.Hh;
  const int P = 4 * Hh + (L - 1) * (Hh * Hh + Hh) + Hh + 1;  // parameters of this net
  const int base = fn.base;
  for (int i = threadIdx.x; i < P; i += kBT) red[i] = 0.0f;
  __syncthreads();
  // row sums of the per-lane partials (every wavefront, no LDS)
#pragma unroll
  for (int q = 0; q < 4 * NB; ++q) {
    G.dwo[q] = row16_sum(G.dwo[q]);
  }
  G.dbo = row16_sum(G.dbo);
  // the delta of hidden layer l carries kTanhExp2Scale^(L-1-l) (weights packed times the factor): inv[l]
  // undoes it for layer l's weight and bias gradients (layer 0: its weights and bias, rows 0..3 of dW0)
  float inv[L];
  inv[L - 1] = 1.0f;
#pragma unroll
  for (int l = L - 2; l >= 0; --l) inv[l] = inv[l + 1] * (1.0f / kTanhExp2Scale);
  // wavefronts add in order (fixed order: deterministic); weight blocks have one owner each
  for (int w = 0; w < kBW; ++w) {
    if (wave == w) {
#pragma unroll
      for (int l = 1; l < L; ++l) {
#pragma unroll
        for (int j = 0; j < CA::kW; ++j) {
          const int q = wave * CA::kW + j;
          if (q < NB * NB) {
            const int ib = q / NB, ob = q % NB;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int in = 16 * ib + 4 * (lane / 16) + r, o = 16 * ob + (lane & 15);
              if (in < Hh && o < Hh) red[fn.w(l) - base + in * Hh + o] = G.dW[l - 1][j][r] * inv[l];
            }
          }
        }
      }
      if (wave < NB) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 4 * (lane / 16) + r, o = 16 * wave + (lane & 15);
          if (row < 3 && o < Hh) red[fn.w(0) - base + row * Hh + o] = G.dW0[r] * inv[0];
          if (row == 3 && o < Hh) red[fn.b(0) - base + o] = G.dW0[r] * inv[0];
        }
        if (lane < 16 && 16 * wave + lane < Hh) {
#pragma unroll
          for (int l = 1; l < L; ++l) red[fn.b(l) - base + 16 * wave + lane] = G.dB[l - 1][0] * inv[l];
        }
      }
      if ((lane & 15) == 15) {
#pragma unroll
        for (int fb = 0; fb < NB; ++fb) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int h = 16 * fb + 4 * (lane / 16) + r;
            if (h < Hh) {
              red[fn.wout() - base + h] += G.dwo[fb * 4 + r];
            }
          }
        }
        if (lane == 15) red[fn.bout() - base] += G.dbo;
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < P; i += kBT) out[i] = red[i];
}

// kBf: W is the TransBf copy (W_l d_l as split bf16 products), the output weights come in wor and the
// stage rows are 16 floats with rotated row groups; else W is the float32 Packed copy (float32 MFMA) and
// the stage rows are padded to kTP floats
// prefetch() is called once per tile, as soon as act is no longer read (after the backprop through
// layer 1, or at the start for L = 1), so that loads for the next tile can take its registers and their
// latency overlaps the remaining contractions
template <int L, int NB, bool kBf, class Prefetch>
__device__ __forceinline__ void coop_backward(const float* __restrict__ W, const float (&wor)[4 * NB], float seed,
                                              float xin, const float (&act)[L][4 * NB], float* __restrict__ stage,
                                              int stride_w, CoopAcc<L, NB>& G, Prefetch&& prefetch) {
  using PK = Packed<L, NB>;
  using CA = CoopAcc<L, NB>;
  using SG = std::conditional_t<kBf, StageRot, StagePad>;
  const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
  float* bufA = stage + wave * stride_w;
  float* bufD = bufA + stride_w / 2;
  // output layer
  float d[4 * NB];
#pragma unroll
  for (int fb = 0; fb < NB; ++fb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float a = act[L - 1][fb * 4 + r];
      float wo;
      if constexpr (kBf)
        wo = wor[fb * 4 + r];
      else
        wo = W[PK::wout() + 16 * fb + 4 * (lane / 16) + r];
      G.dwo[fb * 4 + r] = fmaf(a, seed, G.dwo[fb * 4 + r]);
      d[fb * 4 + r] = seed * wo * (1.0f - a * a);
    }
  }
  if (lane < 16) G.dbo += seed;
  if constexpr (L == 1) prefetch();
#pragma unroll
  for (int l = L - 1; l >= 0; --l) {
    // stage this tile's operands: delta rows, and activation rows (or the inputs x, y, z, 1)
    to_lds<NB, SG>(bufD, d);
    if (l > 0)
      to_lds<NB, SG>(bufA, act[l - 1]);
    else
      bufA[SG::row(lane / 16) * SG::kTPs + (lane & 15)] = xin;
    if (l > 0) {
      // d_{l-1} = (W_l d_l) * (1 - a_{l-1}^2), from this tile's registers (the contraction below reads the
      // staged copy of d_l), before the barrier: a wavefront that arrives early does this matrix work while
      // it waits.
      f4 c[NB];
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) c[ib] = f4{0.f, 0.f, 0.f, 0.f};
      if constexpr (kBf) {
        // split-bf16: chunk cc of d_l (blocks 2cc, 2cc+1) is the B operand, split like the forward activations
        using TB = TransBf<L, NB>;
        const u4* __restrict__ tb = reinterpret_cast<const u4*>(W);
#pragma unroll
        for (int cc = 0; cc < TB::NC; ++cc) {
          u4 dh, dm, dl;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int fb = 2 * cc + u / 2, r = (2 * u) % 4;
            uint32_t h, mm, lo;
            bf_split_pair(fb < NB ? d[fb * 4 + r] : 0.0f, fb < NB ? d[fb * 4 + r + 1] : 0.0f, h, mm, lo);
            dh[u] = h;
            dm[u] = mm;
            dl[u] = lo;
          }
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) {
            const u4 w0 = tb[TB::set(l, ib, cc, 0) * 64 + lane];
            const u4 w1 = tb[TB::set(l, ib, cc, 1) * 64 + lane];
            const u4 w2 = tb[TB::set(l, ib, cc, 2) * 64 + lane];
            c[ib] = mfma_bf(w2, dh, c[ib]);
            c[ib] = mfma_bf(w1, dm, c[ib]);
            c[ib] = mfma_bf(w0, dl, c[ib]);
            c[ib] = mfma_bf(w1, dh, c[ib]);
            c[ib] = mfma_bf(w0, dm, c[ib]);
            c[ib] = mfma_bf(w0, dh, c[ib]);
          }
        }
      } else {
        // float32: the A operand of step r over output block fb is W_l[16 ib + m][16 fb + 4 kq + r]
        // (m = lane%16, kq = lane/16), the float m%4 of unit 16 (m/4) + 4 kq + r of block (fb, ib) -- 4
        // floats 4 apart from this lane's offset
        const int m = lane & 15, kq = lane >> 4;
        const int tofs = 4 * (17 * (m >> 2) + 4 * kq) + (m & 3);
#pragma unroll
        for (int fb = 0; fb < NB; ++fb) {
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) {
            const float* wt = W + PK::fwd(l) + (fb * NB + ib) * PK::kBlk + tofs;
#pragma unroll
            for (int r = 0; r < 4; ++r) c[ib] = mfma4(wt[4 * r], d[fb * 4 + r], c[ib]);
          }
        }
      }
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float a = act[l - 1][ib * 4 + r];
          d[ib * 4 + r] = c[ib][r] * (1.0f - a * a);
        }
      }
      if (l == 1) prefetch();
    }
    __syncthreads();
    if (l > 0) {
#pragma unroll
      for (int j = 0; j < CA::kW; ++j) {
        const int q = wave * CA::kW + j;
        if (q < NB * NB)
          G.dW[l - 1][j] = coop_block<SG>(stage, stride_w, 16 * (q / NB), 16 * (q % NB), false, G.dW[l - 1][j]);
      }
      // the bias gradient of out block `wave`: the staged deltas against a row of ones
      if (wave < NB) G.dB[l - 1] = coop_block<SG>(stage, stride_w, -1, 16 * wave, false, G.dB[l - 1]);
    } else if (wave < NB) {  // b_0 is row 3 of the layer-0 weight gradient
      G.dW0 = coop_block<SG>(stage, stride_w, 0, 16 * wave, true, G.dW0);
    }
    __syncthreads();  // operands consumed: the buffers may be restaged
  }
}

// LDS of the backward kernel for one phase net (floats): its backprop weights (the TransBf copy from two
// 16-wide blocks on, the float32 Packed copy below), with kStored false the forward pass's split-bf16 copy
// (PackedBf, where the forward pass uses one), then the operand stage of the kBW wavefronts (rows of 16
// floats, rotated, with TransBf; of kTP floats with Packed).  3x64 with kStored: 48 KB + 32 KB = 80 KB,
// two workgroups per CU.
template <int L, int NB, bool kStored>
struct BwdPhase {
  static constexpr bool kBf = FwdPack<L, NB>::kBf;
  static constexpr int wr() { return ((kBf ? TransBf<L, NB>::size() : Packed<L, NB>::size()) + 3) & ~3; }
  static constexpr int fp() { return (!kStored && kBf) ? ((PackedBf<L, NB>::size() + 3) & ~3) : 0; }
  static constexpr int stride() { return 2 * 16 * NB * (kBf ? StageRot::kTPs : StagePad::kTPs); }
  static constexpr int floats() { return wr() + fp() + kBW * stride(); }
};
template <int Lm, int NBm, int Lp, int NBp, bool kStored>
struct BwdLds {
  static constexpr int kP = BwdPhase<Lp, NBp, kStored>::floats(), kM = BwdPhase<Lm, NBm, kStored>::floats();
  static constexpr int floats() { return kP > kM ? kP : kM; }
};

// kStored: the activations of every tile come from the forward kernel (act_in, act_f4 f4s per lane and
// tile, + tiles first); else each wavefront recomputes them with the forward kernel's arithmetic (the
// same split-bf16 or float32 pass, so both variants give the same bits)
template <int Lm, int NBm, int Lp, int NBp, bool kStored>
__global__ __launch_bounds__(kBT, 2) void mfma_bwd_kernel(MfmaTileArgs T, const float* __restrict__ params, int Hm, int Hp,
                                                       int64_t e0, int64_t np_, int64_t nm_,
                                                       const float* __restrict__ seed_item, float* __restrict__ gpart,
                                                       int nblk_p, int stride, const f4* __restrict__ act_in,
                                                       int act_f4) {
  extern __shared__ __align__(16) float lds[];
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  const bool plus = (int)blockIdx.x < nblk_p;
  const int PPp = 4 * Hp + (Lp - 1) * (Hp * Hp + Hp) + Hp + 1;
  const int64_t ntp = (np_ + 15) / 16;
  const int64_t ntiles = plus ? (np_ + 15) / 16 : (nm_ + 15) / 16;
  const int nblk = plus ? nblk_p : (int)gridDim.x - nblk_p;
  const int bid = plus ? (int)blockIdx.x : (int)blockIdx.x - nblk_p;
  const int64_t eb = plus ? e0 : e0 + np_;
  const int64_t ee = plus ? e0 + np_ : e0 + np_ + nm_;
  float* out = gpart + (int64_t)blockIdx.x * stride;
  auto run = [&](auto tagL, auto tagNB, FlatNet fn) {
    constexpr int L = decltype(tagL)::value;
    constexpr int NB = decltype(tagNB)::value;
    using CA = CoopAcc<L, NB>;
    using BP = BwdPhase<L, NB, kStored>;
    // the forward copy and the operand stage start on 16-byte boundaries after the backprop weights
    float* fwdW = lds + BP::wr();
    float* stage = lds + BP::wr() + BP::fp();
    if constexpr (BP::kBf) {
      pack_trans_bf<L, NB>(params, fn, lds);
      if constexpr (!kStored) pack_weights_bf<L, NB>(params, fn, fwdW);
    } else {
      pack_weights<L, NB>(params, fn, lds);
    }
    __syncthreads();
    CA G;
#pragma unroll
    for (int l = 0; l < (L > 1 ? L - 1 : 1); ++l)
#pragma unroll
      for (int q = 0; q < CA::kW; ++q) G.dW[l][q] = f4{0.f, 0.f, 0.f, 0.f};
    G.dW0 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int l = 0; l < (L > 1 ? L - 1 : 1); ++l) G.dB[l] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 4 * NB; ++q) G.dwo[q] = 0.0f;
    G.dbo = 0.0f;
    // every wavefront runs the same number of rounds (the barriers inside);
    // tiles past the end carry no entries (zero seeds, zero inputs)
    const int64_t per_round = (int64_t)nblk * kBW;
    const int64_t rounds = (ntiles + per_round - 1) / per_round;
    const int kq = lane / 16;
    // input and seed of a round's item are loaded one round ahead, so that their
    // latency overlaps the previous round's matrix work
    auto item_of = [&](int64_t rd) { return eb + ((rd * nblk + bid) * kBW + wave) * 16 + (lane & 15); };
    auto seed_of = [&](int64_t item) { return item < ee ? seed_item[item - e0] : 0.0f; };
    float xin_next = item_input(T, item_of(0), ee, kq);
    float seed_next = seed_of(item_of(0));
    // the stored activations of a round's tile (layers 1..L-1), loaded for round rd + 1 inside round rd
    // once its activations are dead
    float act_n[L][4 * NB];
    auto load_round = [&](int64_t rd) {
      if constexpr (kStored) {
        const int64_t tile = (rd * nblk + bid) * kBW + wave;
        if (tile < ntiles) {
          load_act<L, NB>(act_in + ((plus ? tile : ntp + tile) * act_f4) * 64 + lane, act_n);
        } else {
#pragma unroll
          for (int l = 1; l < L; ++l)
#pragma unroll
            for (int q = 0; q < 4 * NB; ++q) act_n[l][q] = 0.0f;
        }
      }
    };
    load_round(0);
    for (int64_t rd = 0; rd < rounds; ++rd) {
      const float xin = xin_next, seed = seed_next;
      if (rd + 1 < rounds) {
        const int64_t nx = item_of(rd + 1);
        xin_next = item_input(T, nx, ee, kq);
        seed_next = seed_of(nx);
      }
      float act[L][4 * NB];
      // with TransBf: layer 0's operands and the output weights of this lane, loaded per round (cached
      // reads of the parameter vector; held across rounds they would spill)
      float w0r[NB], wor[4 * NB];
      if constexpr (BP::kBf) lane_weights<NB>(params, fn, w0r, wor);
      if constexpr (kStored) {
#pragma unroll
        for (int l = 1; l < L; ++l)
#pragma unroll
          for (int q = 0; q < 4 * NB; ++q) act[l][q] = act_n[l][q];
        if constexpr (BP::kBf)
          tile_layer0_regs<L, NB>(w0r, xin, act);
        else
          tile_layer0<L, NB>(lds, xin, act);
      } else if constexpr (BP::kBf) {
        tile_forward_bf<L, NB>(fwdW, xin, act);
      } else {
        tile_forward<L, NB>(lds, xin, act);
      }
      coop_backward<L, NB, BP::kBf>(lds, wor, seed, xin, act, stage, BP::stride(), G, [&] {
        if (rd + 1 < rounds) load_round(rd + 1);
      });
    }
    __syncthreads();  // all waves done with the packed weights: reuse LDS for the reduction
    flush_coop<L, NB>(G, fn, lds, out);
  };
  if (plus)
    run(std::integral_constant<int, Lp>{}, std::integral_constant<int, NBp>{}, FlatNet{Hp, Lp, 0});
  else
    run(std::integral_constant<int, Lm>{}, std::integral_constant<int, NBm>{}, FlatNet{Hm, Lm, PPp});
  (void)lane;
}

// ---- kernel 4/5: reduction of gradient partials and optimizer -------------------
// g[p] = inv_b * sum over the workgroups of its phase (fixed order); sumsq partial per workgroup
// 64 parameters per workgroup, each summed over its phase's partial rows by 4
// wavefronts (a quarter of the rows each, in order), then the quarters in order;
// sq_part[workgroup] = sum of g^2 over its 64 parameters.  With count_io, the
// first thread copies the step counter to count_snap (read by the optimizer)
// and advances it, in place of a separate launch.
constexpr int kRedParams = 64;
__global__ __launch_bounds__(256) void mfma_reduce_kernel(const float* __restrict__ gpart, int stride, int nblk_p,
                                                          int nblk_m, int PPp, int PPm, float inv_b,
                                                          float* __restrict__ g, double* __restrict__ sq_part,
                                                          int32_t* __restrict__ count_io,
                                                          int32_t* __restrict__ count_snap) {
  const int j = threadIdx.x & 63, k = threadIdx.x >> 6;
  const int p = blockIdx.x * kRedParams + j;
  double acc = 0.0;
  if (p < PPp + PPm) {
    const bool plus = p < PPp;
    const int b0 = plus ? 0 : nblk_p, nb = plus ? nblk_p : nblk_m;
    const int q = plus ? p : p - PPp;
    const int per = (nb + 3) / 4;
    const int lo = b0 + k * per, hi = b0 + min(nb, (k + 1) * per);
#pragma unroll 8
    for (int b = lo; b < hi; ++b) acc += (double)gpart[(int64_t)b * stride + q];
  }
  __shared__ double part[4][kRedParams];
  part[k][j] = acc;
  __syncthreads();
  if (k == 0) {
    double v = 0.0;
    if (p < PPp + PPm) {
      const float gg = (float)((((part[0][j] + part[1][j]) + part[2][j]) + part[3][j]) * (double)inv_b);
      g[p] = gg;
      v = (double)gg * (double)gg;
    }
    v = dips_wave_sum_d(v);
    if (j == 0) sq_part[blockIdx.x] = v;
  }
  if (count_io && blockIdx.x == 0 && threadIdx.x == 0) {
    const int32_t c = *count_io;
    *count_snap = c;
    *count_io = c + 1;
  }
}

// Parameters [0, split) take their gradient from g, [split, P) from
// g2[p - split] * g2_scale (the hashgrid table's accumulated gradient, read in
// place); the grid strides over P.
__global__ __launch_bounds__(256) void mfma_apply_kernel(const float* __restrict__ g, int P,
                                                         const float* __restrict__ g2, int split, float g2_scale,
                                                         const double* __restrict__ sq_part, int nsq,
                                                         const double* __restrict__ loss_part, int nloss, float inv_b,
                                                         dips_opt_t opt, float* __restrict__ params,
                                                         float* __restrict__ m, float* __restrict__ v,
                                                         const int32_t* __restrict__ count_in,
                                                         float* __restrict__ loss_hist, int loss_cap) {
  // fixed-order float64 sums spread over the workgroup: per thread a strided
  // subset in order, then the DPP wave tree, then the 4 waves in order
  __shared__ double red[4];
  __shared__ float gn_s;
  const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
  double t = 0.0;
  for (int b = threadIdx.x; b < nsq; b += 256) t += sq_part[b];
  t = dips_wave_sum_d(t);
  if (lane == 0) red[wave] = t;
  __syncthreads();
  if (threadIdx.x == 0) gn_s = (float)sqrt(((red[0] + red[1]) + red[2]) + red[3]);
  __syncthreads();
  const int c = *count_in;
  const float gn = gn_s;
  const bool clip = (opt.kind == 0) && !(gn < opt.max_norm);
  float lr_t = opt.lr;
  if (opt.kind == 0) {
    if (opt.sched == 0)
      lr_t = (c <= 0) ? opt.lr : opt.lr * powf(opt.decay_rate, (float)c / opt.transition_steps);
    else
      lr_t = opt.lr * (1.0f - fminf(fmaxf((float)c, 0.0f), opt.transition_steps) / opt.transition_steps);
  }
  const float t1 = (float)(c + 1);
  const float bc1 = 1.0f - powf(0.9f, t1), bc2 = 1.0f - powf(0.999f, t1);
  for (int p = blockIdx.x * 256 + threadIdx.x; p < P; p += gridDim.x * 256) {
    float gg = p < split ? g[p] : g2[p - split] * g2_scale;
    if (clip) gg = (gg / gn) * opt.max_norm;
    float u;
    if (opt.kind == 2) {
      const float nv = 0.1f * (gg * gg) + 0.9f * v[p];
      v[p] = nv;
      u = -(gg / sqrtf(nv + 1e-8f)) * opt.lr;
    } else {
      const float nm = 0.1f * gg + 0.9f * m[p];
      const float nv = 0.001f * (gg * gg) + 0.999f * v[p];
      m[p] = nm;
      v[p] = nv;
      u = (nm / bc1) / (sqrtf(nv / bc2) + 1e-8f);
      u = (opt.kind == 0) ? -(lr_t * u) : -(u * opt.lr);
    }
    params[p] = params[p] + u;
  }
  if (blockIdx.x == 0) {
    double l = 0.0;
    for (int b = threadIdx.x; b < nloss; b += 256) l += loss_part[b];
    l = dips_wave_sum_d(l);
    __syncthreads();
    if (lane == 0) red[wave] = l;
    __syncthreads();
    if (threadIdx.x == 0 && c >= 0 && c < loss_cap)
      loss_hist[c] = (float)((((red[0] + red[1]) + red[2]) + red[3]) * (double)inv_b);
  }
}

__global__ void count_inc_kernel(int32_t* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = *count + 1;
}

// ---- evaluation at arbitrary points: u and grad_x u (forward mode) ----------------
// Value and the three input tangents of one 16-position tile, propagated
// together: a_0 = tanh(W_0 [x,1]), t_0^k = (1 - a_0^2) W_0[k,:]; hidden layers
// a_l = tanh(W_l a_{l-1} + b_l), t_l^k = (1 - a_l^2) (W_l t_{l-1}^k), sharing each
// weight operand between the four MFMA chains.  The packed forward weights carry
// the factor kTanhExp2Scale, which the tangents divide out layer by layer.
// Returns u and du/dx_k of the tile position lane%16 (identical in the 4 lanes
// of a position).
template <int L, int NB>
__device__ __forceinline__ void tile_value_grad(const float* __restrict__ W, float xin, float& u, float (&du)[3]) {
  using PK = Packed<L, NB>;
  const int lane = threadIdx.x & 63;
  float a[4 * NB], t[3][4 * NB];
#pragma unroll
  for (int ob = 0; ob < NB; ++ob) {
    f4 c = {0.f, 0.f, 0.f, 0.f};
    c = mfma4(W[PK::l0() + ob * 64 + lane], xin, c);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = dips_tanh_exp2(c[r]);
      const float g = (1.0f - v * v) * (1.0f / kTanhExp2Scale);
      a[ob * 4 + r] = v;
#pragma unroll
      for (int k = 0; k < 3; ++k) t[k][ob * 4 + r] = g * W[PK::l0() + ob * 64 + 16 * k + 4 * (lane / 16) + r];
    }
  }
#pragma unroll
  for (int l = 1; l < L; ++l) {
    f4 c[NB], ct[3][NB];
#pragma unroll
    for (int ob = 0; ob < NB; ++ob) {
#pragma unroll
      for (int r = 0; r < 4; ++r) c[ob][r] = W[PK::bias(l) + 16 * ob + 4 * (lane / 16) + r];
#pragma unroll
      for (int k = 0; k < 3; ++k) ct[k][ob] = f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int fb = 0; fb < NB; ++fb) {
#pragma unroll
      for (int ob = 0; ob < NB; ++ob) {
        const f4 w4 = *reinterpret_cast<const f4*>(&W[PK::fwd(l) + (ob * NB + fb) * PK::kBlk + 4 * PK::unit(lane)]);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          c[ob] = mfma4(w4[r], a[fb * 4 + r], c[ob]);
#pragma unroll
          for (int k = 0; k < 3; ++k) ct[k][ob] = mfma4(w4[r], t[k][fb * 4 + r], ct[k][ob]);
        }
      }
    }
#pragma unroll
    for (int ob = 0; ob < NB; ++ob) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float v = dips_tanh_exp2(c[ob][r]);
        const float g = (1.0f - v * v) * (1.0f / kTanhExp2Scale);
        a[ob * 4 + r] = v;
#pragma unroll
        for (int k = 0; k < 3; ++k) t[k][ob * 4 + r] = g * ct[k][ob][r];
      }
    }
  }
  float s = 0.0f, d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
#pragma unroll
  for (int fb = 0; fb < NB; ++fb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float wo = W[PK::wout() + 16 * fb + 4 * (lane / 16) + r];
      s = fmaf(a[fb * 4 + r], wo, s);
      d0 = fmaf(t[0][fb * 4 + r], wo, d0);
      d1 = fmaf(t[1][fb * 4 + r], wo, d1);
      d2 = fmaf(t[2][fb * 4 + r], wo, d2);
    }
  }
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  d0 += __shfl_xor(d0, 16, 64);
  d0 += __shfl_xor(d0, 32, 64);
  d1 += __shfl_xor(d1, 16, 64);
  d1 += __shfl_xor(d1, 32, 64);
  d2 += __shfl_xor(d2, 16, 64);
  d2 += __shfl_xor(d2, 32, 64);
  u = s + W[PK::bout()];
  du[0] = d0;
  du[1] = d1;
  du[2] = d2;
}

// u, grad u and grad u . n at m points; both networks packed once per
// workgroup (+ net first); a wavefront takes 16 consecutive points and runs
// each network only if one of its points selects it.
template <int Lm, int NBm, int Lp, int NBp>
__global__ __launch_bounds__(kMT) void mfma_eval_kernel(const float* __restrict__ params, int Hm, int Hp,
                                                        const float* __restrict__ pts,
                                                        const uint8_t* __restrict__ phase,
                                                        const float* __restrict__ nrm, int64_t m,
                                                        float* __restrict__ u, float* __restrict__ grad,
                                                        float* __restrict__ gradn) {
  extern __shared__ __align__(16) float lds[];
  constexpr int kOffM = (Packed<Lp, NBp>::size() + 3) & ~3;  // the - net starts 16-byte aligned
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  const int PPp = 4 * Hp + (Lp - 1) * (Hp * Hp + Hp) + Hp + 1;
  pack_weights<Lp, NBp>(params, FlatNet{Hp, Lp, 0}, lds);
  pack_weights<Lm, NBm>(params, FlatNet{Hm, Lm, PPp}, lds + kOffM);
  __syncthreads();
  const int64_t ntiles = (m + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < ntiles; tile += (int64_t)gridDim.x * 4) {
    const int64_t p = tile * 16 + (lane & 15);
    const bool valid = p < m;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    bool plus = false;
    if (valid) {
      x = pts[3 * p];
      y = pts[3 * p + 1];
      z = pts[3 * p + 2];
      plus = phase[p] != 0;
    }
    const int k = lane / 16;
    const float xin = (k == 0) ? x : (k == 1) ? y : (k == 2) ? z : 1.0f;
    float up = 0.0f, um = 0.0f, gp[3] = {0.f, 0.f, 0.f}, gm[3] = {0.f, 0.f, 0.f};
    if (__any(valid && plus)) tile_value_grad<Lp, NBp>(lds, xin, up, gp);
    if (__any(valid && !plus)) tile_value_grad<Lm, NBm>(lds + kOffM, xin, um, gm);
    if (valid && lane < 16) {
      const float v = plus ? up : um;
      const float g0 = plus ? gp[0] : gm[0], g1 = plus ? gp[1] : gm[1], g2 = plus ? gp[2] : gm[2];
      u[p] = v;
      if (grad) {
        grad[3 * p] = g0;
        grad[3 * p + 1] = g1;
        grad[3 * p + 2] = g2;
      }
      if (gradn && nrm) gradn[p] = g0 * nrm[3 * p] + g1 * nrm[3 * p + 1] + g2 * nrm[3 * p + 2];
    }
  }
}

// loss = inv_b * sum of the row-sum partials (fixed order), gradient copy-out
__global__ __launch_bounds__(256) void mfma_loss_out_kernel(const double* __restrict__ loss_part, int nloss,
                                                            float inv_b, double* __restrict__ loss_out) {
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
  double l = 0.0;
  for (int b = threadIdx.x; b < nloss; b += 256) l += loss_part[b];
  l = dips_wave_sum_d(l);
  if (lane == 0) red[wave] = l;
  __syncthreads();
  if (threadIdx.x == 0) *loss_out = (((red[0] + red[1]) + red[2]) + red[3]) * (double)inv_b;
}

// ---- host side ----------------------------------------------------------------
// launchers of the row-sum, optimizer and counter kernels for other entry-parallel
// networks (the hashgrid network)
void rowsum_launch(const int32_t* row_ptr, const float* row_c, int64_t row_begin, int64_t nrows, const float* ent_w,
                   int64_t nnz, const float* u, float* r, double* loss_part, hipStream_t st) {
  const int nrb = (int)((nrows + kMT - 1) / kMT);
  hipLaunchKernelGGL(mfma_rowsum_kernel, dim3(xcd_grid(nrb)), dim3(kRST), 0, st, row_ptr, row_c, row_begin, nrows, ent_w,
                     nnz, u, (const int32_t*)nullptr, r, loss_part);
}

int rowsum_blocks(int64_t nrows) { return (int)((nrows + kMT - 1) / kMT); }

void apply_launch(const float* g, int64_t P, const float* g2, int64_t split, float g2_scale, const double* sq_part,
                  int nsq, const double* loss_part, int nloss, float inv_b, const dips_opt_t& opt, float* params,
                  float* m, float* v, int32_t* count, float* loss_hist, int loss_cap, hipStream_t st) {
  const int64_t nb = std::min<int64_t>((P + 255) / 256, kApplyBlocks);
  hipLaunchKernelGGL(mfma_apply_kernel, dim3((unsigned)nb), dim3(256), 0, st, g, (int)P, g2, (int)split, g2_scale,
                     sq_part, nsq, loss_part, nloss, inv_b, opt, params, m, v, count, loss_hist, loss_cap);
  hipLaunchKernelGGL(count_inc_kernel, dim3(1), dim3(64), 0, st, count);
}

void loss_out_launch(const double* loss_part, int nloss, float inv_b, double* loss_out, hipStream_t st) {
  hipLaunchKernelGGL(mfma_loss_out_kernel, dim3(1), dim3(256), 0, st, loss_part, nloss, inv_b, loss_out);
}

static int nb_of(int H) { return (H + 15) / 16; }

#define DIPS_MFMA_SHAPES(X) \
  X(1, 1, 2, 1)             \
  X(2, 1, 2, 1)             \
  X(3, 4, 3, 4)             \
  X(1, 1, 3, 4)             \
  X(2, 2, 2, 2)             \
  X(3, 2, 3, 2)

int mfma_supported(const dips_net_t* n) {
  if (n->Hm > 64 || n->Hp > 64 || n->pc_w1 > 0 || n->act != 0) return 0;
  const int a = n->Lm, b = nb_of(n->Hm), c = n->Lp, d = nb_of(n->Hp);
#define X(A, B, C, D) \
  if (a == A && b == B && c == C && d == D) return 1;
  DIPS_MFMA_SHAPES(X)
#undef X
  return 0;
}

static int net_params(int L, int H) { return 4 * H + (L - 1) * (H * H + H) + H + 1; }

// f4s per lane of one tile's stored activations (hidden layers 1..L-1 of the wider phase net: (L - 1) x NB)
static int act_f4_of(const dips_net_t* n) {
  return std::max((n->Lm - 1) * nb_of(n->Hm), (n->Lp - 1) * nb_of(n->Hp));
}

// the workspace without the stored activations: u per entry, r per row, gradient partials (bwd
// workgroups x max net size), g, the step counter's snapshot
static int64_t mfma_base_floats(const dips_net_t* n, int64_t max_rows, int64_t max_nnz) {
  const int P = net_params(n->Lm, n->Hm) + net_params(n->Lp, n->Hp);
  const int stride = std::max(net_params(n->Lm, n->Hm), net_params(n->Lp, n->Hp));
  return max_nnz + max_rows + (int64_t)kMfmaBwdBlocks * stride + P + 64;
}

// ... plus, 256-byte aligned, the hidden activations of every forward tile (at most max_nnz / 16 + 2 tiles:
// an item is at least one entry, and each phase rounds up to whole tiles), which the backward kernel reads
// instead of recomputing them
int64_t mfma_workspace_floats(const dips_net_t* n, int64_t max_rows, int64_t max_nnz) {
  return mfma_base_floats(n, max_rows, max_nnz) + 64 + (max_nnz / 16 + 2) * (int64_t)act_f4_of(n) * 256;
}

// dips_eval for the network shapes of the MFMA path
int mfma_eval_launch(const dips_net_t* net, const float* params, const float* pts, const uint8_t* phase,
                     const float* nrm, int64_t m, float* u, float* grad, float* gradn, hipStream_t st) {
  if (!mfma_supported(net)) {
    set_error("dips_eval: network shape has no compiled kernel");
    return DIPS_EUNSUPPORTED;
  }
  if (m <= 0) return DIPS_OK;
  const int Hm = net->Hm, Hp = net->Hp;
  const int a = net->Lm, b = nb_of(Hm), c = net->Lp, d = nb_of(Hp);
  const int64_t ntiles = (m + 15) / 16;
  const int nblk = (int)std::min<int64_t>((ntiles + 3) / 4, 1024);
#define X(A, B, C, D)                                                                                            \
  if (a == A && b == B && c == C && d == D) {                                                                   \
    const size_t lds = sizeof(float) * (((Packed<C, D>::size() + 3) & ~3) + Packed<A, B>::size());             \
    hipFuncSetAttribute((const void*)mfma_eval_kernel<A, B, C, D>, hipFuncAttributeMaxDynamicSharedMemorySize,  \
                        (int)lds);                                                                             \
    hipLaunchKernelGGL((mfma_eval_kernel<A, B, C, D>), dim3(nblk), dim3(kMT), lds, st, params, Hm, Hp, pts, phase, \
                       nrm, m, u, grad, gradn);                                                                 \
    return hip_status(hipGetLastError(), "mfma_eval_kernel launch");                                            \
  }
  DIPS_MFMA_SHAPES(X)
#undef X
  set_error("dips_eval: network shape has no compiled kernel");
  return DIPS_EUNSUPPORTED;
}

// Workgroups of a persistent kernel resident at once on the current device
// (compute units x occupancy), looked up once per kernel and device; fallback
// if the runtime cannot say.  A persistent grid of exactly this size runs in
// one wave of workgroups, so no workgroup starts after others have finished.
int resident_blocks(const void* fn, int threads, size_t lds, int fallback) {
  static std::mutex mu;
  static std::vector<std::pair<std::pair<const void*, int>, int>> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fallback;
  std::lock_guard<std::mutex> lock(mu);
  for (const auto& e : cache)
    if (e.first.first == fn && e.first.second == dev) return e.second;
  int per_cu = 0, cus = 0;
  int n = fallback;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && per_cu > 0 && cus > 0)
    n = per_cu * cus;
  cache.push_back({{fn, dev}, n});
  return n;
}

// opt == nullptr: stop after the reduction (gradient in grad_out, loss in loss_out)
int mfma_step(const dips_grid_t* g, const dips_records_t* rec, const dips_mfma_batch_t* bt, const dips_net_t* net,
              const dips_opt_t* opt, float inv_b, float* params, float* m, float* v, int32_t* count, float* loss_hist,
              int loss_cap, float* work, int64_t work_floats, double* dwork, int64_t dwork_doubles, float* grad_out,
              double* loss_out, hipStream_t st) {
  if (net->pc_w1 > 0 || net->act != 0) {
    set_error("dips_mfma_train_step: the preconditioner and CELU networks run in the tile kernel only (dips_tile_*)");
    return DIPS_EUNSUPPORTED;
  }
  const int Lm = net->Lm, Hm = net->Hm, Lp = net->Lp, Hp = net->Hp;
  const int PPp = net_params(Lp, Hp), PPm = net_params(Lm, Hm);
  const int P = PPp + PPm;
  const int stride = std::max(PPp, PPm);
  const int64_t nrows = bt->nrows, np_ = bt->n_plus, nm_ = bt->n_minus;
  const int64_t nent = bt->ent_end - bt->ent_begin;
  if (np_ + nm_ != bt->t_count || bt->t_count < 0 || bt->t_count != nent) {
    // every entry of the batch's CSR range must be listed once: the row sums read all of their values
    set_error("dips_mfma_train_step: invalid batch (t_count must equal ent_end - ent_begin and n_plus + n_minus)");
    return DIPS_EINVAL;
  }
  // workspace layout (floats): u per item, then seed per item (at most one item per entry) | r rows |
  // gradient partials | g | the step counter's value for the optimizer (one int32)
  float* u = work;  // indexed by item - i0
  float* r = u + nent;
  float* gpart = r + nrows;
  float* gvec = gpart + (int64_t)kMfmaBwdBlocks * stride;
  int32_t* count_snap = reinterpret_cast<int32_t*>(gvec + P);
  if ((int64_t)(gvec + P + 1 - work) > work_floats) {
    set_error("dips_mfma_train_step: workspace too small");
    return DIPS_EINVAL;
  }
  const int nrb = (int)((nrows + kMT - 1) / kMT);
  const int nred = (P + kRedParams - 1) / kRedParams;
  double* loss_part = dwork;
  double* sq_part = dwork + nrb;
  if (nrb + nred + 1 > dwork_doubles) {
    set_error("dips_mfma_train_step: float64 workspace too small");
    return DIPS_EINVAL;
  }
  // items of the forward and backward tiles: the batch's groups of entries at one position when it lists
  // them (each evaluated once), else its entries
  const bool grouped = bt->t_group != nullptr;
  if (grouped && (bt->g_begin < 0 || bt->g_plus < 0 || bt->g_minus < 0 || bt->g_plus + bt->g_minus > bt->t_count)) {
    set_error("dips_mfma_train_step: invalid entry groups");
    return DIPS_EINVAL;
  }
  const int64_t i0 = grouped ? bt->g_begin : bt->t_begin;
  const int64_t ip = grouped ? bt->g_plus : np_, im = grouped ? bt->g_minus : nm_;
  if (!bt->i_pos || !bt->e_item || bt->i_count < i0 + ip + im) {
    set_error("dips_mfma_train_step: the batch needs its level's item positions and entry -> item map "
              "(i_pos, i_count, e_item; dips_mfma_items)");
    return DIPS_EINVAL;
  }
  MfmaTileArgs T{point_set(g), rec->hx, rec->hy, rec->hz, rec->row_point, bt->t_row, bt->t_code, bt->t_w,
                 bt->t_csr, bt->t_group, bt->i_pos, bt->i_count};
  const int a = Lm, b = nb_of(Hm), c = Lp, d = nb_of(Hp);
  const int64_t ntp = (ip + 15) / 16, ntm = (im + 15) / 16;
  // the forward tiles' activations for the backward, 256-byte aligned after the counter snapshot, when the
  // workspace holds them (dips_mfma_workspace sizes it so); else the backward recomputes them
  const int act_f4 = act_f4_of(net);
  const int64_t act_off = ((int64_t)(gvec + P + 1 - work) + 63) & ~(int64_t)63;
  const bool stored = act_off + (ntp + ntm) * act_f4 * 256 <= work_floats;
  f4* act = stored ? reinterpret_cast<f4*>(work + act_off) : nullptr;
  // workgroups split by phase in proportion to the tiles (each at least 1 if it has tiles)
  auto split = [&](int total, int& p, int& m) {
    p = (ntp > 0) ? (int)std::max<int64_t>(1, (int64_t)((double)total * ntp / std::max<int64_t>(1, ntp + ntm))) : 0;
    m = (ntm > 0) ? std::max(1, total - p) : 0;
    if (p + m > total) p = total - m;
  };
  int nbp, nbm, nfp, nfm;
  split(kMfmaBwdBlocks, nbp, nbm);
  bool launched = false;
#define X(A, B, C, D)                                                                                              \
  if (!launched && a == A && b == B && c == C && d == D) {                                                      \
    launched = true;                                                                                               \
    constexpr int PKF = FwdPack<C, D>::size() > FwdPack<A, B>::size() ? FwdPack<C, D>::size() : FwdPack<A, B>::size(); \
    const size_t lds_f = sizeof(float) * PKF;                                                                      \
    const size_t red_b = sizeof(float) * std::max(PPp, PPm);                                                       \
    const size_t bwd_lds = std::max(sizeof(float) * (stored ? BwdLds<A, B, C, D, true>::floats()                   \
                                                            : BwdLds<A, B, C, D, false>::floats()), red_b);        \
    if (bwd_lds > 160 * 1024 || lds_f > 160 * 1024) {                                                              \
      set_error("dips_mfma_train_step: network too large for the LDS of one workgroup");                            \
      return DIPS_EUNSUPPORTED;                                                                                    \
    }                                                                                                              \
    /* LDS limits set once per instantiation (for its widest networks), so that a step is launches only */      \
    static const bool attrs_set = [] {                                                                             \
      const size_t red_max = sizeof(float) * std::max(net_params(C, 16 * D), net_params(A, 16 * B));             \
      hipFuncSetAttribute((const void*)mfma_fwd_kernel<A, B, C, D>, hipFuncAttributeMaxDynamicSharedMemorySize,    \
                          (int)(sizeof(float) * PKF));                                                             \
      hipFuncSetAttribute((const void*)mfma_bwd_kernel<A, B, C, D, true>,                                          \
                          hipFuncAttributeMaxDynamicSharedMemorySize,                                              \
                          (int)std::max(sizeof(float) * BwdLds<A, B, C, D, true>::floats(), red_max));             \
      hipFuncSetAttribute((const void*)mfma_bwd_kernel<A, B, C, D, false>,                                         \
                          hipFuncAttributeMaxDynamicSharedMemorySize,                                              \
                          (int)std::max(sizeof(float) * BwdLds<A, B, C, D, false>::floats(), red_max));            \
      return true;                                                                                                 \
    }();                                                                                                           \
    (void)attrs_set;                                                                                               \
    split(resident_blocks((const void*)mfma_fwd_kernel<A, B, C, D>, kFT, lds_f, kMfmaFwdBlocks), nfp, nfm);       \
    if (nfp + nfm > 0)                                                                                             \
      hipLaunchKernelGGL((mfma_fwd_kernel<A, B, C, D>), dim3(nfp + nfm), dim3(kFT), lds_f, st, T, params, Hm, Hp,   \
                         i0, ip, im, u, nfp, act, act_f4);                                                         \
    hipLaunchKernelGGL(mfma_rowsum_kernel, dim3(xcd_grid(nrb)), dim3(kRST), 0, st, rec->row_ptr, rec->row_c,      \
                       bt->row_begin, nrows, rec->ent_w, rec->nnz, u - i0, bt->e_item, r, loss_part);              \
    if (ip + im > 0)                                                                                               \
      hipLaunchKernelGGL(mfma_seed_kernel, dim3(xcd_grid((ip + im + 255) / 256)), dim3(256), 0, st, T, i0,          \
                         ip + im, bt->row_begin, r, u);                                                            \
    if (nbp + nbm > 0 && stored)                                                                                   \
      hipLaunchKernelGGL((mfma_bwd_kernel<A, B, C, D, true>), dim3(nbp + nbm), dim3(kBT), bwd_lds, st, T,           \
                         params, Hm, Hp, i0, ip, im, u, gpart, nbp, stride, act, act_f4);                          \
    if (nbp + nbm > 0 && !stored)                                                                                  \
      hipLaunchKernelGGL((mfma_bwd_kernel<A, B, C, D, false>), dim3(nbp + nbm), dim3(kBT), bwd_lds, st, T,          \
                         params, Hm, Hp, i0, ip, im, u, gpart, nbp, stride, nullptr, act_f4);                      \
  }
  DIPS_MFMA_SHAPES(X)
#undef X
  if (!launched) {
    set_error("dips_mfma_train_step: network shape has no MFMA kernel");
    return DIPS_EUNSUPPORTED;
  }
  int status = hip_status(hipGetLastError(), "mfma kernels launch");
  if (status) return status;
  // the reduction reads the + net partials from workgroups [0, nbp) and the - net from [nbp, nbp+nbm);
  // parameters of the + net come first in the flat vector
  // (with the optimizer) the reduction also hands the step counter over to count_snap and advances it
  hipLaunchKernelGGL(mfma_reduce_kernel, dim3(nred), dim3(256), 0, st, gpart, stride, nbp, nbm, PPp, PPm, inv_b,
                     grad_out ? grad_out : gvec, sq_part, opt ? count : nullptr, count_snap);
  if (!opt) {
    hipLaunchKernelGGL(mfma_loss_out_kernel, dim3(1), dim3(256), 0, st, loss_part, nrb, inv_b, loss_out);
    return hip_status(hipGetLastError(), "mfma reduce launch");
  }
  hipLaunchKernelGGL(mfma_apply_kernel, dim3((unsigned)std::min(nred, kApplyBlocks)), dim3(256), 0, st, gvec, P,
                     nullptr, P, 1.0f, sq_part, nred, loss_part, nrb, inv_b, *opt, params, m, v, count_snap, loss_hist,
                     loss_cap);
  return hip_status(hipGetLastError(), "mfma reduce/apply launch");
}

// item positions (SoA, stride bt->i_count) and entry -> item map of one batch, once per level
int mfma_items_launch(const dips_grid_t* g, const dips_records_t* rec, const dips_mfma_batch_t* bt, float* i_pos,
                      int32_t* e_item, hipStream_t st) {
  const bool grouped = bt->t_group != nullptr;
  const int64_t i0 = grouped ? bt->g_begin : bt->t_begin;
  const int64_t n = grouped ? bt->g_plus + bt->g_minus : bt->n_plus + bt->n_minus;
  if (!i_pos || !e_item || i0 < 0 || n < 0 || bt->i_count < i0 + n || bt->t_count != bt->ent_end - bt->ent_begin ||
      (grouped && (bt->g_plus < 0 || bt->g_minus < 0 || n > bt->t_count))) {
    set_error("dips_mfma_items: invalid batch or outputs (i_count must cover the batch's items)");
    return DIPS_EINVAL;
  }
  if (n == 0) return DIPS_OK;
  MfmaTileArgs T{point_set(g), rec->hx, rec->hy, rec->hz, rec->row_point, bt->t_row, bt->t_code, bt->t_w,
                 bt->t_csr, bt->t_group, i_pos, bt->i_count};
  hipLaunchKernelGGL(mfma_items_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, T, i0, n, e_item, i_pos);
  return hip_status(hipGetLastError(), "mfma_items_kernel launch");
}

// ---- the same step without a caller-given workspace ------------------------------
// The float workspace of mfma_step (u per entry, r per row, per-workgroup
// gradient partials, g) and its float64 partials live in one file-scope device
// array, kNoAllocFloats floats followed by kNoAllocDoubles doubles; each device
// holds its own copy until the process ends.  Every value a step reads from it
// is written earlier in the same step, so nothing is cleared between calls;
// calls that overlap on one device share the array and race.
constexpr int64_t kNoAllocFloats = int64_t(8) << 20;  // 32 MiB
constexpr int64_t kNoAllocDoubles = int64_t(1) << 13;  // 64 KiB
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocFloats + 2 * kNoAllocDoubles];

static std::mutex g_scratch_mu;
static float* g_scratch_dev[kNoAllocMaxDevices] = {};

// device address of the current device's copy of scratch, looked up once per device
static int scratch_of_current_device(float** out) {
  int dev = -1;
  int s = hip_status(hipGetDevice(&dev), "dips_mfma_train_step_noalloc: hipGetDevice");
  if (s) return s;
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error("dips_mfma_train_step_noalloc: device index out of range");
    return DIPS_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    s = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                   "dips_mfma_train_step_noalloc: hipGetSymbolAddress(scratch)");
    if (s) return s;
    g_scratch_dev[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_dev[dev];
  return DIPS_OK;
}

int mfma_step_noalloc(const dips_grid_t* g, const dips_records_t* rec, const dips_mfma_batch_t* bt,
                      const dips_net_t* net, const dips_opt_t* opt, float inv_b, float* params, float* m, float* v,
                      int32_t* count, float* loss_hist, int loss_cap, hipStream_t st) {
  const int64_t nrows = bt->nrows, nent = bt->ent_end - bt->ent_begin;
  const int64_t floats = mfma_base_floats(net, nrows, nent);
  const int64_t P = net_params(net->Lm, net->Hm) + net_params(net->Lp, net->Hp);
  const int64_t doubles = (nrows + kMT - 1) / kMT + (P + kRedParams - 1) / kRedParams + 1;
  if (floats > kNoAllocFloats || doubles > kNoAllocDoubles) {
    set_error("dips_mfma_train_step_noalloc: the batch needs " + std::to_string(floats) + " floats and " +
              std::to_string(doubles) + " doubles of scratch, beyond the fixed " + std::to_string(kNoAllocFloats) +
              " and " + std::to_string(kNoAllocDoubles) + " (use dips_mfma_train_step with a workspace)");
    return DIPS_EINVAL;
  }
  float* base = nullptr;
  const int s = scratch_of_current_device(&base);
  if (s) return s;
  return mfma_step(g, rec, bt, net, opt, inv_b, params, m, v, count, loss_hist, loss_cap, base, kNoAllocFloats,
                   reinterpret_cast<double*>(base + kNoAllocFloats), kNoAllocDoubles, nullptr, nullptr, st);
}

}  // namespace dips
