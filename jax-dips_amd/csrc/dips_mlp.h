// Per-lane two-phase tanh MLP forward / backward used by the loss and eval kernels.
// The weight pointer type is a template parameter: kernels pass a pointer in
// the constant address space (dips_weights) so that the uniform weight reads
// become scalar loads through the scalar cache, with the weights used as
// scalar operands; activations and gradient accumulators live in registers.
#pragma once
#include "dips_common.h"

namespace dips {

// Weights seen through the constant address space: uniform reads are scalar
// loads.  Valid while the kernel does not write the parameters before reading
// them (the optimizer writes them only after the last read of the step).
typedef const __attribute__((address_space(4))) float dips_weight_t;

__device__ __forceinline__ dips_weight_t* dips_weights(const float* p) { return (dips_weight_t*)p; }

// tanh in float32 without branches: x + x^3 g(x^2) for |x| < 0.55 (degree-4
// minimax polynomial g fitted for this library, <= 0.9 ulp) and
// 1 - 2/(1 + exp(2|x|)) with the sign of x otherwise (<= 2 ulp with the
// hardware exp2 and reciprocal); the result saturates to +-1.
__device__ __forceinline__ float dips_tanh(float x) {
  const float ax = fabsf(x);
  const float t = x * x;
  float g = fmaf(-0.00624905f, t, 0.02107461f);
  g = fmaf(g, t, -0.05386142f);
  g = fmaf(g, t, 0.13332817f);
  g = fmaf(g, t, -0.3333333f);
  const float small = fmaf(x * t, g, x);
  const float e = __builtin_amdgcn_exp2f(ax * 2.8853900817779268f);  // exp(2|x|)
  const float big = fmaf(-2.0f, __builtin_amdgcn_rcpf(1.0f + e), 1.0f);
  return ax < 0.55f ? small : copysignf(big, x);
}

// tanh as 1 - 2/(1 + exp(2x)) with the hardware exp2 and reciprocal: 3 VALU
// instructions and 2 transcendentals (dips_tanh: 14 and 2).  Absolute error
// below 4e-7 for every x with both hardware ops 1 ulp off (relative accuracy
// degrades only where |tanh x| is small); saturates to +-1, propagates NaN.  The
// hidden activation of every network kernel (tanh is their VALU bottleneck).
constexpr float kTanhExp2Scale = 2.8853900817779268f;  // 2 log2(e): exp(2x) = exp2(kTanhExp2Scale x)

// tanh(s / kTanhExp2Scale): the same form for an argument already scaled, as the
// MFMA kernels have it when the scale is folded into the packed weights.
__device__ __forceinline__ float dips_tanh_exp2(float s) {
  return fmaf(-2.0f, __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(s)), 1.0f);
}

__device__ __forceinline__ float dips_tanh_fast(float x) { return dips_tanh_exp2(x * kTanhExp2Scale); }

// Hidden activations: 0 = tanh (DoubleMLP, nn/mlp/MLP.py:64-69), 1 = CELU with
// alpha 1 (the coordinate MLPs of the hash network, nn/hash_encoding/blocks/
// nerfs_flax.py:164: x above 0, expm1(x) below).  act_grad is the derivative
// written in the activation value a: 1 - a^2 for tanh; 1 above 0 and a + 1
// (= exp(x)) below for CELU.
constexpr int kActTanh = 0;
constexpr int kActCelu = 1;

template <int Act>
__device__ __forceinline__ float act_f(float s) {
  if constexpr (Act == kActCelu)
    return s > 0.0f ? s : expm1f(s);
  else
    return dips_tanh_fast(s);
}

template <int Act>
__device__ __forceinline__ float act_grad(float a) {
  if constexpr (Act == kActCelu)
    return a > 0.0f ? 1.0f : a + 1.0f;
  else
    return 1.0f - a * a;
}

// Network shapes with a compiled per-lane kernel.  (Lm, Hm, Lp, Hp)
#define DIPS_SHAPES(X) \
  X(1, 1, 2, 10)       \
  X(1, 3, 2, 10)       \
  X(1, 1, 1, 1)        \
  X(2, 10, 2, 10)      \
  X(2, 16, 2, 16)


template <int L, int H, int Act = kActTanh, typename WP>
__device__ __forceinline__ float mlp_forward(WP P, float x, float y, float z) {
  using NL = NetLayout<L, H>;
  float a[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float s = P[NL::b(0) + j];
    s = fmaf(x, P[NL::w(0) + 0 * H + j], s);
    s = fmaf(y, P[NL::w(0) + 1 * H + j], s);
    s = fmaf(z, P[NL::w(0) + 2 * H + j], s);
    a[j] = act_f<Act>(s);
  }
#pragma unroll
  for (int l = 1; l < L; ++l) {
    float t[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float s = P[NL::b(l) + j];
#pragma unroll
      for (int i = 0; i < H; ++i) s = fmaf(a[i], P[NL::w(l) + i * H + j], s);
      t[j] = act_f<Act>(s);
    }
#pragma unroll
    for (int j = 0; j < H; ++j) a[j] = t[j];
  }
  float o = P[NL::bout()];
#pragma unroll
  for (int i = 0; i < H; ++i) o = fmaf(a[i], P[NL::wout() + i], o);
  return o;
}

// Forward pass that keeps the hidden activations (for a backward pass without
// recomputation); returns the network output.
template <int L, int H, int Act = kActTanh, typename WP>
__device__ __forceinline__ float mlp_forward_keep(WP P, float x, float y, float z, float (&act)[L][H]) {
  using NL = NetLayout<L, H>;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float s = P[NL::b(0) + j];
    s = fmaf(x, P[NL::w(0) + 0 * H + j], s);
    s = fmaf(y, P[NL::w(0) + 1 * H + j], s);
    s = fmaf(z, P[NL::w(0) + 2 * H + j], s);
    act[0][j] = act_f<Act>(s);
  }
#pragma unroll
  for (int l = 1; l < L; ++l) {
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float s = P[NL::b(l) + j];
#pragma unroll
      for (int i = 0; i < H; ++i) s = fmaf(act[l - 1][i], P[NL::w(l) + i * H + j], s);
      act[l][j] = act_f<Act>(s);
    }
  }
  float o = P[NL::bout()];
#pragma unroll
  for (int i = 0; i < H; ++i) o = fmaf(act[L - 1][i], P[NL::wout() + i], o);
  return o;
}

// Adds seed * d NN(x,y,z) / d theta into G (register array of NL::size()),
// given the activations of mlp_forward_keep at the same input.
template <int L, int H, int Act = kActTanh, typename WP>
__device__ __forceinline__ void mlp_backward_kept(WP P, float x, float y, float z, float seed,
                                                  const float (&act)[L][H], float* G) {
  using NL = NetLayout<L, H>;
  float d[H];
  G[NL::bout()] += seed;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    G[NL::wout() + i] = fmaf(seed, act[L - 1][i], G[NL::wout() + i]);
    float a = act[L - 1][i];
    d[i] = seed * P[NL::wout() + i] * act_grad<Act>(a);
  }
#pragma unroll
  for (int l = L - 1; l >= 1; --l) {
    float dn[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
      float acc = 0.0f;
#pragma unroll
      for (int j = 0; j < H; ++j) {
        G[NL::w(l) + i * H + j] = fmaf(act[l - 1][i], d[j], G[NL::w(l) + i * H + j]);
        acc = fmaf(P[NL::w(l) + i * H + j], d[j], acc);
      }
      float a = act[l - 1][i];
      dn[i] = acc * act_grad<Act>(a);
    }
#pragma unroll
    for (int j = 0; j < H; ++j) G[NL::b(l) + j] += d[j];
#pragma unroll
    for (int i = 0; i < H; ++i) d[i] = dn[i];
  }
#pragma unroll
  for (int j = 0; j < H; ++j) {
    G[NL::w(0) + 0 * H + j] = fmaf(x, d[j], G[NL::w(0) + 0 * H + j]);
    G[NL::w(0) + 1 * H + j] = fmaf(y, d[j], G[NL::w(0) + 1 * H + j]);
    G[NL::w(0) + 2 * H + j] = fmaf(z, d[j], G[NL::w(0) + 2 * H + j]);
    G[NL::b(0) + j] += d[j];
  }
}

// Pre-activation gradients of every hidden layer for the output seed:
// dl[L-1] = seed * Wout * (1 - a^2), dl[l-1] = (W_l dl[l]) * (1 - a_{l-1}^2).
template <int L, int H, int Act = kActTanh, typename WP>
__device__ __forceinline__ void mlp_deltas(WP P, float seed, const float (&act)[L][H], float (&dl)[L][H]) {
  using NL = NetLayout<L, H>;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const float a = act[L - 1][i];
    dl[L - 1][i] = seed * P[NL::wout() + i] * act_grad<Act>(a);
  }
#pragma unroll
  for (int l = L - 1; l >= 1; --l) {
#pragma unroll
    for (int i = 0; i < H; ++i) {
      float acc = 0.0f;
#pragma unroll
      for (int j = 0; j < H; ++j) acc = fmaf(P[NL::w(l) + i * H + j], dl[l][j], acc);
      const float a = act[l - 1][i];
      dl[l - 1][i] = acc * act_grad<Act>(a);
    }
  }
}

// Adds seed * d NN(x,y,z) / d theta into G, recomputing the forward activations.
template <int L, int H, int Act = kActTanh, typename WP>
__device__ __forceinline__ void mlp_backward(WP P, float x, float y, float z, float seed, float* G) {
  float act[L][H];
  (void)mlp_forward_keep<L, H, Act>(P, x, y, z, act);
  mlp_backward_kept<L, H, Act>(P, x, y, z, seed, act, G);
}


// Forward value and input gradient d NN / d (x, y, z).
template <int L, int H, int Act = kActTanh, typename WP>
__device__ __forceinline__ float mlp_value_grad(WP P, float x, float y, float z, float g[3]) {
  using NL = NetLayout<L, H>;
  float act[L][H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float s = P[NL::b(0) + j];
    s = fmaf(x, P[NL::w(0) + 0 * H + j], s);
    s = fmaf(y, P[NL::w(0) + 1 * H + j], s);
    s = fmaf(z, P[NL::w(0) + 2 * H + j], s);
    act[0][j] = act_f<Act>(s);
  }
#pragma unroll
  for (int l = 1; l < L; ++l) {
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float s = P[NL::b(l) + j];
#pragma unroll
      for (int i = 0; i < H; ++i) s = fmaf(act[l - 1][i], P[NL::w(l) + i * H + j], s);
      act[l][j] = act_f<Act>(s);
    }
  }
  float o = P[NL::bout()];
  float d[H];
#pragma unroll
  for (int i = 0; i < H; ++i) {
    o = fmaf(act[L - 1][i], P[NL::wout() + i], o);
    const float a = act[L - 1][i];
    d[i] = P[NL::wout() + i] * act_grad<Act>(a);
  }
#pragma unroll
  for (int l = L - 1; l >= 1; --l) {
    float dn[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
      float acc = 0.0f;
#pragma unroll
      for (int j = 0; j < H; ++j) acc = fmaf(P[NL::w(l) + i * H + j], d[j], acc);
      const float a = act[l - 1][i];
      dn[i] = acc * act_grad<Act>(a);
    }
#pragma unroll
    for (int i = 0; i < H; ++i) d[i] = dn[i];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < H; ++j) acc = fmaf(P[NL::w(0) + c * H + j], d[j], acc);
    g[c] = acc;
  }
  return o;
}

}  // namespace dips
