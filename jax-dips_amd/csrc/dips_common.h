// Shared device helpers for the DIPS HIP library (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DIPS_WAVE 64

// Per-entry code of a record: bits 0..6 = lattice offset o in [-2,2]^3 encoded
// as 25*(oz+2) + 5*(oy+2) + (ox+2); bit 7 = network phase (1 -> Omega+ net).
__device__ __forceinline__ void dips_decode_offset(uint32_t code, int& ox, int& oy, int& oz) {
  int c = (int)(code & 127u);
  oz = c / 25 - 2;
  int rem = c - (oz + 2) * 25;
  oy = rem / 5 - 2;
  ox = rem - (oy + 2) * 5 - 2;
}

__device__ __forceinline__ float dips_wave_sum(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, DIPS_WAVE);
  return v;
}

__device__ __forceinline__ double dips_wave_sum_d(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, DIPS_WAVE);
  return v;
}

// Flat parameter layout of one tanh MLP: for each layer W[fan_in][fan_out]
// row-major, then b[fan_out]; layers 3->H, (H->H)*(L-1), H->1.
template <int L, int H>
struct NetLayout {
  static constexpr int kIn = 3;
  static constexpr int size() { return (kIn * H + H) + (L - 1) * (H * H + H) + (H + 1); }
  static constexpr int w(int layer) {
    return layer == 0 ? 0 : (kIn * H + H) + (layer - 1) * (H * H + H);
  }
  static constexpr int b(int layer) { return w(layer) + (layer == 0 ? kIn * H : H * H); }
  static constexpr int wout() { return (kIn * H + H) + (L - 1) * (H * H + H); }
  static constexpr int bout() { return wout() + H; }
};
