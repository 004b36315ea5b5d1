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

// Sum over the 64 lanes of a wavefront with DPP row shifts and row broadcasts
// (no LDS traffic); every lane receives the total.
__device__ __forceinline__ float dips_wave_sum(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xF, 0xF, true));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x112, 0xF, 0xF, true));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x114, 0xF, 0xF, true));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x118, 0xF, 0xF, true));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x142, 0xA, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x143, 0xC, 0xF, false));
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
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
