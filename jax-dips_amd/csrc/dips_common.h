// Shared device helpers for the DIPS HIP library (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DIPS_WAVE 64

// Training points: grid nodes (index (i*ny + j)*nz + k) followed by extra
// points stored as float32 triples (index ngrid + q).
struct PointSet {
  const float* xs;
  const float* ys;
  const float* zs;
  int ny, nz;
  int64_t ngrid;
  const float* extra;
  int64_t nextra;
};

// Loads and stores through explicitly global pointers: for pointers read from
// memory the compiler would otherwise emit flat accesses (which also count
// against the scalar-load wait counter).
template <typename T>
__device__ __forceinline__ T gld(const T* p, int64_t i) {
  return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <typename T>
__device__ __forceinline__ void gst(T* p, int64_t i, T v) {
  ((__attribute__((address_space(1))) T*)p)[i] = v;
}

// Coordinates of point pt; false if pt is not a point of the set.
__device__ __forceinline__ bool point_xyz(const PointSet& s, int64_t pt, float& x, float& y, float& z) {
  if (pt < 0) return false;
  if (pt < s.ngrid) {
    const int k = (int)(pt % s.nz);
    const int64_t ij = pt / s.nz;
    const int j = (int)(ij % s.ny);
    const int i = (int)(ij / s.ny);
    x = gld(s.xs, i);
    y = gld(s.ys, j);
    z = gld(s.zs, k);
    return true;
  }
  const int64_t q = pt - s.ngrid;
  if (q >= s.nextra) return false;
  x = gld(s.extra, 3 * q);
  y = gld(s.extra, 3 * q + 1);
  z = gld(s.extra, 3 * q + 2);
  return true;
}

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

template <int Ctrl, int RowMask, bool BoundCtrl>
__device__ __forceinline__ double dips_dpp_d(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, Ctrl, RowMask, 0xF, BoundCtrl);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), Ctrl, RowMask, 0xF, BoundCtrl);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// float64 wavefront sum with the same DPP tree as dips_wave_sum (fixed order,
// no LDS traffic); every lane receives the total.
__device__ __forceinline__ double dips_wave_sum_d(double v) {
  v += dips_dpp_d<0x111, 0xF, true>(v);
  v += dips_dpp_d<0x112, 0xF, true>(v);
  v += dips_dpp_d<0x114, 0xF, true>(v);
  v += dips_dpp_d<0x118, 0xF, true>(v);
  v += dips_dpp_d<0x142, 0xA, false>(v);
  v += dips_dpp_d<0x143, 0xC, false>(v);
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

template <int Ctrl, int RowMask, bool BoundCtrl>
__device__ __forceinline__ float dips_dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), Ctrl, RowMask, 0xF,
                                                               BoundCtrl));
}

// Segmented inclusive scan over the wavefront: lane l ends with the sum of
// lanes head..l, where head <= l is the first lane of l's segment (segments
// are contiguous lane ranges; head == l starts one).  Row-local Hillis-Steele
// steps (row_shr 1, 2, 4, 8) followed by the two row broadcasts (lane 15 of a
// row to the next row, lane 31 to rows 2 and 3), each added only where it lies
// inside the lane's segment: a fixed summation tree without LDS traffic.
__device__ __forceinline__ float dips_segmented_scan(float v, int lane, int head) {
  float t = dips_dpp_f<0x111, 0xF, true>(v);
  if (lane - 1 >= head) v += t;
  t = dips_dpp_f<0x112, 0xF, true>(v);
  if (lane - 2 >= head) v += t;
  t = dips_dpp_f<0x114, 0xF, true>(v);
  if (lane - 4 >= head) v += t;
  t = dips_dpp_f<0x118, 0xF, true>(v);
  if (lane - 8 >= head) v += t;
  t = dips_dpp_f<0x142, 0xA, false>(v);  // rows 1 and 3 receive lanes 15 and 47
  if (head < (lane & ~15)) v += t;
  t = dips_dpp_f<0x143, 0xC, false>(v);  // rows 2 and 3 receive lane 31
  if (head < 32) v += t;
  return v;
}

__device__ __forceinline__ int dips_wave_sum_i(int v) {
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
