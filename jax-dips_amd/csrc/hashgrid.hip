// Multiresolution hash network with the hashgrid encoder: training step and
// evaluation over record entries.
//
// Encoder (nn/hash_encoding/blocks/encoders_flax.py:67-254): pos = (x + bound)
// / (2 bound); per level l the scale s_l = N_min b^l - 1 (host float64, stored
// float32), pos_scaled = pos * s_l + 0.5 (two float32 roundings), corners
// floor(pos_scaled) + {0,1}^3 in the order (0,0,0), (0,0,1), ..., (1,1,1),
// table row = (dense ? x + y res + z res^2 : x ^ y*2654435761 ^ z*805459861)
// mod T + the level's first row (uint32 arithmetic), trilinear weights; the
// encoding is level-major (l * 2 + f).  Network (model.py:53-94,
// blocks/nerfs_flax.py:142-170): per phase W0[E][16], b0, CELU, Wout[16], bout.
//
// Step: forward (one lane per phase-sorted entry: encoding + MLP, both kept
// for the backward) -> row sums (the MFMA path's kernel) -> backward (one
// wavefront per 64 entries: the MLP weight gradient W0 = enc^T delta is a
// K = 64 contraction on the matrix cores through LDS staging, biases and
// output weights per lane; the table gradient d enc = W0 delta goes to the 8
// corners of every level as (row, value) records in buckets of 8192 rows, one
// fixed-capacity slot per workgroup and bucket) -> bucket sums (every table
// row written) -> reduction (fixed order for the
// MLPs; float64 norm partials) -> Adam over every parameter, the table's
// gradient read in place.  Tables above 2^25 rows use atomic adds instead of
// the buckets.
#include <math.h>

#include <algorithm>
#include <cstdlib>

#include "dips_common.h"
#include "dips_hip.h"
#include "dips_mlp.h"

namespace dips {

constexpr int kHL = 16;               // levels (maximum)
constexpr int kHF = 2;                // features per table row
constexpr int kHE = kHL * kHF;        // encoding width (maximum)
constexpr int kHW = 16;               // hidden width
constexpr int kHashWaves = 8;
constexpr int kHashThreads = kHashWaves * DIPS_WAVE;
constexpr int kHashBlocks = 256;      // forward / backward workgroups (persistent), a multiple of kXcds
constexpr int kXcds = 8;              // accelerator dies (L2 domains) of the GPU
constexpr int kHStride = 68;          // staged operand row: 64 entries + 4
constexpr int kHStage = 2 * 16 * kHStride;
// the backward's staging area holds d enc (kHE rows of a wavefront)
static_assert(kHE * DIPS_WAVE <= kHStage, "staging area too small");

constexpr int kSHMax = 8;  // highest order of the spherical-harmonics encoder (SphericalHarmonicsEncoder, L <= 8)

constexpr uint32_t kNoMask = 0xffffffffu;
struct HashLevels {
  int L;    // levels (hashgrid), frequencies (frequency encoding) or highest order (spherical harmonics)
  int enc;  // 0 = hashgrid, 1 = frequency encoding, 2 = spherical harmonics
  uint32_t T;
  uint32_t tmask;  // T - 1 when T is a power of two (idx % T = idx & tmask), else kNoMask: the modulo
  int64_t rows;
  float bound, two_bound;
  float scale[kHL];
  uint32_t res[kHL];     // resolution of a dense level, 0 for a hashed level
  uint32_t offset[kHL];  // first table row of the level (tables of < 2^31 parameters)
  float shc[kSHMax * kSHMax];  // spherical harmonics: (-1)^|m| c_l^|m| of feature l^2 + l + m
};

__host__ __device__ inline int hash_mlp_size(int E) { return E * kHW + kHW + kHW + 1; }

constexpr int kFL = 10;          // frequencies of the frequency encoding (FrequencyEncoder(L=10), model.py:122-123)
constexpr int kFE = 6 * kFL;     // its encoding width
constexpr int kFEPad = 64;       // ... padded to whole 16-row MFMA blocks

static int hash_levels(const dips_hashgrid_t* hg, HashLevels& H) {
  if (hg && hg->encoder == 2) {
    if (hg->levels < 1 || hg->levels > kSHMax) {
      set_error("spherical harmonics encoding: invalid description (levels = highest order, 1..8)");
      return DIPS_EINVAL;
    }
    H = HashLevels{};
    H.L = hg->levels;
    H.enc = 2;
    H.T = 1;
    H.rows = 0;
    H.bound = hg->bound;
    H.two_bound = 2.0f * hg->bound;
    // c_l^k = sqrt((2l+1)/(4 pi) (l-k)!/(l+k)!), times sqrt 2 for k > 0, with the sign (-1)^k
    for (int l = 0; l < H.L; ++l)
      for (int m = -l; m <= l; ++m) {
        const int k = m < 0 ? -m : m;
        double ratio = 1.0;  // (l-k)! / (l+k)!
        for (int j = l - k + 1; j <= l + k; ++j) ratio /= (double)j;
        double c = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * ratio) * (k > 0 ? sqrt(2.0) : 1.0);
        if (k & 1) c = -c;
        H.shc[l * l + l + m] = (float)c;
      }
    return DIPS_OK;
  }
  if (hg && hg->encoder == 1) {
    if (hg->levels != kFL || !(hg->bound > 0.0f)) {
      set_error("frequency encoding: invalid description (levels = 10 frequencies, bound > 0)");
      return DIPS_EINVAL;
    }
    H = HashLevels{};
    H.L = hg->levels;
    H.enc = 1;
    H.T = 1;
    H.rows = 0;
    H.bound = hg->bound;
    H.two_bound = 2.0f * hg->bound;
    return DIPS_OK;
  }
  if (!hg || hg->encoder != 0 || hg->levels < 2 || hg->levels > kHL || hg->features != kHF || hg->table_size < 1 ||
      hg->table_size > (int64_t)1 << 31 || hg->n_min < 1 || hg->n_max < hg->n_min || !(hg->bound > 0.0f)) {
    set_error("hashgrid: invalid encoder description (levels 2..16, features 2, table_size 1..2^31, n_min <= n_max)");
    return DIPS_EINVAL;
  }
  H.L = hg->levels;
  H.enc = 0;
  H.T = (uint32_t)hg->table_size;
  H.tmask = (H.T & (H.T - 1)) == 0 ? H.T - 1 : kNoMask;
  H.bound = hg->bound;
  H.two_bound = 2.0f * hg->bound;
  const double b = exp((log((double)hg->n_max) - log((double)hg->n_min)) / (double)(hg->levels - 1));
  int64_t off = 0;
  for (int l = 0; l < kHL; ++l) {
    H.scale[l] = 0.0f;
    H.res[l] = 0;
    H.offset[l] = 0;
  }
  for (int l = 0; l < H.L; ++l) {
    const double scale = (double)hg->n_min * pow(b, (double)l) - 1.0;
    const int64_t res = (int64_t)ceil(scale) + 1;
    const double n3 = (double)res * (double)res * (double)res;
    const bool dense = n3 <= (double)hg->table_size;
    H.scale[l] = (float)scale;
    H.res[l] = dense ? (uint32_t)res : 0u;
    H.offset[l] = (uint32_t)off;
    off += dense ? res * res * res : hg->table_size;
  }
  H.rows = off;
  return DIPS_OK;
}

// position in the unit cube, pos = (x + bound) / (2 bound), correctly rounded
__device__ __forceinline__ void hash_pos(const HashLevels& H, float x, float y, float z, float (&p)[3]) {
  p[0] = (x + H.bound) / H.two_bound;
  p[1] = (y + H.bound) / H.two_bound;
  p[2] = (z + H.bound) / H.two_bound;
}

// cell of position p on level l: the corner coordinates (lower and upper per
// axis, float -> uint32 saturating at 0) and the fractions modf(pos_scaled)
// with the reference's weight clip folded in: a negative pos_scaled gives the
// lower corner weight 1 and the upper weight 0, i.e. fraction 0; live[d] is
// false there (the weights do not depend on that coordinate).
__device__ __forceinline__ void hash_cell(const HashLevels& H, int l, const float (&p)[3], uint32_t (&lo)[3],
                                          uint32_t (&hi)[3], float (&frac)[3], bool (&live)[3]) {
  // pos * scale + 0.5 rounds twice, as written in the reference: the operators
  // stand in this scope so that contraction into a fused multiply-add is off
#pragma clang fp contract(off)
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float ps = p[d] * H.scale[l] + 0.5f;
    const float fl = floorf(ps);
    const float fh = fl + 1.0f;
    live[d] = ps >= 0.0f;
    frac[d] = live[d] ? ps - fl : 0.0f;
    lo[d] = fl > 0.0f ? (uint32_t)fl : 0u;
    hi[d] = fh > 0.0f ? (uint32_t)fh : 0u;
  }
}

__device__ __forceinline__ int64_t hash_row(const HashLevels& H, int l, const uint32_t (&lo)[3],
                                            const uint32_t (&hi)[3], int c) {
  const uint32_t vx = ((c >> 2) & 1) ? hi[0] : lo[0];
  const uint32_t vy = ((c >> 1) & 1) ? hi[1] : lo[1];
  const uint32_t vz = (c & 1) ? hi[2] : lo[2];
  uint32_t idx;
  const uint32_t r = H.res[l];
  if (r) {
    idx = vx + vy * r + vz * (r * r);
  } else {
    idx = vx ^ (vy * 2654435761u) ^ (vz * 805459861u);
  }
  // the reference's tables are powers of two (lpbe.yaml:54, 2^14 to 2^24): a mask, not a division
  if (H.tmask != kNoMask)
    idx &= H.tmask;
  else
    idx %= H.T;
  const int64_t row = H.offset[l] + (int64_t)idx;
  return row < H.rows ? row : H.rows - 1;
}

__device__ __forceinline__ float hash_weight(const float (&frac)[3], int c) {
  const float wx = ((c >> 2) & 1) ? frac[0] : 1.0f - frac[0];
  const float wy = ((c >> 1) & 1) ? frac[1] : 1.0f - frac[1];
  const float wz = (c & 1) ? frac[2] : 1.0f - frac[2];
  return wx * wy * wz;
}

__device__ __forceinline__ void hash_encode(const HashLevels& H, const float* __restrict__ table, const float (&p)[3],
                                            float (&enc)[kHE]) {
#pragma unroll
  for (int l = 0; l < kHL; ++l) {
    enc[2 * l] = 0.0f;
    enc[2 * l + 1] = 0.0f;
    if (l < H.L) {
      uint32_t lo[3], hi[3];
      float frac[3];
      bool live[3];
      hash_cell(H, l, p, lo, hi, frac, live);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int64_t row = hash_row(H, l, lo, hi, c);
        const float w = hash_weight(frac, c);
        const float2 t = reinterpret_cast<const float2*>(table)[row];
        enc[2 * l] = fmaf(w, t.x, enc[2 * l]);
        enc[2 * l + 1] = fmaf(w, t.y, enc[2 * l + 1]);
      }
    }
  }
}

// MLP of one phase: act (CELU hidden), returns u; W = flat MLP (W0[E][16], b0, Wout, bout)
template <int N, typename WP>
__device__ __forceinline__ float hash_mlp(WP W, int E, const float (&enc)[N], float (&act)[kHW]) {
#pragma unroll
  for (int j = 0; j < kHW; ++j) {
    float s = W[E * kHW + j];
#pragma unroll
    for (int k = 0; k < N; ++k)
      if (k < E) s = fmaf(enc[k], W[k * kHW + j], s);
    act[j] = act_f<kActCelu>(s);
  }
  float u = W[E * kHW + 2 * kHW];
#pragma unroll
  for (int j = 0; j < kHW; ++j) u = fmaf(act[j], W[E * kHW + kHW + j], u);
  return u;
}

// hidden deltas dl_j = seed Wout_j celu'(a_j) and the encoding gradient W0 dl
template <typename WP>
__device__ __forceinline__ void hash_delta(WP W, int E, float seed, const float (&act)[kHW], float (&dl)[kHW],
                                           float (&denc)[kHE]) {
#pragma unroll
  for (int j = 0; j < kHW; ++j) dl[j] = seed * W[E * kHW + kHW + j] * act_grad<kActCelu>(act[j]);
#pragma unroll
  for (int k = 0; k < kHE; ++k) {
    float s = 0.0f;
    if (k < E) {
#pragma unroll
      for (int j = 0; j < kHW; ++j) s = fmaf(W[k * kHW + j], dl[j], s);
    }
    denc[k] = s;
  }
}

struct HashArgs {
  PointSet ps;
  float hx, hy, hz;
  const int32_t* row_point;
  const int32_t* t_row;
  const uint32_t* t_code;
  const float* t_w;
  const int32_t* t_csr;
  int64_t t_begin, t_count, n_plus;
};

__device__ __forceinline__ bool hash_entry(const HashArgs& A, int64_t t, float& x, float& y, float& z, int& row) {
  x = y = z = 0.0f;
  row = -1;
  if (t < 0 || t >= A.t_count) return false;
  const int64_t e = A.t_begin + t;
  row = A.t_row[e];
  float px, py, pz;
  if (!point_xyz(A.ps, A.row_point[row], px, py, pz)) return false;
  int ox, oy, oz;
  dips_decode_offset(A.t_code[e], ox, oy, oz);
  // fl32(p + o h): o h is exact for o in [-2, 2], so the fused form rounds once, as the
  // separate product and sum do; written fused so that every kernel computes it alike
  x = fmaf((float)ox, A.hx, px);
  y = fmaf((float)oy, A.hy, py);
  z = fmaf((float)oz, A.hz, pz);
  return true;
}

// ---- backward -----------------------------------------------------------------
typedef float hf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void hash_stage_mfma(const float* AT, const float* BT, int lane, hf4& acc) {
  const int m = lane & 15, kq = lane >> 4;
  const hf4* ap = reinterpret_cast<const hf4*>(AT + m * kHStride + 16 * kq);
  const hf4* bp = reinterpret_cast<const hf4*>(BT + m * kHStride + 16 * kq);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const hf4 av = ap[q], bv = bp[q];
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bv.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bv.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bv.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bv.w, acc, 0, 0, 0);
  }
}

// gradient accumulators of one phase MLP within a wavefront (encodings of up
// to 16 NB features)
template <int NB>
struct HashGrad {
  hf4 acc[NB];      // W0 rows 16 b .. 16 b + 15 (x the 16 hidden units)
  float gb0[kHW];   // per-lane partial sums of delta
  float gwo[kHW];   // per-lane partial sums of seed * act
  float gbo;

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = hf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < kHW; ++j) gb0[j] = gwo[j] = 0.0f;
    gbo = 0.0f;
  }

  // lanes that are not of this phase carry seed 0 and zero deltas; wave-uniform call
  __device__ __forceinline__ void add(float* AT, float* BT, int lane, int E, float seed,
                                      const float (&enc)[16 * NB], const float (&act)[kHW],
                                      const float (&dl)[kHW]) {
#pragma unroll
    for (int j = 0; j < kHW; ++j) BT[j * kHStride + lane] = dl[j];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (E > 16 * b) {
#pragma unroll
        for (int i = 0; i < 16; ++i) AT[i * kHStride + lane] = enc[16 * b + i];
        __builtin_amdgcn_wave_barrier();
        hash_stage_mfma(AT, BT, lane, acc[b]);
        __builtin_amdgcn_wave_barrier();
      }
    }
#pragma unroll
    for (int j = 0; j < kHW; ++j) {
      gb0[j] += dl[j];
      gwo[j] = fmaf(seed, act[j], gwo[j]);
    }
    gbo += seed;
  }

  // this wavefront's sums into red[0 .. PM)
  __device__ __forceinline__ void flush(float* red, int lane, int E) {
#pragma unroll
    for (int blk = 0; blk < NB; ++blk)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * blk + 4 * (lane >> 4) + r, j = lane & 15;
        if (i < E) red[i * kHW + j] = acc[blk][r];
      }
#pragma unroll
    for (int j = 0; j < kHW; ++j) {
      const float a = dips_wave_sum(gb0[j]);
      const float b = dips_wave_sum(gwo[j]);
      if (lane == 0) {
        red[E * kHW + j] = a;
        red[E * kHW + kHW + j] = b;
      }
    }
    const float c = dips_wave_sum(gbo);
    if (lane == 0) red[E * kHW + 2 * kHW] = c;
  }
};

// Table-gradient sinks of hash_scatter.  Atomic: adds into the table gradient.
// Buckets: appends (row, v0, v1) to the workgroup's slot of the row's bucket at
// the workgroup's LDS cursor.
//
// Buckets are block-cyclic over groups of 64 consecutive rows: group G = row >>
// 6 belongs to bucket G mod nb (nb a power of two), at local rows ((G / nb) <<
// 6) + (row & 63) < kBRows.  The hot rows of the small dense levels thus spread
// over many buckets, and a bucket still writes whole 512-byte row groups.
constexpr int kGShift = 6;
constexpr int kBRows = 8192;       // local rows of a bucket (LDS accumulator)
constexpr int kMaxBuckets = 4096;  // tables of up to 2^25 rows use the buckets

struct BucketMap {
  int shift;  // log2(nb)
  __device__ __forceinline__ int bucket(int32_t row) const { return (row >> kGShift) & ((1 << shift) - 1); }
  __device__ __forceinline__ int local(uint32_t row) const {
    return (int)(((row >> kGShift) >> shift) << kGShift) | (int)(row & ((1u << kGShift) - 1));
  }
};

struct AtomicSink {
  float* gtab;
  __device__ __forceinline__ void put(int32_t row, float v0, float v1) const {
    unsafeAtomicAdd(gtab + 2 * (int64_t)row, v0);
    unsafeAtomicAdd(gtab + 2 * (int64_t)row + 1, v1);
  }
};

// Records of one workgroup: bucket b's slot holds records [b * cap, (b + 1) * cap)
// of `slots`; records past a slot's capacity go to gtab by atomic adds (gtab is
// zeroed before the backward and the bucket sums add to it).
struct BucketSink {
  int* cursor;   // LDS, per bucket: records this workgroup has sent to it
  uint4* slots;  // records (row, v0 bits, v1 bits, 0): one 16-byte store each
  int cap;       // records a slot holds
  BucketMap map;
  float* gtab;
  __device__ __forceinline__ void put(int32_t row, float v0, float v1) const {
    const int b = map.bucket(row);
    const int pos = atomicAdd(cursor + b, 1);
    if (pos < cap) {
      slots[(int64_t)b * cap + pos] = make_uint4((uint32_t)row, __float_as_uint(v0), __float_as_uint(v1), 0u);
    } else {
      unsafeAtomicAdd(gtab + 2 * (int64_t)row, v0);
      unsafeAtomicAdd(gtab + 2 * (int64_t)row + 1, v1);
    }
  }
};


// Sends (v0, v1) of every lane in live for table row `row` to the sink.  Lanes
// that share a row are summed over the wavefront first: up to kAggIters rows of
// at least kAggMin lanes each, the other lanes sending their own values.
// Wave-uniform call.
constexpr int kAggIters = 8;
constexpr int kAggMin = 4;
template <typename Sink>
__device__ __forceinline__ void hash_scatter(const Sink& sink, int32_t row, float v0, float v1, uint64_t live,
                                             int lane) {
  uint64_t rem = live;
#pragma unroll 1
  for (int it = 0; it < kAggIters && rem != 0; ++it) {
    const int leader = __builtin_ctzll(rem);
    const int32_t key = __builtin_amdgcn_readlane(row, leader);
    const uint64_t same = rem & __ballot(row == key);
    if (__builtin_popcountll(same) < kAggMin) break;
    const bool in = (same >> lane) & 1;
    const float s0 = dips_wave_sum(in ? v0 : 0.0f);
    const float s1 = dips_wave_sum(in ? v1 : 0.0f);
    if (lane == leader) sink.put(key, s0, s1);
    rem &= ~same;
  }
  if ((rem >> lane) & 1) sink.put(row, v0, v1);
}

// The 8 corner records of every lane in live on a dense (coarse) level, where a
// wavefront's entries fall into a few cells.  The entries are in table order
// (x fastest), so the lanes of one cell form runs of adjacent lanes: a run is a
// maximal range of live lanes whose corner-0 row (which fixes the cell on a
// dense level) equals the previous lane's.  One segmented scan per value sums
// every run at once (fixed DPP tree), and the last lane of each run sends the
// run's 8 records, all runs in parallel.  A cell met in two runs (the wavefront
// wrapping to the next row of cells) sends two sets of records, summed later
// like any others.  Wave-uniform call.
template <typename Sink>
__device__ __forceinline__ void hash_scatter_cell(const Sink& sink, const int32_t (&rows)[8], const float (&v0)[8],
                                                  const float (&v1)[8], uint64_t live, int lane) {
  const bool on = (live >> lane) & 1;
  const int32_t prev_key = __shfl(rows[0], lane > 0 ? lane - 1 : 0, DIPS_WAVE);
  const bool prev_on = lane > 0 && ((live >> (lane - 1)) & 1);
  const uint64_t starts = __ballot(on && !(prev_on && prev_key == rows[0]));
  // the run of this lane begins at the last start at or below it (lanes outside live carry zeros and
  // close the run before them; they send nothing)
  const uint64_t below = starts & (~0ull >> (63 - lane));
  const int head = below ? 63 - __builtin_clzll(below) : lane;
  const bool last = on && (lane == 63 || !((live >> (lane + 1)) & 1) || ((starts >> (lane + 1)) & 1));
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float s0 = dips_segmented_scan(on ? v0[c] : 0.0f, lane, head);
    const float s1 = dips_segmented_scan(on ? v1[c] : 0.0f, lane, head);
    if (last) sink.put(rows[c], s0, s1);
  }
}

// Lanes of live at bit-identical positions (a stencil neighbour of one point
// that rounds onto the next point; the host orders each batch's entries by
// position, so such entries are adjacent) share every corner row and weight on
// every level, so their table-gradient values can be summed before the scatter.
// A live lane whose position equals the previous lane's (also live) folds into
// the first lane of that run, rep; rep is the lane itself for every other lane.
// Returns the mask of folding lanes.  Wave-uniform call.
__device__ __forceinline__ uint64_t hash_dups(float x, float y, float z, uint64_t live, int lane, int& rep) {
  const int prev = lane > 0 ? lane - 1 : 0;
  const uint32_t px = __float_as_uint(__shfl(x, prev)), py = __float_as_uint(__shfl(y, prev)),
                 pz = __float_as_uint(__shfl(z, prev));
  const bool fold = lane > 0 && ((live >> lane) & 1) && ((live >> prev) & 1) && px == __float_as_uint(x) &&
                    py == __float_as_uint(y) && pz == __float_as_uint(z);
  const uint64_t folded = __ballot(fold);
  // the last run start at or below this lane (lane 0 always starts one)
  const uint64_t upto = ~folded & (lane == DIPS_WAVE - 1 ? ~0ull : ((1ull << (lane + 1)) - 1));
  rep = DIPS_WAVE - 1 - __builtin_clzll(upto);
  return folded;
}

// The 64-entry chunks a wavefront of the persistent forward / backward grid
// takes: each XCD (workgroups blockIdx % 8, dispatched round-robin) takes a
// contiguous eighth of the entries, so that the table rows of a region stay in
// one L2; both kernels use the same assignment (the backward reads what the
// forward stored for the same entries from the same XCD's L2).
// kSplit: workgroups of kSplit * kHashWaves wavefronts cover the same chunks
// (wavefront w takes every kSplit-th round of slot w % kHashWaves), so that a
// kernel with more wavefronts per workgroup keeps the per-workgroup chunk sets.
template <int kSplit = 1>
struct ChunkRange {
  int64_t first, end, step;
  __device__ __forceinline__ ChunkRange(int64_t t_count, int wave) {
    const int64_t nchunks = (t_count + 63) / 64;
    const int xcd = blockIdx.x % kXcds, xblk = blockIdx.x / kXcds;
    end = nchunks * (xcd + 1) / kXcds;
    const int64_t round = (int64_t)(gridDim.x / kXcds) * kHashWaves;
    step = round * kSplit;
    first = nchunks * xcd / kXcds + (int64_t)xblk * kHashWaves + wave % kHashWaves + (wave / kHashWaves) * round;
  }
};
constexpr int kFwdSplit = 2;  // the forward runs 16 wavefronts per workgroup (its gathers need them)
constexpr int kFwdThreads = kFwdSplit * kHashThreads;

// ---- forward: u per entry --------------------------------------------------------
// Persistent, on the chunks of ChunkRange.
__global__ __launch_bounds__(kFwdThreads) void hash_fwd_kernel(HashArgs A, HashLevels H,
                                                                const float* __restrict__ params,
                                                                float* __restrict__ u_out,
                                                                float* __restrict__ encbuf) {
  // LDS: each wavefront's encodings (level loop not unrolled: an unrolled one
  // overflows the instruction cache)
  extern __shared__ float hash_lenc[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / DIPS_WAVE);
  float* const lenc = hash_lenc + wave * (kHE * DIPS_WAVE);
  const int E = 2 * H.L, PM = hash_mlp_size(E);
  const float* table = params + 2 * (int64_t)PM;
  const ChunkRange<kFwdSplit> cr(A.t_count, wave);
  for (int64_t ch = cr.first; ch < cr.end; ch += cr.step) {
    const int64_t t = ch * 64 + lane;
    float x, y, z;
    int row;
    const bool valid = hash_entry(A, t, x, y, z, row);
    float p[3];
    hash_pos(H, x, y, z, p);
#pragma unroll 1
    for (int l = 0; l < H.L; ++l) {
      uint32_t lo[3], hi[3];
      float frac[3];
      bool live[3];
      hash_cell(H, l, p, lo, hi, frac, live);
      float e0 = 0.0f, e1 = 0.0f;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float w = hash_weight(frac, c);
        const float2 tv = reinterpret_cast<const float2*>(table)[hash_row(H, l, lo, hi, c)];
        e0 = fmaf(w, tv.x, e0);
        e1 = fmaf(w, tv.y, e1);
      }
      lenc[(2 * l) * DIPS_WAVE + lane] = e0;
      lenc[(2 * l + 1) * DIPS_WAVE + lane] = e1;
      if (valid) {
        // the encoding, feature-major, for the backward
        encbuf[(int64_t)(2 * l) * A.t_count + t] = e0;
        encbuf[(int64_t)(2 * l + 1) * A.t_count + t] = e1;
      }
    }
    __builtin_amdgcn_wave_barrier();
    float enc[kHE];
#pragma unroll
    for (int k = 0; k < kHE; ++k) enc[k] = k < E ? lenc[k * DIPS_WAVE + lane] : 0.0f;
    __builtin_amdgcn_wave_barrier();
    const bool plus = t < A.n_plus;
    float act[kHW];
    const float* Praw = params;
    asm volatile("" : "+s"(Praw));
    dips_weight_t* Wp = dips_weights(Praw);
    dips_weight_t* Wm = Wp + PM;
    float up = 0.0f, um = 0.0f;
    if (plus) up = hash_mlp(Wp, E, enc, act);
    asm volatile("" : "+s"(Wm));
    if (!plus) um = hash_mlp(Wm, E, enc, act);
    if (valid) {
      u_out[A.t_csr[A.t_begin + t]] = plus ? up : um;
      // the hidden activations of the entry's network, after the encoding, for the backward
#pragma unroll
      for (int j = 0; j < kHW; ++j) encbuf[(int64_t)(E + j) * A.t_count + t] = act[j];
    }
  }
}

// Table gradient rows of one bucket (one workgroup per bucket): the bucket's
// records summed in LDS, then every row of the bucket written (rows no record
// touches get 0).  The LDS sums are float64: on gfx950 an LDS float64 atomic
// add runs ~9x the rate of a float32 one (measured: 1.8 T against 0.2 T lane
// adds/s chip-wide, conflict-free float32 adds no faster), and it rounds less.
constexpr int kBThreads = 1024;
constexpr size_t kBucketLds = sizeof(double) * 2 * kBRows;  // 128 KB: one workgroup per CU
__global__ __launch_bounds__(kBThreads) void hash_bucket_kernel(const uint4* __restrict__ recs,
                                                                const int* __restrict__ cnt, int cap, int nb,
                                                                BucketMap map, int64_t rows,
                                                                float* __restrict__ gtab, float inv_b,
                                                                float* __restrict__ gout,
                                                                double* __restrict__ sq_part) {
  extern __shared__ double bucket_acc[];  // 2 * kBRows
  double* const acc = bucket_acc;
  __shared__ double wsum[kBThreads / DIPS_WAVE];
  const int b = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x / DIPS_WAVE;
  for (int i = threadIdx.x; i < 2 * kBRows; i += kBThreads) acc[i] = 0.0;
  __syncthreads();
  // the slots of bucket b, one per backward workgroup, a wavefront each in turn
  for (int blk = wave; blk < kHashBlocks; blk += kBThreads / DIPS_WAVE) {
    const int n = cnt[(int64_t)b * kHashBlocks + blk];
    const uint4* seg = recs + ((int64_t)blk * nb + b) * cap;
    for (int i = lane; i < n; i += DIPS_WAVE) {
      const uint4 rc = seg[i];
      const int k = map.local(rc.x);
      atomicAdd(acc + 2 * k, (double)__uint_as_float(rc.y));
      atomicAdd(acc + 2 * k + 1, (double)__uint_as_float(rc.z));
    }
  }
  __syncthreads();
  // local row k of group lg is table row ((lg * nb + b) << 6) + (k & 63); gtab holds
  // the records past a slot's capacity.  Also the bucket's float64 sum of squares
  // of the mean gradient inv_b * gtab (fixed order) for the clipping norm, and,
  // when gout is given, that gradient.
  double v = 0.0;
  for (int i = threadIdx.x; i < 2 * kBRows; i += kBThreads) {
    const int k = i >> 1;
    const int64_t g = ((int64_t)(k >> kGShift) << map.shift) + b;
    const int64_t row = (g << kGShift) + (k & ((1 << kGShift) - 1));
    if (row < rows) {
      const int64_t idx = 2 * row + (i & 1);
      const float tot = (float)((double)gtab[idx] + acc[i]);
      gtab[idx] = tot;
      const float gg = tot * inv_b;
      if (gout) gout[idx] = gg;
      v += (double)gg * (double)gg;
    }
  }
  v = dips_wave_sum_d(v);
  if (lane == 0) wsum[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBThreads / DIPS_WAVE; ++w) t += wsum[w];
    sq_part[b] = t;
  }
}

// kBuckets: the table gradient goes to bucket records (the workgroup's slot of
// each bucket; the records per slot to cnt, bucket-major), else by atomic adds
// into gtab
template <bool kBuckets>
__global__ __launch_bounds__(kHashThreads) void hash_bwd_kernel(HashArgs A, HashLevels H,
                                                                const float* __restrict__ params,
                                                                const float* __restrict__ encbuf, int64_t row_begin,
                                                                const float* __restrict__ r_row,
                                                                float* __restrict__ gtab,
                                                                float* __restrict__ partial, int nb,
                                                                int* __restrict__ cnt, uint4* __restrict__ recs,
                                                                int cap, BucketMap map) {
  extern __shared__ float hash_lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / DIPS_WAVE);
  int* const cursor = reinterpret_cast<int*>(hash_lds + kHashWaves * kHStage);
  if constexpr (kBuckets) {
    for (int b = threadIdx.x; b < nb; b += kHashThreads) cursor[b] = 0;
    __syncthreads();
  }
  float* const AT = hash_lds + wave * kHStage;
  float* const BT = AT + 16 * kHStride;
  const int E = 2 * H.L, PM = hash_mlp_size(E);
  HashGrad<2> gp, gm;
  gp.zero();
  gm.zero();
  const ChunkRange<> cr(A.t_count, wave);
  for (int64_t ch = cr.first; ch < cr.end; ch += cr.step) {
    const int64_t t = ch * 64 + lane;
    float x, y, z;
    int row;
    const bool valid = hash_entry(A, t, x, y, z, row);
    const float seed = valid ? r_row[row - row_begin] * A.t_w[A.t_begin + t] : 0.0f;
    const bool plus = t < A.n_plus;
    // the forward's encoding of this entry (0 for lanes past the batch)
    float enc[kHE];
#pragma unroll
    for (int k = 0; k < kHE; ++k) enc[k] = (valid && k < E) ? encbuf[(int64_t)k * A.t_count + t] : 0.0f;
    float p[3];
    hash_pos(H, x, y, z, p);
    const float* Praw = params;
    asm volatile("" : "+s"(Praw));
    dips_weight_t* Wp = dips_weights(Praw);
    dips_weight_t* Wm = Wp + PM;
    // the forward's hidden activations of the entry's own network
    float act[kHW];
#pragma unroll
    for (int j = 0; j < kHW; ++j) act[j] = valid ? encbuf[(int64_t)(E + j) * A.t_count + t] : 0.0f;
    float dl[kHW];
    float denc[kHE];
    if (plus) hash_delta(Wp, E, seed, act, dl, denc);
    asm volatile("" : "+s"(Wm));
    if (!plus) hash_delta(Wm, E, seed, act, dl, denc);
    // MLP weight gradients, one network at a time (the other's lanes contribute zeros)
    if (__any(plus && valid)) {
      float d2[kHW];
      const float s2 = plus ? seed : 0.0f;
#pragma unroll
      for (int j = 0; j < kHW; ++j) d2[j] = plus ? dl[j] : 0.0f;
      gp.add(AT, BT, lane, E, s2, enc, act, d2);
    }
    if (__any(!plus && valid)) {
      float d2[kHW];
      const float s2 = plus ? 0.0f : seed;
#pragma unroll
      for (int j = 0; j < kHW; ++j) d2[j] = plus ? 0.0f : dl[j];
      gm.add(AT, BT, lane, E, s2, enc, act, d2);
    }
    // table gradient: scatter d enc to the corners (wave-uniform from here on;
    // the bucket records of every valid entry were counted, zero seeds included)
    uint64_t live_lanes = __ballot(kBuckets ? valid : (valid && seed != 0.0f));
    if (live_lanes) {
      // d enc through the wavefront's (now free) staging area, so that the level
      // loop need not be unrolled (an unrolled loop overflows the instruction cache)
      __builtin_amdgcn_wave_barrier();
      int rep;
      const uint64_t folded = hash_dups(x, y, z, live_lanes, lane, rep);
#pragma unroll
      for (int k = 0; k < kHE; ++k) AT[k * DIPS_WAVE + lane] = denc[k];
      __builtin_amdgcn_wave_barrier();
      if (folded) {
        // lanes at a position another lane holds add their d enc to that lane's
        if ((folded >> lane) & 1) {
#pragma unroll
          for (int k = 0; k < kHE; ++k)
            if (k < E) atomicAdd(AT + k * DIPS_WAVE + rep, denc[k]);
        }
        __builtin_amdgcn_wave_barrier();
        live_lanes &= ~folded;
      }
#pragma unroll 1
      for (int l = 0; l < H.L; ++l) {
        {
          uint32_t lo[3], hi[3];
          float frac[3];
          bool live[3];
          hash_cell(H, l, p, lo, hi, frac, live);
          const float d0 = AT[(2 * l) * DIPS_WAVE + lane], d1 = AT[(2 * l + 1) * DIPS_WAVE + lane];
          int32_t rows[8];
          float v0[8], v1[8];
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            rows[c] = (int32_t)hash_row(H, l, lo, hi, c);
            const float w = hash_weight(frac, c);
            v0[c] = w * d0;
            v1[c] = w * d1;
          }
          const bool dense = H.res[l] != 0;
          if constexpr (kBuckets) {
            const BucketSink sink{cursor, recs + (int64_t)blockIdx.x * nb * cap, cap, map, gtab};
            if (dense) {
              hash_scatter_cell(sink, rows, v0, v1, live_lanes, lane);
            } else {
#pragma unroll
              for (int c = 0; c < 8; ++c) hash_scatter(sink, rows[c], v0[c], v1[c], live_lanes, lane);
            }
          } else {
            const AtomicSink sink{gtab};
            if (dense) {
              hash_scatter_cell(sink, rows, v0, v1, live_lanes, lane);
            } else {
#pragma unroll
              for (int c = 0; c < 8; ++c) hash_scatter(sink, rows[c], v0[c], v1[c], live_lanes, lane);
            }
          }
        }
      }
    }
  }
  // workgroup reduction: per-wavefront rows in LDS, then the wavefronts in order
  const int stride = 2 * PM;
  __syncthreads();
  if constexpr (kBuckets) {
    // records in this workgroup's slot of each bucket, bucket-major for the bucket sums
    for (int b = threadIdx.x; b < nb; b += kHashThreads)
      cnt[(int64_t)b * kHashBlocks + blockIdx.x] = min(cursor[b], cap);
  }
  float* red = hash_lds + wave * stride;
  gp.flush(red, lane, E);
  gm.flush(red + PM, lane, E);
  __syncthreads();
  for (int p = threadIdx.x; p < stride; p += kHashThreads) {
    float s = 0.0f;
    for (int w = 0; w < kHashWaves; ++w) s += hash_lds[w * stride + p];
    partial[(int64_t)blockIdx.x * stride + p] = s;
  }
}

// Gradient of the MLPs, g[p] = inv_b * (sum of the backward's partial rows),
// one wavefront per parameter: lanes take workgroups lane, lane + 64, ... in
// order, then the DPP tree (float64, fixed order).
__global__ __launch_bounds__(256) void hash_mlp_reduce_kernel(const float* __restrict__ partial, int nblk, int stride,
                                                              float inv_b, float* __restrict__ g) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * 4 + threadIdx.x / 64;
  if (p >= stride) return;  // whole wavefronts
  double acc = 0.0;
  for (int b = lane; b < nblk; b += 64) acc += (double)partial[(int64_t)b * stride + p];
  acc = dips_wave_sum_d(acc);
  if (lane == 0) g[p] = (float)(acc * (double)inv_b);
}

// Table gradient inv_b * gtab, written to g only when write_table (the
// loss/gradient entry point; the optimizer reads gtab in place), and float64
// sums of squares of the whole gradient per workgroup (fixed grid, grid-stride
// order) for the clipping norm; g[0, stride) holds the MLPs' already.
constexpr int kRedBlocks = 1024;
__global__ __launch_bounds__(256) void hash_reduce_kernel(int stride, const float* __restrict__ gtab, int64_t P,
                                                          float inv_b, int write_table, float* __restrict__ g,
                                                          double* __restrict__ sq_part) {
  double v = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < P; p += (int64_t)gridDim.x * 256) {
    float gg;
    if (p < stride) {
      gg = g[p];
    } else {
      gg = gtab[p - stride] * inv_b;
      if (write_table) g[p] = gg;
    }
    v += (double)gg * (double)gg;
  }
  v = dips_wave_sum_d(v);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) sq_part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// p[0, n) = 0: 16-byte stores from the first 16-byte boundary, single floats for
// the (at most 3) floats before it and after the last whole 16 bytes.  Grid:
// hash_zero_blocks(n) workgroups of 256.
__global__ __launch_bounds__(256) void hash_zero_kernel(float* __restrict__ p, int64_t n) {
  const int64_t head = std::min<int64_t>(n, (int64_t)((16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15) / 4);
  const int64_t n4 = (n - head) / 4, tail = head + 4 * n4;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n4) reinterpret_cast<float4*>(p + head)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (i < head) p[i] = 0.0f;
  if (i < n - tail) p[tail + i] = 0.0f;
}
static unsigned hash_zero_blocks(int64_t n) { return (unsigned)((n / 4 + 4 + 255) / 256); }

// ---- evaluation: u and d u / d x ----------------------------------------------
__global__ __launch_bounds__(256) void hash_eval_kernel(HashLevels H, const float* __restrict__ params,
                                                        const float* __restrict__ pts,
                                                        const uint8_t* __restrict__ phase, int64_t m,
                                                        float* __restrict__ u, float* __restrict__ grad) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= m) return;
  const float x = pts[3 * t], y = pts[3 * t + 1], z = pts[3 * t + 2];
  const int E = 2 * H.L, PM = hash_mlp_size(E);
  const float* table = params + 2 * (int64_t)PM;
  float enc[kHE], act[kHW];
  float p[3];
  hash_pos(H, x, y, z, p);
  hash_encode(H, table, p, enc);
  const float* W = params + (phase[t] ? 0 : PM);
  const float v = hash_mlp(W, E, enc, act);
  u[t] = v;
  if (!grad) return;
  // d u / d enc_k = sum_j W0[k][j] Wout_j celu'(a_j)
  float dj[kHW];
#pragma unroll
  for (int j = 0; j < kHW; ++j) dj[j] = W[E * kHW + kHW + j] * act_grad<kActCelu>(act[j]);
  float g[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int l = 0; l < kHL; ++l) {
    if (l < H.L) {
      float du0 = 0.0f, du1 = 0.0f;
#pragma unroll
      for (int j = 0; j < kHW; ++j) {
        du0 = fmaf(W[(2 * l) * kHW + j], dj[j], du0);
        du1 = fmaf(W[(2 * l + 1) * kHW + j], dj[j], du1);
      }
      uint32_t lo[3], hi[3];
      float frac[3];
      bool live[3];
      hash_cell(H, l, p, lo, hi, frac, live);
      const float dps = H.scale[l] / H.two_bound;  // d pos_scaled / d x
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int64_t row = hash_row(H, l, lo, hi, c);
        const float2 tv = reinterpret_cast<const float2*>(table)[row];
        const float val = du0 * tv.x + du1 * tv.y;
        const float wx = ((c >> 2) & 1) ? frac[0] : 1.0f - frac[0];
        const float wy = ((c >> 1) & 1) ? frac[1] : 1.0f - frac[1];
        const float wz = (c & 1) ? frac[2] : 1.0f - frac[2];
        // d weight / d pos_scaled per axis: +-1 where the fraction follows the position
        const float sx = live[0] ? (((c >> 2) & 1) ? 1.0f : -1.0f) : 0.0f;
        const float sy = live[1] ? (((c >> 1) & 1) ? 1.0f : -1.0f) : 0.0f;
        const float sz = live[2] ? ((c & 1) ? 1.0f : -1.0f) : 0.0f;
        g[0] = fmaf(val * sx * wy * wz, dps, g[0]);
        g[1] = fmaf(val * wx * sy * wz, dps, g[1]);
        g[2] = fmaf(val * wx * wy * sz, dps, g[2]);
      }
    }
  }
  grad[3 * t] = g[0];
  grad[3 * t + 1] = g[1];
  grad[3 * t + 2] = g[2];
}

// ---- frequency encoding ---------------------------------------------------------
// FrequencyEncoder (encoders_flax.py:269-307): pos = (x + bound) / (2 bound);
// A[d][l] = (2^l pi) pos_d (the power-of-two multiple of pi is exact, then one
// rounding), feature (d L + l) 2 = sin A, + 1 = cos A, L = 10.  Width 60,
// padded with zeros to 64.
__device__ __forceinline__ void freq_encode(const HashLevels& H, const float (&p)[3], float (&enc)[kFEPad]) {
#pragma unroll
  for (int k = 0; k < kFEPad; ++k) enc[k] = 0.0f;
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int l = 0; l < kFL; ++l) {
      const float a = ldexpf(3.14159265358979323846f, l) * p[d];
      float sn, cs;
      sincosf(a, &sn, &cs);
      enc[(d * kFL + l) * 2] = sn;
      enc[(d * kFL + l) * 2 + 1] = cs;
    }
}

// ---- spherical-harmonics encoding -------------------------------------------------
// SphericalHarmonicsEncoder (encoders_flax.py:311-538) of highest order L <= 8 on
// the raw position (model.py:84; the bound is not used): feature l^2 + l + m is
// shc[l^2 + l + m] Q_l^k(z) T_m(x, y), k = |m|, with Q_l^k = d^k P_l / dz^k from
// the three-term recurrence (l - k) Q_l^k = (2l - 1) z Q_{l-1}^k - (l + k - 1)
// Q_{l-2}^k (Q_k^k = (2k - 1)!!, Q_{k+1}^k = (2k + 1) z Q_k^k) and T_m = Re (x +
// iy)^k (m > 0), Im (x + iy)^k (m < 0), 1 (m = 0).  Features past L^2 are zero.
__device__ __forceinline__ void sh_encode(const HashLevels& H, float x, float y, float z, float (&enc)[kFEPad]) {
#pragma unroll
  for (int f = 0; f < kFEPad; ++f) enc[f] = 0.0f;
  float ck = 1.0f, sk = 0.0f, dfact = 1.0f;  // Re / Im (x + iy)^k, (2k - 1)!!
#pragma unroll
  for (int k = 0; k < kSHMax; ++k) {
    if (k > 0) {
      const float c1 = x * ck - y * sk;
      sk = x * sk + y * ck;
      ck = c1;
      dfact *= (float)(2 * k - 1);
    }
    float q2 = 0.0f, q1 = dfact;  // Q_{l-1}^k, Q_l^k
#pragma unroll
    for (int l = k; l < kSHMax; ++l) {
      if (l > k) {
        const float q = (l == k + 1) ? (float)(2 * k + 1) * z * q1
                                     : ((float)(2 * l - 1) * z * q1 - (float)(l + k - 1) * q2) / (float)(l - k);
        q2 = q1;
        q1 = q;
      }
      if (l < H.L) {
        const int f = l * l + l;
        if (k == 0) {
          enc[f] = H.shc[f] * q1;
        } else {
          enc[f + k] = H.shc[f + k] * q1 * ck;
          enc[f - k] = H.shc[f - k] * q1 * sk;
        }
      }
    }
  }
}

// the dense encoders' features of a position (frequency: of pos = (x + bound) / 2 bound; spherical harmonics:
// of x itself)
__device__ __forceinline__ void dense_encode(const HashLevels& H, float x, float y, float z, float (&enc)[kFEPad]) {
  if (H.enc == 2) {
    sh_encode(H, x, y, z, enc);
  } else {
    float p[3];
    hash_pos(H, x, y, z, p);
    freq_encode(H, p, enc);
  }
}

__device__ __forceinline__ int dense_width(const HashLevels& H) { return H.enc == 2 ? H.L * H.L : 6 * H.L; }

// forward: u per entry (one lane per phase-sorted entry) and the encoding for the backward
__global__ __launch_bounds__(256) void freq_fwd_kernel(HashArgs A, HashLevels H, const float* __restrict__ params,
                                                       float* __restrict__ u_out, float* __restrict__ encbuf) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  float x, y, z;
  int row;
  const bool valid = hash_entry(A, t, x, y, z, row);
  const int E = dense_width(H), PM = hash_mlp_size(E);
  float enc[kFEPad];
  dense_encode(H, x, y, z, enc);
  const bool plus = t < A.n_plus;
  float act[kHW];
  const float* Praw = params;
  asm volatile("" : "+s"(Praw));
  dips_weight_t* Wp = dips_weights(Praw);
  dips_weight_t* Wm = Wp + PM;
  float up = 0.0f, um = 0.0f;
  if (plus) up = hash_mlp(Wp, E, enc, act);
  asm volatile("" : "+s"(Wm));
  if (!plus) um = hash_mlp(Wm, E, enc, act);
  if (valid) {
    u_out[A.t_csr[A.t_begin + t]] = plus ? up : um;
#pragma unroll
    for (int k = 0; k < kFEPad; ++k)
      if (k < E) encbuf[(int64_t)k * A.t_count + t] = enc[k];
  }
}

// backward: the MLP weight gradients (no table), partial rows per workgroup as
// in hash_bwd_kernel
__global__ __launch_bounds__(kHashThreads) void freq_bwd_kernel(HashArgs A, HashLevels H,
                                                                const float* __restrict__ params,
                                                                const float* __restrict__ encbuf, int64_t row_begin,
                                                                const float* __restrict__ r_row,
                                                                float* __restrict__ partial) {
  extern __shared__ float hash_lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / DIPS_WAVE);
  float* const AT = hash_lds + wave * kHStage;
  float* const BT = AT + 16 * kHStride;
  const int E = dense_width(H), PM = hash_mlp_size(E);
  HashGrad<kFEPad / 16> gp, gm;
  gp.zero();
  gm.zero();
  const ChunkRange<> cr(A.t_count, wave);
  for (int64_t ch = cr.first; ch < cr.end; ch += cr.step) {
    const int64_t t = ch * 64 + lane;
    float x, y, z;
    int row;
    const bool valid = hash_entry(A, t, x, y, z, row);
    const float seed = valid ? r_row[row - row_begin] * A.t_w[A.t_begin + t] : 0.0f;
    const bool plus = t < A.n_plus;
    float enc[kFEPad];
#pragma unroll
    for (int k = 0; k < kFEPad; ++k) enc[k] = (valid && k < E) ? encbuf[(int64_t)k * A.t_count + t] : 0.0f;
    const float* Praw = params;
    asm volatile("" : "+s"(Praw));
    dips_weight_t* Wp = dips_weights(Praw);
    dips_weight_t* Wm = Wp + PM;
    float act[kHW] = {};
    float dl[kHW];
    if (plus) {
      (void)hash_mlp(Wp, E, enc, act);
#pragma unroll
      for (int j = 0; j < kHW; ++j) dl[j] = seed * Wp[E * kHW + kHW + j] * act_grad<kActCelu>(act[j]);
    }
    asm volatile("" : "+s"(Wm));
    if (!plus) {
      (void)hash_mlp(Wm, E, enc, act);
#pragma unroll
      for (int j = 0; j < kHW; ++j) dl[j] = seed * Wm[E * kHW + kHW + j] * act_grad<kActCelu>(act[j]);
    }
    if (__any(plus && valid)) {
      float d2[kHW];
#pragma unroll
      for (int j = 0; j < kHW; ++j) d2[j] = plus ? dl[j] : 0.0f;
      gp.add(AT, BT, lane, E, plus ? seed : 0.0f, enc, act, d2);
    }
    if (__any(!plus && valid)) {
      float d2[kHW];
#pragma unroll
      for (int j = 0; j < kHW; ++j) d2[j] = plus ? 0.0f : dl[j];
      gm.add(AT, BT, lane, E, plus ? 0.0f : seed, enc, act, d2);
    }
  }
  const int stride = 2 * PM;
  __syncthreads();
  float* red = hash_lds + wave * stride;
  gp.flush(red, lane, E);
  gm.flush(red + PM, lane, E);
  __syncthreads();
  for (int q = threadIdx.x; q < stride; q += kHashThreads) {
    float sum = 0.0f;
    for (int w = 0; w < kHashWaves; ++w) sum += hash_lds[w * stride + q];
    partial[(int64_t)blockIdx.x * stride + q] = sum;
  }
}

// evaluation: u and d u / d x.  Frequency: d enc / d x through sin' = cos,
// cos' = -sin.  Spherical harmonics: d/dz of Q_l^k is Q_l^{k+1}; d/dx, d/dy of
// (x + iy)^k are k (x + iy)^{k-1} and i k (x + iy)^{k-1}.
__global__ __launch_bounds__(256) void freq_eval_kernel(HashLevels H, const float* __restrict__ params,
                                                        const float* __restrict__ pts,
                                                        const uint8_t* __restrict__ phase, int64_t m,
                                                        float* __restrict__ u, float* __restrict__ grad) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= m) return;
  const float x = pts[3 * t], y = pts[3 * t + 1], z = pts[3 * t + 2];
  const int E = dense_width(H), PM = hash_mlp_size(E);
  float enc[kFEPad], act[kHW];
  dense_encode(H, x, y, z, enc);
  const float* W = params + (phase[t] ? 0 : PM);
  u[t] = hash_mlp(W, E, enc, act);
  if (!grad) return;
  float dj[kHW];
#pragma unroll
  for (int j = 0; j < kHW; ++j) dj[j] = W[E * kHW + kHW + j] * act_grad<kActCelu>(act[j]);
  auto du_of = [&](int f) {  // d u / d enc_f
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < kHW; ++j) s = fmaf(W[f * kHW + j], dj[j], s);
    return s;
  };
  float g[3] = {0.f, 0.f, 0.f};
  if (H.enc == 2) {
    // Q[l][k] = d^k P_l / dz^k (zero for k > l), T from the complex powers
    float Q[kSHMax][kSHMax + 1];
#pragma unroll
    for (int l = 0; l < kSHMax; ++l)
#pragma unroll
      for (int k = 0; k <= kSHMax; ++k) Q[l][k] = 0.0f;
    float dfact = 1.0f;
#pragma unroll
    for (int k = 0; k < kSHMax; ++k) {
      if (k > 0) dfact *= (float)(2 * k - 1);
      Q[k][k] = dfact;
#pragma unroll
      for (int l = k + 1; l < kSHMax; ++l)
        Q[l][k] = (l == k + 1) ? (float)(2 * k + 1) * z * Q[k][k]
                               : ((float)(2 * l - 1) * z * Q[l - 1][k] - (float)(l + k - 1) * Q[l - 2][k]) / (float)(l - k);
    }
    float C[kSHMax], S[kSHMax];
    C[0] = 1.0f;
    S[0] = 0.0f;
#pragma unroll
    for (int k = 1; k < kSHMax; ++k) {
      C[k] = x * C[k - 1] - y * S[k - 1];
      S[k] = x * S[k - 1] + y * C[k - 1];
    }
#pragma unroll
    for (int l = 0; l < kSHMax; ++l) {
      if (l >= H.L) break;
#pragma unroll
      for (int k = 0; k <= l; ++k) {
        const int f = l * l + l;
        const float q = Q[l][k], dq = Q[l][k + 1];
        if (k == 0) {
          const float c = du_of(f) * H.shc[f];
          g[2] = fmaf(c, dq, g[2]);
        } else {
          const float kf = (float)k;
          const float cp = du_of(f + k) * H.shc[f + k], cm = du_of(f - k) * H.shc[f - k];
          // m = +k: q C_k;  m = -k: q S_k
          g[0] = fmaf(q * kf, cp * C[k - 1] + cm * S[k - 1], g[0]);
          g[1] = fmaf(q * kf, cm * C[k - 1] - cp * S[k - 1], g[1]);
          g[2] = fmaf(dq, cp * C[k] + cm * S[k], g[2]);
        }
      }
    }
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int l = 0; l < kFL; ++l) {
        const int ks = (d * kFL + l) * 2;
        const float du_s = du_of(ks), du_c = du_of(ks + 1);
        const float da = ldexpf(3.14159265358979323846f, l) / H.two_bound;  // d A / d x_d
        g[d] = fmaf(du_s * enc[ks + 1] - du_c * enc[ks], da, g[d]);
      }
  }
  grad[3 * t] = g[0];
  grad[3 * t + 1] = g[1];
  grad[3 * t + 2] = g[2];
}

// ---- host side ------------------------------------------------------------------
static int hash_width(const HashLevels& H) { return H.enc == 1 ? 6 * H.L : H.enc == 2 ? H.L * H.L : 2 * H.L; }

int64_t hash_table_rows(const dips_hashgrid_t* hg) {
  HashLevels H;
  if (hash_levels(hg, H)) return -1;
  return H.rows;
}

int64_t hash_param_count(const dips_hashgrid_t* hg) {
  HashLevels H;
  if (hash_levels(hg, H)) return -1;
  return 2 * (int64_t)hash_mlp_size(hash_width(H)) + H.rows * kHF;
}

// float workspace: u (max_nnz) | r (max_rows) | partial rows | table gradient | g
// table-gradient records of a batch of nent entries (one per entry, level and corner)
static int64_t hash_records(const HashLevels& H, int64_t nent) { return nent * H.L * 8; }

// the bucketed table gradient serves tables of up to kMaxBuckets buckets and
// batches of fewer than 2^31 records; others use atomic adds
static bool hash_bucketed(const HashLevels& H, int64_t nent, int* nb, int* shift = nullptr) {
  // smallest power of two nb with nb * (kBRows / 64) row groups covering the table
  const int64_t groups = (H.rows + (1 << kGShift) - 1) >> kGShift;
  int sh = 0;
  while (((int64_t)kBRows >> kGShift << sh) < groups) ++sh;
  *nb = 1 << sh;
  if (shift) *shift = sh;
  return *nb <= kMaxBuckets && hash_records(H, nent) < INT_MAX;
}

// records a (workgroup, bucket) slot holds: the mean per slot if no entries
// were aggregated (aggregation leaves about half), at least 64; records past it
// take atomic adds.  DIPS_HASH_SLOT_CAP (read once per process) lowers it, so
// that tests can run the atomic overflow path.
static int hash_slot_cap(const HashLevels& H, int64_t nent, int nb) {
  static const int64_t forced = [] {
    const char* e = std::getenv("DIPS_HASH_SLOT_CAP");
    return e ? std::max<int64_t>(1, std::atoll(e)) : (int64_t)0;
  }();
  const int64_t slots = (int64_t)kHashBlocks * nb;
  const int64_t mean = (hash_records(H, nent) + slots - 1) / slots;
  const int64_t cap = std::max<int64_t>(64, (mean + 15) / 16 * 16);
  return (int)(forced > 0 ? std::min(cap, forced) : cap);
}

int64_t hash_workspace(const dips_hashgrid_t* hg, int64_t max_rows, int64_t max_nnz, int64_t* doubles) {
  HashLevels H;
  if (hash_levels(hg, H) || max_rows < 1 || max_nnz < 0) return -1;
  const int64_t PM2 = 2 * (int64_t)hash_mlp_size(hash_width(H));
  const int64_t P = PM2 + H.rows * kHF;
  // norm partials: kRedBlocks, or one per bucket and one for the MLPs
  if (doubles) *doubles = rowsum_blocks(max_rows) + std::max(kRedBlocks, kMaxBuckets + 1) + 2;
  int64_t n = max_nnz + max_rows + (int64_t)kHashBlocks * PM2 + H.rows * kHF + P + 64;
  int nb;
  n += max_nnz * (hash_width(H) + (H.enc == 0 ? kHW : 0));  // the forward's encodings (and activations)
  if (H.enc == 0 && hash_bucketed(H, max_nnz, &nb))
    n += 4 * (int64_t)kHashBlocks * nb * hash_slot_cap(H, max_nnz, nb) + 4 + (int64_t)kHashBlocks * nb;
  return n;
}

int hash_step(const dips_grid_t* g, const dips_records_t* rec, const dips_mfma_batch_t* bt, const dips_hashgrid_t* hg,
              const dips_opt_t* opt, float inv_b, float* params, float* m, float* v, int32_t* count, float* loss_hist,
              int loss_cap, float* work, int64_t work_floats, double* dwork, int64_t dwork_doubles, float* grad_out,
              double* loss_out, hipStream_t st) {
  HashLevels H;
  int s = hash_levels(hg, H);
  if (s) return s;
  if (!bt || bt->t_count < 0 || bt->n_plus + bt->n_minus != bt->t_count ||
      bt->t_count != bt->ent_end - bt->ent_begin) {
    // every entry of the batch's CSR range must be listed once: the row sums read all of their values
    set_error("hash step: invalid batch (t_count must equal ent_end - ent_begin and n_plus + n_minus)");
    return DIPS_EINVAL;
  }
  const int PM = hash_mlp_size(hash_width(H));
  const int64_t stride = 2 * (int64_t)PM;
  const int64_t P = stride + H.rows * kHF;
  const int64_t nrows = bt->nrows;
  const int64_t nent = bt->ent_end - bt->ent_begin;
  int64_t need_d = 0;
  const int64_t need = hash_workspace(hg, std::max<int64_t>(nrows, 1), nent, &need_d);
  if (need < 0 || work_floats < need || dwork_doubles < need_d || !work || !dwork) {
    set_error("hash step: workspace too small");
    return DIPS_EINVAL;
  }
  if (P > INT_MAX) {
    set_error("hash step: too many parameters");
    return DIPS_EINVAL;
  }
  // float workspace: u | r | MLP partial rows | table gradient | g | encodings | [record slots (16 B records,
  // kHashBlocks x nb slots of cap) | records per slot]
  float* u = work;
  float* r = u + nent;
  float* partial = r + nrows;
  float* gtab = partial + (int64_t)kHashBlocks * stride;
  float* gvec = gtab + H.rows * kHF;
  int nb = 0, nb_shift = 0;
  const bool buckets = H.enc == 0 && hash_bucketed(H, nent, &nb, &nb_shift);
  const BucketMap bmap{nb_shift};
  const int cap = buckets ? hash_slot_cap(H, nent, nb) : 0;
  const int64_t nrec = buckets ? (int64_t)kHashBlocks * nb * cap : 0;
  float* encbuf = gvec + P;
  float* rbase = encbuf + nent * (hash_width(H) + (H.enc == 0 ? kHW : 0));
  rbase += (4 - (reinterpret_cast<uintptr_t>(rbase) / sizeof(float)) % 4) % 4;  // 16-byte alignment
  uint4* recs = reinterpret_cast<uint4*>(rbase);
  int* cnt = reinterpret_cast<int*>(rbase + 4 * nrec);
  const int nrb = rowsum_blocks(nrows);
  double* loss_part = dwork;
  double* sq_part = dwork + nrb;
  const int nsq = buckets ? nb + 1 : (int)std::min<int64_t>((P + 255) / 256, kRedBlocks);
  HashArgs A{point_set(g), rec->hx,  rec->hy,    rec->hz,      rec->row_point, bt->t_row, bt->t_code,
             bt->t_w,      bt->t_csr, bt->t_begin, bt->t_count, bt->n_plus};
  if (H.enc != 0) {
    // frequency and spherical-harmonics encodings: no table, so no records and no table gradient
    if (bt->t_count > 0)
      hipLaunchKernelGGL(freq_fwd_kernel, dim3((unsigned)((bt->t_count + 255) / 256)), dim3(256), 0, st, A, H,
                         params, u - bt->ent_begin, encbuf);
    rowsum_launch(rec->row_ptr, rec->row_c, bt->row_begin, nrows, rec->ent_w, rec->nnz, u - bt->ent_begin, r,
                  loss_part, st);
    const size_t flds = sizeof(float) * std::max<int64_t>((int64_t)kHashWaves * kHStage, (int64_t)kHashWaves * stride);
    static const bool freq_set = hipFuncSetAttribute((const void*)freq_bwd_kernel,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     160 * 1024) == hipSuccess;
    (void)freq_set;
    if (flds > 160 * 1024) {
      set_error("hash step: encoder too wide for the workgroup's LDS");
      return DIPS_EUNSUPPORTED;
    }
    hipLaunchKernelGGL(freq_bwd_kernel, dim3(kHashBlocks), dim3(kHashThreads), flds, st, A, H, params, encbuf,
                       bt->row_begin, r, partial);
  } else {
  constexpr size_t fwd_lds = sizeof(float) * kFwdSplit * kHashWaves * kHE * DIPS_WAVE;
  static const bool fwd_set = hipFuncSetAttribute((const void*)hash_fwd_kernel,
                                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (int)fwd_lds) == hipSuccess;
  (void)fwd_set;
  if (bt->t_count > 0)
    hipLaunchKernelGGL(hash_fwd_kernel, dim3(kHashBlocks), dim3(kFwdThreads), fwd_lds, st, A, H, params,
                       u - bt->ent_begin, encbuf);
  rowsum_launch(rec->row_ptr, rec->row_c, bt->row_begin, nrows, rec->ent_w, rec->nnz, u - bt->ent_begin, r, loss_part,
                st);
  const size_t lds = sizeof(float) * std::max<int64_t>((int64_t)kHashWaves * kHStage + (buckets ? nb : 0),
                                                       (int64_t)kHashWaves * stride);
  static const bool lds_set =
      hipFuncSetAttribute((const void*)hash_bwd_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          160 * 1024) == hipSuccess &&
      hipFuncSetAttribute((const void*)hash_bwd_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          160 * 1024) == hipSuccess &&
      hipFuncSetAttribute((const void*)hash_bucket_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)kBucketLds) == hipSuccess;
  (void)lds_set;
  if (lds > 160 * 1024) {
    set_error("hash step: encoder too wide for the workgroup's LDS");
    return DIPS_EUNSUPPORTED;
  }
  // gtab is zeroed first: the atomic path adds every record to it, the bucketed
  // path the records past a slot's capacity (the bucket sums then add theirs)
  hipLaunchKernelGGL(hash_zero_kernel, dim3(hash_zero_blocks(H.rows * kHF)), dim3(256), 0, st, gtab, H.rows * kHF);
  if (buckets) {
    hipLaunchKernelGGL(hash_bwd_kernel<true>, dim3(kHashBlocks), dim3(kHashThreads), lds, st, A, H, params, encbuf,
                       bt->row_begin, r, gtab, partial, nb, cnt, recs, cap, bmap);
    hipLaunchKernelGGL(hash_bucket_kernel, dim3(nb), dim3(kBThreads), kBucketLds, st, recs, cnt, cap, nb, bmap, H.rows, gtab,
                       inv_b, grad_out ? grad_out + stride : nullptr, sq_part);
  } else {
    hipLaunchKernelGGL(hash_bwd_kernel<false>, dim3(kHashBlocks), dim3(kHashThreads), lds, st, A, H, params, encbuf,
                       bt->row_begin, r, gtab, partial, 0, nullptr, nullptr, 0, bmap);
  }
  }
  float* const gout = grad_out ? grad_out : gvec;
  hipLaunchKernelGGL(hash_mlp_reduce_kernel, dim3((unsigned)((stride + 3) / 4)), dim3(256), 0, st, partial,
                     kHashBlocks, (int)stride, inv_b, gout);
  if (buckets)  // the bucket sums gave the table's norm partials: the MLPs' alone, after them
    hipLaunchKernelGGL(hash_reduce_kernel, dim3(1), dim3(256), 0, st, (int)stride, gtab, stride, inv_b, 0, gout,
                       sq_part + nb);
  else
    hipLaunchKernelGGL(hash_reduce_kernel, dim3(nsq), dim3(256), 0, st, (int)stride, gtab, P, inv_b,
                       grad_out ? 1 : 0, gout, sq_part);
  if (!opt) {
    loss_out_launch(loss_part, nrb, inv_b, loss_out, st);
    return hip_status(hipGetLastError(), "hash step launch");
  }
  apply_launch(gvec, P, gtab, stride, inv_b, sq_part, nsq, loss_part, nrb, inv_b, *opt, params, m, v, count, loss_hist,
               loss_cap, st);
  return hip_status(hipGetLastError(), "hash step launch");
}

int hash_eval(const dips_hashgrid_t* hg, const float* params, const float* pts, const uint8_t* phase, int64_t m,
              float* u, float* grad, hipStream_t st) {
  HashLevels H;
  int s = hash_levels(hg, H);
  if (s) return s;
  if (2 * (int64_t)hash_mlp_size(hash_width(H)) + H.rows * kHF > INT_MAX) {
    set_error("hash eval: too many parameters");
    return DIPS_EINVAL;
  }
  if (m <= 0) return DIPS_OK;
  if (H.enc != 0) {
    hipLaunchKernelGGL(freq_eval_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, H, params, pts, phase,
                       m, u, grad);
    return hip_status(hipGetLastError(), "freq_eval_kernel launch");
  }
  hipLaunchKernelGGL(hash_eval_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, H, params, pts, phase, m,
                     u, grad);
  return hip_status(hipGetLastError(), "hash_eval_kernel launch");
}

}  // namespace dips
