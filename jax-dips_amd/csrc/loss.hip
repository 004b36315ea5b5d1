// Fused residual + loss + parameter-gradient kernel (the loss forward and its VJP).
//
// One lane owns one record row (training point).  Pass 1 evaluates the
// selected phase network at each entry position and accumulates the affine
// residual r = c + sum w_e NN(pos_e).  Pass 2 recomputes each entry's forward
// activations (nothing is stored between passes) and back-propagates the
// seed r * w_e (the 1/B of the mean is applied by the optimizer kernel),
// accumulating the parameter gradient of both phase networks in registers.
// Each block then reduces its lanes with wave shuffles plus LDS and writes one
// partial row (gradient, then 0.5*sum r^2); the reduction over blocks happens
// in a fixed order in optim.hip, so results are bitwise reproducible.
//
// Reference: loss = mean(optax.l2_loss(lhs, rhs)) over vmap(compute_Ax_and_b_
// preconditioned_fn) (jax_dips/solvers/poisson/trainer.py:892-912) and its
// value_and_grad (trainer.py:786); the network is DoubleMLP (nn/mlp/MLP.py:93-139).
#include "dips_common.h"
#include "dips_hip.h"
#include "dips_mlp.h"
#include "dips_step.h"

namespace dips {

constexpr int kLossThreads = 256;
template <int Lm, int Hm, int Lp, int Hp, bool kGrad, bool kFused>
__global__ __launch_bounds__(kLossThreads) void loss_grad_kernel(
    PointSet ps,
    float hx, float hy, float hz, const int32_t* __restrict__ row_point, const int32_t* __restrict__ row_ptr,
    const float* __restrict__ row_c, int64_t row_begin, int64_t nrows, const uint32_t* __restrict__ ent_code,
    const float* __restrict__ ent_w, int64_t nnz, float* params, float* __restrict__ partial,
    FusedOpt fo, int64_t nbulk, int nblk_bulk) {
  using NP = NetLayout<Lp, Hp>;
  using NM = NetLayout<Lm, Hm>;
  constexpr int PP = NP::size();
  constexpr int PM = NM::size();
  constexpr int PT = PP + PM;
  dips_weight_t* Pp = dips_weights(params);
  dips_weight_t* Pm = Pp + PP;

  float Gp[kGrad ? PP : 1];
  float Gm[kGrad ? PM : 1];
  if (kGrad) {
#pragma unroll
    for (int i = 0; i < PP; ++i) Gp[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < PM; ++i) Gm[i] = 0.0f;
  }
  float lsum = 0.0f;

  const int lane = threadIdx.x & (DIPS_WAVE - 1);
  const int wave = threadIdx.x / DIPS_WAVE;
  if ((int)blockIdx.x < nblk_bulk) {
    // bulk rows (few entries): one lane per row
    const int64_t local = (int64_t)blockIdx.x * kLossThreads + threadIdx.x;
    const int64_t row = row_begin + local;
    const int pt = (local < nbulk) ? row_point[row] : -1;
    float px, py, pz;
    if (point_xyz(ps, pt, px, py, pz)) {
      // entry range clamped to the entry arrays (rows are well formed by construction)
      const int beg = max(row_ptr[row], 0);
      const int end = (int)min((int64_t)row_ptr[row + 1], nnz);
      float r = row_c[row];
      for (int e = beg; e < end; ++e) {
        const uint32_t code = ent_code[e];
        int ox, oy, oz;
        dips_decode_offset(code, ox, oy, oz);
        const float x = px + (float)ox * hx;
        const float y = py + (float)oy * hy;
        const float z = pz + (float)oz * hz;
        const float u = (code & 128u) ? mlp_forward<Lp, Hp>(Pp, x, y, z) : mlp_forward<Lm, Hm>(Pm, x, y, z);
        r = fmaf(ent_w[e], u, r);
      }
      lsum = 0.5f * r * r;
      if (kGrad) {
        for (int e = beg; e < end; ++e) {
          const uint32_t code = ent_code[e];
          int ox, oy, oz;
          dips_decode_offset(code, ox, oy, oz);
          const float x = px + (float)ox * hx;
          const float y = py + (float)oy * hy;
          const float z = pz + (float)oz * hz;
          const float seed = r * ent_w[e];
          if (code & 128u)
            mlp_backward<Lp, Hp>(Pp, x, y, z, seed, Gp);
          else
            mlp_backward<Lm, Hm>(Pm, x, y, z, seed, Gm);
        }
      }
    }
  } else {
    // interface rows (many entries): one wavefront per row, entries spread over the lanes
    const int64_t wrow = ((int64_t)blockIdx.x - nblk_bulk) * (kLossThreads / DIPS_WAVE) + wave;
    const int64_t row = row_begin + nbulk + wrow;
    const int pt = (nbulk + wrow < nrows) ? row_point[row] : -1;  // uniform across the wave
    float px, py, pz;
    if (point_xyz(ps, pt, px, py, pz)) {
      const int beg = max(row_ptr[row], 0);
      const int end = (int)min((int64_t)row_ptr[row + 1], nnz);
      float acc = 0.0f;
      for (int e = beg + lane; e < end; e += DIPS_WAVE) {
        const uint32_t code = ent_code[e];
        int ox, oy, oz;
        dips_decode_offset(code, ox, oy, oz);
        const float x = px + (float)ox * hx;
        const float y = py + (float)oy * hy;
        const float z = pz + (float)oz * hz;
        const float u = (code & 128u) ? mlp_forward<Lp, Hp>(Pp, x, y, z) : mlp_forward<Lm, Hm>(Pm, x, y, z);
        acc = fmaf(ent_w[e], u, acc);
      }
      const float r = dips_wave_sum(acc) + row_c[row];
      lsum = (lane == 0) ? 0.5f * r * r : 0.0f;
      if (kGrad) {
        for (int e = beg + lane; e < end; e += DIPS_WAVE) {
          const uint32_t code = ent_code[e];
          int ox, oy, oz;
          dips_decode_offset(code, ox, oy, oz);
          const float x = px + (float)ox * hx;
          const float y = py + (float)oy * hy;
          const float z = pz + (float)oz * hz;
          const float seed = r * ent_w[e];
          if (code & 128u)
            mlp_backward<Lp, Hp>(Pp, x, y, z, seed, Gp);
          else
            mlp_backward<Lm, Hm>(Pm, x, y, z, seed, Gm);
        }
      }
    }
  }

  // block reduction: DPP wave sums, then LDS across the 4 waves
  __shared__ float red[kLossThreads / DIPS_WAVE][PT + 1];
  if (kGrad) {
#pragma unroll
    for (int p = 0; p < PP; ++p) {
      const float s = dips_wave_sum(Gp[p]);
      if (lane == 0) red[wave][p] = s;
    }
#pragma unroll
    for (int p = 0; p < PM; ++p) {
      const float s = dips_wave_sum(Gm[p]);
      if (lane == 0) red[wave][PP + p] = s;
    }
  }
  {
    const float s = dips_wave_sum(lsum);
    if (lane == 0) red[wave][PT] = s;
  }
  __syncthreads();
  constexpr int stride = PT + 1;
  const int nblk = (int)gridDim.x;
  for (int p = threadIdx.x; p < stride; p += kLossThreads) {
    float s = 0.0f;
    if (kGrad || p == PT) {
#pragma unroll
      for (int w = 0; w < kLossThreads / DIPS_WAVE; ++w) s += red[w][p];
    }
    // fused: parameter-major rows (p * nblk + block) so the last workgroup reads contiguous runs
    if (kFused)
      partial[(int64_t)p * nblk + blockIdx.x] = s;
    else
      partial[(int64_t)blockIdx.x * stride + p] = s;
  }
  if (!kFused) return;

  // the last workgroup to arrive reduces all partial rows in block order and
  // applies the optimizer; one thread per workgroup fences (release after the
  // barrier, acquire in the last workgroup) and counts
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
  __shared__ double dred[kLossThreads / DIPS_WAVE];
  __shared__ float gnorm_s;
  __shared__ int count_s;
  constexpr int kPer = (stride + kLossThreads - 1) / kLossThreads;
  float gl[kPer];
  double sq = 0.0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int p = threadIdx.x + q * kLossThreads;
    gl[q] = 0.0f;
    if (p < stride) {
      const float* src = partial + (int64_t)p * nblk;
      double acc = 0.0;
      // plain loads after the acquire fence above, unrolled so several are in flight
#pragma unroll 8
      for (int b = 0; b < nblk; ++b) acc += (double)src[b];
      const float g = (float)(acc * (double)fo.inv_b);
      gl[q] = g;
      if (p < PT) sq += (double)g * (double)g;
    }
  }
  sq = dips_wave_sum_d(sq);
  if (lane == 0) dred[wave] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kLossThreads / DIPS_WAVE; ++w) t += dred[w];
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
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int p = threadIdx.x + q * kLossThreads;
    if (p < PT) apply_update(fo.opt, c, lr_t, bc1, bc2, clip, gn, p, gl[q], params, fo.m, fo.v);
    if (p == PT && c >= 0 && c < fo.loss_cap) fo.loss_hist[c] = gl[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    *fo.count = c + 1;
    *fo.sync = 0u;
  }
}

constexpr int kRowsPerWaveBlock = kLossThreads / DIPS_WAVE;  // interface rows per workgroup

// workgroups for nrows rows of which the first nbulk are lane-per-row rows
int64_t loss_blocks(int64_t nrows, int64_t nbulk) {
  return (nbulk + kLossThreads - 1) / kLossThreads + (nrows - nbulk + kRowsPerWaveBlock - 1) / kRowsPerWaveBlock;
}

// upper bound of loss_blocks over all splits (every row an interface row)
int64_t partial_blocks(int64_t nrows) { return (nrows + kRowsPerWaveBlock - 1) / kRowsPerWaveBlock + 1; }

int precond_params(const dips_net_t* net) {
  const int w1 = net->pc_w1, w2 = net->pc_w2;
  if (w1 <= 0) return 0;
  return 26 * w1 + w1 + (w2 > 0 ? w1 * w2 + w2 + w2 + 1 : w1 + 1);
}

int partial_stride(const dips_net_t* net) {
  auto sz = [](int L, int H) { return (3 * H + H) + (L - 1) * (H * H + H) + (H + 1); };
  return sz(net->Lp, net->Hp) + sz(net->Lm, net->Hm) + precond_params(net) + 1;
}

template <int Lm, int Hm, int Lp, int Hp>
static void launch_shape(const dips_grid_t* g, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                         float* params, float* partial, int mode, const FusedOpt& fo, int nblk, int64_t nbulk,
                         int nblk_bulk, hipStream_t st) {
  dim3 grid(nblk), block(kLossThreads);
#define DIPS_LAUNCH(G, F)                                                                                         \
  hipLaunchKernelGGL((loss_grad_kernel<Lm, Hm, Lp, Hp, G, F>), grid, block, 0, st, point_set(g), rec->hx, rec->hy, \
                     rec->hz, rec->row_point, rec->row_ptr, rec->row_c, row_begin, nrows, \
                     rec->ent_code, rec->ent_w, rec->nnz, params, partial, fo, nbulk, \
                     nblk_bulk)
  if (mode == 2)
    DIPS_LAUNCH(true, true);
  else if (mode == 1)
    DIPS_LAUNCH(true, false);
  else
    DIPS_LAUNCH(false, false);
#undef DIPS_LAUNCH
}

int net_supported(const dips_net_t* n) {
  if (n->pc_w1 > 0 || n->act != 0) return 0;  // the preconditioner and CELU run in the tile kernel only
#define X(a, b, c, d) \
  if (n->Lm == a && n->Hm == b && n->Lp == c && n->Hp == d) return 1;
  DIPS_SHAPES(X)
#undef X
  return 0;
}

int loss_grad_launch(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                     int64_t nbulk, const dips_net_t* net, const float* params, float* partial, int64_t partial_cap,
                     int want_grad, int* nblk_out, hipStream_t stream, const FusedOpt* fused) {
  if (net->pc_w1 > 0) {
    set_error("dips_loss_grad: the preconditioner runs in the tile kernel only (dips_tile_*)");
    return DIPS_EUNSUPPORTED;
  }
  if (nrows <= 0) {
    set_error("dips_loss_grad: nrows must be positive");
    return DIPS_EINVAL;
  }
  if (row_begin < 0 || row_begin + nrows > rec->nrows) {
    set_error("dips_loss_grad: row range outside the records");
    return DIPS_EINVAL;
  }
  if (nbulk < 0 || nbulk > nrows) {
    set_error("dips_loss_grad: nbulk outside [0, nrows]");
    return DIPS_EINVAL;
  }
  const int64_t nblk_bulk = (nbulk + kLossThreads - 1) / kLossThreads;
  const int64_t nblk = loss_blocks(nrows, nbulk);
  const int64_t need = nblk * (int64_t)partial_stride(net);
  if (partial_cap < need) {
    set_error("dips_loss_grad: partial buffer too small");
    return DIPS_EINVAL;
  }
  if (nblk > 0x7fffffff) {
    set_error("dips_loss_grad: too many rows for one launch");
    return DIPS_EINVAL;
  }
  *nblk_out = (int)nblk;
  FusedOpt fo{};
  if (fused) fo = *fused;
  const int mode = fused ? 2 : (want_grad ? 1 : 0);
#define X(a, b, c, d)                                                                                              \
  if (net->Lm == a && net->Hm == b && net->Lp == c && net->Hp == d) {                                             \
    launch_shape<a, b, c, d>(grid, rec, row_begin, nrows, const_cast<float*>(params), partial, mode, fo, (int)nblk, \
                             nbulk, (int)nblk_bulk, stream);                                                      \
    return hip_status(hipGetLastError(), "loss_grad_kernel launch");                                              \
  }
  DIPS_SHAPES(X)
#undef X
  set_error("dips_loss_grad: network shape has no compiled kernel");
  return DIPS_EUNSUPPORTED;
}

int train_step_launch(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                      int64_t nbulk, const dips_net_t* net, const dips_opt_t* opt, float inv_batch, float* params,
                      float* m, float* v, int32_t* count, float* loss_hist, int loss_cap, float* partial,
                      int64_t partial_cap, uint32_t* sync, hipStream_t stream) {
  FusedOpt fo{*opt, inv_batch, m, v, count, loss_hist, loss_cap, sync};
  int nb = 0;
  return loss_grad_launch(grid, rec, row_begin, nrows, nbulk, net, params, partial, partial_cap, 1, &nb, stream, &fo);
}

}  // namespace dips
