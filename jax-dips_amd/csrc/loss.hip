// Fused residual + loss + parameter-gradient kernel (forward K2 and its VJP K3).
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

namespace dips {

constexpr int kLossThreads = 256;

template <int Lm, int Hm, int Lp, int Hp, bool kGrad>
__global__ __launch_bounds__(kLossThreads) void loss_grad_kernel(
    const float* __restrict__ xs, const float* __restrict__ ys, const float* __restrict__ zs, int ny, int nz,
    float hx, float hy, float hz, const int32_t* __restrict__ row_point, const int32_t* __restrict__ row_ptr,
    const float* __restrict__ row_c, int64_t row_begin, int64_t nrows, const uint32_t* __restrict__ ent_code,
    const float* __restrict__ ent_w, int64_t nnz, int npts, const float* __restrict__ params,
    float* __restrict__ partial) {
  using NP = NetLayout<Lp, Hp>;
  using NM = NetLayout<Lm, Hm>;
  constexpr int PP = NP::size();
  constexpr int PM = NM::size();
  constexpr int PT = PP + PM;
  const float* __restrict__ Pp = params;
  const float* __restrict__ Pm = params + PP;

  float Gp[kGrad ? PP : 1];
  float Gm[kGrad ? PM : 1];
  if (kGrad) {
#pragma unroll
    for (int i = 0; i < PP; ++i) Gp[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < PM; ++i) Gm[i] = 0.0f;
  }
  float lsum = 0.0f;

  const int64_t local = (int64_t)blockIdx.x * kLossThreads + threadIdx.x;
  const int64_t row = row_begin + local;
  const int pt = (local < nrows) ? row_point[row] : -1;
  if (pt >= 0 && pt < npts) {
    const int k = pt % nz;
    const int ij = pt / nz;
    const int j = ij % ny;
    const int i = ij / ny;
    const float px = xs[i], py = ys[j], pz = zs[k];
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

  // block reduction: wave shuffles, then LDS across the 4 waves
  __shared__ float red[kLossThreads / DIPS_WAVE][PT + 1];
  const int lane = threadIdx.x & (DIPS_WAVE - 1);
  const int wave = threadIdx.x / DIPS_WAVE;
  if (kGrad) {
#pragma unroll
    for (int p = 0; p < PP; ++p) {
      float s = dips_wave_sum(Gp[p]);
      if (lane == 0) red[wave][p] = s;
    }
#pragma unroll
    for (int p = 0; p < PM; ++p) {
      float s = dips_wave_sum(Gm[p]);
      if (lane == 0) red[wave][PP + p] = s;
    }
  }
  {
    float s = dips_wave_sum(lsum);
    if (lane == 0) red[wave][PT] = s;
  }
  __syncthreads();
  const int stride = PT + 1;
  float* out = partial + (int64_t)blockIdx.x * stride;
  for (int p = threadIdx.x; p < stride; p += kLossThreads) {
    if (!kGrad && p < PT) {
      out[p] = 0.0f;
      continue;
    }
    float s = 0.0f;
#pragma unroll
    for (int w = 0; w < kLossThreads / DIPS_WAVE; ++w) s += red[w][p];
    out[p] = s;
  }
}

int64_t partial_blocks(int64_t nrows) { return (nrows + kLossThreads - 1) / kLossThreads; }

int partial_stride(const dips_net_t* net) {
  auto sz = [](int L, int H) { return (3 * H + H) + (L - 1) * (H * H + H) + (H + 1); };
  return sz(net->Lp, net->Hp) + sz(net->Lm, net->Hm) + 1;
}

template <int Lm, int Hm, int Lp, int Hp>
static void launch_shape(const dips_grid_t* g, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                         const float* params, float* partial, int want_grad, int nblk, hipStream_t st) {
  dim3 grid(nblk), block(kLossThreads);
  if (want_grad)
    hipLaunchKernelGGL((loss_grad_kernel<Lm, Hm, Lp, Hp, true>), grid, block, 0, st, g->xs, g->ys, g->zs, g->ny,
                       g->nz, rec->hx, rec->hy, rec->hz, rec->row_point, rec->row_ptr, rec->row_c, row_begin, nrows,
                       rec->ent_code, rec->ent_w, rec->nnz, g->nx * g->ny * g->nz, params, partial);
  else
    hipLaunchKernelGGL((loss_grad_kernel<Lm, Hm, Lp, Hp, false>), grid, block, 0, st, g->xs, g->ys, g->zs, g->ny,
                       g->nz, rec->hx, rec->hy, rec->hz, rec->row_point, rec->row_ptr, rec->row_c, row_begin, nrows,
                       rec->ent_code, rec->ent_w, rec->nnz, g->nx * g->ny * g->nz, params, partial);
}

int net_supported(const dips_net_t* n) {
#define X(a, b, c, d) \
  if (n->Lm == a && n->Hm == b && n->Lp == c && n->Hp == d) return 1;
  DIPS_SHAPES(X)
#undef X
  return 0;
}

int loss_grad_launch(const dips_grid_t* grid, const dips_records_t* rec, int64_t row_begin, int64_t nrows,
                     const dips_net_t* net, const float* params, float* partial, int64_t partial_cap,
                     int want_grad, int* nblk_out, hipStream_t stream) {
  if (nrows <= 0) {
    set_error("dips_loss_grad: nrows must be positive");
    return DIPS_EINVAL;
  }
  if (row_begin < 0 || row_begin + nrows > rec->nrows) {
    set_error("dips_loss_grad: row range outside the records");
    return DIPS_EINVAL;
  }
  const int64_t nblk = partial_blocks(nrows);
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
#define X(a, b, c, d)                                                                                  \
  if (net->Lm == a && net->Hm == b && net->Lp == c && net->Hp == d) {                                 \
    launch_shape<a, b, c, d>(grid, rec, row_begin, nrows, params, partial, want_grad, (int)nblk, stream); \
    return hip_status(hipGetLastError(), "loss_grad_kernel launch");                                  \
  }
  DIPS_SHAPES(X)
#undef X
  set_error("dips_loss_grad: network shape has no compiled kernel");
  return DIPS_EUNSUPPORTED;
}

}  // namespace dips
