// Gradient reduction + fused optimizer step.
//
// One workgroup of 1024 lanes reduces the per-block partial rows written by
// the loss kernel in a fixed block order with float64 accumulation, scales by
// 1/B (the batch mean of trainer.py:901), applies the optax chain of
// jax_dips/solvers/optimizers.py:33-97 in place and records the batch loss.
//   custom : g <- g * min(1, max_norm/||g||) (clip_by_global_norm: g/||g||*max_norm
//            when ||g|| >= max_norm); Adam(b1=.9, b2=.999, eps=1e-8) with bias
//            correction at t = count+1; * schedule(count); * -1
//   adam   : Adam then * -lr
//   rmsprop: nu = .9 nu + .1 g^2; g / sqrt(nu + 1e-8) * -lr
// schedule: exponential lr * decay^(count/transition_steps) (count = 0 -> lr);
//           polynomial lr * (1 - min(count, T)/T).
// The step counter lives on the device so the sequence can be replayed from a
// captured graph.
#include "dips_common.h"
#include "dips_hip.h"

namespace dips {

constexpr int kOptThreads = 1024;
constexpr int kOptPer = 32;  // gradient values held in registers per lane
constexpr int kOptMaxParams = kOptThreads * kOptPer;

__global__ __launch_bounds__(kOptThreads) void opt_step_kernel(const float* __restrict__ partial, int nblk,
                                                               int P, float inv_b, dips_opt_t opt,
                                                               float* __restrict__ params, float* __restrict__ m,
                                                               float* __restrict__ v, int32_t* __restrict__ count,
                                                               float* __restrict__ loss_hist, int loss_cap) {
  float g_r[kOptPer];
  __shared__ double red[kOptThreads / DIPS_WAVE];
  __shared__ float gnorm_s;
  const int stride = P + 1;
  const int lane = threadIdx.x & (DIPS_WAVE - 1);
  const int wave = threadIdx.x / DIPS_WAVE;

  double sq = 0.0;
#pragma unroll
  for (int q = 0; q < kOptPer; ++q) {
    const int p = threadIdx.x + q * kOptThreads;
    g_r[q] = 0.0f;
    if (p < P) {
      double acc = 0.0;
#pragma unroll 8
      for (int b = 0; b < nblk; ++b) acc += (double)partial[(int64_t)b * stride + p];
      const float g = (float)(acc * (double)inv_b);
      g_r[q] = g;
      sq += (double)g * (double)g;
    }
  }
  sq = dips_wave_sum_d(sq);
  if (lane == 0) red[wave] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kOptThreads / DIPS_WAVE; ++w) t += red[w];
    gnorm_s = (float)sqrt(t);
    double l = 0.0;
#pragma unroll 8
    for (int b = 0; b < nblk; ++b) l += (double)partial[(int64_t)b * stride + P];
    const int c = *count;
    if (c >= 0 && c < loss_cap) loss_hist[c] = (float)(l * (double)inv_b);
  }
  __syncthreads();
  const int c = *count;
  const float gn = gnorm_s;
  const bool clip = (opt.kind == 0) && !(gn < opt.max_norm);
  float lr_t = opt.lr;
  if (opt.kind == 0) {
    if (opt.sched == 0) {
      lr_t = (c <= 0) ? opt.lr : opt.lr * powf(opt.decay_rate, (float)c / opt.transition_steps);
    } else {
      const float cc = fminf(fmaxf((float)c, 0.0f), opt.transition_steps);
      lr_t = opt.lr * (1.0f - cc / opt.transition_steps);
    }
  }
  const float t1 = (float)(c + 1);
  const float bc1 = 1.0f - powf(0.9f, t1);
  const float bc2 = 1.0f - powf(0.999f, t1);
#pragma unroll
  for (int q = 0; q < kOptPer; ++q) {
    const int p = threadIdx.x + q * kOptThreads;
    if (p >= P) continue;
    float g = g_r[q];
    if (clip) g = (g / gn) * opt.max_norm;
    float u;
    if (opt.kind == 2) {
      const float nv = 0.1f * (g * g) + 0.9f * v[p];
      v[p] = nv;
      u = -(g / sqrtf(nv + 1e-8f)) * opt.lr;
    } else {
      const float nm = 0.1f * g + 0.9f * m[p];
      const float nv = 0.001f * (g * g) + 0.999f * v[p];  // (1 - 0.999) rounded once, as optax
      m[p] = nm;
      v[p] = nv;
      const float mh = nm / bc1;
      const float vh = nv / bc2;
      u = mh / (sqrtf(vh) + 1e-8f);
      u = (opt.kind == 0) ? -(lr_t * u) : -(u * opt.lr);
    }
    params[p] = params[p] + u;
  }
  __syncthreads();
  if (threadIdx.x == 0) *count = c + 1;
}

int opt_step_launch(const dips_net_t* net, const dips_opt_t* opt, const float* partial, int nblk, float inv_batch,
                    float* params, float* m, float* v, int32_t* count, float* loss_hist, int loss_cap,
                    hipStream_t stream) {
  const int P = partial_stride(net) - 1;
  if (P > kOptMaxParams) {
    set_error("dips_opt_step: more parameters than the single-workgroup optimizer holds");
    return DIPS_EUNSUPPORTED;
  }
  if (nblk <= 0) {
    set_error("dips_opt_step: nblk must be positive");
    return DIPS_EINVAL;
  }
  if (opt->kind < 0 || opt->kind > 2 || opt->sched < 0 || opt->sched > 1 || !(opt->transition_steps > 0.0f)) {
    set_error("dips_opt_step: invalid optimizer configuration");
    return DIPS_EINVAL;
  }
  hipLaunchKernelGGL(opt_step_kernel, dim3(1), dim3(kOptThreads), 0, stream, partial, nblk, P, inv_batch, *opt,
                     params, m, v, count, loss_hist, loss_cap);
  return hip_status(hipGetLastError(), "opt_step_kernel launch");
}

}  // namespace dips
