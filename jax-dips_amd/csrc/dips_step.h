// Optimizer state and update shared by the fused single-launch training steps.
#pragma once
#include <math.h>

#include "dips_common.h"
#include "dips_hip.h"

namespace dips {

// Optimizer state consumed by the fused step (last workgroup of the launch).
struct FusedOpt {
  dips_opt_t opt;
  float inv_b;
  float* m;
  float* v;
  int32_t* count;
  float* loss_hist;
  int loss_cap;
  uint32_t* sync;  // arrival counter, zero between launches
};

// Optimizer update of one parameter value with gradient g (optax chain of
// solvers/optimizers.py:33-97; see optim.hip for the float32 constants).
__device__ __forceinline__ void apply_update_value(const dips_opt_t& opt, float lr_t, float bc1, float bc2, bool clip,
                                                   float gn, float g, float& param, float& m, float& v) {
  if (clip) g = (g / gn) * opt.max_norm;
  float u;
  if (opt.kind == 2) {
    const float nv = 0.1f * (g * g) + 0.9f * v;
    v = nv;
    u = -(g / sqrtf(nv + 1e-8f)) * opt.lr;
  } else {
    const float nm = 0.1f * g + 0.9f * m;
    const float nv = 0.001f * (g * g) + 0.999f * v;
    m = nm;
    v = nv;
    u = (nm / bc1) / (sqrtf(nv / bc2) + 1e-8f);
    u = (opt.kind == 0) ? -(lr_t * u) : -(u * opt.lr);
  }
  param = param + u;
}

// The same update of parameter p in memory.
__device__ __forceinline__ void apply_update(const dips_opt_t& opt, int c, float lr_t, float bc1, float bc2,
                                             bool clip, float gn, int p, float g, float* params, float* m, float* v) {
  (void)c;
  float pv = params[p], mv = m[p], vv = v[p];
  apply_update_value(opt, lr_t, bc1, bc2, clip, gn, g, pv, mv, vv);
  if (opt.kind != 2) m[p] = mv;
  v[p] = vv;
  params[p] = pv;
}

__device__ __forceinline__ float schedule_lr(const dips_opt_t& opt, int c) {
  if (opt.kind != 0) return opt.lr;
  if (opt.sched == 0) return (c <= 0) ? opt.lr : opt.lr * powf(opt.decay_rate, (float)c / opt.transition_steps);
  const float cc = fminf(fmaxf((float)c, 0.0f), opt.transition_steps);
  return opt.lr * (1.0f - cc / opt.transition_steps);
}

}  // namespace dips
