// Network evaluation at arbitrary points: u, grad_x u and the normal derivative.
//
// Reference: Trainer.evaluate_solution_fn (jax_dips/solvers/poisson/trainer.py:
// 836-844) and the post-processing of evaluate_solution_and_gradients
// (trainer.py:914-977): grad u is jax.grad of the selected network with respect
// to the input point, d_n u = grad u . n with n the central-difference normal of
// phi at the eval spacing (discretization.py:199-218), supplied by the caller.
#include "dips_common.h"
#include "dips_hip.h"
#include "dips_mlp.h"

namespace dips {

template <int Lm, int Hm, int Lp, int Hp, int Act>
__global__ __launch_bounds__(256) void eval_kernel(const float* __restrict__ params, const float* __restrict__ pts,
                                                   const uint8_t* __restrict__ phase, const float* __restrict__ nrm,
                                                   int64_t m, float* __restrict__ u, float* __restrict__ grad,
                                                   float* __restrict__ gradn) {
  using NP = NetLayout<Lp, Hp>;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const float x = pts[t * 3 + 0], y = pts[t * 3 + 1], z = pts[t * 3 + 2];
  float g[3];
  dips_weight_t* W = dips_weights(params);
  const float v = phase[t] ? mlp_value_grad<Lp, Hp, Act>(W, x, y, z, g)
                           : mlp_value_grad<Lm, Hm, Act>(W + NP::size(), x, y, z, g);
  u[t] = v;
  if (grad) {
    grad[t * 3 + 0] = g[0];
    grad[t * 3 + 1] = g[1];
    grad[t * 3 + 2] = g[2];
  }
  if (gradn && nrm) gradn[t] = g[0] * nrm[t * 3 + 0] + g[1] * nrm[t * 3 + 1] + g[2] * nrm[t * 3 + 2];
}

int eval_launch(const dips_net_t* net, const float* params, const float* pts, const uint8_t* phase, const float* nrm,
                int64_t m, float* u, float* grad, float* gradn, hipStream_t st) {
  if (m <= 0) return DIPS_OK;
  const int64_t nb = (m + 255) / 256;
  if (nb > 0x7fffffff) {
    set_error("dips_eval: too many points for one launch");
    return DIPS_EINVAL;
  }
#define X(a, b, c, d, e)                                                                                        \
  if (net->Lm == a && net->Hm == b && net->Lp == c && net->Hp == d && net->act == e) {                         \
    hipLaunchKernelGGL((eval_kernel<a, b, c, d, e>), dim3((unsigned)nb), dim3(256), 0, st, params, pts, phase, nrm, \
                       m, u, grad, gradn);                                                                      \
    return hip_status(hipGetLastError(), "eval_kernel launch");                                                 \
  }
#define XT(a, b, c, d) X(a, b, c, d, 0)
  DIPS_SHAPES(XT)
  X(1, 16, 1, 16, 1)
  X(2, 16, 2, 16, 1)
#undef XT
#undef X
  // wide networks: MFMA tiles of 16 points
  return mfma_eval_launch(net, params, pts, phase, nrm, m, u, grad, gradn, st);
}

}  // namespace dips
