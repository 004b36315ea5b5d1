// Matrix-free BiCGSTAB on grid values of the same cut-cell FV discretisation.
//
// The grid-value operator of the reference (Trainer.residual_vector with
// compute_Ax_and_b_discrete_fn, jax_dips/solvers/poisson/trainer.py:883-890,
// discretization.py:522-741) is the per-point residual of the training step with
// the network replaced by values of a grid function u, read through a ghost layer
// that extrapolates linearly beyond the box (u[-1] = 2 u[0] - u[1] per axis,
// domain/interpolate.py:762-816).  The zoom-0 records of the eval grid therefore
// define it: r(p) = c_p + sum_e w_e u(p + o_e); the phase bit of an entry is
// irrelevant here because the grid function carries one value per node.
// BiCGSTAB solves A u = -c with float32 operator applications and float64
// inner products and norms, reduced deterministically (fixed block order).
#include <math.h>

#include "dips_common.h"
#include "dips_hip.h"

namespace dips {

constexpr int kJudgeThreads = 256;

// grid value with the linear ghost layer: indices -1 and n are extrapolated
__device__ __forceinline__ float u_ghost(const float* __restrict__ u, int i, int j, int k, int nx, int ny, int nz) {
  // tensor product of 1-D rules: idx in [0,n) -> itself; -1 -> 2*[0] - [1]; n -> 2*[n-1] - [n-2]
  int ia[2], ja[2], ka[2];
  float wi[2], wj[2], wk[2];
  int ni = 1, nj = 1, nk = 1;
  auto rule = [](int idx, int n, int* a, float* w, int& cnt) {
    if (idx < 0) {
      a[0] = 0, a[1] = 1, w[0] = 2.0f, w[1] = -1.0f, cnt = 2;
    } else if (idx >= n) {
      a[0] = n - 1, a[1] = n - 2, w[0] = 2.0f, w[1] = -1.0f, cnt = 2;
    } else {
      a[0] = idx, w[0] = 1.0f, cnt = 1;
    }
  };
  rule(i, nx, ia, wi, ni);
  rule(j, ny, ja, wj, nj);
  rule(k, nz, ka, wk, nk);
  float s = 0.0f;
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b < nj; ++b)
      for (int c = 0; c < nk; ++c)
        s = fmaf(wi[a] * wj[b] * wk[c], u[((int64_t)ia[a] * ny + ja[b]) * nz + ka[c]], s);
  return s;
}

// y[pt] = (with_c ? c_row : 0) + sum_e w_e u(pt + o_e)   for every record row
__global__ __launch_bounds__(kJudgeThreads) void grid_apply_kernel(
    int nx, int ny, int nz, const int32_t* __restrict__ row_point, const int32_t* __restrict__ row_ptr,
    const float* __restrict__ row_c, int64_t nrows, const uint32_t* __restrict__ ent_code,
    const float* __restrict__ ent_w, int64_t nnz, const float* __restrict__ u, float* __restrict__ y, int with_c) {
  const int64_t row = (int64_t)blockIdx.x * kJudgeThreads + threadIdx.x;
  if (row >= nrows) return;
  const int pt = row_point[row];
  const int64_t npts = (int64_t)nx * ny * nz;
  if (pt < 0 || pt >= npts) return;
  const int k = pt % nz, ij = pt / nz, j = ij % ny, i = ij / ny;
  const int beg = max(row_ptr[row], 0);
  const int end = (int)min((int64_t)row_ptr[row + 1], nnz);
  float acc = with_c ? row_c[row] : 0.0f;
  for (int e = beg; e < end; ++e) {
    int ox, oy, oz;
    dips_decode_offset(ent_code[e], ox, oy, oz);
    const int ii = i + ox, jj = j + oy, kk = k + oz;
    float v;
    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz)
      v = u[((int64_t)ii * ny + jj) * nz + kk];
    else if (ii >= -1 && ii <= nx && jj >= -1 && jj <= ny && kk >= -1 && kk <= nz)
      v = u_ghost(u, ii, jj, kk, nx, ny, nz);
    else
      v = 0.0f;  // beyond the single ghost layer: cannot occur for stencils of radius 2 around interior points
    acc = fmaf(ent_w[e], v, acc);
  }
  y[pt] = acc;
}

// dst = a*x + b*y + c*z (any pointer may alias dst; null operands count as zero)
__global__ void axpbypcz_kernel(int64_t n, float* dst, double a, const float* x, double b, const float* y,
                                double c, const float* z) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double v = 0.0;
  if (x) v += a * (double)x[t];
  if (y) v += b * (double)y[t];
  if (z) v += c * (double)z[t];
  dst[t] = (float)v;
}

// partial[blk] = sum over the block's slice of x*y (float64)
__global__ __launch_bounds__(kJudgeThreads) void dot_partial_kernel(int64_t n, const float* __restrict__ x,
                                                                    const float* __restrict__ y,
                                                                    double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * kJudgeThreads + threadIdx.x; t < n; t += (int64_t)gridDim.x * kJudgeThreads)
    s += (double)x[t] * (double)y[t];
  s = dips_wave_sum_d(s);
  __shared__ double red[kJudgeThreads / DIPS_WAVE];
  if ((threadIdx.x & (DIPS_WAVE - 1)) == 0) red[threadIdx.x / DIPS_WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kJudgeThreads / DIPS_WAVE; ++w) t += red[w];
    partial[blockIdx.x] = t;
  }
}

// out[slot] = sum of partial[0..nblk) in block order
__global__ void dot_final_kernel(const double* __restrict__ partial, int nblk, double* __restrict__ out, int slot) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double t = 0.0;
  for (int b = 0; b < nblk; ++b) t += partial[b];
  out[slot] = t;
}

constexpr int kDotBlocks = 512;

struct JudgeCtx {
  const dips_grid_t* g;
  const dips_records_t* rec;
  hipStream_t st;
  double* dpart;  // kDotBlocks
  double* dres;   // 8 result slots
};

static int apply(const JudgeCtx& C, const float* u, float* y, int with_c) {
  const int64_t nb = (C.rec->nrows + kJudgeThreads - 1) / kJudgeThreads;
  hipLaunchKernelGGL(grid_apply_kernel, dim3((unsigned)nb), dim3(kJudgeThreads), 0, C.st, C.g->nx, C.g->ny, C.g->nz,
                     C.rec->row_point, C.rec->row_ptr, C.rec->row_c, C.rec->nrows, C.rec->ent_code, C.rec->ent_w,
                     C.rec->nnz, u, y, with_c);
  return hip_status(hipGetLastError(), "grid_apply_kernel launch");
}

static void lin(const JudgeCtx& C, int64_t n, float* dst, double a, const float* x, double b, const float* y, double c,
                const float* z) {
  hipLaunchKernelGGL(axpbypcz_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, C.st, n, dst, a, x, b, y, c, z);
}

static void dot(const JudgeCtx& C, int64_t n, const float* x, const float* y, int slot) {
  hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(kJudgeThreads), 0, C.st, n, x, y, C.dpart);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(64), 0, C.st, C.dpart, kDotBlocks, C.dres, slot);
}

static int fetch(const JudgeCtx& C, double* host, int count) {
  hipError_t e = hipMemcpyAsync(host, C.dres, sizeof(double) * count, hipMemcpyDeviceToHost, C.st);
  if (e == hipSuccess) e = hipStreamSynchronize(C.st);
  return hip_status(e, "judge scalar readback");
}

int grid_apply(const dips_grid_t* g, const dips_records_t* rec, const float* u, float* y, int with_c,
               hipStream_t st) {
  JudgeCtx C{g, rec, st, nullptr, nullptr};
  return apply(C, u, y, with_c);
}

int bicgstab(const dips_grid_t* g, const dips_records_t* rec, float* x, float* work, double* dwork, int max_iter,
             double rtol, int* iters, double* relres, hipStream_t st) {
  const int64_t n = (int64_t)g->nx * g->ny * g->nz;
  float* r = work;
  float* rh = work + n;
  float* p = work + 2 * n;
  float* v = work + 3 * n;
  float* s = work + 4 * n;
  float* t = work + 5 * n;
  JudgeCtx C{g, rec, st, dwork, dwork + kDotBlocks};
  double h[4];
  int status;
  // r = b - A x = -(A x + c); b = -c
  if ((status = apply(C, x, r, 1))) return status;
  lin(C, n, r, -1.0, r, 0.0, nullptr, 0.0, nullptr);
  // ||b||
  if ((status = apply(C, x, t, 0))) return status;      // t = A x
  lin(C, n, s, -1.0, r, -1.0, t, 0.0, nullptr);          // s = -(r + A x) = c
  dot(C, n, s, s, 0);
  dot(C, n, r, r, 1);
  if ((status = fetch(C, h, 2))) return status;
  const double bnorm = sqrt(h[0]) > 0 ? sqrt(h[0]) : 1.0;
  double res = sqrt(h[1]) / bnorm;
  lin(C, n, rh, 1.0, r, 0.0, nullptr, 0.0, nullptr);
  lin(C, n, p, 0.0, nullptr, 0.0, nullptr, 0.0, nullptr);
  lin(C, n, v, 0.0, nullptr, 0.0, nullptr, 0.0, nullptr);
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  int it = 0;
  while (it < max_iter && res > rtol) {
    dot(C, n, rh, r, 0);
    if ((status = fetch(C, h, 1))) return status;
    const double rho_new = h[0];
    if (rho_new == 0.0 || omega == 0.0) break;  // breakdown
    const double beta = (rho_new / rho) * (alpha / omega);
    rho = rho_new;
    lin(C, n, p, 1.0, r, beta, p, -beta * omega, v);  // p = r + beta (p - omega v)
    if ((status = apply(C, p, v, 0))) return status;  // v = A p
    dot(C, n, rh, v, 0);
    if ((status = fetch(C, h, 1))) return status;
    if (h[0] == 0.0) break;
    alpha = rho / h[0];
    lin(C, n, s, 1.0, r, -alpha, v, 0.0, nullptr);    // s = r - alpha v
    if ((status = apply(C, s, t, 0))) return status;  // t = A s
    dot(C, n, t, s, 0);
    dot(C, n, t, t, 1);
    if ((status = fetch(C, h, 2))) return status;
    omega = (h[1] > 0.0) ? h[0] / h[1] : 0.0;
    lin(C, n, x, 1.0, x, alpha, p, omega, s);          // x += alpha p + omega s
    lin(C, n, r, 1.0, s, -omega, t, 0.0, nullptr);     // r = s - omega t
    dot(C, n, r, r, 0);
    if ((status = fetch(C, h, 1))) return status;
    res = sqrt(h[0]) / bnorm;
    ++it;
  }
  // true residual of the float32 iterate
  if ((status = apply(C, x, r, 1))) return status;
  dot(C, n, r, r, 0);
  if ((status = fetch(C, h, 1))) return status;
  *iters = it;
  *relres = sqrt(h[0]) / bnorm;
  return DIPS_OK;
}

}  // namespace dips
