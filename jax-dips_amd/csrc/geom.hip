// Cut-cell geometry, regression rows and record assembly in FP64.
//
// All kernels are per-work-item (one lane per cell, stencil point or training
// point), read callback samples prepared by the host (phi at cell corners and
// on 27-cubes, mu/alpha/beta/f/k/g at the positions the reference uses) and
// compute in double precision; records leave as float32.
//
// Reference (paths relative to the JAX-DIPS repository):
//   crossing test, 5-tet split, Gamma/Omega- pieces, volumes, apertures and the
//   interface integral: jax_dips/geometry/geometric_integrations_per_point.py:32-1009
//   regression rows (pinv with cutoff 10*max(M,N)*eps_f32), normals, zeta/gamma
//   and the ghost-value branches: jax_dips/solvers/poisson/discretization.py:164-519
//   residual assembly (lhs/diag - rhs/diag, nan_to_num, box points):
//   discretization.py:299-423
//
// Volumes and areas use |det M|/6 and |a x b|/2, which equal the reference's
// Gram forms sqrt|det(M M^T)|/6 and sqrt|det(M M^T)|/2 in exact arithmetic and
// keep their relative accuracy for thin sub-cells.
#include <math.h>

#include "dips_common.h"
#include "dips_hip.h"

namespace dips {

constexpr int kGeomThreads = 128;

// P000, P100, P101, P001, P010, P110, P011, P111
__constant__ int kCorner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1},
                                  {-1, 1, -1},  {1, 1, -1},  {-1, 1, 1}, {1, 1, 1}};
__constant__ int kTet[5][4] = {{0, 1, 4, 3}, {5, 1, 4, 7}, {2, 1, 7, 3}, {6, 7, 4, 3}, {7, 1, 4, 3}};
__constant__ int kFaceTets[6][2] = {{0, 3}, {1, 2}, {0, 2}, {1, 3}, {0, 1}, {2, 3}};
__constant__ int kStencil[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 v3(double a, double b, double c) { return V3{a, b, c}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ double comp(const V3& a, int ax) { return ax == 0 ? a.x : (ax == 1 ? a.y : a.z); }

__device__ __forceinline__ double tet_vol(V3 a, V3 b, V3 c, V3 d) {
  V3 u = sub(b, a), v = sub(c, a), w = sub(d, a);
  double det = u.x * (v.y * w.z - v.z * w.y) - u.y * (v.x * w.z - v.z * w.x) + u.z * (v.x * w.y - v.y * w.x);
  return fabs(det) * (1.0 / 6.0);
}

__device__ __forceinline__ double tri_area(V3 a, V3 b, V3 c) {
  V3 u = sub(b, a), v = sub(c, a);
  double cx = u.y * v.z - u.z * v.y, cy = u.z * v.x - u.x * v.z, cz = u.x * v.y - u.y * v.x;
  return 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

// Edge point Q(a,b) = (phi_a S_b - phi_b S_a)/(phi_a - phi_b)
__device__ __forceinline__ V3 edge_pt(V3 Sa, V3 Sb, double pa, double pb) {
  double inv = 1.0 / (pa - pb);
  return v3((pa * Sb.x - pb * Sa.x) * inv, (pa * Sb.y - pb * Sa.y) * inv, (pa * Sb.z - pb * Sa.z) * inv);
}

// stable ascending order of 4 keys
__device__ __forceinline__ void stable_order4(const double k[4], int o[4]) {
  o[0] = 0, o[1] = 1, o[2] = 2, o[3] = 3;
#pragma unroll
  for (int i = 1; i < 4; ++i) {
#pragma unroll
    for (int j = i; j > 0; --j) {
      if (k[o[j]] < k[o[j - 1]]) {
        int t = o[j];
        o[j] = o[j - 1];
        o[j - 1] = t;
      }
    }
  }
}

struct Pieces {
  int ntet;      // number of Omega- sub-tets (0..3)
  V3 om[3][4];   // Omega- sub-tets
  int ntri;      // number of Gamma triangles (0..2)
  V3 tri[2][3];  // Gamma triangles
};

// Omega- and Gamma pieces of one tet (per-point file :52-198)
__device__ void tet_pieces(const V3 S[4], const double ph[4], Pieces& P) {
  int eta = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) eta += (ph[i] < 0.0) ? 1 : 0;
  P.ntet = 0;
  P.ntri = 0;
  if (eta == 0) return;
  if (eta == 4) {
    P.ntet = 1;
#pragma unroll
    for (int i = 0; i < 4; ++i) P.om[0][i] = S[i];
    return;
  }
  int o[4];
  stable_order4(ph, o);
  V3 s0 = S[o[0]], s1 = S[o[1]], s2 = S[o[2]], s3 = S[o[3]];
  double p0 = ph[o[0]], p1 = ph[o[1]], p2 = ph[o[2]], p3 = ph[o[3]];
  if (eta == 1) {
    V3 q01 = edge_pt(s0, s1, p0, p1), q02 = edge_pt(s0, s2, p0, p2), q03 = edge_pt(s0, s3, p0, p3);
    P.ntet = 1;
    P.om[0][0] = s0, P.om[0][1] = q01, P.om[0][2] = q02, P.om[0][3] = q03;
    P.ntri = 1;
    P.tri[0][0] = q01, P.tri[0][1] = q02, P.tri[0][2] = q03;
  } else if (eta == 2) {
    V3 q02 = edge_pt(s0, s2, p0, p2), q03 = edge_pt(s0, s3, p0, p3);
    V3 q12 = edge_pt(s1, s2, p1, p2), q13 = edge_pt(s1, s3, p1, p3);
    P.ntet = 3;
    P.om[0][0] = s0, P.om[0][1] = s1, P.om[0][2] = q02, P.om[0][3] = q13;
    P.om[1][0] = q12, P.om[1][1] = s1, P.om[1][2] = q02, P.om[1][3] = q13;
    P.om[2][0] = s0, P.om[2][1] = q03, P.om[2][2] = q02, P.om[2][3] = q13;
    P.ntri = 2;
    P.tri[0][0] = q02, P.tri[0][1] = q03, P.tri[0][2] = q13;
    P.tri[1][0] = q02, P.tri[1][1] = q12, P.tri[1][2] = q13;
  } else {  // eta == 3
    V3 q13 = edge_pt(s1, s3, p1, p3), q03 = edge_pt(s0, s3, p0, p3), q23 = edge_pt(s2, s3, p2, p3);
    P.ntet = 3;
    P.om[0][0] = s0, P.om[0][1] = s1, P.om[0][2] = s2, P.om[0][3] = q13;
    P.om[1][0] = s0, P.om[1][1] = q03, P.om[1][2] = s2, P.om[1][3] = q13;
    P.om[2][0] = q23, P.om[2][1] = q03, P.om[2][2] = s2, P.om[2][3] = q13;
    // Gamma piece uses the order of -phi (the single positive vertex first)
    double nk[4] = {-ph[0], -ph[1], -ph[2], -ph[3]};
    int on[4];
    stable_order4(nk, on);
    V3 t0 = S[on[0]], t1 = S[on[1]], t2 = S[on[2]], t3 = S[on[3]];
    double n0 = nk[on[0]], n1 = nk[on[1]], n2 = nk[on[2]], n3 = nk[on[3]];
    P.ntri = 1;
    P.tri[0][0] = edge_pt(t0, t1, n0, n1), P.tri[0][1] = edge_pt(t0, t2, n0, n2), P.tri[0][2] = edge_pt(t0, t3, n0, n3);
  }
}

__device__ __forceinline__ bool on_face(double a, double face, double atol) {
  return fabs(a - face) <= atol + 1e-5 * fabs(face);
}

// corner k of the cell centred at c (centre already rounded to float32)
__device__ __forceinline__ V3 corner_pos(const float c[3], const float half[3], int k) {
  return v3((double)(c[0] + (float)kCorner[k][0] * half[0]), (double)(c[1] + (float)kCorner[k][1] * half[1]),
            (double)(c[2] + (float)kCorner[k][2] * half[2]));
}

// ----------------------------------------------------------------------------
// Regions of the 7 stencil cells of every point: 0 crossed, else sign(phi(P000)).
__global__ void regions_kernel(const double* __restrict__ phic, int64_t n7, int8_t* __restrict__ region) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n7) return;
  const double* p = phic + t * 8;
  int eta = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) eta += (p[k] < 0.0) ? 1 : 0;
  int8_t r;
  if (eta > 0 && eta < 8)
    r = 0;
  else
    r = (p[0] > 0.0) ? 1 : ((p[0] < 0.0) ? -1 : 0);
  region[t] = r;
}

// ----------------------------------------------------------------------------
// Full geometry of crossed centre cells.
__global__ void cells_kernel(PointSet ps, float hx, float hy, float hz, const int32_t* __restrict__ cell_pt,
                             int64_t ncell, const double* __restrict__ phic_pt /* (npts,7,8), centre = e 0 */,
                             double* __restrict__ vol_m, double* __restrict__ area_m /* (ncell,6) */,
                             double* __restrict__ tri /* (ncell,10,3,3) */, double* __restrict__ tri_a /* (ncell,10) */) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int pt = cell_pt[c];
  float ctr[3];
  if (!point_xyz(ps, pt, ctr[0], ctr[1], ctr[2])) return;
  const float half[3] = {hx * 0.5f, hy * 0.5f, hz * 0.5f};
  const double* ph = phic_pt + (int64_t)pt * 56;
  V3 C[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) C[q] = corner_pos(ctr, half, q);

  double vm = 0.0;
  double am[6] = {0, 0, 0, 0, 0, 0};
  const double hd[3] = {hx, hy, hz};
  for (int t = 0; t < 5; ++t) {
    V3 S[4];
    double p4[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      S[q] = C[kTet[t][q]];
      p4[q] = ph[kTet[t][q]];
    }
    Pieces P;
    tet_pieces(S, p4, P);
    for (int s = 0; s < P.ntet; ++s) vm += tet_vol(P.om[s][0], P.om[s][1], P.om[s][2], P.om[s][3]);
    // apertures: faces whose tet list contains t
    for (int f = 0; f < 6; ++f) {
      if (kFaceTets[f][0] != t && kFaceTets[f][1] != t) continue;
      const int ax = f >> 1;
      const double face = (double)(ctr[ax] + ((f & 1) ? 1.0f : -1.0f) * half[ax]);
      const double atol = 1e-10 * hd[ax];
      for (int s = 0; s < P.ntet; ++s) {
        int cnt = 0;
        V3 tv[3];
        for (int q = 0; q < 4; ++q) {
          if (on_face(comp(P.om[s][q], ax), face, atol)) {
            if (cnt < 3) tv[cnt] = P.om[s][q];
            ++cnt;
          }
        }
        if (cnt == 3) am[f] += tri_area(tv[0], tv[1], tv[2]);
      }
    }
    // Gamma triangles, 2 slots per tet
    for (int s = 0; s < 2; ++s) {
      double* tp = tri + (c * 10 + t * 2 + s) * 9;
      if (s < P.ntri) {
        const V3* T = P.tri[s];
        for (int q = 0; q < 3; ++q) {
          tp[q * 3 + 0] = T[q].x;
          tp[q * 3 + 1] = T[q].y;
          tp[q * 3 + 2] = T[q].z;
        }
        tri_a[c * 10 + t * 2 + s] = tri_area(T[0], T[1], T[2]);
      } else {
        for (int q = 0; q < 9; ++q) tp[q] = 0.0;
        tri_a[c * 10 + t * 2 + s] = 0.0;
      }
    }
  }
  vol_m[c] = vm;
  for (int f = 0; f < 6; ++f) area_m[c * 6 + f] = am[f];
}

// ----------------------------------------------------------------------------
// Symmetric 3x3 pseudo-inverse by cyclic Jacobi; singular values <= rcond*max are dropped.
__device__ void pinv_sym3(const double Min[3][3], double rcond, double Pi[3][3]) {
  double A[3][3], V[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      A[r][c] = Min[r][c];
      V[r][c] = (r == c) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 32; ++sweep) {
    double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
    double diag = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]);
    if (off <= 1e-18 * diag || off == 0.0) break;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        if (A[p][q] == 0.0) continue;
        double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < 3; ++k) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = cs * akp - sn * akq;
          A[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = cs * apk - sn * aqk;
          A[q][k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = cs * vkp - sn * vkq;
          V[k][q] = sn * vkp + cs * vkq;
        }
      }
    }
  }
  double ev[3] = {A[0][0], A[1][1], A[2][2]};
  double smax = fmax(fabs(ev[0]), fmax(fabs(ev[1]), fabs(ev[2])));
  double cutoff = rcond * smax;
  double inv[3];
  for (int k = 0; k < 3; ++k) inv[k] = (fabs(ev[k]) > cutoff) ? 1.0 / ev[k] : 0.0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += V[r][k] * inv[k] * V[c][k];
      Pi[r][c] = s;
    }
}

// Regression rows and ghost-value forms at crossed stencil points.
// out_a: (n,2,27) coefficients of U_q for u- and u+; out_k: (n,2) multipliers of
// J = alpha(rG) + delta*beta(rG)/mu_sel(rG) in u- and u+; out_sel: 1 -> mu_sel = mu+
// (mu- > mu+ branch), 0 -> mu_sel = mu-; out_r: (n,3) r_Gamma; out_delta: delta.
__global__ void regression_kernel(PointSet ps, float hx, float hy, float hz, const int32_t* __restrict__ cs_pt,
                                  const int8_t* __restrict__ cs_e, int64_t ncs,
                                  const double* __restrict__ phi_cube /* (ncs,27) */,
                                  const double* __restrict__ mu_s /* (ncs,2): mu-, mu+ at s */, double rcond,
                                  double* __restrict__ out_a, double* __restrict__ out_k, int8_t* __restrict__ out_sel,
                                  double* __restrict__ out_r, double* __restrict__ out_delta) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncs) return;
  const int pt = cs_pt[t];
  const int e = cs_e[t];
  float px, py, pz;
  if (e < 0 || e > 6 || !point_xyz(ps, pt, px, py, pz)) return;
  const float s32[3] = {px + (float)kStencil[e][0] * hx, py + (float)kStencil[e][1] * hy,
                        pz + (float)kStencil[e][2] * hz};
  const double h[3] = {hx, hy, hz};
  const double* ph = phi_cube + t * 27;
  // cube offsets X_q (x fastest) in float32 then widened, as get_Xijk builds them
  double X[27][3];
  for (int q = 0; q < 27; ++q) {
    X[q][0] = (double)((float)(q % 3 - 1) * hx);
    X[q][1] = (double)((float)((q / 3) % 3 - 1) * hy);
    X[q][2] = (double)((float)(q / 9 - 1) * hz);
  }
  double D[2][3][27];  // D-, D+
  for (int side = 0; side < 2; ++side) {
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int q = 0; q < 27; ++q) {
      const bool w = side == 0 ? (ph[q] < 0.0) : (ph[q] > 0.0);
      if (!w) continue;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[r][c] += X[q][r] * X[q][c];
    }
    double Pi[3][3];
    pinv_sym3(M, rcond, Pi);
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 27; ++q) {
        const bool w = side == 0 ? (ph[q] < 0.0) : (ph[q] > 0.0);
        double s = 0.0;
        if (w)
          for (int c = 0; c < 3; ++c) s += Pi[r][c] * X[q][c];
        D[side][r][q] = isfinite(s) ? s : 0.0;
      }
  }
  const double gx = (ph[14] - ph[12]) / (2.0 * h[0]);
  const double gy = (ph[16] - ph[10]) / (2.0 * h[1]);
  const double gz = (ph[22] - ph[4]) / (2.0 * h[2]);
  const double nrm = sqrt(gx * gx + gy * gy + gz * gz);
  const double n[3] = {gx / nrm, gy / nrm, gz / nrm};
  const double delta = ph[13];
  const double mum = mu_s[t * 2 + 0], mup = mu_s[t * 2 + 1];
  double zm_q[27], zp_q[27];
  const double fp = ((mup - mum) / mum) * delta;
  const double fm = ((mup - mum) / mup) * delta;
  double zm = 0.0, zp = 0.0;
  for (int q = 0; q < 27; ++q) {
    const double cm = n[0] * D[0][0][q] + n[1] * D[0][1][q] + n[2] * D[0][2][q];
    const double cp = n[0] * D[1][0][q] + n[1] * D[1][1][q] + n[2] * D[1][2][q];
    zm_q[q] = fm * cm;
    zp_q[q] = fp * cp;
    if (q != 13) {
      zm += zm_q[q];
      zp += zp_q[q];
    }
  }
  zm = -zm;
  zp = -zp;
  const bool mbig = mum > mup;
  const bool pos = delta > 0.0;
  double* am = out_a + t * 54;
  double* ap = am + 27;
  for (int q = 0; q < 54; ++q) am[q] = 0.0;
  double km = 0.0, kp = 0.0;
  // extrapolated form: -row.U + (1 - row_s + row_13) U_13 ; identity form: U_13
  if (mbig) {
    if (pos) {  // u- from gamma-, u+ = u_c
      const double den = 1.0 - zm;
      double gsum = 0.0;
      for (int q = 0; q < 27; ++q) {
        const double g = zm_q[q] / den;
        am[q] = -g;
        if (q != 13) gsum += g;
      }
      const double gm = -gsum;
      am[13] += 1.0 - gm + zm_q[13] / den;
      km = -(1.0 - gm);
      ap[13] = 1.0;
    } else {  // u- = u_c, u+ from zeta-
      am[13] = 1.0;
      for (int q = 0; q < 27; ++q) ap[q] = -zm_q[q];
      ap[13] += 1.0 - zm + zm_q[13];
      kp = 1.0;
    }
  } else {
    if (pos) {  // u- from zeta+, u+ = u_c
      for (int q = 0; q < 27; ++q) am[q] = -zp_q[q];
      am[13] += 1.0 - zp + zp_q[13];
      km = -1.0;
      ap[13] = 1.0;
    } else {  // u- = u_c, u+ from gamma+
      const double den = 1.0 + zp;
      double gsum = 0.0;
      for (int q = 0; q < 27; ++q) {
        const double g = zp_q[q] / den;
        ap[q] = -g;
        if (q != 13) gsum += g;
      }
      const double gp = -gsum;
      ap[13] += 1.0 - gp + zp_q[13] / den;
      kp = 1.0 - gp;
      am[13] = 1.0;
    }
  }
  out_k[t * 2 + 0] = km;
  out_k[t * 2 + 1] = kp;
  out_sel[t] = mbig ? 1 : 0;
  out_delta[t] = delta;
  out_r[t * 3 + 0] = (double)s32[0] - delta * n[0];
  out_r[t * 3 + 1] = (double)s32[1] - delta * n[1];
  out_r[t * 3 + 2] = (double)s32[2] - delta * n[2];
}

// ----------------------------------------------------------------------------
// Record assembly.  mode 0: count entries per point into nnz[pt]; mode 1: write
// entries at ent_off[pt] (row order chosen by the host) and row_c.
using AssembleArgs = dips_assemble_t;

__device__ __forceinline__ int offset_code(int ox, int oy, int oz) { return 25 * (oz + 2) + 5 * (oy + 2) + (ox + 2); }

__global__ void assemble_kernel(AssembleArgs A) {
  const int64_t pt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= A.npts) return;
  const PointSet ps{A.xs, A.ys, A.zs, A.ny, A.nz, (int64_t)A.nx * A.ny * A.nz, A.extra, A.extra ? A.nextra : 0};
  float p[3];
  if (!point_xyz(ps, pt, p[0], p[1], p[2])) return;
  const float h[3] = {A.hx, A.hy, A.hz};
  const double vol = (double)A.hx * (double)A.hy * (double)A.hz;
  bool boundary = false;
  for (int a = 0; a < 3; ++a) {
    const float tol = 1e-6f * h[a];
    if (fabsf(p[a] - A.bnd[2 * a]) < tol || fabsf(p[a] - A.bnd[2 * a + 1]) < tol) boundary = true;
  }
  const double* pv = A.pt_vals + pt * 5;
  int64_t wpos = 0;
  if (A.mode == 1) wpos = A.ent_off[pt];
  const int row = (A.mode == 1) ? A.row_of[pt] : 0;

  // the 26 cell coefficients of the point (preconditioner input), for box
  // points too: mu*A/h of the 12 (face, phase) pairs, V-, V+, the 12 face
  // areas (compute_interface_faces / compute_domain_faces,
  // geometric_integrations_per_point.py:873-900, :966-996), stored float32
  if (A.mode == 1 && A.row_coeffs) {
    const int8_t rc0 = A.region[pt * 7];
    const double vol0 = vol;
    const double ff[6] = {(double)A.hy * A.hz, (double)A.hy * A.hz, (double)A.hx * A.hz,
                          (double)A.hx * A.hz, (double)A.hx * A.hy, (double)A.hx * A.hy};
    const double* mf = A.mu_face + pt * 12;
    float* out = A.row_coeffs + (int64_t)row * 26;
    const int s0 = (rc0 == 0) ? A.cc_slot[pt] : -1;
    if (rc0 == 0 && s0 < 0) {
      for (int q = 0; q < 26; ++q) out[q] = __builtin_nanf("");  // crossed cell without geometry
    } else {
      double vm, vp;
      double am[6], ap[6];
      if (rc0 == 0) {
        vm = A.vol_m[s0];
        vp = vol0 - vm;
        if (vp < 0) vp = 0;
        for (int f = 0; f < 6; ++f) {
          am[f] = A.area_m[s0 * 6 + f];
          ap[f] = ff[f] - am[f];
          if (ap[f] < 0) ap[f] = 0;
        }
      } else {
        const double mm = (rc0 == -1) ? 1.0 : 0.0, pm = (rc0 == 1) ? 1.0 : 0.0;
        vm = vol0 * mm;
        vp = vol0 * pm;
        for (int f = 0; f < 6; ++f) {
          am[f] = ff[f] * mm;
          ap[f] = ff[f] * pm;
        }
      }
      for (int f = 0; f < 6; ++f) {
        const double hf = (double)h[f >> 1];
        out[2 * f] = (float)(am[f] * mf[f * 2 + 0] / hf);
        out[2 * f + 1] = (float)(ap[f] * mf[f * 2 + 1] / hf);
        out[14 + 2 * f] = (float)am[f];
        out[15 + 2 * f] = (float)ap[f];
      }
      out[12] = (float)vm;
      out[13] = (float)vp;
    }
  }

  if (boundary) {
    // (u h^3)/h^3 - (g h^3)/h^3 : weight 1 on NN(p), phase from phi(p)
    if (A.mode == 0) {
      A.nnz[pt] = 1;
    } else {
      const double g = pv[4];
      double c = (g * vol) / vol;
      if (!isfinite(c)) c = isnan(c) ? 0.0 : (c > 0 ? 3.4028234663852886e38 : -3.4028234663852886e38);
      A.ent_code[wpos] = (uint32_t)offset_code(0, 0, 0) | ((A.phis[pt * 7] >= 0.0) ? 128u : 0u);
      A.ent_w[wpos] = 1.0f;
      A.row_c[row] = (float)(-c);
    }
    return;
  }

  // centre-cell coefficients
  const int8_t* reg = A.region + pt * 7;
  const int8_t rc = reg[0];
  double cm[6], cp[6], Vm, Vp, ibeta = 0.0;
  const double face_full[6] = {(double)A.hy * A.hz, (double)A.hy * A.hz, (double)A.hx * A.hz,
                               (double)A.hx * A.hz, (double)A.hx * A.hy, (double)A.hx * A.hy};
  const double* muf = A.mu_face + pt * 12;
  if (rc == 0) {
    const int s = A.cc_slot[pt];
    if (s < 0) {  // inconsistent slot table: emit an empty row instead of reading out of bounds
      if (A.mode == 0) A.nnz[pt] = 0;
      return;
    }
    Vm = A.vol_m[s];
    Vp = vol - Vm;
    if (Vp < 0) Vp = 0;
    for (int f = 0; f < 6; ++f) {
      const double amf = A.area_m[s * 6 + f];
      double apf = face_full[f] - amf;
      if (apf < 0) apf = 0;
      const double hf = (double)h[f >> 1];
      cm[f] = amf * muf[f * 2 + 0] / hf;
      cp[f] = apf * muf[f * 2 + 1] / hf;
    }
    // all 10 slots, including zero-area padding triangles at the origin, so a
    // beta that is not finite there poisons the sum exactly as in the reference
    for (int t = 0; t < 10; ++t) {
      const double* b = A.beta_tri + (s * 10 + t) * 3;
      ibeta += A.tri_a[s * 10 + t] * ((b[0] + b[1] + b[2]) / 3.0);
    }
  } else {
    const double mm = (rc == -1) ? 1.0 : 0.0, pm = (rc == 1) ? 1.0 : 0.0;
    Vm = vol * mm;
    Vp = vol * pm;
    for (int f = 0; f < 6; ++f) {
      const double hf = (double)h[f >> 1];
      cm[f] = face_full[f] * (muf[f * 2 + 0] * mm) / hf;
      cp[f] = face_full[f] * (muf[f * 2 + 1] * pm) / hf;
    }
  }
  const double sum_cm = cm[0] + cm[1] + cm[2] + cm[3] + cm[4] + cm[5];
  const double sum_cp = cp[0] + cp[1] + cp[2] + cp[3] + cp[4] + cp[5];
  const double kmv = pv[0], kpv = pv[1];
  const double diag = kpv * Vp + kmv * Vm + sum_cm + sum_cp;
  const double rhs = pv[2] * Vm + pv[3] * Vp + ibeta;
  double rd = rhs / diag;
  if (!isfinite(rd)) rd = isnan(rd) ? 0.0 : (rd > 0 ? 3.4028234663852886e38 : -3.4028234663852886e38);

  // stencil coefficients of u-_e and u+_e
  double coef_m[7], coef_p[7];
  coef_m[0] = kmv * Vm + sum_cm;
  coef_p[0] = kpv * Vp + sum_cp;
  for (int e = 1; e < 7; ++e) {
    coef_m[e] = -cm[e - 1];
    coef_p[e] = -cp[e - 1];
  }
  bool any_cut = false;
  for (int e = 0; e < 7; ++e) any_cut |= (reg[e] == 0);

  if (!any_cut) {
    // 7 bulk entries (zero weights are dropped)
    double w[7];
    bool finite = true;
    for (int e = 0; e < 7; ++e) {
      const double c = (reg[e] == -1) ? coef_m[e] : ((reg[e] == 1) ? coef_p[e] : 0.0);
      w[e] = c / diag;
      finite &= isfinite(w[e]);
    }
    int cnt = 0;
    if (finite) {
      for (int e = 0; e < 7; ++e) {
        if (w[e] == 0.0) continue;
        if (A.mode == 1) {
          const uint32_t code = (uint32_t)offset_code(kStencil[e][0], kStencil[e][1], kStencil[e][2]) |
                                ((A.phis[pt * 7 + e] >= 0.0) ? 128u : 0u);
          A.ent_code[wpos + cnt] = code;
          A.ent_w[wpos + cnt] = (float)w[e];
        }
        ++cnt;
      }
    }
    if (A.mode == 0)
      A.nnz[pt] = cnt;
    else
      A.row_c[row] = (float)(0.0 - rd);
    return;
  }

  // interface point: accumulate 125 offsets x 2 phases in the scratch row
  const int ip = A.ip_slot[pt];
  if (ip < 0) {
    if (A.mode == 0) A.nnz[pt] = 0;
    return;
  }
  double* W = A.scratch + (int64_t)ip * 250;
  if (A.mode == 0) {
    for (int q = 0; q < 250; ++q) W[q] = 0.0;
    double lhs_c = 0.0;
    for (int e = 0; e < 7; ++e) {
      const int r = reg[e];
      if (r != 0) {
        const double c = (r == -1) ? coef_m[e] : coef_p[e];
        const int code = offset_code(kStencil[e][0], kStencil[e][1], kStencil[e][2]);
        const int ph = (A.phis[pt * 7 + e] >= 0.0) ? 1 : 0;
        W[code * 2 + ph] += c;
        continue;
      }
      const int s = A.cs_slot[pt * 7 + e];
      if (s < 0) continue;
      const double* a = A.cs_a + (int64_t)s * 54;
      const double* kk = A.cs_k + (int64_t)s * 2;
      const double* cv = A.cs_vals + (int64_t)s * 4;
      const double dl = A.cs_delta[s];
      const double musel = A.cs_sel[s] ? cv[3] : cv[2];
      const double J = cv[0] + dl * (cv[1] / musel);
      const double* pc = A.phi_cube + (int64_t)s * 27;
      for (int q = 0; q < 27; ++q) {
        const double wq = coef_m[e] * a[q] + coef_p[e] * a[27 + q];
        if (wq == 0.0 && isfinite(wq)) continue;
        const int ox = kStencil[e][0] + (q % 3 - 1), oy = kStencil[e][1] + ((q / 3) % 3 - 1),
                  oz = kStencil[e][2] + (q / 9 - 1);
        const int ph = (pc[q] >= 0.0) ? 1 : 0;
        W[offset_code(ox, oy, oz) * 2 + ph] += wq;
      }
      lhs_c += coef_m[e] * (kk[0] * J) + coef_p[e] * (kk[1] * J);
    }
    bool finite = isfinite(lhs_c / diag);
    for (int q = 0; q < 250; ++q) {
      W[q] = W[q] / diag;
      finite &= isfinite(W[q]);
    }
    int cnt = 0;
    if (finite) {
      for (int q = 0; q < 250; ++q) cnt += (W[q] != 0.0) ? 1 : 0;
    } else {
      for (int q = 0; q < 250; ++q) W[q] = 0.0;
      lhs_c = 0.0;
    }
    A.nnz[pt] = cnt;  // the row constant is recomputed in mode 1 from the same inputs
    return;
  }
  // mode 1: emit entries in (code, phase) order and the row constant
  double lhs_c = 0.0;
  for (int e = 0; e < 7; ++e) {
    if (reg[e] != 0) continue;
    const int s = A.cs_slot[pt * 7 + e];
    if (s < 0) continue;
    const double* kk = A.cs_k + (int64_t)s * 2;
    const double* cv = A.cs_vals + (int64_t)s * 4;
    const double dl = A.cs_delta[s];
    const double musel = A.cs_sel[s] ? cv[3] : cv[2];
    const double J = cv[0] + dl * (cv[1] / musel);
    lhs_c += coef_m[e] * (kk[0] * J) + coef_p[e] * (kk[1] * J);
  }
  bool any = false;
  for (int q = 0; q < 250; ++q) any |= (W[q] != 0.0);
  double lc = lhs_c / diag;
  if (!any || !isfinite(lc)) lc = 0.0;  // dead row: nan_to_num(lhs/diag) = 0
  int cnt = 0;
  for (int q = 0; q < 250; ++q) {
    if (W[q] == 0.0) continue;
    A.ent_code[wpos + cnt] = (uint32_t)(q >> 1) | ((q & 1) ? 128u : 0u);
    A.ent_w[wpos + cnt] = (float)W[q];
    ++cnt;
  }
  A.row_c[row] = (float)(lc - rd);
}

int geom_regions(const double* phic, int64_t n7, int8_t* region, hipStream_t st) {
  if (n7 <= 0) return DIPS_OK;
  const int64_t nb = (n7 + 255) / 256;
  hipLaunchKernelGGL(regions_kernel, dim3((unsigned)nb), dim3(256), 0, st, phic, n7, region);
  return hip_status(hipGetLastError(), "regions_kernel launch");
}

int geom_cells(const dips_grid_t* g, float hx, float hy, float hz, const int32_t* cell_pt, int64_t ncell,
               const double* phic, double* vol_m, double* area_m, double* tri, double* tri_a, hipStream_t st) {
  if (ncell <= 0) return DIPS_OK;
  const int64_t nb = (ncell + kGeomThreads - 1) / kGeomThreads;
  hipLaunchKernelGGL(cells_kernel, dim3((unsigned)nb), dim3(kGeomThreads), 0, st, point_set(g), hx, hy, hz, cell_pt,
                     ncell, phic, vol_m, area_m, tri, tri_a);
  return hip_status(hipGetLastError(), "cells_kernel launch");
}

int geom_regression(const dips_grid_t* g, float hx, float hy, float hz, const int32_t* cs_pt, const int8_t* cs_e,
                    int64_t ncs, const double* phi_cube, const double* mu_s, double rcond, double* out_a,
                    double* out_k, int8_t* out_sel, double* out_r, double* out_delta, hipStream_t st) {
  if (ncs <= 0) return DIPS_OK;
  const int64_t nb = (ncs + 63) / 64;
  hipLaunchKernelGGL(regression_kernel, dim3((unsigned)nb), dim3(64), 0, st, point_set(g), hx, hy, hz, cs_pt, cs_e,
                     ncs, phi_cube, mu_s, rcond, out_a, out_k, out_sel, out_r, out_delta);
  return hip_status(hipGetLastError(), "regression_kernel launch");
}

int geom_assemble(const AssembleArgs& a, hipStream_t st) {
  if (a.npts <= 0) return DIPS_OK;
  const int64_t nb = (a.npts + kGeomThreads - 1) / kGeomThreads;
  hipLaunchKernelGGL(assemble_kernel, dim3((unsigned)nb), dim3(kGeomThreads), 0, st, a);
  return hip_status(hipGetLastError(), "assemble_kernel launch");
}

}  // namespace dips
