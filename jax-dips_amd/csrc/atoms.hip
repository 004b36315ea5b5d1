// Molecular level set and charge density: the callbacks of the Poisson-Boltzmann
// problems (union of atomic spheres, Gaussian-smeared point charges).
//
// The reference evaluates the level set phi(x) = min_i |x - c_i| - r_i by a scan
// over every atom per query (examples/biomolecules/geometry.py:10-39).  Here the
// atoms are binned in a cell list (cells of size >= cap + max radius) and a query
// reads only the atoms of its 27 neighbouring cells: min over them, capped at cap,
// equals the global minimum wherever that is below cap (every other atom is at
// least cap away from its surface term).  The charge density sums
// q_i exp(-|x - c_i|^2 / (2 sigma^2)) / (2 pi sigma^2)^1.5 over the same atoms.
// One thread per query, float64; the atom table stays in the caches.
#include <math.h>

#include "dips_common.h"
#include "dips_hip.h"

namespace dips {

#pragma clang fp contract(off)

__global__ __launch_bounds__(256) void atoms_kernel(const double* __restrict__ atoms, const int32_t* __restrict__ cstart,
                                                    const int32_t* __restrict__ catoms, int nc, double lo, double cell,
                                                    double cap, double inv2s2, double norm,
                                                    const double* __restrict__ pos, int64_t m, double* __restrict__ phi,
                                                    double* __restrict__ chg) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const double x = pos[3 * t], y = pos[3 * t + 1], z = pos[3 * t + 2];
  auto cell_of = [&](double v) {
    const double c = floor((v - lo) / cell);
    return (int)fmin(fmax(c, 0.0), (double)(nc - 1));
  };
  const int cx = cell_of(x), cy = cell_of(y), cz = cell_of(z);
  double best = INFINITY, q = 0.0;
  for (int dx = -1; dx <= 1; ++dx) {
    const int ix = cx + dx;
    if (ix < 0 || ix >= nc) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int iy = cy + dy;
      if (iy < 0 || iy >= nc) continue;
      for (int dz = -1; dz <= 1; ++dz) {
        const int iz = cz + dz;
        if (iz < 0 || iz >= nc) continue;
        const int c = (ix * nc + iy) * nc + iz;
        for (int k = cstart[c]; k < cstart[c + 1]; ++k) {
          const double* a = atoms + 5 * (int64_t)catoms[k];
          const double ddx = a[0] - x, ddy = a[1] - y, ddz = a[2] - z;
          const double d2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
          best = fmin(best, sqrt(d2) - a[3]);
          if (chg) q += a[4] * exp(-d2 * inv2s2);
        }
      }
    }
  }
  if (phi) phi[t] = fmin(best, cap);
  if (chg) chg[t] = q * norm;
}

int atoms_launch(const double* atoms, int32_t natoms, const int32_t* cell_start, const int32_t* cell_atoms, int32_t nc,
                 double lo, double cell, double cap, double sigma, const double* pos, int64_t m, double* phi,
                 double* charge, hipStream_t st) {
  if (natoms < 0 || nc < 1 || m < 0 || !(cell > 0.0) || !(sigma > 0.0) || (m > 0 && !pos) ||
      (natoms > 0 && (!atoms || !cell_start || !cell_atoms)) || (!phi && !charge)) {
    set_error("dips_atoms_eval: invalid arguments");
    return DIPS_EINVAL;
  }
  if (m == 0) return DIPS_OK;
  const double s2 = sigma * sigma;
  const double norm = 1.0 / pow(2.0 * M_PI * s2, 1.5);
  const int64_t nb = (m + 255) / 256;
  hipLaunchKernelGGL(atoms_kernel, dim3((unsigned)nb), dim3(256), 0, st, atoms, cell_start, cell_atoms, nc, lo, cell,
                     cap, 1.0 / (2.0 * s2), norm, pos, m, phi, charge);
  return hip_status(hipGetLastError(), "atoms_kernel launch");
}

}  // namespace dips
