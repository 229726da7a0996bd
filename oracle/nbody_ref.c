/*
 * CPU restatement of the reference's numba pair sums, used as the CPU
 * baseline and as a second oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * Loop order and arithmetic follow python/kernels/kernels.py:
 *   oseen_kernel_source_target_numba      :794-846
 *   stresslet_kernel_source_target_numba  :1186-1217
 * i.e. @njit(parallel=True, fastmath=True): a parallel loop over targets
 * (prange), a serial loop over sources, FP64, AoS (N,3) arrays, and the
 * coincident-point regularisation r -> sqrt(r^2 + reg^2) for r < eps.
 * Built with -O3 -ffast-math -fopenmp to mirror fastmath/parallel.
 */
#include <math.h>
#include <stdint.h>

void nbody_oseen_f64(int64_t ns, const double* rs, int64_t nt, const double* rt,
                     const double* f, double eta, double reg, double eps, double* u) {
  const double factor = 1.0 / (8.0 * M_PI * eta);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; ++i) {
    double ux = 0.0, uy = 0.0, uz = 0.0;
    const double tx = rt[3 * i], ty = rt[3 * i + 1], tz = rt[3 * i + 2];
    for (int64_t j = 0; j < ns; ++j) {
      const double x = tx - rs[3 * j], y = ty - rs[3 * j + 1], z = tz - rs[3 * j + 2];
      double r = sqrt(x * x + y * y + z * z);
      if (r < eps) r = sqrt(r * r + reg * reg);
      const double fr = factor / r;
      const double gr = factor / (r * r * r);
      const double fx = f[3 * j], fy = f[3 * j + 1], fz = f[3 * j + 2];
      const double Mxx = fr + gr * x * x, Mxy = gr * x * y, Mxz = gr * x * z;
      const double Myy = fr + gr * y * y, Myz = gr * y * z, Mzz = fr + gr * z * z;
      ux += Mxx * fx + Mxy * fy + Mxz * fz;
      uy += Mxy * fx + Myy * fy + Myz * fz;
      uz += Mxz * fx + Myz * fy + Mzz * fz;
    }
    u[3 * i] = ux;
    u[3 * i + 1] = uy;
    u[3 * i + 2] = uz;
  }
}

void nbody_stresslet_f64(int64_t ns, const double* rs, int64_t nt, const double* rt,
                         const double* nrm, const double* q, double eta, double reg, double eps,
                         double* u) {
  const double factor = -3.0 / (4.0 * M_PI * eta);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; ++i) {
    double ux = 0.0, uy = 0.0, uz = 0.0;
    const double tx = rt[3 * i], ty = rt[3 * i + 1], tz = rt[3 * i + 2];
    for (int64_t j = 0; j < ns; ++j) {
      const double x = tx - rs[3 * j], y = ty - rs[3 * j + 1], z = tz - rs[3 * j + 2];
      double r = sqrt(x * x + y * y + z * z);
      if (r < eps) r = sqrt(r * r + reg * reg);
      const double r2 = r * r;
      const double ri5 = 1.0 / (r2 * r2 * r);
      const double f0 = factor * (x * q[3 * j] + y * q[3 * j + 1] + z * q[3 * j + 2]) *
                        (x * nrm[3 * j] + y * nrm[3 * j + 1] + z * nrm[3 * j + 2]) * ri5;
      ux += f0 * x;
      uy += f0 * y;
      uz += f0 * z;
    }
    u[3 * i] = ux;
    u[3 * i + 1] = uy;
    u[3 * i + 2] = uz;
  }
}
