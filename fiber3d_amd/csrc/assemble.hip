// Batched per-fiber assembly of the implicit fiber system (FP64), one
// 256-thread workgroup per fiber, for fibers of one node count N (5..64).
//
// Reference: fiber.form_linear_operator (python/fiber/fiber.py:322-447),
// apply_BC_rectangular with force- and torque-free ends (:810-961),
// force_operator (:1399-1431) and the tension projection of
// tstep_utils.py:1238-1336; derivative matrices of
// tstep/fiber_matrices.py:38-153 scaled by (2/L)^k.
//
// With D_k = (2/L)^k D0_k, xs = D_1 X, xss = D_2 X, xsss = D_3 X and the
// block-diagonal projection P = diag(PX, PX, PX, PT):
//   A (4N x 4N), blocks a, b in {x, y, z, T}:
//     A_ab (a, b < 3) = delta_ab I/dt + E diag(g_ab) D_4,
//        g_aa = c0 (1 + xs_a^2) + c1 (1 - xs_a^2),  g_ab = (c0 - c1) xs_a xs_b
//     A_aT = -2 c0 diag(xs_a) D_1 - (c0 + c1) diag(xss_a)
//     A_Tb = -(c1 + 7 c0) E diag(xss_b) D_4 - 6 c0 E diag(xsss_b) D_3
//            - p diag(xs_b) D_1
//     A_TT = -2 c0 D_2 + (c0 + c1) diag(|xss|^2)
//   A_sq = [P A ; 14 boundary rows]            (column-major, 4N x 4N)
//   F_a  = [-E D_4 in block a | diag(xss_a) + diag(xs_a) D_1 in block T]
//   Q_c  = P [e_c ; diag(xs_c) D_1]            ((4N - 14) x N)
// Thread j owns column j of A_sq and of the three F_a; it forms the column
// m_k of each A block on the fly and accumulates P m with the projection
// rows broadcast from LDS.
#include "common.h"

namespace f3d {

constexpr int kAsmNmax = 64;
constexpr int kAsmThreads = 256;

// Coefficient of column jj, node k, of block (a, b) of A (before the
// projection): the formulas of the file header.
template <int A_BLK>
__device__ __forceinline__ double a_coef(int b, int k, int jj, int N, const double* sxs,
                                         const double* sxss, const double* sxsss,
                                         const double* D01, const double* D02,
                                         const double* D03, const double* D04, double s1,
                                         double s2, double s3, double s4, double E, double dt,
                                         double c0, double c1, double pen) {
  if (A_BLK < 3) {
    if (b < 3) {
      const double xa = sxs[k * 3 + A_BLK], xb = sxs[k * 3 + b];
      const double g = A_BLK == b ? c0 * (1.0 + xa * xa) + c1 * (1.0 - xa * xa)
                                  : (c0 - c1) * xa * xb;
      return E * g * s4 * D04[k * N + jj] + ((A_BLK == b && k == jj) ? 1.0 / dt : 0.0);
    }
    return -2.0 * c0 * sxs[k * 3 + A_BLK] * s1 * D01[k * N + jj] -
           ((k == jj) ? (c0 + c1) * sxss[k * 3 + A_BLK] : 0.0);
  }
  if (b < 3)
    return -(c1 + 7.0 * c0) * E * sxss[k * 3 + b] * s4 * D04[k * N + jj] -
           6.0 * c0 * E * sxsss[k * 3 + b] * s3 * D03[k * N + jj] -
           pen * sxs[k * 3 + b] * s1 * D01[k * N + jj];
  const double x0 = sxss[k * 3], x1 = sxss[k * 3 + 1], x2 = sxss[k * 3 + 2];
  return -2.0 * c0 * s2 * D02[k * N + jj] +
         ((k == jj) ? (c1 + c0) * (x0 * x0 + x1 * x1 + x2 * x2) : 0.0);
}

// acc[i] = sum_k Pt[k][i] m_k over 64 (zero-padded) projection rows; the
// transposed projection is read two rows at a time from LDS.
#define F3D_ACC_ROWS(Pt, MEXPR)                                                  \
  _Pragma("unroll") for (int i = 0; i < kAsmNmax; ++i) acc[i] = 0.0;           \
  for (int k = 0; k < N; ++k) {                                                \
    const double m = (MEXPR);                                                  \
    const double2* __restrict__ p2 = reinterpret_cast<const double2*>((Pt) + k * kAsmNmax); \
    _Pragma("unroll") for (int i2 = 0; i2 < kAsmNmax / 2; ++i2) {              \
      const double2 pp = p2[i2];                                               \
      acc[2 * i2] = fma(pp.x, m, acc[2 * i2]);                                 \
      acc[2 * i2 + 1] = fma(pp.y, m, acc[2 * i2 + 1]);                         \
    }                                                                          \
  }

template <int A_BLK>
__device__ __forceinline__ void a_block(double (&acc)[kAsmNmax], const double* sPt, int b, int jj,
                                        int N, int rows, double* __restrict__ Acol, int r0,
                                        const double* sxs, const double* sxss,
                                        const double* sxsss, const double* D01,
                                        const double* D02, const double* D03,
                                        const double* D04, double s1, double s2, double s3,
                                        double s4, double E, double dt, double c0, double c1,
                                        double pen) {
  F3D_ACC_ROWS(sPt, (a_coef<A_BLK>(b, k, jj, N, sxs, sxss, sxsss, D01, D02, D03, D04, s1, s2,
                                   s3, s4, E, dt, c0, c1, pen)))
#pragma unroll
  for (int i = 0; i < kAsmNmax; ++i)
    if (i < rows) Acol[r0 + i] = acc[i];
}

__global__ void __launch_bounds__(kAsmThreads)
fiber_assemble_kernel(int nfib, int N, const long* __restrict__ node0,
                      const double* __restrict__ length, const long* __restrict__ a_off,
                      const long* __restrict__ f_off, const long* __restrict__ q_off,
                      const double* __restrict__ X, const double* __restrict__ D0,
                      const double* __restrict__ PX, const double* __restrict__ PT,
                      const double* __restrict__ w0, double E, double dt, double c0, double c1,
                      double pen, double* __restrict__ Aout, double* __restrict__ Fout,
                      double* __restrict__ Q0, double* __restrict__ Q1, double* __restrict__ Q2,
                      double* __restrict__ xs_out, double* __restrict__ xss_out,
                      double* __restrict__ w_out) {
  // projections transposed ([k][i]) and zero-padded to 64 rows
  __shared__ __align__(16) double sPXt[kAsmNmax * kAsmNmax];
  __shared__ __align__(16) double sPTt[kAsmNmax * kAsmNmax];
  __shared__ double sX[kAsmNmax * 3], sxs[kAsmNmax * 3], sxss[kAsmNmax * 3], sxsss[kAsmNmax * 3];

  const int t = threadIdx.x;
  const int nx = N - 4, nt = N - 2, n4 = 4 * N;
  for (int e = t; e < kAsmNmax * kAsmNmax; e += kAsmThreads) {
    const int k = e / kAsmNmax, i = e % kAsmNmax;
    sPXt[e] = (k < N && i < nx) ? PX[i * N + k] : 0.0;
    sPTt[e] = (k < N && i < nt) ? PT[i * N + k] : 0.0;
  }
  const double* __restrict__ D01 = D0;
  const double* __restrict__ D02 = D0 + (long)N * N;
  const double* __restrict__ D03 = D0 + 2l * N * N;
  const double* __restrict__ D04 = D0 + 3l * N * N;

  for (int f = blockIdx.x; f < nfib; f += gridDim.x) {
    const long nd = node0[f];
    const double sc = 2.0 / length[f];
    const double s1 = sc, s2 = sc * sc, s3 = s2 * sc, s4 = s2 * s2;
    __syncthreads();  // LDS reuse across fibers (and the projection load)
    if (t < 3 * N) sX[t] = X[nd * 3 + t];
    __syncthreads();
    // derivatives of the centre line: thread (k, c) for k < N, c < 3
    if (t < 3 * N) {
      const int k = t / 3, c = t % 3;
      double d1 = 0.0, d2 = 0.0, d3 = 0.0;
      for (int q = 0; q < N; ++q) {
        const double xq = sX[q * 3 + c];
        d1 = fma(D01[k * N + q], xq, d1);
        d2 = fma(D02[k * N + q], xq, d2);
        d3 = fma(D03[k * N + q], xq, d3);
      }
      sxs[t] = s1 * d1;
      sxss[t] = s2 * d2;
      sxsss[t] = s3 * d3;
      xs_out[nd * 3 + t] = s1 * d1;
      xss_out[nd * 3 + t] = s2 * d2;
    }
    if (t < N) w_out[nd + t] = w0[t] * (0.5 * length[f]);
    __syncthreads();

    double acc[kAsmNmax];
    // ---- A_sq and F, column j -------------------------------------------
    if (t < n4) {
      const int j = t, b = j / N, jj = j % N;
      double* __restrict__ Acol = Aout + a_off[f] + (long)j * n4;
#define F3D_ARGS sxs, sxss, sxsss, D01, D02, D03, D04, s1, s2, s3, s4, E, dt, c0, c1, pen
      a_block<0>(acc, sPXt, b, jj, N, nx, Acol, 0, F3D_ARGS);
      a_block<1>(acc, sPXt, b, jj, N, nx, Acol, nx, F3D_ARGS);
      a_block<2>(acc, sPXt, b, jj, N, nx, Acol, 2 * nx, F3D_ARGS);
      a_block<3>(acc, sPTt, b, jj, N, nt, Acol, 3 * nx, F3D_ARGS);
#undef F3D_ARGS
      // 14 boundary rows (force- and torque-free ends)
      double br[14];
#pragma unroll
      for (int r = 0; r < 14; ++r) br[r] = 0.0;
      if (b < 3) {
        const double d3f = s3 * D03[jj], d2f = s2 * D02[jj];
        const double d3l = s3 * D03[(N - 1) * N + jj], d2l = s2 * D02[(N - 1) * N + jj];
        br[b] = E * d3f;
        br[3] = -E * d2f * sxss[b];
        br[4 + b] = d2f;
        br[7 + b] = -E * d3l;
        br[10] = E * d2l * sxss[(N - 1) * 3 + b];
        br[11 + b] = d2l;
      } else {
        if (jj == 0) {
          br[0] = -sxs[0];
          br[1] = -sxs[1];
          br[2] = -sxs[2];
          br[3] = -1.0;
        }
        if (jj == N - 1) {
          br[7] = sxs[(N - 1) * 3];
          br[8] = sxs[(N - 1) * 3 + 1];
          br[9] = sxs[(N - 1) * 3 + 2];
          br[10] = 1.0;
        }
      }
#pragma unroll
      for (int r = 0; r < 14; ++r) Acol[n4 - 14 + r] = br[r];

      // force operator: three row blocks N x 4N, column-major, adjacent
      double* __restrict__ Fb = Fout + f_off[f];
      for (int a = 0; a < 3; ++a) {
        double* __restrict__ Fcol = Fb + (long)a * N * n4 + (long)j * N;
        if (b == a) {
          for (int i = 0; i < N; ++i) Fcol[i] = -E * s4 * D04[i * N + jj];
        } else if (b == 3) {
          for (int i = 0; i < N; ++i)
            Fcol[i] = sxs[i * 3 + a] * s1 * D01[i * N + jj] + (i == jj ? sxss[i * 3 + a] : 0.0);
        } else {
          for (int i = 0; i < N; ++i) Fcol[i] = 0.0;
        }
      }
    }
    // ---- tension projections Q_c = P [e_c ; diag(xs_c) D_1], column jj ----
    if (t < 3 * N) {
      const int c = t / N, jj = t % N;
      double* __restrict__ Q = c == 0 ? Q0 : (c == 1 ? Q1 : Q2);
      double* __restrict__ Qcol = Q + q_off[3 * f + c] + (long)jj * (n4 - 14);
      for (int r = 0; r < 3 * nx; ++r) Qcol[r] = 0.0;
      for (int i = 0; i < nx; ++i) Qcol[c * nx + i] = sPXt[jj * kAsmNmax + i];
      F3D_ACC_ROWS(sPTt, sxs[k * 3 + c] * s1 * D01[k * N + jj])
#pragma unroll
      for (int i = 0; i < kAsmNmax; ++i)
        if (i < nt) Qcol[3 * nx + i] = acc[i];
    }
  }
}
#undef F3D_ACC_ROWS

}  // namespace f3d

using namespace f3d;

extern "C" int f3d_fiber_assemble(int64_t nfib, int32_t N, const int64_t* node0,
                                  const double* length, const int64_t* a_off,
                                  const int64_t* f_off, const int64_t* q_off, const double* X,
                                  const double* D0, const double* PX, const double* PT,
                                  const double* w0, double E, double dt, double c0, double c1,
                                  double penalty, double* A_out, double* F_out, double* Q0,
                                  double* Q1, double* Q2, double* xs_out, double* xss_out,
                                  double* w_out, void* stream) {
  F3D_REQUIRE(N >= 5 && N <= kAsmNmax, F3D_EINVAL, "fiber_assemble: N=%d outside [5,%d]", N,
              kAsmNmax);
  F3D_REQUIRE(nfib >= 0 && nfib < (1ll << 31), F3D_EINVAL, "fiber_assemble: bad fiber count");
  F3D_REQUIRE(dt > 0.0, F3D_EINVAL, "fiber_assemble: dt must be positive");
  if (nfib == 0) return F3D_OK;
  F3D_REQUIRE(node0 && length && a_off && f_off && q_off && X && D0 && PX && PT && w0 && A_out &&
                  F_out && Q0 && Q1 && Q2 && xs_out && xss_out && w_out,
              F3D_EINVAL, "fiber_assemble: NULL argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const unsigned grid = (unsigned)(nfib < 4096 ? nfib : 4096);
  F3D_LAUNCH(fiber_assemble_kernel, dim3(grid), dim3(kAsmThreads), 0, st, (int)nfib, (int)N,
             (const long*)node0, length, (const long*)a_off, (const long*)f_off,
             (const long*)q_off, X, D0, PX, PT, w0, E, dt, c0, c1, penalty, A_out, F_out, Q0, Q1,
             Q2, xs_out, xss_out, w_out);
  return F3D_OK;
}
