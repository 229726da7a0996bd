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
// Layout of the work (one workgroup of four waves per fiber): the block
// products P_a M_ab (A_sq's projected rows) and P_T (diag(xs_c) D_1) (the
// tension part of Q_c) run on the FP64 matrix cores as their transposes,
// M^T P^T, one 16-column tile of M per wave at a time: each lane forms its
// own entry of M on the fly (no LDS round trip), the projection comes from
// LDS, and the transposed output puts 16 consecutive rows of a column of
// A_sq in consecutive lanes, so every store is a run of 128 bytes.  The
// remaining entries (boundary rows, the force rows F_a, Q's identity part)
// are written element by element with consecutive threads on consecutive
// addresses.
#include "common.h"

namespace f3d {

constexpr int kAsmNmax = 64;
constexpr int kAsmThreads = 256;

// Entry (k, jj) of block (a, b) of A before the projection (the formulas of
// the file header); k is the row (node), jj the column within block b.
__device__ __forceinline__ double a_coef(int a, int b, int k, int jj, int N, const double* sxs,
                                         const double* sxss, const double* sxsss,
                                         const double* D01, const double* D02,
                                         const double* D03, const double* D04, double s1,
                                         double s2, double s3, double s4, double E, double dt,
                                         double c0, double c1, double pen) {
  if (a < 3) {
    if (b < 3) {
      const double xa = sxs[k * 3 + a], xb = sxs[k * 3 + b];
      const double g = a == b ? c0 * (1.0 + xa * xa) + c1 * (1.0 - xa * xa)
                              : (c0 - c1) * xa * xb;
      return E * g * s4 * D04[k * N + jj] + ((a == b && k == jj) ? 1.0 / dt : 0.0);
    }
    return -2.0 * c0 * sxs[k * 3 + a] * s1 * D01[k * N + jj] -
           ((k == jj) ? (c0 + c1) * sxss[k * 3 + a] : 0.0);
  }
  if (b < 3)
    return -(c1 + 7.0 * c0) * E * sxss[k * 3 + b] * s4 * D04[k * N + jj] -
           6.0 * c0 * E * sxsss[k * 3 + b] * s3 * D03[k * N + jj] -
           pen * sxs[k * 3 + b] * s1 * D01[k * N + jj];
  const double x0 = sxss[k * 3], x1 = sxss[k * 3 + 1], x2 = sxss[k * 3 + 2];
  return -2.0 * c0 * s2 * D02[k * N + jj] +
         ((k == jj) ? (c1 + c0) * (x0 * x0 + x1 * x1 + x2 * x2) : 0.0);
}

template <int NMAX>
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
  typedef double v4d __attribute__((ext_vector_type(4)));
  constexpr int RT = NMAX / 16;  // row tiles of a projection
  __shared__ double sX[NMAX * 3], sxs[NMAX * 3], sxss[NMAX * 3], sxsss[NMAX * 3];
  // projections transposed, sP[k * NMAX + r] = P(r, k), zero-padded to NMAX
  __shared__ double sPX[NMAX * NMAX], sPT[NMAX * NMAX];
  // the unscaled derivative matrices, the same for every fiber of the
  // launch: in LDS at N <= 32, read through the caches at N <= 64
  constexpr int kLdsD = NMAX <= 32 ? 4 : 0;
  __shared__ double sD[kLdsD > 0 ? kLdsD * NMAX * NMAX : 1];

  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int g = lane >> 4, c16 = lane & 15;
  const int nx = N - 4, nt = N - 2, n4 = 4 * N, nq = n4 - 14;
  for (int e = t; e < kLdsD * N * N; e += kAsmThreads) sD[e] = D0[e];
  for (int e = t; e < NMAX * NMAX; e += kAsmThreads) {
    const int k = e / NMAX, r = e % NMAX;
    sPX[e] = (k < N && r < nx) ? PX[r * N + k] : 0.0;
    sPT[e] = (k < N && r < nt) ? PT[r * N + k] : 0.0;
  }
  const double* __restrict__ D01 = kLdsD ? sD : D0;
  const double* __restrict__ D02 = kLdsD ? sD + N * N : D0 + (long)N * N;
  const double* __restrict__ D03 = kLdsD ? sD + 2 * N * N : D0 + 2l * N * N;
  const double* __restrict__ D04 = kLdsD ? sD + 3 * N * N : D0 + 3l * N * N;
  const int ntj = (N + 15) / 16, nks = (N + 3) / 4;

  for (int f = blockIdx.x; f < nfib; f += gridDim.x) {
    const long nd = node0[f];
    const double sc = 2.0 / length[f];
    const double s1 = sc, s2 = sc * sc, s3 = s2 * sc, s4 = s2 * s2;
    __syncthreads();  // LDS reuse across fibers (and the set-up above)
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

    double* __restrict__ Af = Aout + a_off[f];
    // ---- projected products, work item (product p, column tile jt):
    // p = 4 a + b < 16 is block (a, b) of A_sq's projected rows; p = 16 + c
    // is the tension part of Q_c.
    for (int it = wv; it < 19 * ntj; it += 4) {
      const int p = it / ntj, jt = it - p * ntj;
      const int a = p < 16 ? p >> 2 : 4 + (p - 16), b = p < 16 ? p & 3 : 0;
      const bool tens = a >= 3;
      const int rows = tens ? nt : nx;
      const double* __restrict__ sP = tens ? sPT : sPX;
      const int jj = 16 * jt + c16;
      v4d acc[RT];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) acc[rt] = v4d{0.0, 0.0, 0.0, 0.0};
      for (int st = 0; st < nks; ++st) {
        const int k = 4 * st + g;
        double mv = 0.0;
        if (k < N && jj < N) {
          mv = a < 4 ? a_coef(a, b, k, jj, N, sxs, sxss, sxsss, D01, D02, D03, D04, s1, s2, s3,
                              s4, E, dt, c0, c1, pen)
                     : sxs[k * 3 + (a - 4)] * s1 * D01[k * N + jj];
        }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          if (16 * rt < rows) {
            const double pv = sP[k * NMAX + 16 * rt + c16];
            acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(mv, pv, acc[rt], 0, 0, 0);
          }
        }
      }
      // element e of lane (g, c16): row 16 rt + c16 of column 16 jt + 4 e + g
      double* __restrict__ dst;
      long col_stride;
      if (a < 4) {
        dst = Af + (long)b * N * n4 + (a < 3 ? a * nx : 3 * nx);
        col_stride = n4;
      } else {
        double* __restrict__ Q = a == 4 ? Q0 : (a == 5 ? Q1 : Q2);
        dst = Q + q_off[3 * f + (a - 4)] + 3 * nx;
        col_stride = nq;
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const int r = 16 * rt + c16;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int jc = 16 * jt + 4 * e + g;
          if (r < rows && jc < N) dst[(long)jc * col_stride + r] = acc[rt][e];
        }
      }
    }

    // ---- 14 boundary rows of every column (force- and torque-free ends) --
    for (int e = t; e < n4 * 14; e += kAsmThreads) {
      const int j = e / 14, r = e % 14, b = j / N, jj = j % N;
      double v = 0.0;
      if (b < 3) {
        if (r == b) v = E * s3 * D03[jj];
        else if (r == 3) v = -E * s2 * D02[jj] * sxss[b];
        else if (r == 4 + b) v = s2 * D02[jj];
        else if (r == 7 + b) v = -E * s3 * D03[(N - 1) * N + jj];
        else if (r == 10) v = E * s2 * D02[(N - 1) * N + jj] * sxss[(N - 1) * 3 + b];
        else if (r == 11 + b) v = s2 * D02[(N - 1) * N + jj];
      } else {
        if (jj == 0 && r < 3) v = -sxs[r];
        else if (jj == 0 && r == 3) v = -1.0;
        else if (jj == N - 1 && r >= 7 && r < 10) v = sxs[(N - 1) * 3 + (r - 7)];
        else if (jj == N - 1 && r == 10) v = 1.0;
      }
      Af[(long)j * n4 + (n4 - 14) + r] = v;
    }
    // ---- force operator: three row blocks N x 4N, column-major, adjacent --
    double* __restrict__ Fb = Fout + f_off[f];
    for (int e = t; e < 3 * n4 * N; e += kAsmThreads) {
      const int i = e % N, j = (e / N) % n4, a = e / (N * n4);
      const int b = j / N, jj = j % N;
      double v = 0.0;
      if (b == a) v = -E * s4 * D04[i * N + jj];
      else if (b == 3) v = sxs[i * 3 + a] * s1 * D01[i * N + jj] + (i == jj ? sxss[i * 3 + a] : 0.0);
      Fb[e] = v;
    }
    // ---- identity part of Q_c: rows [0, 3 nx), P_X in row block c --------
    for (int e = t; e < 3 * N * 3 * nx; e += kAsmThreads) {
      const int r = e % (3 * nx), jj = (e / (3 * nx)) % N, c = e / (3 * nx * N);
      double* __restrict__ Q = c == 0 ? Q0 : (c == 1 ? Q1 : Q2);
      Q[q_off[3 * f + c] + (long)jj * nq + r] = (r / nx == c) ? PX[(r % nx) * N + jj] : 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Per-fiber right-hand side with the plus-end force f = F_end x_s(L) at the
// last node (zero elsewhere) and the explicit flow v at the nodes:
//   R_a = X_a / dt + v_a + c0 (f_a + x_s,a (x_s . f)) + c1 (f_a - x_s,a (x_s . f))
//   R_T = -p + x_s . (D_1 v) + 2 c0 x_s . (D_1 f) + (c0 - c1) x_ss . f
//   rhs = [P R ; 14 end rows: 0 ... F_end at rows 7..9, F_end . x_s(L) at 10]
// One workgroup per fiber: R in LDS, then one thread per output row.
constexpr int kRhsNmax = 64;

__global__ void __launch_bounds__(256)
fiber_rhs_kernel(int nfib, int N, const long* __restrict__ node0,
                 const double* __restrict__ length, const long* __restrict__ rhs_off,
                 const double* __restrict__ X, const double* __restrict__ xs,
                 const double* __restrict__ xss, const double* __restrict__ v,
                 const double* __restrict__ D0, const double* __restrict__ PX,
                 const double* __restrict__ PT, double dt, double c0, double c1, double pen,
                 double end_force, double* __restrict__ rhs) {
  __shared__ double sX[kRhsNmax * 3], sxs[kRhsNmax * 3], sxss[kRhsNmax * 3], sv[kRhsNmax * 3];
  __shared__ double sR[4 * kRhsNmax];
  const int t = threadIdx.x;
  const int nx = N - 4, nt = N - 2;
  const double* __restrict__ D01 = D0;  // unscaled first derivative, row-major
  for (int f = blockIdx.x; f < nfib; f += gridDim.x) {
    const long nd = node0[f];
    const double s1 = 2.0 / length[f];
    __syncthreads();  // LDS reuse across fibers
    if (t < 3 * N) {
      sX[t] = X[nd * 3 + t];
      sxs[t] = xs[nd * 3 + t];
      sxss[t] = xss[nd * 3 + t];
      sv[t] = v[nd * 3 + t];
    }
    __syncthreads();
    const int L = N - 1;
    const double fl0 = end_force * sxs[L * 3], fl1 = end_force * sxs[L * 3 + 1],
                 fl2 = end_force * sxs[L * 3 + 2];
    if (t < 3 * N) {
      const int a = t / N, k = t % N;
      const bool last = k == L;
      const double xa = sxs[k * 3 + a];
      const double fa = last ? (a == 0 ? fl0 : (a == 1 ? fl1 : fl2)) : 0.0;
      const double xf = last ? sxs[k * 3] * fl0 + sxs[k * 3 + 1] * fl1 + sxs[k * 3 + 2] * fl2
                             : 0.0;
      sR[a * N + k] = sX[k * 3 + a] / dt + sv[k * 3 + a] + c0 * (fa + xa * xf) +
                      c1 * (fa - xa * xf);
    } else if (t < 4 * N) {
      const int k = t - 3 * N;
      double dv = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double d = 0.0;
        for (int q = 0; q < N; ++q) d = fma(D01[k * N + q], sv[q * 3 + a], d);
        dv = fma(sxs[k * 3 + a], s1 * d, dv);
      }
      const double d1l = s1 * D01[k * N + L];
      const double df = (sxs[k * 3] * fl0 + sxs[k * 3 + 1] * fl1 + sxs[k * 3 + 2] * fl2) * d1l;
      const double xssf = k == L ? sxss[k * 3] * fl0 + sxss[k * 3 + 1] * fl1 +
                                       sxss[k * 3 + 2] * fl2
                                 : 0.0;
      sR[3 * N + k] = -pen + dv + 2.0 * c0 * df + (c0 - c1) * xssf;
    }
    __syncthreads();
    double* __restrict__ out = rhs + rhs_off[f];
    for (int r = t; r < 4 * N; r += 256) {
      double val = 0.0;
      if (r < 3 * nx) {
        const int a = r / nx, i = r - a * nx;
        for (int k = 0; k < N; ++k) val = fma(PX[i * N + k], sR[a * N + k], val);
      } else if (r < 3 * nx + nt) {
        const int i = r - 3 * nx;
        for (int k = 0; k < N; ++k) val = fma(PT[i * N + k], sR[3 * N + k], val);
      } else {
        const int br = r - (4 * N - 14);
        if (br == 7) val = fl0;
        else if (br == 8) val = fl1;
        else if (br == 9) val = fl2;
        else if (br == 10) val = fl0 * sxs[L * 3] + fl1 * sxs[L * 3 + 1] + fl2 * sxs[L * 3 + 2];
      }
      out[r] = val;
    }
  }
}

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
#define F3D_ASM(NM)                                                                            \
  F3D_LAUNCH(fiber_assemble_kernel<NM>, dim3(grid), dim3(kAsmThreads), 0, st, (int)nfib, (int)N, \
             (const long*)node0, length, (const long*)a_off, (const long*)f_off,                \
             (const long*)q_off, X, D0, PX, PT, w0, E, dt, c0, c1, penalty, A_out, F_out, Q0, Q1, \
             Q2, xs_out, xss_out, w_out)
  if (N <= 32) F3D_ASM(32);
  else F3D_ASM(64);
#undef F3D_ASM
  return F3D_OK;
}

extern "C" int f3d_fiber_rhs(int64_t nfib, int32_t N, const int64_t* node0, const double* length,
                             const int64_t* rhs_off, const double* X, const double* xs,
                             const double* xss, const double* v, const double* D0,
                             const double* PX, const double* PT, double dt, double c0, double c1,
                             double penalty, double end_force, double* rhs, void* stream) {
  F3D_REQUIRE(N >= 5 && N <= kRhsNmax, F3D_EINVAL, "fiber_rhs: N=%d outside [5,%d]", N,
              kRhsNmax);
  F3D_REQUIRE(nfib >= 0 && nfib < (1ll << 31), F3D_EINVAL, "fiber_rhs: bad fiber count");
  F3D_REQUIRE(dt > 0.0, F3D_EINVAL, "fiber_rhs: dt must be positive");
  if (nfib == 0) return F3D_OK;
  F3D_REQUIRE(node0 && length && rhs_off && X && xs && xss && v && D0 && PX && PT && rhs,
              F3D_EINVAL, "fiber_rhs: NULL argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const unsigned grid = (unsigned)(nfib < 4096 ? nfib : 4096);
  F3D_LAUNCH(fiber_rhs_kernel, dim3(grid), dim3(256), 0, st, (int)nfib, (int)N,
             (const long*)node0, length, (const long*)rhs_off, X, xs, xss, v, D0, PX, PT, dt, c0,
             c1, penalty, end_force, rhs);
  return F3D_OK;
}
