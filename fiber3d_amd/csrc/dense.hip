// Batched dense kernels over per-fiber blocks of uneven order (n <= 256):
//   * LU factorisation with partial pivoting (LAPACK dgetrf semantics):
//     FP64 register/LDS-resident, one persistent 1024-thread workgroup per
//     block; equilibrated FP32 blocked on MFMA (lu_factor_f32_mfma_kernel
//     below), two 256-thread workgroups per CU;
//   * forward/backward substitution with optional FP64 iterative refinement,
//     one wavefront per block;
//   * column-major block GEMV, one wavefront per block.
//
// Reference: scipy.linalg.lu_factor / lu_solve per fiber in
// python/tstep/tstep_utils.py:1405-1419 and python/tstep/tstep.py:985-988;
// the block-diagonal apply As_fibers.dot(XT) in python/tstep/tstep.py:1173.
//
// LU layout on the CU: the 1024 threads form a 32 x 32 grid (tr = row lane,
// tc = column lane); thread (tr, tc) owns rows tr + 32a and columns tc + 32b,
// a, b in [0, 8), i.e. a 2-D cyclic distribution that keeps every thread busy
// as the trailing matrix shrinks.  FP32 blocks (256 KB at n = 256) live wholly
// in registers.  FP64 blocks (512 KB) do not fit the register file, so the
// last 64 columns live in LDS (128 KB) and the rest in registers.  Each
// elimination step costs two workgroup barriers: pivot search (column owners,
// half-wave shuffles) -> publish pivot row / old row k / column k through LDS
// -> every thread swaps, forms its multipliers and updates its tile.
#include "common.h"

#include <type_traits>

#include <stdlib.h>
#include <string.h>

namespace f3d {

constexpr int kLuG = 32;     // thread grid is kLuG x kLuG
constexpr int kLuTL = 8;     // tile slots per thread and dimension
constexpr int kLuNmax = kLuG * kLuTL;  // 256

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffu), lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Power of two s with s * amax in [0.5, 1); 1 for amax == 0 or non-finite.
__device__ __forceinline__ double pow2_scale(double amax) {
  if (!(amax > 0.0) || !isfinite(amax)) return 1.0;
  int e;
  frexp(amax, &e);
  return ldexp(1.0, -e);
}

template <typename T, int TLC_LDS>
struct LuTile {
  static constexpr int TLC_REG = kLuTL - TLC_LDS;
  T r[kLuTL][TLC_REG > 0 ? TLC_REG : 1];
  T* lds;  // column-major [32*TLC_LDS][256] for columns >= 32*TLC_REG
  int tr, tc;
  __device__ __forceinline__ T get(int a, int b) const {
    if (b < TLC_REG) return r[a][b];
    return lds[(tc + kLuG * (b - TLC_REG)) * kLuNmax + tr + kLuG * a];
  }
  __device__ __forceinline__ void set(int a, int b, T v) {
    if (b < TLC_REG) r[a][b] = v;
    else lds[(tc + kLuG * (b - TLC_REG)) * kLuNmax + tr + kLuG * a] = v;
  }
};

template <typename T, int TLC_LDS>
__global__ void __launch_bounds__(1024)
lu_factor_kernel(int nb, const int* __restrict__ n_arr, const long* __restrict__ a_off,
                 const long* __restrict__ x_off, const int* __restrict__ order,
                 const double* __restrict__ A, T* __restrict__ LU, double* __restrict__ rs,
                 double* __restrict__ cs, int* __restrict__ piv, int* __restrict__ perm,
                 int* __restrict__ info, int equilibrate) {
  __shared__ T lds_mat[TLC_LDS > 0 ? kLuG * TLC_LDS * kLuNmax : 1];
  __shared__ T rowU[kLuNmax], rowK[kLuNmax], colL[kLuNmax];
  __shared__ unsigned long long amax_bits[kLuNmax];
  __shared__ double scl_r[kLuNmax], scl_c[kLuNmax];
  __shared__ int sh_perm[kLuNmax];
  __shared__ int sh_piv, sh_info;

  LuTile<T, TLC_LDS> t;
  t.lds = lds_mat;
  t.tr = threadIdx.x % kLuG;
  t.tc = threadIdx.x / kLuG;
  const int tr = t.tr, tc = t.tc;

  for (int w = blockIdx.x; w < nb; w += gridDim.x) {
    const int b = order ? order[w] : w;
    const int n = n_arr[b];
    if (n < 1 || n > kLuNmax) {  // uniform per block: every thread skips
      if (threadIdx.x == 0) info[b] = -1;
      continue;
    }
    const int mc = (n + kLuG - 1) / kLuG;  // occupied slots per dimension
    const double* __restrict__ Ab = A + a_off[b];
    const long xo = x_off[b];
    __syncthreads();  // LDS reuse across blocks

    // ---- equilibration scales (powers of two, so scaling is exact) -------
    if (threadIdx.x < kLuNmax) {
      sh_perm[threadIdx.x] = threadIdx.x;
      amax_bits[threadIdx.x] = 0ull;
      scl_r[threadIdx.x] = 1.0;
      scl_c[threadIdx.x] = 1.0;
    }
    if (threadIdx.x == 0) sh_info = 0;
    __syncthreads();
    if (equilibrate) {
      // row maxima
#pragma unroll
      for (int a = 0; a < kLuTL; ++a) {
        const int i = tr + kLuG * a;
        if (a >= mc || i >= n) continue;
        double m = 0.0;
#pragma unroll
        for (int bb = 0; bb < kLuTL; ++bb) {
          const int j = tc + kLuG * bb;
          if (bb < mc && j < n) m = fmax(m, fabs(Ab[i + (long)j * n]));
        }
        atomicMax(&amax_bits[i], (unsigned long long)__double_as_longlong(m));
      }
      __syncthreads();
      if (threadIdx.x < n) {
        scl_r[threadIdx.x] = pow2_scale(__longlong_as_double((long long)amax_bits[threadIdx.x]));
        amax_bits[threadIdx.x] = 0ull;
      }
      __syncthreads();
      // column maxima of the row-scaled matrix
#pragma unroll
      for (int bb = 0; bb < kLuTL; ++bb) {
        const int j = tc + kLuG * bb;
        if (bb >= mc || j >= n) continue;
        double m = 0.0;
#pragma unroll
        for (int a = 0; a < kLuTL; ++a) {
          const int i = tr + kLuG * a;
          if (a < mc && i < n) m = fmax(m, fabs(Ab[i + (long)j * n]) * scl_r[i]);
        }
        atomicMax(&amax_bits[j], (unsigned long long)__double_as_longlong(m));
      }
      __syncthreads();
      if (threadIdx.x < n)
        scl_c[threadIdx.x] = pow2_scale(__longlong_as_double((long long)amax_bits[threadIdx.x]));
      __syncthreads();
    }

    // ---- load the (scaled) block into the tile ---------------------------
#pragma unroll
    for (int a = 0; a < kLuTL; ++a) {
#pragma unroll
      for (int bb = 0; bb < kLuTL; ++bb) {
        const int i = tr + kLuG * a, j = tc + kLuG * bb;
        T v = 0;
        if (a < mc && bb < mc && i < n && j < n) v = (T)(Ab[i + (long)j * n] * scl_r[i] * scl_c[j]);
        t.set(a, bb, v);
      }
    }
    __syncthreads();

    // ---- right-looking elimination --------------------------------------
    for (int k = 0; k < n; ++k) {
      const int kb = k / kLuG, kc = k % kLuG;
      if (tc == kc) {
        double best = -1.0;
        int bi = n;
#pragma unroll
        for (int a = 0; a < kLuTL; ++a) {
          const int i = tr + kLuG * a;
          T v = 0;
#pragma unroll
          for (int bb = 0; bb < kLuTL; ++bb)
            if (bb == kb) v = t.get(a, bb);
          const double av = fabs((double)v);
          if (a < mc && i >= k && i < n && av > best) {
            best = av;
            bi = i;
          }
        }
#pragma unroll
        for (int o = 16; o >= 1; o >>= 1) {
          const double ob = __shfl_xor(best, o, 32);
          const int oi = __shfl_xor(bi, o, 32);
          if (ob > best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
          }
        }
        if (tr == 0) {
          sh_piv = bi;
          if (best == 0.0 && sh_info == 0) sh_info = k + 1;
        }
      }
      __syncthreads();
      const int p = sh_piv;
      const int ka = k / kLuG, kr = k % kLuG, pa = p / kLuG, pr = p % kLuG;
      // publish pivot row (old row p), old row k and old column k
#pragma unroll
      for (int bb = 0; bb < kLuTL; ++bb) {
        const int j = tc + kLuG * bb;
        if (bb < mc && j < n) {
          if (tr == pr) {
            T v = 0;
#pragma unroll
            for (int a = 0; a < kLuTL; ++a)
              if (a == pa) v = t.get(a, bb);
            rowU[j] = v;
          }
          if (tr == kr) {
            T v = 0;
#pragma unroll
            for (int a = 0; a < kLuTL; ++a)
              if (a == ka) v = t.get(a, bb);
            rowK[j] = v;
          }
        }
      }
      if (tc == kc) {
#pragma unroll
        for (int a = 0; a < kLuTL; ++a) {
          const int i = tr + kLuG * a;
          T v = 0;
#pragma unroll
          for (int bb = 0; bb < kLuTL; ++bb)
            if (bb == kb) v = t.get(a, bb);
          if (a < mc && i < n) colL[i] = v;
        }
      }
      if (threadIdx.x == 0) {
        piv[xo + k] = p;
        const int tmp = sh_perm[k];
        sh_perm[k] = sh_perm[p];
        sh_perm[p] = tmp;
      }
      __syncthreads();
      const T ukk = rowU[k];
      const T inv = (ukk != (T)0) ? (T)1 / ukk : (T)0;
      T u[kLuTL];
#pragma unroll
      for (int bb = 0; bb < kLuTL; ++bb) {
        const int j = tc + kLuG * bb;
        u[bb] = (bb < mc && j > k && j < n) ? rowU[j] : (T)0;
      }
#pragma unroll
      for (int a = 0; a < kLuTL; ++a) {
        const int i = tr + kLuG * a;
        if (a >= mc || i >= n || i < k) continue;
        if (i == k || i == p) {
#pragma unroll
          for (int bb = 0; bb < kLuTL; ++bb) {
            const int j = tc + kLuG * bb;
            if (bb < mc && j < n) t.set(a, bb, i == k ? rowU[j] : rowK[j]);
          }
        }
        if (i == k) continue;
        const T raw = (i == p) ? rowK[k] : colL[i];
        const T l = raw * inv;
        if (tc == kc) {
#pragma unroll
          for (int bb = 0; bb < kLuTL; ++bb)
            if (bb == kb) t.set(a, bb, l);
        }
#pragma unroll
        for (int bb = 0; bb < kLuTL; ++bb)
          if (bb < mc) t.set(a, bb, t.get(a, bb) - l * u[bb]);
      }
      // The next step's pivot search reads only registers/LDS tile entries
      // that this thread itself updated; the next LDS publish happens after
      // that step's first barrier, which every thread reaches only once it
      // is done with rowU/rowK/colL here.
    }
    __syncthreads();

    // ---- write factors, permutation, scales ----------------------------
    T* __restrict__ Lb = LU + a_off[b];
#pragma unroll
    for (int a = 0; a < kLuTL; ++a) {
#pragma unroll
      for (int bb = 0; bb < kLuTL; ++bb) {
        const int i = tr + kLuG * a, j = tc + kLuG * bb;
        if (a < mc && bb < mc && i < n && j < n) Lb[i + (long)j * n] = t.get(a, bb);
      }
    }
    if (threadIdx.x < n) {
      perm[xo + threadIdx.x] = sh_perm[threadIdx.x];
      if (rs) rs[xo + threadIdx.x] = scl_r[threadIdx.x];
      if (cs) cs[xo + threadIdx.x] = scl_c[threadIdx.x];
    }
    if (threadIdx.x == 0) info[b] = sh_info;
  }
}

// ---------------------------------------------------------------------------
// FP32 blocked LU on MFMA (right-looking, panel width 32, LAPACK getrf
// semantics on the equilibrated matrix), one 256-thread workgroup per block,
// two workgroups per CU.  The equilibrated FP32 matrix is written to the
// output buffer and factored in place there (L2-resident: 256 KB per block).
// Per panel k (columns c0 = 32k .. c0 + 31, local rows 0 .. m-1 from c0 on):
//   1. the panel is loaded into LDS (sL, column-major, one row per thread)
//      and factored column by column with partial pivoting: block argmax
//      (first maximum), row swap, multipliers, rank-1 update of the panel;
//      two barriers per column;
//   2. each thread owns one matrix column outside the panel: it applies the
//      panel's 32 row swaps in order and, right of the panel, solves
//      U12 = L11^-1 A12 for its column (copy kept in LDS, sU, row-major);
//   3. A22 -= L21 U12 on 32 x 32 tiles, one tile per wave at a time:
//      v_mfma_f32_32x32x2f32 with A = -L21 and B = U12 from LDS, C / D
//      loaded from and stored to the factor buffer.
// Orders that are not a multiple of 32 are padded with an identity block
// that is never stored: virtual rows and columns stay decoupled, so the
// factors of the real part are exactly those of the unpadded matrix.
constexpr int kPw = 32;               // panel width = MFMA tile edge
constexpr int kMfThreads = 256;       // 4 waves

typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ void __launch_bounds__(kMfThreads, 2)
lu_factor_f32_mfma_kernel(int nb, const int* __restrict__ n_arr, const long* __restrict__ a_off,
                          const long* __restrict__ x_off, const int* __restrict__ order,
                          const double* __restrict__ A, float* __restrict__ LU,
                          double* __restrict__ rs, double* __restrict__ cs,
                          int* __restrict__ piv, int* __restrict__ perm, int* __restrict__ info) {
  __shared__ float sL[kPw * kLuNmax];   // panel, column c at c * 256 + local row
  __shared__ float sU[kPw * kLuNmax];   // U12, row q at q * 256 + (col - c0 - 32)
  __shared__ double scl_r[kLuNmax], scl_c[kLuNmax];
  __shared__ float red_v[kMfThreads / 64];
  __shared__ int red_i[kMfThreads / 64];
  __shared__ int sh_piv[kPw];
  __shared__ int sh_perm[kLuNmax];
  __shared__ int sh_info;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;

  for (int w = blockIdx.x; w < nb; w += gridDim.x) {
    const int b = order ? order[w] : w;
    const int n = n_arr[b];
    if (n < 1 || n > kLuNmax) {  // uniform per block
      if (t == 0) info[b] = -1;
      continue;
    }
    const int npad = (n + kPw - 1) / kPw * kPw;
    const double* __restrict__ Ab = A + a_off[b];
    float* __restrict__ Lb = LU + a_off[b];
    const long xo = x_off[b];
    __syncthreads();  // LDS reuse across blocks

    // ---- power-of-two equilibration, scaled FP32 copy into Lb -------------
    sh_perm[t] = t;
    if (t == 0) sh_info = 0;
    if (t < n) {
      double mx = 0.0;
      for (int c = 0; c < n; ++c) mx = fmax(mx, fabs(Ab[t + (long)c * n]));
      scl_r[t] = pow2_scale(mx);
    }
    __syncthreads();
    if (t < n) {
      double mx = 0.0;
      const double* __restrict__ col = Ab + (long)t * n;
      for (int r = 0; r < n; ++r) mx = fmax(mx, fabs(col[r]) * scl_r[r]);
      scl_c[t] = pow2_scale(mx);
    }
    __syncthreads();
    if (t < n) {
      const double sr = scl_r[t];
      for (int c = 0; c < n; ++c)
        Lb[t + (long)c * n] = (float)(Ab[t + (long)c * n] * sr * scl_c[c]);
    }
    __syncthreads();

    for (int c0 = 0; c0 < npad; c0 += kPw) {
      const int m = npad - c0;  // panel rows, a multiple of 32, <= 256
      // ---- 1. panel into LDS (identity on the padding) --------------------
      {
        const int R = c0 + t;
        for (int c = 0; c < kPw; ++c) {
          const int C = c0 + c;
          float v = 0.f;
          if (t < m) v = (R < n && C < n) ? Lb[R + (long)C * n] : (R == C ? 1.f : 0.f);
          sL[c * kLuNmax + t] = v;
        }
      }
      __syncthreads();
      for (int j = 0; j < kPw; ++j) {
        // argmax |a(r, j)| over local rows j <= r < m, first maximum wins
        // (barrier B1 below also orders this step's swap after every thread's
        // last update of its own row; B2 orders the update after the swap)
        float best = (t >= j && t < m) ? fabsf(sL[j * kLuNmax + t]) : -1.f;
        int bi = t;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
          const float ob = __shfl_xor(best, o, 64);
          const int oi = __shfl_xor(bi, o, 64);
          if (ob > best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
          }
        }
        if (lane == 0) {
          red_v[wave] = best;
          red_i[wave] = bi;
        }
        __syncthreads();  // B1
        float pb = red_v[0];
        int p = red_i[0];
#pragma unroll
        for (int q = 1; q < kMfThreads / 64; ++q) {
          const float vq = red_v[q];
          if (vq > pb || (vq == pb && red_i[q] < p)) {
            pb = vq;
            p = red_i[q];
          }
        }
        if (!(pb >= 0.f) || p < j || p >= m) p = j;  // NaN column: keep the diagonal
        // row swap, threads 0..31 one panel column each
        if (t < kPw && p != j) {
          const float x = sL[t * kLuNmax + j];
          sL[t * kLuNmax + j] = sL[t * kLuNmax + p];
          sL[t * kLuNmax + p] = x;
        }
        if (t == 0) sh_piv[j] = p;
        __syncthreads();  // B2
        const float pivot = sL[j * kLuNmax + j];
        if (t == 0 && pivot == 0.f && sh_info == 0 && c0 + j < n) sh_info = c0 + j + 1;
        if (t > j && t < m) {
          const float inv = pivot != 0.f ? 1.f / pivot : 0.f;
          const float l = sL[j * kLuNmax + t] * inv;
          sL[j * kLuNmax + t] = l;
          for (int c = j + 1; c < kPw; ++c)
            sL[c * kLuNmax + t] = fmaf(-l, sL[c * kLuNmax + j], sL[c * kLuNmax + t]);
        }
      }
      __syncthreads();
      // panel back to the factor buffer; pivots and the row permutation
      {
        const int R = c0 + t;
        if (t < m && R < n)
          for (int c = 0; c < kPw && c0 + c < n; ++c) Lb[R + (long)(c0 + c) * n] = sL[c * kLuNmax + t];
      }
      if (t == 0) {
        for (int j = 0; j < kPw; ++j) {
          const int p = sh_piv[j];
          if (c0 + j < n) piv[xo + c0 + j] = c0 + p;
          const int tmp = sh_perm[c0 + j];
          sh_perm[c0 + j] = sh_perm[c0 + p];
          sh_perm[c0 + p] = tmp;
        }
      }
      // ---- 2. row swaps outside the panel, TRSM right of it ---------------
      if (t < npad && (t < c0 || t >= c0 + kPw)) {
        float* __restrict__ colp = Lb + (long)t * n;
        if (t < n) {
          for (int j = 0; j < kPw; ++j) {
            const int p = sh_piv[j];
            if (p != j && c0 + p < n) {
              const float x = colp[c0 + j];
              colp[c0 + j] = colp[c0 + p];
              colp[c0 + p] = x;
            }
          }
        }
        if (t >= c0 + kPw) {
          const int nv = min(kPw, n - c0);  // real rows of the panel (uniform, >= 1)
          const bool full = t < n && nv == kPw;
          float x[kPw];
          if (full) {
#pragma unroll
            for (int i = 0; i < kPw; ++i) x[i] = colp[c0 + i];
          } else {
            const bool live = t < n;
#pragma unroll
            for (int i = 0; i < kPw; ++i) x[i] = (live && i < nv) ? colp[c0 + i] : 0.f;
          }
          // column-oriented forward substitution with the unit-lower L11
#pragma unroll
          for (int q = 0; q < kPw - 1; ++q) {
            const float xq = x[q];
#pragma unroll
            for (int i = q + 1; i < kPw; ++i) x[i] = fmaf(-sL[q * kLuNmax + i], xq, x[i]);
          }
          const int cu = t - c0 - kPw;
#pragma unroll
          for (int i = 0; i < kPw; ++i) sU[i * kLuNmax + cu] = x[i];
          if (full) {
#pragma unroll
            for (int i = 0; i < kPw; ++i) colp[c0 + i] = x[i];
          } else if (t < n) {
#pragma unroll
            for (int i = 0; i < kPw; ++i)
              if (i < nv) colp[c0 + i] = x[i];
          }
        }
      }
      __syncthreads();
      // ---- 3. trailing update on MFMA ---------------------------------------
      const int nr = (npad - c0 - kPw) / kPw;  // trailing tiles per dimension
      for (int tile = wave; tile < nr * nr; tile += kMfThreads / 64) {
        const int ti = tile / nr, tj = tile % nr;
        const int R0 = c0 + kPw + kPw * ti, C0 = c0 + kPw + kPw * tj;
        const int h = lane >> 5, cl = lane & 31;
        const int C = C0 + cl;
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int R = R0 + 8 * (r >> 2) + 4 * h + (r & 3);
          acc[r] = (R < n && C < n) ? Lb[R + (long)C * n] : 0.f;
        }
        // A (32 x 2 per step): lane holds L21(R0 + cl, 2s + h);
        // B (2 x 32 per step): lane holds U12(2s + h, C0 + cl)
        const float* __restrict__ la = sL + (R0 - c0) + cl;
        const float* __restrict__ ub = sU + (C0 - c0 - kPw) + cl;
#pragma unroll
        for (int s = 0; s < kPw / 2; ++s) {
          const int q = 2 * s + h;
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(-la[q * kLuNmax], ub[q * kLuNmax], acc, 0,
                                                     0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int R = R0 + 8 * (r >> 2) + 4 * h + (r & 3);
          if (R < n && C < n) Lb[R + (long)C * n] = acc[r];
        }
      }
      __syncthreads();
    }

    if (t < n) {
      perm[xo + t] = sh_perm[t];
      rs[xo + t] = scl_r[t];
      cs[xo + t] = scl_c[t];
    }
    if (t == 0) info[b] = sh_info;
  }
}

// ---------------------------------------------------------------------------
// One wavefront per block: lane owns rows lane + 64 m, m < 4.
constexpr int kSlots = 4;

// In-wave row permutation: w_new[row i] = w_old[row perm[i]], with row r held
// by lane r % 64 in slot r / 64.
__device__ __forceinline__ void permute_rows(double (&w)[kSlots], const int (&src_lane)[kSlots],
                                             const int (&src_slot)[kSlots], int n, int lane) {
  double out[kSlots];
#pragma unroll
  for (int m = 0; m < kSlots; ++m) {
    double acc = 0.0;
#pragma unroll
    for (int ms = 0; ms < kSlots; ++ms) {
      const double t = __shfl(w[ms], src_lane[m], 64);
      if (ms == src_slot[m]) acc = t;
    }
    out[m] = (64 * m + lane < n) ? acc : 0.0;
  }
#pragma unroll
  for (int m = 0; m < kSlots; ++m) w[m] = out[m];
}

// Column streaming for the one-wave solves: column k of a column-major n x n
// block is consumed in order k = k_begin, k_begin + step, ... (count columns),
// each lane holding rows lane + 64 m.  Columns come in rounds of kPf, in two
// register sets that trade roles: the next round loads into one set while
// the other's columns are used, so a wave waits for HBM once per round and
// never for the loads it has just issued.  (A single ring reloaded in place
// let the compiler hoist each reload above the mask that reads the old
// value, copy the loaded value back at the end of the round, and wait there
// for the round's own loads.)  want(i, k) masks the rows a column
// contributes (the triangle).
constexpr int kPf = 4;  // 6 or 8 take more than 256 VGPRs: one wave per SIMD instead of two

// The loads are unconditional (row and column indices clamped into the
// block; rows past n and columns past the stream are loaded and discarded)
// and the triangle mask is applied when a column is used: a load under a
// lane mask would end in a branch whose join waits for every load in flight.
template <typename T, typename Want, typename Use>
__device__ __forceinline__ void stream_cols(const T* __restrict__ M, int n, int k_begin, int step,
                                            int count, int lane, Want&& want, Use&& use) {
  T ra[kPf][kSlots], rb[kPf][kSlots];
  auto load = [&](int idx0, T (&r)[kPf][kSlots]) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < kPf; ++u) {
      int k = k_begin + (idx0 + u) * step;
      k = k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
      const T* __restrict__ col = M + (long)k * n;
#pragma unroll
      for (int m = 0; m < kSlots; ++m) {
        const int i = 64 * m + lane;
        r[u][m] = col[i < n ? i : n - 1];
      }
    }
  };
  // tail: stop at column index count (the whole rounds have no branch, so
  // their loads stay where they are issued)
  auto take = [&](auto tail, int idx0, const T (&r)[kPf][kSlots]) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < kPf; ++u) {
      if constexpr (decltype(tail)::value) {
        if (idx0 + u >= count) break;
      }
      const int k = k_begin + (idx0 + u) * step;
      T v[kSlots];
#pragma unroll
      for (int m = 0; m < kSlots; ++m) {
        const int i = 64 * m + lane;
        v[m] = (i < n && want(i, k)) ? r[u][m] : (T)0;
      }
      use(k, v);
    }
  };
  load(0, ra);
  // whole pairs of rounds without a branch between the loads and their uses
  int i0 = 0;
  for (; i0 + 2 * kPf <= count; i0 += 2 * kPf) {
    load(i0 + kPf, rb);
    // The column index of each round passes through an empty asm placed
    // after the loads: the round's work depends on it, so the compiler can
    // neither sink the loads to their uses nor hoist the uses to the loads.
    int ja = i0;
    asm volatile("" : "+s"(ja) : : "memory");
    take(std::false_type{}, ja, ra);
    load(i0 + 2 * kPf, ra);
    int jb = i0 + kPf;
    asm volatile("" : "+s"(jb) : : "memory");
    take(std::false_type{}, jb, rb);
  }
  if (i0 < count) {
    load(i0 + kPf, rb);
    take(std::true_type{}, i0, ra);
    take(std::true_type{}, i0 + kPf, rb);
  }
}

// v[slot] of lane (k % 64), slot k / 64 (k wave-uniform): one readlane per
// slot and scalar selects.  (A select chain over the slots folds into a
// dynamic index of v, which the compiler places in LDS: a write, a read and a
// wait in the substitution's dependent chain.)
template <typename T>
__device__ __forceinline__ double pick_row(const T (&v)[kSlots], int k) {
  const int ms = k >> 6, l = k & 63;
  double r = readlane_f64((double)v[0], l);
#pragma unroll
  for (int m = 1; m < kSlots; ++m) {
    const double t = readlane_f64((double)v[m], l);
    r = ms == m ? t : r;
  }
  return r;
}

// In-register LU solve of one block held by one wave: w <- U^-1 L^-1 w (the
// factor's L unit lower, U upper, column-major), FP64 arithmetic.
template <typename T>
__device__ __forceinline__ void trsv_lu(const T* __restrict__ L, int n, int lane, double (&w)[kSlots]) {
  // forward: w_i -= L(i, k) w_k for i > k
  stream_cols(L, n, 0, 1, n - 1, lane, [](int i, int k) { return i > k; },
              [&](int k, const T (&c)[kSlots]) {
                const double wk = pick_row(w, k);
#pragma unroll
                for (int m = 0; m < kSlots; ++m) w[m] = fma(-(double)c[m], wk, w[m]);
              });
  // backward: w_k /= U(k, k); w_i -= U(i, k) w_k for i < k
  stream_cols(L, n, n - 1, -1, n, lane, [](int i, int k) { return i <= k; },
              [&](int k, const T (&c)[kSlots]) {
                const double ukk = pick_row(c, k);
                const double wk = pick_row(w, k) / ukk;
#pragma unroll
                for (int m = 0; m < kSlots; ++m) {
                  const int i = 64 * m + lane;
                  w[m] = (i == k) ? wk : ((i < k) ? fma(-(double)c[m], wk, w[m]) : w[m]);
                }
              });
}

template <typename T>
__global__ void __launch_bounds__(64)
lu_solve_kernel(const int* __restrict__ n_arr, const long* __restrict__ a_off,
                const long* __restrict__ x_off, const double* __restrict__ A,
                const T* __restrict__ LU, const double* __restrict__ rs,
                const double* __restrict__ cs, const int* __restrict__ perm,
                const double* x, double* y, int refine, double tol, int* __restrict__ iters) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = n_arr[b];
  if (n < 1 || n > 64 * kSlots) {
    if (iters && lane == 0) iters[b] = -1;
    return;
  }
  const long xo = x_off[b];
  const T* __restrict__ L = LU + a_off[b];
  const double* __restrict__ Ab = A ? A + a_off[b] : nullptr;
  const int* __restrict__ pm = perm + xo;

  double xv[kSlots], yv[kSlots], w[kSlots], csv[kSlots];
#pragma unroll
  for (int m = 0; m < kSlots; ++m) {
    const int i = 64 * m + lane;
    xv[m] = i < n ? x[xo + i] : 0.0;
    csv[m] = (cs && i < n) ? cs[xo + i] : 1.0;
  }
  // w = P (rs .* x): row i of the factor is row perm[i] of A.
  int src_lane[kSlots], src_slot[kSlots];
#pragma unroll
  for (int m = 0; m < kSlots; ++m) {
    const int i = 64 * m + lane;
    const int src = i < n ? pm[i] : 0;
    src_lane[m] = src & 63;
    src_slot[m] = src >> 6;
    const double rsv = (rs && i < n) ? rs[xo + i] : 1.0;
    w[m] = xv[m] * rsv;  // still in A's row order
  }
  permute_rows(w, src_lane, src_slot, n, lane);
  trsv_lu<T>(L, n, lane, w);
#pragma unroll
  for (int m = 0; m < kSlots; ++m) yv[m] = w[m] * csv[m];

  int it = 0;
  double anorm = -1.0;
  double rprev = 0.0;
  for (; it < refine; ++it) {
    // r = x - A y in FP64 (first sweep also measures ||A||_inf)
    double r[kSlots], rowabs[kSlots];
#pragma unroll
    for (int m = 0; m < kSlots; ++m) {
      r[m] = xv[m];
      rowabs[m] = 0.0;
    }
    const bool first = anorm < 0.0;
    stream_cols(Ab, n, 0, 1, n, lane, [](int, int) { return true; },
                [&](int j, const double (&c)[kSlots]) {
                  const double yj = pick_row(yv, j);
#pragma unroll
                  for (int mm = 0; mm < kSlots; ++mm) {
                    r[mm] = fma(-c[mm], yj, r[mm]);
                    if (first) rowabs[mm] += fabs(c[mm]);
                  }
                });
    double rn = 0.0, yn = 0.0, an = 0.0;
#pragma unroll
    for (int m = 0; m < kSlots; ++m) {
      rn = fmax(rn, fabs(r[m]));
      yn = fmax(yn, fabs(yv[m]));
      an = fmax(an, rowabs[m]);
    }
    rn = wave_max_f64(rn);
    yn = wave_max_f64(yn);
    if (anorm < 0.0) anorm = wave_max_f64(an);
    if (rn <= tol * anorm * yn) break;
    // stagnation: FP64 refinement stopped contracting the residual
    if (it > 0 && rn > 0.5 * rprev) break;
    rprev = rn;
    // dy = solve(r)
#pragma unroll
    for (int m = 0; m < kSlots; ++m) {
      const int i = 64 * m + lane;
      w[m] = r[m] * ((rs && i < n) ? rs[xo + i] : 1.0);
    }
    permute_rows(w, src_lane, src_slot, n, lane);
    trsv_lu<T>(L, n, lane, w);
#pragma unroll
    for (int m = 0; m < kSlots; ++m) yv[m] += w[m] * csv[m];
  }
  if (iters) iters[b] = it;
#pragma unroll
  for (int m = 0; m < kSlots; ++m) {
    const int i = 64 * m + lane;
    if (i < n) y[xo + i] = yv[m];
  }
}

// ---------------------------------------------------------------------------
// Solve without refinement with the factor staged in LDS: the workgroup's
// four waves copy the block's packed factor (n^2 entries, at most
// kLdsSolveBytes) into LDS in 16-byte pieces, all of it in flight at once,
// and one wave then runs both substitutions out of LDS, where a column costs
// an LDS read instead of an HBM round trip.  Same arithmetic as
// lu_solve_kernel with refine = 0 (FP64 substitutions in the same order).
constexpr int kLdsSolveBytes = 64 * 1024;

// Launch choices set by the autotuner through f3d_lu_set_launch, per
// precision (0 = FP64, 1 = FP32): waves per CU of the one-wave blocked
// factorisation, and whether solves without refinement stage the factor in
// LDS when it fits.
static int g_lu_per_cu[2] = {4, 8};
static int g_lu_solve_staged[2] = {1, 1};

// Substitutions out of LDS.  The loops run over the 64-row slot ms of the
// pivot row k as an unrolled constant, so every register of w is addressed
// statically, and the columns of the next kLdsGroup pivots are read while the
// current group is applied (LDS latency off the dependent chain).
constexpr int kLdsGroup = 8;

template <typename T>
__device__ __forceinline__ void trsv_lu_lds(const T* __restrict__ sL, int n, int lane,
                                            double (&w)[kSlots]) {
  constexpr int G = kLdsGroup;
  // forward: w_i -= L(i, k) w_k for i > k (rows i > k sit in slots >= ms)
#pragma unroll
  for (int ms = 0; ms < kSlots; ++ms) {
    const int k0 = 64 * ms, kend = min(64 * ms + 64, n - 1);
    if (k0 >= kend) continue;
    T cur[G][kSlots], nxt[G][kSlots];
    auto load = [&](int kg, T (&buf)[G][kSlots]) __attribute__((always_inline)) {
#pragma unroll
      for (int u = 0; u < G; ++u)
#pragma unroll
        for (int m = ms; m < kSlots; ++m) {
          const int k = kg + u, i = 64 * m + lane;
          buf[u][m] = (k < kend && i > k && i < n) ? sL[(long)k * n + i] : (T)0;
        }
    };
    load(k0, cur);
    for (int kg = k0; kg < kend; kg += G) {
      if (kg + G < kend) load(kg + G, nxt);
#pragma unroll
      for (int u = 0; u < G; ++u) {
        if (kg + u < kend) {
          const double wk = readlane_f64(w[ms], (kg + u) & 63);
#pragma unroll
          for (int m = ms; m < kSlots; ++m) w[m] = fma(-(double)cur[u][m], wk, w[m]);
        }
      }
#pragma unroll
      for (int u = 0; u < G; ++u)
#pragma unroll
        for (int m = ms; m < kSlots; ++m) cur[u][m] = nxt[u][m];
    }
  }
  // backward: w_k /= U(k, k); w_i -= U(i, k) w_k for i < k (slots <= ms)
#pragma unroll
  for (int ms = kSlots - 1; ms >= 0; --ms) {
    const int khi = min(64 * ms + 63, n - 1), klo = 64 * ms;
    if (khi < klo) continue;
    T cur[G][kSlots], nxt[G][kSlots];
    double dg[G], dn[G];
    auto load = [&](int kg, T (&buf)[G][kSlots], double (&d)[G]) __attribute__((always_inline)) {
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int k = kg - u;
#pragma unroll
        for (int m = 0; m <= ms; ++m) {
          const int i = 64 * m + lane;
          buf[u][m] = (k >= klo && i < k) ? sL[(long)k * n + i] : (T)0;
        }
        d[u] = k >= klo ? (double)sL[(long)k * n + k] : 1.0;
      }
    };
    load(khi, cur, dg);
    for (int kg = khi; kg >= klo; kg -= G) {
      if (kg - G >= klo) load(kg - G, nxt, dn);
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int k = kg - u;
        if (k >= klo) {
          const double wk = readlane_f64(w[ms], k & 63) / dg[u];
#pragma unroll
          for (int m = 0; m <= ms; ++m) {
            const int i = 64 * m + lane;
            w[m] = (i == k) ? wk : ((i < k) ? fma(-(double)cur[u][m], wk, w[m]) : w[m]);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < G; ++u) {
        dg[u] = dn[u];
#pragma unroll
        for (int m = 0; m <= ms; ++m) cur[u][m] = nxt[u][m];
      }
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
lu_solve_lds_kernel(int nmax, const int* __restrict__ n_arr, const long* __restrict__ a_off,
                    const long* __restrict__ x_off, const T* __restrict__ LU,
                    const double* __restrict__ rs, const double* __restrict__ cs,
                    const int* __restrict__ perm, const double* __restrict__ x,
                    double* __restrict__ y, int* __restrict__ iters) {
  extern __shared__ __align__(16) unsigned char lds_solve[];
  T* __restrict__ sL = reinterpret_cast<T*>(lds_solve);
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63;
  const int n = n_arr[b];
  // the whole workgroup leaves together (n is uniform), before any barrier;
  // the launch sized the LDS for nmax
  if (n < 1 || n > nmax) {
    if (iters && t == 0) iters[b] = -1;
    return;
  }
  const T* __restrict__ L = LU + a_off[b];
  const long nn = (long)n * n;
  if ((reinterpret_cast<uintptr_t>(L) & 15) == 0) {
    constexpr int per = 16 / (int)sizeof(T);
    const long nv = nn / per;
    const uint4* __restrict__ src = reinterpret_cast<const uint4*>(L);
    uint4* __restrict__ dst = reinterpret_cast<uint4*>(sL);
    long e0 = t;
    for (; e0 + 256 * 7 < nv; e0 += 256 * 8) {
      uint4 v0 = src[e0], v1 = src[e0 + 256], v2 = src[e0 + 512], v3 = src[e0 + 768];
      uint4 v4 = src[e0 + 1024], v5 = src[e0 + 1280], v6 = src[e0 + 1536], v7 = src[e0 + 1792];
      dst[e0] = v0;
      dst[e0 + 256] = v1;
      dst[e0 + 512] = v2;
      dst[e0 + 768] = v3;
      dst[e0 + 1024] = v4;
      dst[e0 + 1280] = v5;
      dst[e0 + 1536] = v6;
      dst[e0 + 1792] = v7;
    }
    for (; e0 < nv; e0 += 256) dst[e0] = src[e0];
    for (long e = nv * per + t; e < nn; e += 256) sL[e] = L[e];
  } else {
    for (long e = t; e < nn; e += 256) sL[e] = L[e];
  }
  __syncthreads();
  if (t >= 64) return;

  const long xo = x_off[b];
  const int* __restrict__ pm = perm + xo;
  double w[kSlots];
  int src_lane[kSlots], src_slot[kSlots];
#pragma unroll
  for (int m = 0; m < kSlots; ++m) {
    const int i = 64 * m + lane;
    const int sidx = i < n ? pm[i] : 0;
    src_lane[m] = sidx & 63;
    src_slot[m] = sidx >> 6;
    const double rsv = (rs && i < n) ? rs[xo + i] : 1.0;
    w[m] = (i < n ? x[xo + i] : 0.0) * rsv;  // still in A's row order
  }
  permute_rows(w, src_lane, src_slot, n, lane);
  trsv_lu_lds<T>(sL, n, lane, w);
#pragma unroll
  for (int m = 0; m < kSlots; ++m) {
    const int i = 64 * m + lane;
    if (i < n) y[xo + i] = w[m] * ((cs && i < n) ? cs[xo + i] : 1.0);
  }
  if (iters && lane == 0) iters[b] = 0;
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
block_gemv_kernel(const int* __restrict__ m_arr, const int* __restrict__ k_arr,
                  const long* __restrict__ m_off, const long* __restrict__ x_off,
                  const long* __restrict__ y_off, const double* __restrict__ M,
                  const double* __restrict__ x, double* __restrict__ y, double alpha,
                  double beta) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const int m = m_arr[b], kk = k_arr[b];
  if (m < 0 || m > 64 * kSlots || kk < 0) return;
  const double* __restrict__ Mb = M + m_off[b];
  const double* __restrict__ xb = x + x_off[b];
  double acc[kSlots];
#pragma unroll
  for (int s = 0; s < kSlots; ++s) acc[s] = 0.0;
  const int mc = (m + 63) / 64;
  int j = 0;
  for (; j + 4 <= kk; j += 4) {
    double xj[4], v[4][kSlots];
#pragma unroll
    for (int u = 0; u < 4; ++u) xj[u] = xb[j + u];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int s = 0; s < kSlots; ++s) {
        const int i = 64 * s + lane;
        v[u][s] = (s < mc && i < m) ? Mb[i + (long)(j + u) * m] : 0.0;
      }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int s = 0; s < kSlots; ++s) acc[s] = fma(v[u][s], xj[u], acc[s]);
  }
  for (; j < kk; ++j) {
    const double xj = xb[j];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int i = 64 * s + lane;
      if (s < mc && i < m) acc[s] = fma(Mb[i + (long)j * m], xj, acc[s]);
    }
  }
  double* __restrict__ yb = y + y_off[b];
#pragma unroll
  for (int s = 0; s < kSlots; ++s) {
    const int i = 64 * s + lane;
    if (s < mc && i < m) yb[i] = beta == 0.0 ? alpha * acc[s] : fma(alpha, acc[s], beta * yb[i]);
  }
}

// ---------------------------------------------------------------------------
// y_b = M_b x_b for `count` dense ROW-major m x m matrices (FP64, or FP32
// entries with FP64 products and sums).  One wave computes kGemvRows rows of
// one matrix (fewer at the matrix's end), four waves per 256-thread
// workgroup: each 16-byte piece of x it loads serves all its rows, which
// keeps the x reads (from L2) at a quarter of the matrix bytes.  A wave reads
// each row in 1 KB coalesced steps (16 bytes per lane when `vec` is set: m a
// multiple of 16 / sizeof(TM), M and x 16-byte aligned), accumulates each row
// in two FP64 chains and reduces across the wave in a fixed order, so the
// result is deterministic.  Even one matrix of a few thousand rows puts
// thousands of waves in flight, which an HBM-bound sweep over a large inverse
// needs; the column-major row-block GEMV above runs one wave per 256-row
// block.
constexpr int kGemvRows = 4;

__device__ __forceinline__ void gemv_fma16(const double2 a, const double2* __restrict__ v,
                                           double& acc0, double& acc1) {
  acc0 = fma(a.x, v[0].x, acc0);
  acc1 = fma(a.y, v[0].y, acc1);
}

__device__ __forceinline__ void gemv_fma16(const float4 a, const double2* __restrict__ v,
                                           double& acc0, double& acc1) {
  acc0 = fma((double)a.x, v[0].x, acc0);
  acc1 = fma((double)a.y, v[0].y, acc1);
  acc0 = fma((double)a.z, v[1].x, acc0);
  acc1 = fma((double)a.w, v[1].y, acc1);
}

template <typename TM> struct Gemv16;
template <> struct Gemv16<double> { using V = double2; };
template <> struct Gemv16<float> { using V = float4; };

template <typename TM>
__global__ void __launch_bounds__(256)
dense_gemv_kernel(int64_t waves_total, int32_t m, int32_t groups, int vec,
                  const TM* __restrict__ M, const double* __restrict__ x,
                  double* __restrict__ y) {
  constexpr int R = kGemvRows;
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= waves_total) return;
  const int64_t b = w / groups;
  const int r0 = (int)(w - b * groups) * R;
  const int nr = min(R, m - r0);
  const double* __restrict__ xb = x + b * (int64_t)m;
  // rows past the matrix's end repeat its last row (loaded again from
  // cache, never stored)
  const TM* __restrict__ Mr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) Mr[r] = M + (b * (int64_t)m + r0 + min(r, nr - 1)) * (int64_t)m;
  double acc0[R], acc1[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc0[r] = acc1[r] = 0.0;
  if (vec) {
    using V = typename Gemv16<TM>::V;
    constexpr int per = (int)(16 / sizeof(TM));  // matrix entries per 16 bytes
    constexpr int xv = per / 2;                  // double2 of x per 16 bytes of M
    const int nv = m / per;
    const double2* __restrict__ x2 = reinterpret_cast<const double2*>(xb);
    int p = lane;
    for (; p + 64 < nv; p += 128) {
      V a[R][2];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        a[r][0] = reinterpret_cast<const V*>(Mr[r])[p];
        a[r][1] = reinterpret_cast<const V*>(Mr[r])[p + 64];
      }
      double2 v0[xv], v1[xv];
#pragma unroll
      for (int q = 0; q < xv; ++q) {
        v0[q] = x2[p * xv + q];
        v1[q] = x2[(p + 64) * xv + q];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        gemv_fma16(a[r][0], v0, acc0[r], acc1[r]);
        gemv_fma16(a[r][1], v1, acc0[r], acc1[r]);
      }
    }
    for (; p < nv; p += 64) {
      double2 v0[xv];
#pragma unroll
      for (int q = 0; q < xv; ++q) v0[q] = x2[p * xv + q];
#pragma unroll
      for (int r = 0; r < R; ++r)
        gemv_fma16(reinterpret_cast<const V*>(Mr[r])[p], v0, acc0[r], acc1[r]);
    }
  } else {
    for (int j = lane; j < m; j += 64) {
      const double xj = xb[j];
#pragma unroll
      for (int r = 0; r < R; ++r) acc0[r] = fma((double)Mr[r][j], xj, acc0[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double sum = wave_sum_f64(acc0[r] + acc1[r]);
    if (lane == 0 && r < nr) y[b * (int64_t)m + r0 + r] = sum;
  }
}

static int num_cus() {
  static int cached = 0;
  if (cached == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
    cached = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  }
  return cached;
}

int launch_lu_wave(int precision, int64_t nb, const int32_t* n, const int64_t* a_off,
                   const int64_t* x_off, const int32_t* order, const double* A, double* lu64,
                   float* lu32, double* rs, double* cs, int32_t* piv, int32_t* perm,
                   int32_t* info, int num_cu, int per_cu, hipStream_t st);

}  // namespace f3d

using namespace f3d;

extern "C" int f3d_lu_factor_batched(int precision, int64_t nb, int32_t nmax, const int32_t* n,
                                     const int64_t* a_off, const int64_t* x_off,
                                     const int32_t* order, const double* A, double* lu64,
                                     float* lu32, double* rs, double* cs, int32_t* piv,
                                     int32_t* perm, int32_t* info, void* stream) {
  F3D_REQUIRE(precision == F3D_F64 || precision == F3D_F32C, F3D_EINVAL,
              "lu_factor: unknown precision %d", precision);
  F3D_REQUIRE(nb >= 0 && nb < (1ll << 31), F3D_EINVAL, "lu_factor: bad batch count");
  F3D_REQUIRE(nmax >= 1 && nmax <= kLuNmax, F3D_EINVAL, "lu_factor: nmax=%d outside [1,%d]",
              nmax, kLuNmax);
  if (nb == 0) return F3D_OK;
  F3D_REQUIRE(n && a_off && x_off && A && piv && perm && info, F3D_EINVAL,
              "lu_factor: NULL argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // default: one wavefront per block, blocked, MFMA trailing update
  // (lu_wave.hip); F3D_LU_KERNEL=workgroup selects the one-workgroup-per-block
  // kernels below (FP32: F3D_LU32_KERNEL=vector for the unblocked one)
  const char* sel_k = getenv("F3D_LU_KERNEL");
  if (!(sel_k && strcmp(sel_k, "workgroup") == 0)) {
    if (precision == F3D_F64) {
      F3D_REQUIRE(lu64 != nullptr, F3D_EINVAL, "lu_factor: lu64 is NULL");
    } else {
      F3D_REQUIRE(lu32 && rs && cs, F3D_EINVAL, "lu_factor: lu32/rs/cs is NULL");
    }
    return launch_lu_wave(precision, nb, n, a_off, x_off, order, A, lu64, lu32, rs, cs, piv,
                          perm, info, num_cus(), g_lu_per_cu[precision == F3D_F64 ? 0 : 1], st);
  }
  const unsigned grid = (unsigned)(nb < num_cus() ? nb : num_cus());
  if (precision == F3D_F64) {
    F3D_REQUIRE(lu64 != nullptr, F3D_EINVAL, "lu_factor: lu64 is NULL");
    F3D_LAUNCH((lu_factor_kernel<double, 2>), dim3(grid), dim3(1024), 0, st, (int)nb, n,
               (const long*)a_off, (const long*)x_off, order, A, lu64, (double*)nullptr,
               (double*)nullptr, piv, perm, info, 0);
  } else {
    F3D_REQUIRE(lu32 && rs && cs, F3D_EINVAL, "lu_factor: lu32/rs/cs is NULL");
    // F3D_LU32_KERNEL=vector selects the register-resident VALU factorisation
    const char* sel = getenv("F3D_LU32_KERNEL");
    if (sel && strcmp(sel, "vector") == 0) {
      F3D_LAUNCH((lu_factor_kernel<float, 0>), dim3(grid), dim3(1024), 0, st, (int)nb, n,
                 (const long*)a_off, (const long*)x_off, order, A, lu32, rs, cs, piv, perm, info,
                 1);
    } else {
      const long g2 = 2l * num_cus();
      const unsigned grid2 = (unsigned)(nb < g2 ? nb : g2);
      F3D_LAUNCH(lu_factor_f32_mfma_kernel, dim3(grid2), dim3(kMfThreads), 0, st, (int)nb, n,
                 (const long*)a_off, (const long*)x_off, order, A, lu32, rs, cs, piv, perm,
                 info);
    }
  }
  return F3D_OK;
}

extern "C" int f3d_lu_solve_batched(int precision, int64_t nb, int32_t nmax, const int32_t* n,
                                    const int64_t* a_off, const int64_t* x_off, const double* A,
                                    const double* lu64, const float* lu32, const double* rs,
                                    const double* cs, const int32_t* perm, const double* x,
                                    double* y, int32_t refine, double tol, int32_t* iters,
                                    void* stream) {
  F3D_REQUIRE(precision == F3D_F64 || precision == F3D_F32C, F3D_EINVAL,
              "lu_solve: unknown precision %d", precision);
  F3D_REQUIRE(nb >= 0 && nb < (1ll << 31), F3D_EINVAL, "lu_solve: bad batch count");
  F3D_REQUIRE(nmax >= 1 && nmax <= 64 * kSlots, F3D_EINVAL, "lu_solve: nmax=%d outside [1,%d]",
              nmax, 64 * kSlots);
  F3D_REQUIRE(refine >= 0, F3D_EINVAL, "lu_solve: negative refine");
  if (nb == 0) return F3D_OK;
  F3D_REQUIRE(n && a_off && x_off && perm && x && y, F3D_EINVAL, "lu_solve: NULL argument");
  F3D_REQUIRE(refine == 0 || A != nullptr, F3D_EINVAL, "lu_solve: refinement needs A");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // without refinement, factors that fit the LDS budget are staged there
  const size_t elem = precision == F3D_F64 ? sizeof(double) : sizeof(float);
  const size_t lds = align_up((size_t)nmax * nmax * elem, 16);
  // F3D_LU_SOLVE_STREAM=1 keeps the streaming kernel (for comparison)
  const char* force_stream = getenv("F3D_LU_SOLVE_STREAM");
  const bool staged = refine == 0 && lds <= (size_t)kLdsSolveBytes &&
                      g_lu_solve_staged[precision == F3D_F64 ? 0 : 1] &&
                      !(force_stream && *force_stream && *force_stream != '0');
  if (staged && precision == F3D_F64) {
    F3D_REQUIRE(lu64 != nullptr, F3D_EINVAL, "lu_solve: lu64 is NULL");
    F3D_LAUNCH((lu_solve_lds_kernel<double>), dim3((unsigned)nb), dim3(256), lds, st, nmax,
               n, (const long*)a_off, (const long*)x_off, lu64, rs, cs, perm, x, y, iters);
    return F3D_OK;
  }
  if (staged) {
    F3D_REQUIRE(lu32 && rs && cs, F3D_EINVAL, "lu_solve: lu32/rs/cs is NULL");
    F3D_LAUNCH((lu_solve_lds_kernel<float>), dim3((unsigned)nb), dim3(256), lds, st, nmax,
               n, (const long*)a_off, (const long*)x_off, lu32, rs, cs, perm, x, y, iters);
    return F3D_OK;
  }
  if (precision == F3D_F64) {
    F3D_REQUIRE(lu64 != nullptr, F3D_EINVAL, "lu_solve: lu64 is NULL");
    F3D_LAUNCH((lu_solve_kernel<double>), dim3((unsigned)nb), dim3(64), 0, st, n,
               (const long*)a_off, (const long*)x_off, A, lu64, rs, cs, perm, x, y, refine, tol,
               iters);
  } else {
    F3D_REQUIRE(lu32 && rs && cs, F3D_EINVAL, "lu_solve: lu32/rs/cs is NULL");
    F3D_LAUNCH((lu_solve_kernel<float>), dim3((unsigned)nb), dim3(64), 0, st, n,
               (const long*)a_off, (const long*)x_off, A, lu32, rs, cs, perm, x, y, refine, tol,
               iters);
  }
  return F3D_OK;
}

extern "C" int f3d_lu_set_launch(int precision, int32_t factor_waves_per_cu,
                                 int32_t solve_staged) {
  F3D_REQUIRE(precision == F3D_F64 || precision == F3D_F32C, F3D_EINVAL,
              "lu_set_launch: unknown precision %d", precision);
  const int pp = precision == F3D_F64 ? 0 : 1;
  const int cap = pp == 0 ? 4 : 8;
  F3D_REQUIRE(factor_waves_per_cu >= 0 && factor_waves_per_cu <= cap, F3D_EINVAL,
              "lu_set_launch: waves per CU %d outside [0, %d]", factor_waves_per_cu, cap);
  F3D_REQUIRE(solve_staged == 0 || solve_staged == 1, F3D_EINVAL,
              "lu_set_launch: solve_staged must be 0 or 1");
  g_lu_per_cu[pp] = factor_waves_per_cu ? factor_waves_per_cu : cap;
  g_lu_solve_staged[pp] = solve_staged;
  return F3D_OK;
}

template <typename TM>
static int dense_gemv_launch(int64_t count, int32_t m, const TM* M, const double* x, double* y,
                             void* stream) {
  F3D_REQUIRE(count >= 0 && m >= 0, F3D_EINVAL, "dense_gemv: count=%lld m=%d",
              (long long)count, m);
  if (count == 0 || m == 0) return F3D_OK;
  F3D_REQUIRE(M && x && y, F3D_EINVAL, "dense_gemv: NULL argument");
  const int32_t groups = (m + kGemvRows - 1) / kGemvRows;
  const int64_t waves = count * (int64_t)groups;
  F3D_REQUIRE((waves + 3) / 4 < (1ll << 31), F3D_EINVAL, "dense_gemv: %lld x %d rows",
              (long long)count, m);
  const int per = (int)(16 / sizeof(TM));
  const int vec = (m % per == 0) && (((uintptr_t)M | (uintptr_t)x) % 16 == 0);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  F3D_LAUNCH((dense_gemv_kernel<TM>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st,
             waves, m, groups, vec, M, x, y);
  return F3D_OK;
}

extern "C" int f3d_dense_gemv_batched(int64_t count, int32_t m, const double* M,
                                      const double* x, double* y, void* stream) {
  return dense_gemv_launch<double>(count, m, M, x, y, stream);
}

extern "C" int f3d_dense_gemv_batched_f32(int64_t count, int32_t m, const float* M,
                                          const double* x, double* y, void* stream) {
  return dense_gemv_launch<float>(count, m, M, x, y, stream);
}

extern "C" int f3d_block_gemv_batched(int64_t nb, int32_t mmax, const int32_t* m,
                                      const int32_t* k, const int64_t* m_off,
                                      const int64_t* x_off, const int64_t* y_off,
                                      const double* M, const double* x, double* y, double alpha,
                                      double beta, void* stream) {
  F3D_REQUIRE(nb >= 0 && nb < (1ll << 31), F3D_EINVAL, "block_gemv: bad batch count");
  F3D_REQUIRE(mmax >= 1 && mmax <= 64 * kSlots, F3D_EINVAL,
              "block_gemv: mmax=%d outside [1,%d]", mmax, 64 * kSlots);
  if (nb == 0) return F3D_OK;
  F3D_REQUIRE(m && k && m_off && x_off && y_off && M && x && y, F3D_EINVAL,
              "block_gemv: NULL argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  F3D_LAUNCH(block_gemv_kernel, dim3((unsigned)nb), dim3(64), 0, st, m, k, (const long*)m_off,
             (const long*)x_off, (const long*)y_off, M, x, y, alpha, beta);
  return F3D_OK;
}
