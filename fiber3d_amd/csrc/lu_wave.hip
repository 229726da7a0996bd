// Batched LU with partial pivoting, one wavefront per block, on MFMA.
//
// Reference: scipy.linalg.lu_factor per fiber (LAPACK dgetrf: partial
// pivoting, the first maximal |pivot| wins ties) in
// python/tstep/tstep_utils.py:1405-1419, applied at python/tstep/tstep.py:985-988.
//
// Design (gfx950, wave64).  The batch holds thousands of independent blocks
// of order n <= 256, so each block gets ONE wavefront: no workgroup barrier
// anywhere, and several blocks share a CU.  Blocked right-looking LU with
// panels of 16 columns, factored in place in the output buffer (the L2 /
// Infinity-cache resident working copy):
//   1. panel: the m x 16 panel (m = rows from the diagonal down) is held in
//      registers, four rows per lane, rows never moving: each row carries its
//      current position under the row interchanges.  Per column: argmax over
//      the lane's rows by position order, a DPP butterfly for the wave
//      maximum (ties go to the lowest position), the interchange recorded on
//      the positions, the pivot row read by readlane, multipliers and the
//      rank-1 update.  The panel is then stored with every row at its final
//      position, to global memory and (its L part) to LDS.
//   2. swaps and TRSM outside the panel: each lane owns whole columns; it
//      gathers the rows the panel's permutation moves (independent loads, one
//      round trip), writes them back permuted, and right of the panel solves
//      U12 = L11^-1 A12 against L11 read from LDS (broadcast).
//   3. trailing update A22 -= L21 U12 on 16 x 16 tiles with
//      v_mfma_{f64,f32}_16x16x4, computed transposed (E = D^T) so that each
//      group of 16 lanes loads and stores 16 consecutive rows of one column:
//      A' = -U12^T (global), B' = L21^T (LDS), C/D from/to the factor buffer,
//      (FP64 output rows interleave, so those groups hold columns 4 e + g),
//      one tile column at a time with all its loads in flight before its
//      first MFMA.
// Loads and stores are kept out of lane-masked branches where the order is a
// multiple of 16 (clamped indices, values past the block unused): a memory
// operation under a mask ends in a branch whose join waits for every memory
// operation in flight.
// FP32 blocks are first equilibrated by powers of two (exact) per row and
// column; the solves refine in FP64 against the unscaled FP64 matrix.
// Orders that are not a multiple of 16 are padded with an identity block that
// is never stored: padded rows and columns stay decoupled from the real ones,
// so the real factors are exactly those of the unpadded matrix.
#include "common.h"

#include <type_traits>

namespace f3d {

constexpr int kWp = 16;    // panel width = MFMA tile edge
constexpr int kWs = 272;   // LDS column stride in elements (256 rows + pad)
constexpr int kWmax = 256; // largest order
constexpr int kCb = 8;     // columns per round trip of the copy passes

template <typename T>
struct Mf16;
// Operands: lane l provides A(l % 16, k = l / 16) and B(k = l / 16, l % 16).
// Output element e of lane l is D(row(l / 16, e), l % 16): the FP64
// instruction interleaves the rows (4 e + l / 16), the FP32 one keeps four
// consecutive rows per lane (4 (l / 16) + e).
template <>
struct Mf16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int g, int e) { return 4 * e + g; }
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mf16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int g, int e) { return 4 * g + e; }
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

__device__ __forceinline__ double rl(double v, int lane) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffu), lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float rl(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double pow2_scale_w(double amax) {
  if (!(amax > 0.0) || !isfinite(amax)) return 1.0;
  int e;
  frexp(amax, &e);
  return ldexp(1.0, -e);
}

// Data-parallel-primitive lane moves (DPP): quad_perm xor 1 / xor 2,
// row_half_mirror and row_mirror; a butterfly of them leaves every lane of a
// 16-lane row holding the row's maximum.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffu), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <typename T>
__device__ __forceinline__ T row_max(T v) {
  T o;
  o = dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v = o > v ? o : v;
  o = dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v = o > v ? o : v;
  o = dpp_mov<0x141>(v);  // row_half_mirror
  v = o > v ? o : v;
  o = dpp_mov<0x140>(v);  // row_mirror
  v = o > v ? o : v;
  return v;
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// Element (R, C) of the working matrix, identity outside the real n x n part.
template <typename T>
__device__ __forceinline__ T ld_pad(const T* __restrict__ M, int n, int R, int C) {
  return (R < n && C < n) ? M[R + (long)C * n] : (R == C ? (T)1 : (T)0);
}

template <typename T>
__device__ __forceinline__ void lu_wave_body(int nb, const int* __restrict__ n_arr,
                                             const long* __restrict__ a_off,
                                             const long* __restrict__ x_off,
                                             const int* __restrict__ order,
                                             const double* __restrict__ A, T* __restrict__ LU,
                                             double* __restrict__ rs, double* __restrict__ cs,
                                             int* __restrict__ piv, int* __restrict__ perm,
                                             int* __restrict__ info) {
  typedef typename Mf16<T>::v4 v4;
  constexpr bool kEq = sizeof(T) == 4;  // FP32: equilibrate
  // padded orders: trailing-update tiles with loads in flight
  constexpr int kRing = sizeof(T) == 8 ? 16 : 8;
  __shared__ T sL[kWp * kWs];           // panel, column c at c * kWs + local row
  __shared__ int sperm[kWmax];
  __shared__ int sidx[kWmax];
  __shared__ double scl_c[kEq ? kWmax : 1];  // FP32: column scales
  const int lane = threadIdx.x;

  for (int w = blockIdx.x; w < nb; w += gridDim.x) {
    const int b = order ? order[w] : w;
    const int n = n_arr[b];
    if (n < 1 || n > kWmax) {  // uniform per wave
      if (lane == 0) info[b] = -1;
      continue;
    }
    const int npad = (n + kWp - 1) / kWp * kWp;
    const double* __restrict__ Ab = A + a_off[b];
    T* __restrict__ M = LU + a_off[b];
    const long xo = x_off[b];
    int first_zero = 0;

    // ---- working copy (FP32: power-of-two equilibration) ------------------
    // Column passes over A read kCb columns per round trip (lane rows
    // lane + 64 s), so a lone wave keeps kCb x 4 loads in flight.
    // (The loads are unconditional, row and column clamped into the block,
    // and rows and columns past n are zeroed where the values are used: a
    // load under a lane mask ends in a branch whose join waits for every
    // load in flight.)
    auto pass = [&](auto&& use) __attribute__((always_inline)) {
      for (int cb = 0; cb < n; cb += kCb) {
        double v[kCb][4];
#pragma unroll
        for (int u = 0; u < kCb; ++u)
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int c = min(cb + u, n - 1), r = min(lane + 64 * s, n - 1);
            v[u][s] = Ab[r + (long)c * n];
          }
#pragma unroll
        for (int u = 0; u < kCb; ++u) {
          if (cb + u < n) {
#pragma unroll
            for (int s = 0; s < 4; ++s) v[u][s] = lane + 64 * s < n ? v[u][s] : 0.0;
            use(cb + u, v[u]);
          }
        }
      }
    };
    double sr[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) sr[s] = 1.0;
    if (kEq) {
      double mr[4] = {0.0, 0.0, 0.0, 0.0};
      pass([&](int, const double (&v)[4]) {
#pragma unroll
        for (int s = 0; s < 4; ++s) mr[s] = fmax(mr[s], fabs(v[s]));
      });
#pragma unroll
      for (int s = 0; s < 4; ++s) sr[s] = pow2_scale_w(mr[s]);
      // column maxima of the row-scaled matrix -> scl_c (LDS), and in the
      // same pass the scaled column (its scale is known once its maximum is)
      pass([&](int c, const double (&v)[4]) {
        double m = 0.0;
#pragma unroll
        for (int s = 0; s < 4; ++s) m = fmax(m, fabs(v[s]) * sr[s]);
        m = row_max(m);
        m = fmax(fmax(rl(m, 0), rl(m, 16)), fmax(rl(m, 32), rl(m, 48)));
        const double scol = pow2_scale_w(m);
        if (lane == 0) scl_c[c] = scol;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int r = lane + 64 * s;
          if (r < n) M[r + (long)c * n] = (T)(v[s] * sr[s] * scol);
        }
      });
      __syncthreads();
    } else {
      pass([&](int c, const double (&v)[4]) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int r = lane + 64 * s;
          if (r < n) M[r + (long)c * n] = (T)v[s];
        }
      });
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) sperm[lane + 64 * s] = lane + 64 * s;
    __syncthreads();  // one wave: orders the copy's stores before the loads below

    for (int c0 = 0; c0 < npad; c0 += kWp) {
      const int m = npad - c0;  // panel rows (multiple of 16, <= 256)
      // ---- 1. panel in registers ------------------------------------------
      // Rows stay where they were loaded (lane + 64 s, local row i); pos[s]
      // is the row's current position under LAPACK's row interchanges, which
      // are only recorded here and carried out when the panel is stored.
      T a[4][kWp];
      int pos[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int i = lane + 64 * s;
        pos[s] = i < m ? i : 0x7fff;
        if (npad == n) {
          // rows past the panel (i >= m) load a clamped row and are never
          // used (position 0x7fff): no lane mask around the loads
          const int R = min(c0 + i, n - 1);
#pragma unroll
          for (int c = 0; c < kWp; ++c) a[s][c] = M[R + (long)(c0 + c) * n];
        } else {
#pragma unroll
          for (int c = 0; c < kWp; ++c) a[s][c] = (i < m) ? ld_pad(M, n, c0 + i, c0 + c) : (T)0;
        }
      }
      int pv[kWp];
      sfor<0, kWp>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        // argmax |a| over the rows at positions >= j; the first (lowest
        // position) maximum wins, as LAPACK's idamax
        T best = (T)-1;
        int bpos = 0x7fff, bslot = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const T v = a[s][j] < (T)0 ? -a[s][j] : a[s][j];
          if (pos[s] >= j && pos[s] < m && (v > best || (v == best && pos[s] < bpos))) {
            best = v;
            bpos = pos[s];
            bslot = s;
          }
        }
        const T rmax = row_max(best);
        T wmax = rl(rmax, 0);
#pragma unroll
        for (int r = 1; r < 4; ++r) {
          const T o = rl(rmax, 16 * r);
          wmax = o > wmax ? o : wmax;
        }
        int p, bl, bs;
        if (!(wmax >= (T)0)) {
          p = j;  // no finite candidate (all NaN): keep the row at position j
          bl = -1;
          bs = 0;
        } else {
          unsigned long long tie = __ballot(best == wmax);
          if (__popcll(tie) > 1) {
            // several lanes hold the maximum: the lowest position among them
            int q = (best == wmax) ? bpos : 0x7fff;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) q = min(q, __shfl_xor(q, o, 64));
            tie = __ballot(best == wmax && bpos == q);
          }
          bl = (int)__builtin_ctzll(tie);
          p = __builtin_amdgcn_readlane(bpos, bl);
          bs = __builtin_amdgcn_readlane(bslot, bl);
        }
        if (bl < 0) {  // locate the row at position j
          const unsigned long long at = __ballot(pos[0] == j || pos[1] == j || pos[2] == j ||
                                                 pos[3] == j);
          bl = at ? (int)__builtin_ctzll(at) : 0;
          const int sl = pos[1] == j ? 1 : pos[2] == j ? 2 : pos[3] == j ? 3 : 0;
          bs = __builtin_amdgcn_readlane(sl, bl);
        }
        pv[j] = p;
        // interchange positions j and p
#pragma unroll
        for (int s = 0; s < 4; ++s) pos[s] = pos[s] == j ? p : (pos[s] == p ? j : pos[s]);
        // the pivot row's entries from column j on (lane bl, slot bs)
        // (one uniform branch per slot: a select between slots here would be
        // folded into an indexed load and put the panel in scratch memory)
        T u[kWp];
        auto grab = [&](auto S) __attribute__((always_inline)) {
          constexpr int sl = decltype(S)::value;
#pragma unroll
          for (int c = j; c < kWp; ++c) u[c] = rl(a[sl][c], bl);
        };
        if (bs == 0) grab(std::integral_constant<int, 0>{});
        else if (bs == 1) grab(std::integral_constant<int, 1>{});
        else if (bs == 2) grab(std::integral_constant<int, 2>{});
        else grab(std::integral_constant<int, 3>{});
        const T piv_v = u[j];
        if (piv_v == (T)0 && first_zero == 0 && c0 + j < n) first_zero = c0 + j + 1;
        const T inv = piv_v != (T)0 ? (T)1 / piv_v : (T)0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if (pos[s] > j && pos[s] < m) {
            const T l = a[s][j] * inv;
            a[s][j] = l;
#pragma unroll
            for (int c = j + 1; c < kWp; ++c) a[s][c] = fma(-l, u[c], a[s][c]);
          }
        }
      });
      // panel -> factor buffer (real part) and LDS (all m rows), each row at
      // its final position; sidx[position] = the local row loaded there
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int i = lane + 64 * s;
        if (i < m) {
          const int P = pos[s];
          sidx[P] = i;
          if (npad == n) {
            // every row and column is real: the stores need no guards (a
            // guard per store makes each wait for the one before)
            T* __restrict__ dst = M + c0 + P + (long)c0 * n;
#pragma unroll
            for (int c = 0; c < kWp; ++c) {
              sL[c * kWs + P] = a[s][c];
              dst[(long)c * n] = a[s][c];
            }
          } else {
#pragma unroll
            for (int c = 0; c < kWp; ++c) {
              sL[c * kWs + P] = a[s][c];
              if (c0 + P < n && c0 + c < n) M[c0 + P + (long)(c0 + c) * n] = a[s][c];
            }
          }
        }
      }
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < kWp; ++j)
          if (c0 + j < n) piv[xo + c0 + j] = c0 + pv[j];
      }
      __syncthreads();  // sidx complete
      // position r now holds the row loaded from local row sidx[r]
      int src_top[kWp];
#pragma unroll
      for (int r = 0; r < kWp; ++r) src_top[r] = sidx[r];
      int src_low[kWp];  // for pivot positions below the top 16
#pragma unroll
      for (int j = 0; j < kWp; ++j) src_low[j] = sidx[pv[j]];
      // global row permutation (row k of LU is row sperm[k] of A)
      {
        int g[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int i = lane + 64 * s;
          g[s] = (i < m) ? sperm[c0 + sidx[i]] : 0;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int i = lane + 64 * s;
          if (i < m) sperm[c0 + i] = g[s];
        }
      }
      __syncthreads();  // panel and L11 visible in LDS and memory
      // ---- 2. swaps outside the panel; TRSM right of it ----------------------
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = lane + 64 * q;
        const bool left = col < c0, right = col >= c0 + kWp;
        if ((left || right) && col < n) {
        T* __restrict__ cp = M + (long)col * n;
        T xt[kWp], xl[kWp];
        if (npad == n) {
          // every row is real: unconditional loads (the values of pivots
          // inside the top 16 are not stored)
#pragma unroll
          for (int r = 0; r < kWp; ++r) xt[r] = cp[c0 + src_top[r]];
#pragma unroll
          for (int j = 0; j < kWp; ++j) xl[j] = cp[c0 + src_low[j]];
        } else {
#pragma unroll
          for (int r = 0; r < kWp; ++r) {
            const int R = c0 + src_top[r];
            xt[r] = R < n ? cp[R] : (T)0;
          }
#pragma unroll
          for (int j = 0; j < kWp; ++j) {
            const int R = c0 + src_low[j];
            xl[j] = (pv[j] >= kWp && R < n) ? cp[R] : (T)0;
          }
        }
        if (right) {
          // forward substitution with the unit-lower L11 (LDS, broadcast)
          sfor<0, kWp - 1>([&](auto K) __attribute__((always_inline)) {
            constexpr int k = decltype(K)::value;
            sfor<k + 1, kWp>([&](auto I) __attribute__((always_inline)) {
              constexpr int i = decltype(I)::value;
              xt[i] = fma(-sL[k * kWs + i], xt[k], xt[i]);
            });
          });
        }
        if (npad == n) {
#pragma unroll
          for (int r = 0; r < kWp; ++r) cp[c0 + r] = xt[r];
          // the rows swapped in from below the top 16; a pivot inside the
          // top 16 stores top row j's own value again (no branch per store)
#pragma unroll
          for (int j = 0; j < kWp; ++j) {
            const bool low = pv[j] >= kWp;
            cp[c0 + (low ? pv[j] : j)] = low ? xl[j] : xt[j];
          }
        } else {
#pragma unroll
          for (int r = 0; r < kWp; ++r)
            if (c0 + r < n) cp[c0 + r] = xt[r];
#pragma unroll
          for (int j = 0; j < kWp; ++j)
            if (pv[j] >= kWp && c0 + pv[j] < n) cp[c0 + pv[j]] = xl[j];
        }
        }
      }
      __syncthreads();  // U12 and the swapped rows visible to the tile loads
      // ---- 3. trailing update on MFMA (transposed tiles) --------------------
      // Tiles walk down a tile column, then right.  A lane's four elements of
      // a tile sit at tile base + lc[e]; orders that are multiples of 16 take
      // the branch without padding checks.
      const int nr = (npad - c0 - kWp) / kWp;  // trailing tiles per dimension
      const int T_all = nr * nr;
      if (T_all > 0) {
        const int g = lane >> 4, e16 = lane & 15;
        auto trailing = [&](auto FULLc) __attribute__((always_inline)) {
          constexpr bool FULL = decltype(FULLc)::value;
          int lc[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) lc[e] = e16 + Mf16<T>::row(g, e) * n;
          const int R00 = c0 + kWp;  // first trailing row and column
          auto load = [&](int rt, int ct) __attribute__((always_inline)) {
            v4 v;
            const int R0 = R00 + kWp * rt, C0 = R00 + kWp * ct;
#pragma unroll
            for (int e = 0; e < 4; ++e)
              v[e] = FULL ? M[R0 + C0 * n + lc[e]]
                          : ld_pad(M, n, R0 + e16, C0 + Mf16<T>::row(g, e));
            return v;
          };
          auto next = [&](int& rt, int& ct) __attribute__((always_inline)) {
            if (++rt == nr) {
              rt = 0;
              ++ct;
            }
          };
          if constexpr (FULL) {
            // One tile column at a time: the column's U12 strip, then all
            // its tiles (rows past the last tile reload the last one, in
            // bounds and unused), issue before its first MFMA with no branch
            // between, so a column waits for memory once.
            // tiles loaded together: all 16 >= the 15 trailing tiles of a
            // column in FP64 (accumulators in AGPRs), 4 in FP32 (two waves
            // per SIMD share the register file)
            constexpr int TH = sizeof(T) == 8 ? kWmax / kWp : kWmax / kWp / 4;
            for (int ct = 0; ct < nr; ++ct) {
              const int C0 = R00 + kWp * ct;
              T ua[4];
#pragma unroll
              for (int q = 0; q < 4; ++q) ua[q] = -M[c0 + g + 4 * q + (C0 + e16) * n];
#pragma unroll
              for (int h = 0; h < kWmax / kWp; h += TH) {
                if (h < nr) {
                  v4 tile[TH];
#pragma unroll
                  for (int u = 0; u < TH; ++u)
                    tile[u] = load(h + u < nr ? h + u : nr - 1, ct);
                  // four tiles at a time, their MFMA chains interleaved; a
                  // tile past the last recomputes the last one from the same
                  // loaded values and stores the same result again
#pragma unroll
                  for (int r4 = 0; r4 < TH; r4 += 4) {
                    if (h + r4 < nr) {
                      v4 acc[4];
                      const T* __restrict__ lb[4];
                      int rts[4];
#pragma unroll
                      for (int u = 0; u < 4; ++u) {
                        acc[u] = tile[r4 + u];
                        rts[u] = h + r4 + u < nr ? h + r4 + u : nr - 1;
                        lb[u] = sL + (R00 + kWp * rts[u] - c0) + e16;
                      }
#pragma unroll
                      for (int q = 0; q < 4; ++q)
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                          acc[u] = Mf16<T>::mma(ua[q], lb[u][(g + 4 * q) * kWs], acc[u]);
#pragma unroll
                      for (int u = 0; u < 4; ++u) {
                        const int R0 = R00 + kWp * rts[u];
#pragma unroll
                        for (int e = 0; e < 4; ++e) M[R0 + C0 * n + lc[e]] = acc[u][e];
                      }
                    }
                  }
                }
              }
            }
            return;
          }
          // (orders not a multiple of 16: tiles with loads in flight in a ring)
          v4 ring[kRing];
          int l_rt = 0, l_ct = 0;  // next tile to load
#pragma unroll
          for (int u = 0; u < kRing; ++u)
            if (u < T_all) {
              ring[u] = load(l_rt, l_ct);
              next(l_rt, l_ct);
            }
          int c_rt = 0, c_ct = 0;  // tile being multiplied
          T ua[4];
          for (int t0 = 0; t0 < T_all; t0 += kRing) {
#pragma unroll
            for (int u = 0; u < kRing; ++u) {
              const int t = t0 + u;
              if (t < T_all) {
                v4 acc = ring[u];
                if (t + kRing < T_all) {
                  ring[u] = load(l_rt, l_ct);
                  next(l_rt, l_ct);
                }
                const int R0 = R00 + kWp * c_rt, C0 = R00 + kWp * c_ct;
                if (c_rt == 0) {
                  // A'(a, k) = -U12(k, C0 + a): lane provides a = e16, k = g + 4 s
#pragma unroll
                  for (int s = 0; s < 4; ++s)
                    ua[s] = FULL ? -M[c0 + g + 4 * s + (C0 + e16) * n]
                                 : -ld_pad(M, n, c0 + g + 4 * s, C0 + e16);
                }
                const T* __restrict__ lb = sL + (R0 - c0) + e16;
#pragma unroll
                for (int s = 0; s < 4; ++s) acc = Mf16<T>::mma(ua[s], lb[(g + 4 * s) * kWs], acc);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                  if (FULL) {
                    M[R0 + C0 * n + lc[e]] = acc[e];
                  } else {
                    const int R = R0 + e16, C = C0 + Mf16<T>::row(g, e);
                    if (R < n && C < n) M[R + C * n] = acc[e];
                  }
                }
                next(c_rt, c_ct);
              }
            }
          }
        };
        if (npad == n)
          trailing(std::integral_constant<bool, true>{});
        else
          trailing(std::integral_constant<bool, false>{});
      }
      __syncthreads();  // trailing matrix visible to the next panel's loads
    }

    // ---- outputs: permutation, scales, info --------------------------------
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int r = lane + 64 * s;
      if (r < n) {
        perm[xo + r] = sperm[r];
        if (kEq) {
          rs[xo + r] = sr[s];
          cs[xo + r] = scl_c[r];
        }
      }
    }
    if (lane == 0) info[b] = first_zero;
    __syncthreads();  // LDS reuse by the next block
  }
}

// FP64: one wave per SIMD (registers); FP32: two.
__global__ void __launch_bounds__(64, 1)
lu_wave_f64_kernel(int nb, const int* __restrict__ n_arr, const long* __restrict__ a_off,
                   const long* __restrict__ x_off, const int* __restrict__ order,
                   const double* __restrict__ A, double* __restrict__ LU, int* __restrict__ piv,
                   int* __restrict__ perm, int* __restrict__ info) {
  lu_wave_body<double>(nb, n_arr, a_off, x_off, order, A, LU, nullptr, nullptr, piv, perm, info);
}
__global__ void __launch_bounds__(64, 2)
lu_wave_f32_kernel(int nb, const int* __restrict__ n_arr, const long* __restrict__ a_off,
                   const long* __restrict__ x_off, const int* __restrict__ order,
                   const double* __restrict__ A, float* __restrict__ LU, double* __restrict__ rs,
                   double* __restrict__ cs, int* __restrict__ piv, int* __restrict__ perm,
                   int* __restrict__ info) {
  lu_wave_body<float>(nb, n_arr, a_off, x_off, order, A, LU, rs, cs, piv, perm, info);
}

int launch_lu_wave(int precision, int64_t nb, const int32_t* n, const int64_t* a_off,
                   const int64_t* x_off, const int32_t* order, const double* A, double* lu64,
                   float* lu32, double* rs, double* cs, int32_t* piv, int32_t* perm,
                   int32_t* info, int num_cu, int per_cu, hipStream_t st) {
  // per_cu: waves in flight per CU (LDS holds at most 4 FP64 panels of 35 KB
  // or 9 FP32 ones)
  const long g = (long)per_cu * num_cu;
  const unsigned grid = (unsigned)(nb < g ? nb : g);
  if (precision == F3D_F64) {
    F3D_LAUNCH(lu_wave_f64_kernel, dim3(grid), dim3(64), 0, st, (int)nb, n, (const long*)a_off,
               (const long*)x_off, order, A, lu64, piv, perm, info);
  } else {
    F3D_LAUNCH(lu_wave_f32_kernel, dim3(grid), dim3(64), 0, st, (int)nb, n, (const long*)a_off,
               (const long*)x_off, order, A, lu32, rs, cs, piv, perm, info);
  }
  return F3D_OK;
}

}  // namespace f3d
