// Source-target Stokes pair sums (Oseen / rotlet / stresslet) for gfx950.
//
// Reference semantics: python/kernels/kernels.py:794-846 (Oseen),
// :856-908 (rotlet), :1186-1217 (stresslet source-target), :1123-1183
// (stresslet self-sum).  For a pair, d = t - s, r = |d|; if r < eps the
// distance is regularised, r <- sqrt(r^2 + reg^2), while d is kept.
//
// Design (MI355X / CDNA4, wave64):
//   * pack_kernel converts the caller's sources (any AoS/SoA strides, FP64) into
//     one contiguous record per source (8 or 12 reals: position, density[,
//     normal]), multiplying an optional per-source weight, rounding to the
//     compute type, and zero-padding to a multiple of the 64-source tile.
//   * pairsum_kernel: each lane owns K targets (registers).  The source index
//     is wave-uniform, so source records are fetched with scalar loads into
//     SGPRs and every VALU op takes its source operand straight from an SGPR:
//     no LDS traffic and no VGPRs spent on sources.  Sources are split into
//     P chunks over grid.y so the launch carries thousands of waves.
//   * FP32C: pairs are summed in plain FP32 within a 64-source tile, tile sums
//     are added with Kahan compensation, and (sum, -compensation) are written
//     as FP32 partials per chunk; reduce_kernel combines chunks in FP64.
//   * Exclusion ranges [lo_i, hi_i) (same fiber / self term) are tested only in
//     tiles that intersect the wave's union of ranges; all other tiles run a
//     mask-free path.
#include "common.h"

#include <stdarg.h>
#include <stdlib.h>

namespace f3d {

constexpr int kTile = 64;        // sources per compensated tile
constexpr int kBlock = 256;      // threads per block (4 waves)
constexpr int kPackBlocks = 256; // grid of the pack kernel (sigma partials)

template <int KIND>
struct RecLen {
  static constexpr int value = (KIND == F3D_STRESSLET) ? 12 : 8;
};

__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }
// v_rsq_f64 is accurate to about 2^-22 relative; one Newton step brings it to
// ~1e-14 relative, below the FP64 parity bar of 1e-10.
__device__ __forceinline__ double rsqrt_t(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double h = x * y;
  double e = fma(-h, y, 1.0);
  return fma(0.5 * y, e, y);
}

// ---------------------------------------------------------------------------
// pack: sources -> records.  Also accumulates sigma = sum_j n_j . q_j per
// block when sigma_part != nullptr (periphery complementary term).
template <typename T, int KIND>
__global__ void __launch_bounds__(256)
pack_kernel(int ns, int ns_pad, const double* __restrict__ pos, const double* __restrict__ nrm,
            const double* __restrict__ den, const double* __restrict__ w, long sp, long sc,
            T* __restrict__ rec, double* __restrict__ sigma_part) {
  constexpr int R = RecLen<KIND>::value;
  double sig = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ns_pad; j += gridDim.x * blockDim.x) {
    T* r = rec + (size_t)j * R;
    if (j < ns) {
      const double wj = w ? w[j] : 1.0;
      const long b = (long)j * sp;
      r[0] = (T)pos[b];
      r[1] = (T)pos[b + sc];
      r[2] = (T)pos[b + 2 * sc];
      if (KIND == F3D_STRESSLET) {
        const double n0 = nrm[b], n1 = nrm[b + sc], n2 = nrm[b + 2 * sc];
        const double q0 = den[b] * wj, q1 = den[b + sc] * wj, q2 = den[b + 2 * sc] * wj;
        r[3] = (T)n0; r[4] = (T)n1; r[5] = (T)n2;
        r[6] = (T)q0; r[7] = (T)q1; r[8] = (T)q2;
        r[9] = 0; r[10] = 0; r[11] = 0;
        sig += n0 * q0 + n1 * q1 + n2 * q2;
      } else {
        r[3] = (T)(den[b] * wj);
        r[4] = (T)(den[b + sc] * wj);
        r[5] = (T)(den[b + 2 * sc] * wj);
        r[6] = 0; r[7] = 0;
      }
    } else {
      // Zero-density padding: contributes exactly zero (finite r since the
      // regularisation keeps r >= reg when a target sits at the origin).
#pragma unroll
      for (int c = 0; c < R; ++c) r[c] = 0;
    }
  }
  if (sigma_part) {
    __shared__ double red[4];
    sig = wave_sum_f64(sig);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sig;
    __syncthreads();
    if (threadIdx.x == 0) sigma_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
  }
}

// ---------------------------------------------------------------------------
// One pair: accumulate the un-scaled kernel contribution of record r into
// (ax, ay, az).  MASK: zero the pair when excl is set.  CLAMP: cap 1/r at
// rmax instead of the exact regularisation.
template <typename T, int KIND, bool CLAMP, bool MASK>
__device__ __forceinline__ void pair(const T* r, T tx, T ty, T tz, T reg2, T eps2,
                                     T rmax, bool excl, T& ax, T& ay, T& az) {
  const T dx = tx - r[0];
  const T dy = ty - r[1];
  const T dz = tz - r[2];
  T r2 = dx * dx + dy * dy + dz * dz;
  if (!CLAMP) r2 = (r2 < eps2) ? r2 + reg2 : r2;
  T ri = rsqrt_t(r2);
  if (CLAMP) ri = ri < rmax ? ri : rmax;
  if (MASK) ri = excl ? (T)0 : ri;
  if (KIND == F3D_OSEEN) {
    const T fd = dx * r[3] + dy * r[4] + dz * r[5];
    const T s = fd * (ri * ri);
    ax += ri * (r[3] + s * dx);
    ay += ri * (r[4] + s * dy);
    az += ri * (r[5] + s * dz);
  } else if (KIND == F3D_ROTLET) {
    const T ri3 = ri * ri * ri;
    ax += ri3 * (r[4] * dz - r[5] * dy);
    ay += ri3 * (r[5] * dx - r[3] * dz);
    az += ri3 * (r[3] * dy - r[4] * dx);
  } else {
    const T ri2 = ri * ri;
    const T dn = dx * r[3] + dy * r[4] + dz * r[5];
    const T dq = dx * r[6] + dy * r[7] + dz * r[8];
    const T f0 = (dq * ri2) * (dn * ri2) * ri;
    ax += f0 * dx;
    ay += f0 * dy;
    az += f0 * dz;
  }
}

template <typename T, int KIND, int K, bool CLAMP, bool MASK>
__device__ __forceinline__ void tile(const T* __restrict__ rec, int j0, const T (&tx)[K],
                                     const T (&ty)[K], const T (&tz)[K], const int (&lo)[K],
                                     const int (&hi)[K], T reg2, T eps2, T rmax, T (&ax)[K],
                                     T (&ay)[K], T (&az)[K]) {
  constexpr int R = RecLen<KIND>::value;
  constexpr int U = 4;  // sources whose records are fetched together
  constexpr int NV = (KIND == F3D_STRESSLET) ? 9 : 6;
  for (int jj = 0; jj < kTile; jj += U) {
    // Issue the U scalar record loads before any use so their latencies overlap.
    T s[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T* __restrict__ r = rec + (size_t)(j0 + jj + u) * R;
#pragma unroll
      for (int c = 0; c < NV; ++c) s[u][c] = r[c];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + jj + u;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const bool excl = MASK && ((unsigned)(j - lo[k]) < (unsigned)(hi[k] - lo[k]));
        pair<T, KIND, CLAMP, MASK>(s[u], tx[k], ty[k], tz[k], reg2, eps2, rmax, excl, ax[k],
                                   ay[k], az[k]);
      }
    }
  }
}

// partial layout: part[(p*3 + c)*nt + i]; for T=float a second array holds
// the negated Kahan compensation.
template <typename T, int KIND, int K, bool CLAMP>
__global__ void __launch_bounds__(kBlock)
pairsum_kernel(const T* __restrict__ rec, int chunk, int ns_pad, const double* __restrict__ tpos,
               long t_sp, long t_sc, int nt, const int* __restrict__ excl_lo,
               const int* __restrict__ excl_hi, T reg2, T eps2, T rmax, T* __restrict__ part_hi,
               T* __restrict__ part_lo) {
  const int base = blockIdx.x * (kBlock * K) + threadIdx.x;
  T tx[K], ty[K], tz[K];
  int lo[K], hi[K];
  int wlo = 0x7fffffff, whi = -0x7fffffff;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = base + k * kBlock;
    const int ic = i < nt ? i : nt - 1;
    const long b = (long)ic * t_sp;
    tx[k] = (T)tpos[b];
    ty[k] = (T)tpos[b + t_sc];
    tz[k] = (T)tpos[b + 2 * t_sc];
    lo[k] = 0;
    hi[k] = 0;
    if (excl_lo != nullptr && i < nt) {
      lo[k] = excl_lo[i];
      hi[k] = excl_hi[i];
      if (hi[k] > lo[k]) {
        wlo = min(wlo, lo[k]);
        whi = max(whi, hi[k]);
      }
    }
  }
  // Wave-uniform union of the exclusion ranges (kept in SGPRs).
  wlo = __builtin_amdgcn_readfirstlane(wave_min_i32(wlo));
  whi = __builtin_amdgcn_readfirstlane(wave_max_i32(whi));

  T sx[K], sy[K], sz[K], cx[K], cy[K], cz[K];
#pragma unroll
  for (int k = 0; k < K; ++k) sx[k] = sy[k] = sz[k] = cx[k] = cy[k] = cz[k] = 0;

  const int j_begin = blockIdx.y * chunk;
  const int j_end = min(j_begin + chunk, ns_pad);
  for (int j0 = j_begin; j0 < j_end; j0 += kTile) {
    T ax[K], ay[K], az[K];
#pragma unroll
    for (int k = 0; k < K; ++k) ax[k] = ay[k] = az[k] = 0;
    if (j0 < whi && j0 + kTile > wlo)
      tile<T, KIND, K, CLAMP, true>(rec, j0, tx, ty, tz, lo, hi, reg2, eps2, rmax, ax, ay, az);
    else
      tile<T, KIND, K, CLAMP, false>(rec, j0, tx, ty, tz, lo, hi, reg2, eps2, rmax, ax, ay, az);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (sizeof(T) == 4) {
        // Kahan: add the tile sum a into (s, c).
        T y, t;
        y = ax[k] - cx[k]; t = sx[k] + y; cx[k] = (t - sx[k]) - y; sx[k] = t;
        y = ay[k] - cy[k]; t = sy[k] + y; cy[k] = (t - sy[k]) - y; sy[k] = t;
        y = az[k] - cz[k]; t = sz[k] + y; cz[k] = (t - sz[k]) - y; sz[k] = t;
      } else {
        sx[k] += ax[k];
        sy[k] += ay[k];
        sz[k] += az[k];
      }
    }
  }
  const size_t p = blockIdx.y;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = base + k * kBlock;
    if (i < nt) {
      part_hi[(p * 3 + 0) * nt + i] = sx[k];
      part_hi[(p * 3 + 1) * nt + i] = sy[k];
      part_hi[(p * 3 + 2) * nt + i] = sz[k];
      if (sizeof(T) == 4) {
        part_lo[(p * 3 + 0) * nt + i] = -cx[k];
        part_lo[(p * 3 + 1) * nt + i] = -cy[k];
        part_lo[(p * 3 + 2) * nt + i] = -cz[k];
      }
    }
  }
}

// Periphery epilogue terms (optional): see f3d_periphery_apply.
struct ShellEpi {
  const double* nrm;
  const double* w;
  const double* ex;
  const double* ey;
  const double* ez;
  const double* q;
  const double* add;
  const double* sigma_part;  // kPackBlocks partial sums of n.q
};

template <typename T>
__global__ void __launch_bounds__(256)
reduce_kernel(const T* __restrict__ part_hi, const T* __restrict__ part_lo, int P, int nt,
              double factor, double* __restrict__ out, long o_sp, long o_sc, int accumulate,
              ShellEpi epi) {
  double sigma = 0.0;
  if (epi.sigma_part) {
    __shared__ double red[4];
    double v = epi.sigma_part[threadIdx.x];  // blockDim.x == kPackBlocks
    v = wave_sum_f64(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    sigma = red[0] + red[1] + red[2] + red[3];
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  double u[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int p = 0; p < P; ++p) {
      const size_t idx = ((size_t)p * 3 + c) * nt + i;
      acc += (double)part_hi[idx];
      if (sizeof(T) == 4) acc += (double)part_lo[idx];
    }
    u[c] = factor * acc;
  }
  const long b = (long)i * o_sp;
  if (epi.q) {
    const double wi = epi.w[i];
    const double q0 = epi.q[b], q1 = epi.q[b + o_sc], q2 = epi.q[b + 2 * o_sc];
    const double a0 = q0 / wi, a1 = q1 / wi, a2 = q2 / wi;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long bc = b + c * o_sc;
      u[c] -= a0 * epi.ex[bc] + a1 * epi.ey[bc] + a2 * epi.ez[bc];
      u[c] -= epi.q[bc] / wi;
      u[c] += epi.nrm[bc] * sigma;
      if (epi.add) u[c] += epi.add[bc];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long bc = b + c * o_sc;
    out[bc] = accumulate ? out[bc] + u[c] : u[c];
  }
}

// ---------------------------------------------------------------------------
// Host-side plan: split count and chunk length.
struct Plan {
  int K;
  int ns_pad;
  int P;
  int chunk;
};

static int env_int(const char* name, int dflt) {
  const char* s = getenv(name);
  if (!s || !*s) return dflt;
  return atoi(s);
}

static Plan make_plan(int precision, int64_t ns, int64_t nt) {
  Plan pl;
  pl.K = env_int(precision == F3D_F64 ? "F3D_PAIRSUM_K64" : "F3D_PAIRSUM_K32", 2);
  if (pl.K != 1 && pl.K != 2 && pl.K != 4) pl.K = 2;
  pl.ns_pad = (int)align_up((size_t)(ns > 0 ? ns : 1), kTile);
  const int64_t waves_per_split = ceil_div(nt > 0 ? nt : 1, 64 * pl.K);
  const int target_waves = env_int("F3D_PAIRSUM_WAVES", 8192);
  const int64_t tiles = pl.ns_pad / kTile;
  int64_t P = ceil_div(target_waves, waves_per_split);
  if (P < 1) P = 1;
  if (P > tiles) P = tiles;
  const int64_t tiles_per = ceil_div(tiles, P);
  pl.chunk = (int)(tiles_per * kTile);
  pl.P = (int)ceil_div(pl.ns_pad, pl.chunk);
  return pl;
}

static size_t rec_bytes(int kernel, int precision, int ns_pad) {
  const size_t r = kernel == F3D_STRESSLET ? 12 : 8;
  const size_t t = precision == F3D_F64 ? 8 : 4;
  return align_up((size_t)ns_pad * r * t, 256);
}

static size_t part_bytes(int precision, int P, int64_t nt) {
  const size_t per = precision == F3D_F64 ? 8 : 8;  // f64, or 2 x f32
  return align_up((size_t)P * 3 * (size_t)(nt > 0 ? nt : 1) * per, 256);
}

static size_t pairsum_ws(int kernel, int precision, int64_t ns, int64_t nt) {
  const Plan pl = make_plan(precision, ns, nt);
  return rec_bytes(kernel, precision, pl.ns_pad) + part_bytes(precision, pl.P, nt) +
         align_up(kPackBlocks * sizeof(double), 256);
}

template <typename T, int KIND>
static int launch_pairsum(const Plan& pl, uint32_t flags, int64_t ns, const double* pos,
                          const double* nrm, const double* den, const double* w, int64_t s_sp,
                          int64_t s_sc, int64_t nt, const double* tpos, int64_t t_sp,
                          int64_t t_sc, const int32_t* lo, const int32_t* hi, double eta,
                          double reg, double eps, double* out, int64_t o_sp, int64_t o_sc,
                          char* ws, const ShellEpi* epi_in, hipStream_t st) {
  T* rec = reinterpret_cast<T*>(ws);
  char* p = ws + rec_bytes(KIND, sizeof(T) == 8 ? F3D_F64 : F3D_F32C, pl.ns_pad);
  T* part_hi = reinterpret_cast<T*>(p);
  T* part_lo = part_hi + (size_t)pl.P * 3 * nt;
  double* sigma_part = reinterpret_cast<double*>(
      p + part_bytes(sizeof(T) == 8 ? F3D_F64 : F3D_F32C, pl.P, nt));

  const bool want_sigma = epi_in != nullptr;
  F3D_LAUNCH((pack_kernel<T, KIND>), dim3(kPackBlocks), dim3(256), 0, st, (int)ns, pl.ns_pad,
             pos, nrm, den, w, (long)s_sp, (long)s_sc, rec, want_sigma ? sigma_part : nullptr);

  const T reg2 = (T)(reg * reg);
  const T eps2 = (T)(eps * eps);
  const T rmax = (T)(1.0 / eps);
  const dim3 grid((unsigned)ceil_div(nt, (int64_t)kBlock * pl.K), (unsigned)pl.P);
  const bool clamp = (sizeof(T) == 4) && (flags & F3D_CLAMP_NEAR);
#define F3D_PS(KK, CL)                                                                        \
  F3D_LAUNCH((pairsum_kernel<T, KIND, KK, CL>), grid, dim3(kBlock), 0, st, rec, pl.chunk,     \
             pl.ns_pad, tpos, (long)t_sp, (long)t_sc, (int)nt, lo, hi, reg2, eps2, rmax,      \
             part_hi, part_lo)
  if (clamp) {
    if (pl.K == 1) F3D_PS(1, true);
    else if (pl.K == 2) F3D_PS(2, true);
    else F3D_PS(4, true);
  } else {
    if (pl.K == 1) F3D_PS(1, false);
    else if (pl.K == 2) F3D_PS(2, false);
    else F3D_PS(4, false);
  }
#undef F3D_PS

  double factor;
  if (KIND == F3D_STRESSLET) factor = -3.0 / (4.0 * kPi * eta);
  else factor = 1.0 / (8.0 * kPi * eta);
  ShellEpi epi{};
  if (epi_in) {
    epi = *epi_in;
    epi.sigma_part = sigma_part;
  }
  F3D_LAUNCH((reduce_kernel<T>), dim3((unsigned)ceil_div(nt, 256)), dim3(256), 0, st, part_hi,
             part_lo, pl.P, (int)nt, factor, out, (long)o_sp, (long)o_sc,
             (flags & F3D_ACCUMULATE) ? 1 : 0, epi);
  return F3D_OK;
}

static int pairsum_impl(int kernel, int precision, uint32_t flags, int64_t ns, const double* pos,
                        const double* nrm, const double* den, const double* w, int64_t s_sp,
                        int64_t s_sc, int64_t nt, const double* tpos, int64_t t_sp, int64_t t_sc,
                        const int32_t* lo, const int32_t* hi, double eta, double reg, double eps,
                        double* out, int64_t o_sp, int64_t o_sc, void* ws, size_t ws_bytes,
                        const ShellEpi* epi, void* stream) {
  F3D_REQUIRE(kernel == F3D_OSEEN || kernel == F3D_ROTLET || kernel == F3D_STRESSLET, F3D_EINVAL,
              "pairsum: unknown kernel %d", kernel);
  F3D_REQUIRE(precision == F3D_F64 || precision == F3D_F32C, F3D_EINVAL,
              "pairsum: unknown precision %d", precision);
  F3D_REQUIRE(ns >= 0 && nt >= 0, F3D_EINVAL, "pairsum: negative size");
  F3D_REQUIRE(ns < (1ll << 30) && nt < (1ll << 30), F3D_EINVAL,
              "pairsum: ns=%lld nt=%lld exceed 2^30", (long long)ns, (long long)nt);
  F3D_REQUIRE((lo == nullptr) == (hi == nullptr), F3D_EINVAL,
              "pairsum: excl_lo and excl_hi must both be set or both NULL");
  F3D_REQUIRE(eta > 0.0, F3D_EINVAL, "pairsum: eta must be positive");
  F3D_REQUIRE(eps >= 0.0 && reg >= 0.0, F3D_EINVAL, "pairsum: eps and reg must be >= 0");
  if (nt == 0) return F3D_OK;
  F3D_REQUIRE(tpos != nullptr && out != nullptr, F3D_EINVAL, "pairsum: NULL target/output");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (ns == 0 && epi == nullptr) {
    if (flags & F3D_ACCUMULATE) return F3D_OK;
    // out = 0 on the caller's layout: reuse the reduce kernel with P = 0.
    F3D_LAUNCH((reduce_kernel<double>), dim3((unsigned)ceil_div(nt, 256)), dim3(256), 0, st,
               (const double*)nullptr, (const double*)nullptr, 0, (int)nt, 0.0, out, (long)o_sp,
               (long)o_sc, 0, ShellEpi{});
    return F3D_OK;
  }
  F3D_REQUIRE(pos != nullptr && den != nullptr, F3D_EINVAL, "pairsum: NULL source arrays");
  F3D_REQUIRE(kernel != F3D_STRESSLET || nrm != nullptr, F3D_EINVAL,
              "pairsum: stresslet needs normals");
  const size_t need = pairsum_ws(kernel, precision, ns, nt);
  F3D_REQUIRE(ws != nullptr && ws_bytes >= need, F3D_EWORKSPACE,
              "pairsum: workspace %zu bytes < required %zu", ws_bytes, need);
  const Plan pl = make_plan(precision, ns, nt);
  char* wsc = reinterpret_cast<char*>(ws);
#define F3D_ARGS pl, flags, ns, pos, nrm, den, w, s_sp, s_sc, nt, tpos, t_sp, t_sc, lo, hi, eta, \
                 reg, eps, out, o_sp, o_sc, wsc, epi, st
  if (precision == F3D_F64) {
    if (kernel == F3D_OSEEN) return launch_pairsum<double, F3D_OSEEN>(F3D_ARGS);
    if (kernel == F3D_ROTLET) return launch_pairsum<double, F3D_ROTLET>(F3D_ARGS);
    return launch_pairsum<double, F3D_STRESSLET>(F3D_ARGS);
  } else {
    if (kernel == F3D_OSEEN) return launch_pairsum<float, F3D_OSEEN>(F3D_ARGS);
    if (kernel == F3D_ROTLET) return launch_pairsum<float, F3D_ROTLET>(F3D_ARGS);
    return launch_pairsum<float, F3D_STRESSLET>(F3D_ARGS);
  }
#undef F3D_ARGS
}

}  // namespace f3d

using namespace f3d;

extern "C" size_t f3d_pairsum_workspace_size(int kernel, int precision, int64_t ns, int64_t nt) {
  if (precision != F3D_F64 && precision != F3D_F32C) return 0;
  return pairsum_ws(kernel, precision, ns, nt);
}

extern "C" int f3d_pairsum(int kernel, int precision, uint32_t flags, int64_t ns,
                           const double* src_pos, const double* src_nrm, const double* src_den,
                           const double* src_w, int64_t s_sp, int64_t s_sc, int64_t nt,
                           const double* trg_pos, int64_t t_sp, int64_t t_sc,
                           const int32_t* excl_lo, const int32_t* excl_hi, double eta, double reg,
                           double eps, double* out, int64_t o_sp, int64_t o_sc, void* ws,
                           size_t ws_bytes, void* stream) {
  return pairsum_impl(kernel, precision, flags, ns, src_pos, src_nrm, src_den, src_w, s_sp, s_sc,
                      nt, trg_pos, t_sp, t_sc, excl_lo, excl_hi, eta, reg, eps, out, o_sp, o_sc,
                      ws, ws_bytes, nullptr, stream);
}

extern "C" size_t f3d_periphery_workspace_size(int precision, int64_t n) {
  if (precision != F3D_F64 && precision != F3D_F32C) return 0;
  return pairsum_ws(F3D_STRESSLET, precision, n, n);
}

extern "C" int f3d_periphery_apply(int precision, int64_t n, const double* pos,
                                   const double* nrm, const double* w, const double* ex,
                                   const double* ey, const double* ez, const double* q,
                                   const double* add, int64_t sp, int64_t sc, double eta,
                                   double reg, double eps, double* out, void* ws, size_t ws_bytes,
                                   void* stream) {
  F3D_REQUIRE(n > 0, F3D_EINVAL, "periphery_apply: n must be positive");
  F3D_REQUIRE(nrm && w && ex && ey && ez && q, F3D_EINVAL, "periphery_apply: NULL argument");
  ShellEpi epi{nrm, w, ex, ey, ez, q, add, nullptr};
  return pairsum_impl(F3D_STRESSLET, precision, 0u, n, pos, nrm, q, nullptr, sp, sc, n, pos, sp,
                      sc, nullptr, nullptr, eta, reg, eps, out, sp, sc, ws, ws_bytes, &epi,
                      stream);
}

// ---------------------------------------------------------------------------
// Dense stresslet . normal matrix (python/kernels/kernels.py:1063-1120):
// block (i, j != i) = (-3/(4 pi eta)) (d . n_j) / r^5 d d^T, d = x_i - x_j,
// zero diagonal blocks; row-major 3n x 3n with leading dimension ld.
namespace f3d {
__global__ void __launch_bounds__(256)
stresslet_matrix_kernel(int n, const double* __restrict__ pos, const double* __restrict__ nrm,
                        long sp, long sc, double factor, double reg2, double eps2,
                        double* __restrict__ M, long ld) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const long bi = (long)i * sp, bj = (long)j * sp;
  const double dx = pos[bi] - pos[bj];
  const double dy = pos[bi + sc] - pos[bj + sc];
  const double dz = pos[bi + 2 * sc] - pos[bj + 2 * sc];
  double c = 0.0;
  if (i != j) {
    double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < eps2) r2 += reg2;
    const double r = sqrt(r2);
    const double dn = dx * nrm[bj] + dy * nrm[bj + sc] + dz * nrm[bj + 2 * sc];
    c = factor * dn / (r2 * r2 * r);
  }
  const double d[3] = {dx, dy, dz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double* row = M + (3L * i + a) * ld + 3L * j;
#pragma unroll
    for (int b = 0; b < 3; ++b) row[b] = c * d[a] * d[b];
  }
}
}  // namespace f3d

extern "C" int f3d_stresslet_matrix(int64_t n, const double* pos, const double* nrm, int64_t sp,
                                    int64_t sc, double eta, double reg, double eps, double* M,
                                    int64_t ld, void* stream) {
  F3D_REQUIRE(n >= 0 && n < (1ll << 30), F3D_EINVAL, "stresslet_matrix: bad n");
  F3D_REQUIRE(ld >= 3 * n, F3D_EINVAL, "stresslet_matrix: ld < 3n");
  F3D_REQUIRE(eta > 0.0, F3D_EINVAL, "stresslet_matrix: eta must be positive");
  if (n == 0) return F3D_OK;
  F3D_REQUIRE(pos && nrm && M, F3D_EINVAL, "stresslet_matrix: NULL argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  F3D_LAUNCH(stresslet_matrix_kernel, dim3((unsigned)ceil_div(n, 256), (unsigned)n), dim3(256), 0,
             st, (int)n, pos, nrm, (long)sp, (long)sc, -3.0 / (4.0 * kPi * eta), reg * reg,
             eps * eps, M, (long)ld);
  return F3D_OK;
}
