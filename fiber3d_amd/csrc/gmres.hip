// Restarted GMRES on the device: Arnoldi with classical Gram-Schmidt and one
// re-orthogonalisation (CGS2), the Hessenberg matrix, its Givens rotations
// and the least-squares residual estimate, all in FP64 and all in device
// memory, so that an Arnoldi step needs no host round trip and can be
// recorded in a HIP graph (the step index lives in device memory).
//
// Reference semantics: bpm_utilities/gmres.py:94-108 over SciPy's restarted
// GMRES, called at python/tstep/tstep.py:1209-1216 (right preconditioning,
// relative tolerance on ||b||, restart length m).
//
// State (one per solve, in the caller's workspace; see gmres_layout):
//   H    (m+1) x m rotated Hessenberg, column-major (ld m+1)
//   Hr   (m+1) x m unrotated Hessenberg: A P^-1 V_k = V_{k+1} Hr
//   cs, sn  Givens rotations;  g (m+1) rotated right-hand side
//   est  (m) |g_{k+1}| after step k;  h (m+1) Gram-Schmidt coefficients
//   part dot-product partial sums;  y (m+1) least-squares solution / Hr y
//   ctl  j (current step), breakdown flag, block counters
// Dot products are summed per block and the blocks' partials combined by
// the last block to finish (a device-scope counter), in a fixed order, so
// results do not depend on block scheduling.
#include "common.h"

namespace f3d {

constexpr int kGb = 256;       // threads per block
constexpr int kGchunk = 4096;  // vector elements per block in the dot kernels
constexpr int kGmaxBlocks = 1024;

struct GmresLayout {
  size_t H, Hr, cs, sn, g, est, h, part, y, ctl, total;
};

__host__ __device__ inline int gm_nchunks(int64_t n) {
  const int64_t c = (n + kGchunk - 1) / kGchunk;
  return (int)(c < kGmaxBlocks ? c : kGmaxBlocks);
}

__host__ __device__ inline GmresLayout gmres_layout(int64_t n, int m) {
  GmresLayout L;
  size_t o = 0;
  auto take = [&](size_t count) {
    const size_t at = o;
    o += (count * 8 + 255) / 256 * 256;
    return at;
  };
  L.H = take((size_t)(m + 1) * m);
  L.Hr = take((size_t)(m + 1) * m);
  L.cs = take(m);
  L.sn = take(m);
  L.g = take(m + 1);
  L.est = take(m);
  L.h = take(m + 1);
  // dot partials ((m+1) x chunks) and norm partials (one per update block)
  const size_t dots = (size_t)(m + 1) * gm_nchunks(n);
  L.part = take(dots > 2048 ? dots : 2048);
  L.y = take(2 * (m + 1));
  L.ctl = take(8);
  L.total = o;
  return L;
}

// control words (int32 view of ctl)
enum { CTL_J = 0, CTL_BREAK = 1, CTL_CNT0 = 2, CTL_CNT1 = 3, CTL_CNT2 = 4, CTL_CNT3 = 5 };

struct GmresPtrs {
  double *H, *Hr, *cs, *sn, *g, *est, *h, *part, *y;
  int* ctl;
};

__device__ __host__ inline GmresPtrs gmres_ptrs(void* ws, int64_t n, int m) {
  const GmresLayout L = gmres_layout(n, m);
  char* b = reinterpret_cast<char*>(ws);
  GmresPtrs p;
  p.H = reinterpret_cast<double*>(b + L.H);
  p.Hr = reinterpret_cast<double*>(b + L.Hr);
  p.cs = reinterpret_cast<double*>(b + L.cs);
  p.sn = reinterpret_cast<double*>(b + L.sn);
  p.g = reinterpret_cast<double*>(b + L.g);
  p.est = reinterpret_cast<double*>(b + L.est);
  p.h = reinterpret_cast<double*>(b + L.h);
  p.part = reinterpret_cast<double*>(b + L.part);
  p.y = reinterpret_cast<double*>(b + L.y);
  p.ctl = reinterpret_cast<int*>(b + L.ctl);
  return p;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum_f64(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kGb / 64; ++k) s += red[k];
  return s;
}

// Is this block the last of `total` to arrive at counter ctr?  (The caller
// wrote its partial results before; the fence orders them before the count.)
__device__ __forceinline__ bool last_block(int* ctr, int total) {
  __shared__ int last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int prev = atomicAdd(ctr, 1);
    last = (prev == total - 1);
    if (last) *ctr = 0;  // ready for the next launch (stream-ordered)
  }
  __syncthreads();
  if (last) __threadfence();
  return last;
}

// ---- start of a cycle: beta = ||r||, V_0 = r / beta, g = beta e_1 ----------
__global__ void __launch_bounds__(kGb)
gm_norm_kernel(int64_t n, int m, const double* __restrict__ r, GmresPtrs p) {
  __shared__ double red[kGb / 64];
  const int nb = gridDim.x;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kGb + threadIdx.x; i < n; i += (int64_t)nb * kGb)
    s = fma(r[i], r[i], s);
  s = block_sum(s, red);
  if (threadIdx.x == 0) p.part[blockIdx.x] = s;
  if (last_block(&p.ctl[CTL_CNT0], nb) && threadIdx.x == 0) {
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += p.part[b];
    const double beta = sqrt(t);
    for (int k = 0; k <= m; ++k) p.g[k] = 0.0;
    p.g[0] = beta;
    p.ctl[CTL_J] = 0;
    p.ctl[CTL_BREAK] = 0;
  }
}

__global__ void __launch_bounds__(kGb)
gm_first_kernel(int64_t n, const double* __restrict__ r, double* __restrict__ V, GmresPtrs p) {
  const double beta = p.g[0];
  const double inv = beta > 0.0 ? 1.0 / beta : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kGb + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGb)
    V[i] = r[i] * inv;
}

// ---- one Arnoldi step -------------------------------------------------------
// xin = V_j (the operator's input buffer)
__global__ void __launch_bounds__(kGb)
gm_current_kernel(int64_t n, const double* __restrict__ V, double* __restrict__ xin, GmresPtrs p) {
  const int j = p.ctl[CTL_J];
  const double* __restrict__ vj = V + (int64_t)j * n;
  for (int64_t i = (int64_t)blockIdx.x * kGb + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGb)
    xin[i] = vj[i];
}

// partial dots part[k * nc + c] = sum over chunk c of V_k . w, k = 0..j;
// the last block sums them into h (add: h += the sums)
__global__ void __launch_bounds__(kGb)
gm_dots_kernel(int64_t n, int m, const double* __restrict__ V, const double* __restrict__ w,
               GmresPtrs p, int add, int cnt) {
  __shared__ double red[kGb / 64];
  const int j = p.ctl[CTL_J];
  const int nc = gridDim.x;
  const int c = blockIdx.x;
  const int64_t i0 = (int64_t)c * kGchunk;
  const int64_t i1 = i0 + kGchunk < n ? i0 + kGchunk : n;
  for (int k = blockIdx.y; k <= j; k += gridDim.y) {
    const double* __restrict__ vk = V + (int64_t)k * n;
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kGb) s = fma(vk[i], w[i], s);
    s = block_sum(s, red);
    if (threadIdx.x == 0) p.part[(int64_t)k * nc + c] = s;
  }
  if (last_block(&p.ctl[cnt], nc * gridDim.y)) {
    for (int k = threadIdx.x; k <= j; k += kGb) {
      double t = 0.0;
      for (int b = 0; b < nc; ++b) t += p.part[(int64_t)k * nc + b];
      p.h[k] = add ? p.h[k] + t : t;
      if (!add) p.y[k] = t;       // this pass's coefficients, for the update
      else p.y[m + 1 + k] = t;
    }
  }
}

// w -= sum_k coef[k] V_k (coef = this pass's coefficients); with norm != 0
// also the partial sums of ||w||^2, and the last block finishes the step:
// Hessenberg column j, its Givens rotation, g and the residual estimate.
__global__ void __launch_bounds__(kGb)
gm_update_kernel(int64_t n, int m, const double* __restrict__ V, double* __restrict__ w,
                 GmresPtrs p, int pass, int norm) {
  __shared__ double red[kGb / 64];
  __shared__ double coef[160];
  const int j = p.ctl[CTL_J];
  const double* __restrict__ src = pass == 0 ? p.y : p.y + m + 1;
  for (int k = threadIdx.x; k <= j; k += kGb) coef[k] = src[k];
  __syncthreads();
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kGb + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGb) {
    double v = w[i];
    for (int k = 0; k <= j; ++k) v = fma(-coef[k], V[(int64_t)k * n + i], v);
    w[i] = v;
    s = fma(v, v, s);
  }
  if (!norm) return;
  s = block_sum(s, red);
  if (threadIdx.x == 0) p.part[blockIdx.x] = s;
  if (last_block(&p.ctl[CTL_CNT2], gridDim.x) && threadIdx.x == 0) {
    double t = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) t += p.part[b];
    const double hn = sqrt(t);
    const int ld = m + 1;
    double* __restrict__ Hc = p.H + (int64_t)j * ld;
    double* __restrict__ Hrc = p.Hr + (int64_t)j * ld;
    for (int k = 0; k <= j; ++k) Hc[k] = Hrc[k] = p.h[k];
    Hc[j + 1] = Hrc[j + 1] = hn;
    // earlier rotations, then this column's
    for (int k = 0; k < j; ++k) {
      const double a = Hc[k], b = Hc[k + 1];
      Hc[k] = p.cs[k] * a + p.sn[k] * b;
      Hc[k + 1] = -p.sn[k] * a + p.cs[k] * b;
    }
    const double a = Hc[j], b = Hc[j + 1];
    const double den = hypot(a, b);
    const double c = den > 0.0 ? a / den : 1.0, sn = den > 0.0 ? b / den : 0.0;
    p.cs[j] = c;
    p.sn[j] = sn;
    Hc[j] = den;
    Hc[j + 1] = 0.0;
    const double gj = p.g[j];
    p.g[j + 1] = -sn * gj;
    p.g[j] = c * gj;
    p.est[j] = fabs(p.g[j + 1]);
    if (!(hn > 0.0)) p.ctl[CTL_BREAK] = j + 1;
  }
}

// V_{j+1} = w / h_{j+1,j}; the last block advances j
__global__ void __launch_bounds__(kGb)
gm_next_kernel(int64_t n, int m, double* __restrict__ V, const double* __restrict__ w,
               GmresPtrs p) {
  const int j = p.ctl[CTL_J];
  const double hn = p.Hr[(int64_t)j * (m + 1) + j + 1];
  const double inv = hn > 0.0 ? 1.0 / hn : 0.0;
  double* __restrict__ vn = V + (int64_t)(j + 1) * n;
  if (j + 1 <= m)
    for (int64_t i = (int64_t)blockIdx.x * kGb + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kGb)
      vn[i] = w[i] * inv;
  if (last_block(&p.ctl[CTL_CNT3], gridDim.x) && threadIdx.x == 0)
    p.ctl[CTL_J] = j + 1 <= m ? j + 1 : m;
}

// ---- end of a cycle: y = H_k^-1 g_k, hy = Hr_{k+1,k} y ----------------------
__global__ void __launch_bounds__(64)
gm_lsq_kernel(int m, int k, GmresPtrs p) {
  const int lane = threadIdx.x;
  const int ld = m + 1;
  double* __restrict__ y = p.y;          // y[0..k)
  double* __restrict__ hy = p.y + m + 1;  // hy[0..k]
  for (int i = k - 1; i >= 0; --i) {
    double s = 0.0;
    for (int l = i + 1 + lane; l < k; l += 64) s = fma(p.H[(int64_t)l * ld + i], y[l], s);
    s = wave_sum_f64(s);
    if (lane == 0) y[i] = (p.g[i] - s) / p.H[(int64_t)i * ld + i];
    __syncthreads();
  }
  for (int i = lane; i <= k; i += 64) {
    double s = 0.0;
    for (int l = 0; l < k; ++l) s = fma(p.Hr[(int64_t)l * ld + i], y[l], s);
    hy[i] = s;
  }
}

// out = sign * sum_{i < cnt} c_i V_i (+ base), c = y (which = 0) or hy (1)
__global__ void __launch_bounds__(kGb)
gm_combine_kernel(int64_t n, int m, const double* __restrict__ V, int cnt, int which,
                  const double* __restrict__ base, double sign, double* __restrict__ out,
                  GmresPtrs p) {
  __shared__ double coef[161];
  const double* __restrict__ src = which ? p.y + m + 1 : p.y;
  for (int k = threadIdx.x; k < cnt; k += kGb) coef[k] = src[k];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kGb + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGb) {
    double v = 0.0;
    for (int k = 0; k < cnt; ++k) v = fma(coef[k], V[(int64_t)k * n + i], v);
    out[i] = base ? base[i] + sign * v : sign * v;
  }
}

static unsigned gm_grid(int64_t n) {
  const int64_t b = (n + kGb - 1) / kGb;
  return (unsigned)(b < 2048 ? (b > 0 ? b : 1) : 2048);
}

}  // namespace f3d

using namespace f3d;

extern "C" size_t f3d_gmres_workspace_size(int64_t n, int32_t m) {
  if (n < 1 || m < 1 || m > 159) return 0;
  return gmres_layout(n, m).total;
}

extern "C" int f3d_gmres_begin(int64_t n, int32_t m, const double* r, double* V, void* ws,
                               void* stream) {
  F3D_REQUIRE(n >= 1 && m >= 1 && m <= 159, F3D_EINVAL, "gmres_begin: n=%lld m=%d",
              (long long)n, m);
  F3D_REQUIRE(r && V && ws, F3D_EINVAL, "gmres_begin: NULL argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  GmresPtrs p = gmres_ptrs(ws, n, m);
  const unsigned nb = (unsigned)gm_nchunks(n);
  F3D_LAUNCH(gm_norm_kernel, dim3(nb), dim3(kGb), 0, st, n, (int)m, r, p);
  F3D_LAUNCH(gm_first_kernel, dim3(gm_grid(n)), dim3(kGb), 0, st, n, r, V, p);
  return F3D_OK;
}

extern "C" int f3d_gmres_current(int64_t n, int32_t m, const double* V, double* xin, void* ws,
                                 void* stream) {
  F3D_REQUIRE(n >= 1 && m >= 1 && m <= 159 && V && xin && ws, F3D_EINVAL,
              "gmres_current: bad argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  F3D_LAUNCH(gm_current_kernel, dim3(gm_grid(n)), dim3(kGb), 0, st, n, V, xin,
             gmres_ptrs(ws, n, m));
  return F3D_OK;
}

extern "C" int f3d_gmres_arnoldi(int64_t n, int32_t m, double* V, double* w, void* ws,
                                 void* stream) {
  F3D_REQUIRE(n >= 1 && m >= 1 && m <= 159 && V && w && ws, F3D_EINVAL,
              "gmres_arnoldi: bad argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  GmresPtrs p = gmres_ptrs(ws, n, m);
  const unsigned nc = (unsigned)gm_nchunks(n);
  // enough k-slices to fill the chip for small n; each k is read by one slice
  const unsigned ky = nc >= 256 ? 1u : (unsigned)((256 + nc - 1) / nc);
  const unsigned g = gm_grid(n);
  // CGS2: h1 = V w; w -= V^T h1; h2 = V w; w -= V^T h2 (+ ||w||, Givens)
  F3D_LAUNCH(gm_dots_kernel, dim3(nc, ky), dim3(kGb), 0, st, n, (int)m, (const double*)V,
             (const double*)w, p, 0, (int)CTL_CNT0);
  F3D_LAUNCH(gm_update_kernel, dim3(g), dim3(kGb), 0, st, n, (int)m, (const double*)V, w, p, 0,
             0);
  F3D_LAUNCH(gm_dots_kernel, dim3(nc, ky), dim3(kGb), 0, st, n, (int)m, (const double*)V,
             (const double*)w, p, 1, (int)CTL_CNT1);
  F3D_LAUNCH(gm_update_kernel, dim3(g), dim3(kGb), 0, st, n, (int)m, (const double*)V, w, p, 1,
             1);
  F3D_LAUNCH(gm_next_kernel, dim3(g), dim3(kGb), 0, st, n, (int)m, V, (const double*)w, p);
  return F3D_OK;
}

extern "C" int f3d_gmres_solve(int64_t n, int32_t m, int32_t k, const double* V, double* upd,
                               const double* r, double* r_upd, void* ws, void* stream) {
  F3D_REQUIRE(n >= 1 && m >= 1 && m <= 159 && k >= 0 && k <= m && V && upd && ws, F3D_EINVAL,
              "gmres_solve: bad argument");
  F3D_REQUIRE((r == nullptr) == (r_upd == nullptr), F3D_EINVAL,
              "gmres_solve: r and r_upd both set or both NULL");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  GmresPtrs p = gmres_ptrs(ws, n, m);
  const unsigned g = gm_grid(n);
  F3D_LAUNCH(gm_lsq_kernel, dim3(1), dim3(64), 0, st, (int)m, (int)k, p);
  F3D_LAUNCH(gm_combine_kernel, dim3(g), dim3(kGb), 0, st, n, (int)m, V, (int)k, 0,
             (const double*)nullptr, 1.0, upd, p);
  if (r)
    F3D_LAUNCH(gm_combine_kernel, dim3(g), dim3(kGb), 0, st, n, (int)m, V, (int)k + 1, 1, r,
               -1.0, r_upd, p);
  return F3D_OK;
}
