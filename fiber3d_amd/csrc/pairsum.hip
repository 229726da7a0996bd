// Source-target Stokes pair sums (Oseen / rotlet / stresslet) for gfx950.
//
// Reference semantics: python/kernels/kernels.py:794-846 (Oseen),
// :856-908 (rotlet), :1186-1217 (stresslet source-target), :1123-1183
// (stresslet self-sum).  For a pair, d = t - s, r = |d|; if r < eps the
// distance is regularised, r <- sqrt(r^2 + reg^2), while d is kept.
//
// Design (MI355X / CDNA4, wave64):
//   * pack_kernel converts the caller's sources (any AoS/SoA strides, FP64) into
//     one contiguous record per source (6 or 9 32/64-bit words: position,
//     density[, normal]), multiplying an optional per-source weight, and
//     zero-padding to a multiple of the 64-source tile.  tile_box_kernel
//     records each tile's centre and radius.
//   * pairsum_kernel: each lane owns K targets (registers).  The source index
//     is wave-uniform, so source records are fetched with scalar loads into
//     SGPRs and every VALU op takes its source operand straight from an SGPR:
//     no LDS traffic and no VGPRs spent on sources.  Sources are split into
//     P chunks over grid.y so the launch carries thousands of waves.
//   * FP32C: positions are quantised once per call to 31-bit integers on the
//     bounding box of sources and targets (bbox_partial/final), so d = t - s is an
//     exact integer difference rounded once to FP32; the pair is then
//     evaluated in FP32 in units of the quantum h and the sum rescaled by
//     h^-p (the kernels are homogeneous of degree -1 / -2 in d).  Rounding the
//     coordinates themselves to FP32 would cost ~1e-7 |x| / r relative per
//     pair, which fails 1e-6 for the close pairs of random point sets.
//     Pairs are summed in plain FP32 within a 64-source tile, tile sums are
//     added with Kahan compensation, and (sum, -compensation) are written as
//     FP32 partials per chunk; reduce_kernel combines chunks in FP64.
//   * Near pairs: the regularisation test r < eps is one compare per pair
//     whose result is OR-ed over the tile; only a tile in which some lane saw
//     a near pair is recomputed with the exact regularised form.
//   * Exclusion ranges [lo_i, hi_i) (same fiber / self term) are tested only in
//     tiles that intersect the wave's union of ranges; all other tiles run a
//     mask-free path.
//   * Spatial order: Oseen and rotlet calls of at least 65536 sources
//     and targets, without exclusion ranges, visit both in Morton order
//     (sorted on the device each call), so that tiles and waves are compact.
//   * Far tiles: a tile whose centre is more than twice its radius (and more
//     than its radius + 2 eps) from the box of the wave's targets has no near
//     pair, so it runs without the vote; FP32C evaluates it on FP32
//     coordinates relative to the tile's quantised centre (far records),
//     without the integer subtract and convert.
#include "common.h"

#include <stdarg.h>
#include <stdlib.h>

#include <mutex>
#include <type_traits>

namespace f3d {

constexpr int kTile = 64;        // sources per compensated tile
constexpr int kBlock = 256;      // threads per block (4 waves)
constexpr int kPackBlocks = 256; // grid of the pack kernel (sigma partials)
constexpr double kQuantLevels = 2147483648.0 - 65536.0;  // 2^31 minus margin

// FP64 pairs use w = 2/r, one Newton step on v_rsq_f64 without its factor
// 1/2 (one FP64 operation fewer per pair).  The kernels are homogeneous in
// 1/r, so the sums come out scaled by a power of 2 that the reduction takes
// back: rotlet w^3 = 8/r^3, stresslet w^5 = 32/r^5.  Oseen dots d with f/4
// (a second copy of the force in the record) so that its two terms keep
// their ratio: w (f + w^2 (d.f/4) d) = 2 (f/r + (d.f) d/r^3).
template <typename T, int KIND>
struct TwoOverR {
  static constexpr bool value = std::is_same<T, double>::value;
  static constexpr double scale =
      !value ? 1.0 : KIND == F3D_OSEEN ? 0.5 : KIND == F3D_ROTLET ? 0.125 : 1.0 / 32.0;
};

// Record words per source, unpadded so that a group of records loads with a
// few wide scalar loads: 6 for Oseen and rotlet (x, y, z, f or tau), 9 for
// the stresslet (x, y, z, n, q) and for FP64 Oseen (x, y, z, f, f/4).  FP32C
// records hold the quantised position as int32 bits in words 0..2; their far
// records (same layout) hold the position relative to the source tile's
// quantised centre, in units of the quantum, as FP32.
template <typename T, int KIND>
struct RecLen {
  static constexpr int value =
      (KIND == F3D_STRESSLET || (TwoOverR<T, KIND>::value && KIND == F3D_OSEEN)) ? 9 : 6;
};

// Tile geometry for the near/far split: centre and radius of the sources of
// one 64-source tile (radius < 0: no real source) and, for FP32C, the
// quantised centre.  A tile is far from a wave when the distance from its
// centre to the box of the wave's targets exceeds both the far ratio (FP32C;
// 0 for FP64) times its radius and its radius plus 2 eps: then no pair of
// the two is near (r > eps: no regularisation, no vote), and for FP32C the
// tile-relative FP32 coordinates give d with a relative error of a few ulp
// (|target - centre| and |source - centre| are at most a small multiple of
// |d|).
struct TileBox {
  double c[3];
  double r;
  int q[3];
  int pad;
};
constexpr double kFarRatio32 = 2.0;

// homogeneity degree of the kernel in d (for the FP32C rescaling)
template <int KIND>
struct Degree {
  static constexpr int value = (KIND == F3D_OSEEN) ? 1 : 2;
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
// 2/sqrt(x) to the same accuracy: y (3 - x y^2) = 2 y (1.5 - 0.5 x y^2), three
// operations after v_rsq_f64 where the step above takes four.
__device__ __forceinline__ double rsqrt2_t(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return y * fma(-x, y * y, 3.0);
}

// Quantisation geometry of one FP32C call, computed on the device:
// q = rint((x - origin) / h), h = extent / kQuantLevels.
struct Geo {
  double origin[3];
  double h;
  double inv_h;
  double pad[3];
};

// ---------------------------------------------------------------------------
// Bounding box of sources and targets -> Geo, in two passes: per-block
// min/max (kBoxBlocks blocks, grid-stride), then one block combines them.
constexpr int kBoxBlocks = 128;

__device__ __forceinline__ void block_minmax(double (&lo)[3], double (&hi)[3], double* red) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      lo[a] = fmin(lo[a], __shfl_xor(lo[a], o, 64));
      hi[a] = fmax(hi[a], __shfl_xor(hi[a], o, 64));
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      red[w * 6 + a] = lo[a];
      red[w * 6 + 3 + a] = hi[a];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        lo[a] = fmin(lo[a], red[k * 6 + a]);
        hi[a] = fmax(hi[a], red[k * 6 + 3 + a]);
      }
    }
  }
}

__global__ void __launch_bounds__(256)
bbox_partial_kernel(int ns, const double* __restrict__ pos, long sp, long sc, int nt,
                    const double* __restrict__ tpos, long t_sp, long t_sc,
                    double* __restrict__ part) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ns + nt; j += gridDim.x * blockDim.x) {
    const double* p = j < ns ? pos + (long)j * sp : tpos + (long)(j - ns) * t_sp;
    const long c = j < ns ? sc : t_sc;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double v = p[a * c];
      lo[a] = fmin(lo[a], v);
      hi[a] = fmax(hi[a], v);
    }
  }
  __shared__ double red[4 * 6];
  block_minmax(lo, hi, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      part[blockIdx.x * 6 + a] = lo[a];
      part[blockIdx.x * 6 + 3 + a] = hi[a];
    }
  }
}

__global__ void __launch_bounds__(kBoxBlocks)
bbox_final_kernel(const double* __restrict__ part, double min_extent, Geo* __restrict__ geo) {
  double lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = part[threadIdx.x * 6 + a];
    hi[a] = part[threadIdx.x * 6 + 3 + a];
  }
  __shared__ double red[(kBoxBlocks / 64) * 6];
  block_minmax(lo, hi, red);
  if (threadIdx.x == 0) {
    double ext = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      geo->origin[a] = lo[a];
      ext = fmax(ext, hi[a] - lo[a]);
    }
    ext = fmax(ext, min_extent);
    geo->h = ext / kQuantLevels;
    geo->inv_h = kQuantLevels / ext;
  }
}

__device__ __forceinline__ int quantise(double x, double origin, double inv_h) {
  return (int)rint((x - origin) * inv_h);
}

// ---------------------------------------------------------------------------
// Tile boxes: one wave per 64-source tile, the bounding box's centre and the
// largest distance of a real source from it.
__device__ __forceinline__ double readfirst_f64(double v) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(u & 0xffffffffu));
  const int hi = __builtin_amdgcn_readfirstlane((int)(u >> 32));
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_min_f64(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// ---------------------------------------------------------------------------
// Spatial order (Oseen and rotlet, large calls): sources and targets
// are visited in the Morton order of their quantised positions, 5 bits per
// axis (32^3 cells, points of one cell in their original order), so that the
// 64 sources of a tile and the 64 K targets of a wave are spatially compact
// and most tiles are far from most waves.  A stable LSD radix sort of the
// 15-bit codes in two 8-bit passes: per-block digit counts, one exclusive
// scan, a stable scatter.  Deterministic (no global atomics), and every
// buffer it reads was written earlier in the same call (no memset).
constexpr int kSortBits = 5;          // per axis
constexpr int kSortBlock = 256;       // threads per sort block
constexpr int kSortMaxBlocks = 64;    // blocks per pass (a multiple of 16)
constexpr int kSortScan = 1024;       // threads of the scan
constexpr int64_t kSortMin = 65536;   // sources and targets of a call that sorts

// 5 bits -> every third bit
__device__ __forceinline__ unsigned spread3(unsigned v) {
  v &= 0x1fu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void __launch_bounds__(256)
morton_kernel(int n, const double* __restrict__ pos, long sp, long sc,
              const Geo* __restrict__ geo, unsigned* __restrict__ key, int* __restrict__ val) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    unsigned c = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int q = quantise(pos[(long)i * sp + a * sc], geo->origin[a], geo->inv_h);
      c |= spread3((unsigned)q >> (31 - kSortBits)) << a;
    }
    key[i] = c;
    val[i] = i;
  }
}

// counts[d * gridDim.x + b]: digit d in block b's chunk (blocks past n count 0)
__global__ void __launch_bounds__(kSortBlock)
radix_count_kernel(int n, int chunk, int shift, const unsigned* __restrict__ key,
                   int* __restrict__ counts) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long b0 = (long)blockIdx.x * chunk;
  const long b1 = b0 + chunk < n ? b0 + chunk : n;
  for (long e = b0 + threadIdx.x; e < b1; e += kSortBlock)
    atomicAdd(&h[(key[e] >> shift) & 255u], 1);
  __syncthreads();
  counts[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of the 256 x kSortMaxBlocks counts in place; each thread
// owns a contiguous run of kScanPer counts, loaded at once (one latency)
constexpr int kScanPer = 256 * kSortMaxBlocks / kSortScan;
static_assert(kScanPer % 4 == 0, "scan runs are whole int4s");
__global__ void __launch_bounds__(kSortScan)
radix_scan_kernel(int* __restrict__ counts) {
  __shared__ int part[2 * (kSortScan / 64)];
  int4* __restrict__ c4 = reinterpret_cast<int4*>(counts + (long)threadIdx.x * kScanPer);
  int4 v[kScanPer / 4];
#pragma unroll
  for (int e = 0; e < kScanPer / 4; ++e) v[e] = c4[e];
  int s = 0;
#pragma unroll
  for (int e = 0; e < kScanPer / 4; ++e) s += v[e].x + v[e].y + v[e].z + v[e].w;
  // inclusive scan of the thread sums: within each wave by shuffles, then
  // over the wave totals
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int up = __shfl_up(inc, o, 64);
    if (lane >= o) inc += up;
  }
  if (lane == 63) part[w] = inc;
  __syncthreads();
  if (w == 0) {
    // exclusive prefix of the wave totals into part[16 + wave]
    const int tot = lane < kSortScan / 64 ? part[lane] : 0;
    int t = tot;
#pragma unroll
    for (int o = 1; o < kSortScan / 64; o <<= 1) {
      const int up = __shfl_up(t, o, 64);
      if (lane >= o) t += up;
    }
    if (lane < kSortScan / 64) part[kSortScan / 64 + lane] = t - tot;
  }
  __syncthreads();
  int run = part[kSortScan / 64 + w] + inc - s;
#pragma unroll
  for (int e = 0; e < kScanPer / 4; ++e) {
    int4 o;
    o.x = run; run += v[e].x;
    o.y = run; run += v[e].y;
    o.z = run; run += v[e].z;
    o.w = run; run += v[e].w;
    c4[e] = o;
  }
}

// Stable scatter of block b's chunk, 256 elements at a time: an element's
// place is its digit's offset for the block, plus the elements of the same
// digit in earlier rounds, earlier waves and earlier lanes.
__global__ void __launch_bounds__(kSortBlock)
radix_scatter_kernel(int n, int chunk, int shift, const int* __restrict__ offs,
                     const unsigned* __restrict__ kin, const int* __restrict__ vin,
                     unsigned* __restrict__ kout, int* __restrict__ vout) {
  __shared__ int base[256];
  __shared__ int wc[kSortBlock / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  base[tid] = offs[tid * gridDim.x + blockIdx.x];
  const long b0 = (long)blockIdx.x * chunk;
  const long b1 = b0 + chunk < n ? b0 + chunk : n;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (long t0 = b0; t0 < b1; t0 += kSortBlock) {
#pragma unroll
    for (int q = 0; q < kSortBlock / 64; ++q) wc[q][tid] = 0;
    __syncthreads();
    const long e = t0 + tid;
    const bool valid = e < b1;
    unsigned k = 0;
    int v = 0, d = 0;
    if (valid) {
      k = kin[e];
      v = vin[e];
      d = (int)((k >> shift) & 255u);
    }
    // lanes of this wave holding the same digit
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const unsigned long long bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const int rank = __popcll(m & below);
    if (valid && rank == 0) wc[w][d] = __popcll(m);
    __syncthreads();
    if (valid) {
      int at = base[d] + rank;
      for (int q = 0; q < w; ++q) at += wc[q][d];
      kout[at] = k;
      vout[at] = v;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int q = 0; q < kSortBlock / 64; ++q) add += wc[q][tid];
    base[tid] += add;
    __syncthreads();
  }
}

struct SortBufs {
  unsigned* key[2];
  int* val[2];
};

static int sort_blocks(int64_t n, int* chunk) {
  int64_t c = ceil_div(n > 0 ? n : 1, (int64_t)kSortMaxBlocks);
  c = (int64_t)align_up((size_t)c, kSortBlock);
  if (c < 4 * kSortBlock) c = 4 * kSortBlock;
  *chunk = (int)c;
  return kSortMaxBlocks;  // blocks past n count zeros
}

static size_t sort_set_bytes(int64_t n) {
  return 4 * align_up((size_t)(n > 0 ? n : 1) * 4, 256);
}
static size_t sort_count_bytes() { return align_up((size_t)256 * kSortMaxBlocks * 4, 256); }

static SortBufs sort_set(char* p, int64_t n) {
  const size_t one = align_up((size_t)(n > 0 ? n : 1) * 4, 256);
  SortBufs b;
  b.key[0] = reinterpret_cast<unsigned*>(p);
  b.key[1] = reinterpret_cast<unsigned*>(p + one);
  b.val[0] = reinterpret_cast<int*>(p + 2 * one);
  b.val[1] = reinterpret_cast<int*>(p + 3 * one);
  return b;
}

// Morton order of n points: *perm = the permutation (sorted place -> index).
static int spatial_order(int64_t n, const double* pos, int64_t sp, int64_t sc, const Geo* geo,
                         SortBufs b, int* counts, hipStream_t st, const int** perm) {
  int chunk;
  const int nb = sort_blocks(n, &chunk);
  F3D_LAUNCH(morton_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(n, (int64_t)256), 1024)),
             dim3(256), 0, st, (int)n, pos, (long)sp, (long)sc, geo, b.key[0], b.val[0]);
  int cur = 0;
  for (int shift = 0; shift < 3 * kSortBits; shift += 8) {
    F3D_LAUNCH(radix_count_kernel, dim3(nb), dim3(kSortBlock), 0, st, (int)n, chunk, shift,
               (const unsigned*)b.key[cur], counts);
    F3D_LAUNCH(radix_scan_kernel, dim3(1), dim3(kSortScan), 0, st, counts);
    F3D_LAUNCH(radix_scatter_kernel, dim3(nb), dim3(kSortBlock), 0, st, (int)n, chunk, shift,
               (const int*)counts, (const unsigned*)b.key[cur], (const int*)b.val[cur],
               b.key[cur ^ 1], b.val[cur ^ 1]);
    cur ^= 1;
  }
  *perm = b.val[cur];
  return F3D_OK;
}

// ---------------------------------------------------------------------------
// Tile boxes: one wave per 64-source tile (sources in the order of perm when
// given), the bounding box's centre and the largest distance of a real
// source from it.
__global__ void __launch_bounds__(64)
tile_box_kernel(int ns, int tiles, const double* __restrict__ pos, long sp, long sc,
                const Geo* __restrict__ geo, const int* __restrict__ perm,
                TileBox* __restrict__ box) {
  const int t = blockIdx.x;
  if (t >= tiles) return;
  const int j = t * kTile + (int)threadIdx.x;
  const bool valid = j < ns;
  const int o = valid ? (perm ? perm[j] : j) : 0;
  double x[3], lo[3], hi[3], c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    x[a] = valid ? pos[(long)o * sp + a * sc] : 0.0;
    lo[a] = wave_min_f64(valid ? x[a] : 1e300);
    hi[a] = wave_max_f64(valid ? x[a] : -1e300);
    c[a] = 0.5 * (lo[a] + hi[a]);
  }
  double r2 = 0.0;
  if (valid) {
#pragma unroll
    for (int a = 0; a < 3; ++a) r2 += (x[a] - c[a]) * (x[a] - c[a]);
  }
  const double rmax = sqrt(wave_max_f64(r2));
  const bool any = t * kTile < ns;
  if (threadIdx.x == 0) {
    TileBox b;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      b.c[a] = any ? c[a] : 0.0;
      b.q[a] = (any && geo) ? quantise(c[a], geo->origin[a], geo->inv_h) : 0;
    }
    // rounding of the centre and the root: a relative margin
    b.r = any ? rmax * (1.0 + 1e-12) + 1e-300 : -1.0;
    b.pad = 0;
    box[t] = b;
  }
}

// ---------------------------------------------------------------------------
// pack: sources -> records.  Also accumulates sigma = sum_j n_j . q_j per
// block when sigma_part != nullptr (periphery complementary term).
// FP32C records hold the quantised position as int32 bits in words 0..2.
template <typename T, int KIND>
__global__ void __launch_bounds__(256)
pack_kernel(int ns, int ns_pad, const double* __restrict__ pos, const double* __restrict__ nrm,
            const double* __restrict__ den, const double* __restrict__ w, long sp, long sc,
            const Geo* __restrict__ geo, const TileBox* __restrict__ tbox,
            const int* __restrict__ perm, T* __restrict__ rec, T* __restrict__ rec_far,
            double* __restrict__ sigma_part) {
  constexpr int R = RecLen<T, KIND>::value;
  constexpr bool kQuant = std::is_same<T, float>::value;
  double sig = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ns_pad; j += gridDim.x * blockDim.x) {
    T r[R];
#pragma unroll
    for (int c = 0; c < R; ++c) r[c] = 0;
    // far record position words (FP32C)
    T f[3] = {0, 0, 0};
    // Zero-density padding for j >= ns; the pair kernel masks those sources,
    // so a target at the padding position never meets an unregularised r = 0.
    if (j < ns) {
      const int o = perm ? perm[j] : j;  // original source index
      const double wj = w ? w[o] : 1.0;
      const long b = (long)o * sp;
      if constexpr (KIND == F3D_STRESSLET) {
        const double n0 = nrm[b], n1 = nrm[b + sc], n2 = nrm[b + 2 * sc];
        const double q0 = den[b] * wj, q1 = den[b + sc] * wj, q2 = den[b + 2 * sc] * wj;
        r[3] = (T)n0; r[4] = (T)n1; r[5] = (T)n2;
        r[6] = (T)q0; r[7] = (T)q1; r[8] = (T)q2;
        sig += n0 * q0 + n1 * q1 + n2 * q2;
      } else {
        r[3] = (T)(den[b] * wj);
        r[4] = (T)(den[b + sc] * wj);
        r[5] = (T)(den[b + 2 * sc] * wj);
        if constexpr (TwoOverR<T, KIND>::value && KIND == F3D_OSEEN) {
          r[6] = 0.25 * r[3];
          r[7] = 0.25 * r[4];
          r[8] = 0.25 * r[5];
        }
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double x = pos[b + a * sc];
        if constexpr (kQuant) {
          const int q = quantise(x, geo->origin[a], geo->inv_h);
          r[a] = __int_as_float(q);
          f[a] = (float)(q - tbox[j / kTile].q[a]);
        } else {
          r[a] = x;
        }
      }
    }
    T* out = rec + (size_t)j * R;
#pragma unroll
    for (int c = 0; c < R; ++c) out[c] = r[c];
    if constexpr (kQuant) {
      T* of = rec_far + (size_t)j * R;
#pragma unroll
      for (int c = 0; c < R; ++c) of[c] = c < 3 ? f[c] : r[c];
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
// Target coordinate type: quantised int32 for FP32C, FP64 for F64.
template <typename T>
using TCoord = typename std::conditional<std::is_same<T, float>::value, int, double>::type;

// near-pair handling of one pair
enum { MODE_EXACT = 0, MODE_CLAMP = 1, MODE_VOTE = 2 };

// a b + c for a wave-uniform c.  FP64: one VOP3 v_fma_f64 that reads c from
// its SGPRs; left to itself the compiler turns an FMA whose addend has no
// other use into v_fmac_f64 after two v_mov copies of c.
__device__ __forceinline__ float fma_s(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_s(double a, double b, double c) {
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}

// The kernel's contribution for d = (dx, dy, dz), ri = 1/|d| (FP64: 2/|d|)
// and the source's density words dv (Oseen f, FP64 also f/4; rotlet tau;
// stresslet n, q).
template <typename T, int KIND>
__device__ __forceinline__ void accum(const T* dv, T dx, T dy, T dz, T ri, T& ax, T& ay,
                                      T& az) {
  if (KIND == F3D_OSEEN) {
    constexpr int q = TwoOverR<T, KIND>::value ? 3 : 0;
    const T fd = dx * dv[q] + dy * dv[q + 1] + dz * dv[q + 2];
    const T s = fd * (ri * ri);
    ax += ri * fma_s(s, dx, dv[0]);
    ay += ri * fma_s(s, dy, dv[1]);
    az += ri * fma_s(s, dz, dv[2]);
  } else if (KIND == F3D_ROTLET) {
    const T ri3 = ri * ri * ri;
    ax += ri3 * (dv[1] * dz - dv[2] * dy);
    ay += ri3 * (dv[2] * dx - dv[0] * dz);
    az += ri3 * (dv[0] * dy - dv[1] * dx);
  } else {
    const T ri2 = ri * ri;
    const T dn = dx * dv[0] + dy * dv[1] + dz * dv[2];
    const T dq = dx * dv[3] + dy * dv[4] + dz * dv[5];
    const T f0 = (dq * ri2) * (dn * ri2) * ri;
    ax += f0 * dx;
    ay += f0 * dy;
    az += f0 * dz;
  }
}

// One pair: accumulate the un-scaled kernel contribution of record r into
// (ax, ay, az).  MASK: zero the pair when excl is set.
template <typename T, int KIND, int MODE, bool MASK>
__device__ __forceinline__ void pair(const T* r, TCoord<T> tx, TCoord<T> ty, TCoord<T> tz,
                                     T reg2, T eps2, T rmax, bool excl, T& ax, T& ay, T& az,
                                     unsigned long long& vote) {
  T dx, dy, dz;
  if constexpr (std::is_same<T, float>::value) {
    dx = (float)(tx - __float_as_int(r[0]));
    dy = (float)(ty - __float_as_int(r[1]));
    dz = (float)(tz - __float_as_int(r[2]));
  } else {
    dx = tx - r[0];
    dy = ty - r[1];
    dz = tz - r[2];
  }
  T r2 = dx * dx + dy * dy + dz * dz;
  if (MODE == MODE_EXACT) r2 = (r2 < eps2) ? r2 + reg2 : r2;
  // the near-pair vote lives in SGPRs: one v_cmp into a lane mask, OR-ed on the SALU
  if (MODE == MODE_VOTE) vote |= __ballot(r2 < eps2);
  T ri;
  if constexpr (TwoOverR<T, KIND>::value) ri = rsqrt2_t(r2);
  else ri = rsqrt_t(r2);
  if (MODE == MODE_CLAMP) ri = ri < rmax ? ri : rmax;
  if (MASK) ri = excl ? (T)0 : ri;
  accum<T, KIND>(r + 3, dx, dy, dz, ri, ax, ay, az);
}

// The first NV words of the U records from source j on (wave-uniform: scalar
// loads into SGPRs).
template <typename T, int R, int NV, int U>
__device__ __forceinline__ void load_recs(const T* __restrict__ rec, int j, T (&s)[U][NV]) {
  // one base address and constant offsets, so that adjacent words merge into
  // wide scalar loads
  const T* __restrict__ r = rec + (size_t)j * R;
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int c = 0; c < NV; ++c) s[u][c] = r[u * R + c];
}

// Records fetched together per group.  FP32 Oseen and rotlet load the next
// group's records while the current group computes; the FP64 records, and
// the FP32 stresslet records, of two groups do not fit the scalar registers
// beside the kernel's own.
constexpr int kGroup = 4;
template <typename T, int KIND>
struct Prefetch {
  static constexpr bool value = std::is_same<T, float>::value && KIND != F3D_STRESSLET;
};

// body(s, jj) for each group of U records of the 64-source tile at j0.
template <typename T, int KIND, int U, class Body>
__device__ __forceinline__ void for_groups(const T* __restrict__ rec, int j0, Body&& body) {
  constexpr int R = RecLen<T, KIND>::value, NV = R;
  if constexpr (Prefetch<T, KIND>::value) {
    // two register sets that trade roles (no copies): the next group loads
    // into one while the other computes (the last group reloads itself: no
    // read past the tile)
    static_assert(kTile % (2 * U) == 0, "groups pair up within a tile");
    T s[U][NV], nx[U][NV];
    load_recs<T, R, NV, U>(rec, j0, s);
    for (int jj = 0; jj < kTile; jj += 2 * U) {
      load_recs<T, R, NV, U>(rec, j0 + jj + U, nx);
      body(s, jj);
      load_recs<T, R, NV, U>(rec, j0 + (jj + 2 * U < kTile ? jj + 2 * U : jj + U), s);
      body(nx, jj + U);
    }
  } else {
    for (int jj = 0; jj < kTile; jj += U) {
      // the U scalar record loads issue before any use, so their latencies overlap
      T s[U][NV];
      load_recs<T, R, NV, U>(rec, j0 + jj, s);
      body(s, jj);
    }
  }
}

// One 64-source tile; returns (wave-uniformly) whether any pair was near when
// MODE == MODE_VOTE.  MASK: drop the pairs of excluded ranges and of the
// zero-padding sources j >= ns (a padding record may sit on a target, where
// an unregularised 1/r would be infinite).
template <typename T, int KIND, int K, int MODE, bool MASK>
__device__ __forceinline__ bool tile(const T* __restrict__ rec, int j0, int ns,
                                     const TCoord<T> (&tx)[K],
                                     const TCoord<T> (&ty)[K], const TCoord<T> (&tz)[K],
                                     const int (&lo)[K], const int (&hi)[K], T reg2, T eps2,
                                     T rmax, T (&ax)[K], T (&ay)[K], T (&az)[K]) {
  constexpr int NV = RecLen<T, KIND>::value;
  unsigned long long vote = 0;
  for_groups<T, KIND, kGroup>(rec, j0, [&](const T (&s)[kGroup][NV], int jj) {
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int j = j0 + jj + u;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const bool excl =
            MASK && (((unsigned)(j - lo[k]) < (unsigned)(hi[k] - lo[k])) || j >= ns);
        pair<T, KIND, MODE, MASK>(s[u], tx[k], ty[k], tz[k], reg2, eps2, rmax, excl, ax[k],
                                  ay[k], az[k], vote);
      }
    }
  });
  return MODE == MODE_VOTE && vote != 0ull;
}

// One far 64-source tile (no near pair, no exclusion): d from plain
// coordinates, no vote and no regularisation.  rec: FP64 records, or FP32C
// far records (positions relative to the tile's quantised centre); lx, ly,
// lz: the target in the same coordinates.
template <typename T, int KIND, int K>
__device__ __forceinline__ void tile_far(const T* __restrict__ rec, int j0, const T (&lx)[K],
                                         const T (&ly)[K], const T (&lz)[K], T (&ax)[K],
                                         T (&ay)[K], T (&az)[K]) {
  constexpr int NV = RecLen<T, KIND>::value;
  for_groups<T, KIND, kGroup>(rec, j0, [&](const T (&s)[kGroup][NV], int) {
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const T dx = lx[k] - s[u][0], dy = ly[k] - s[u][1], dz = lz[k] - s[u][2];
        const T r2 = dx * dx + dy * dy + dz * dz;
        T ri;
        if constexpr (TwoOverR<T, KIND>::value) ri = rsqrt2_t(r2);
        else ri = rsqrt_t(r2);
        accum<T, KIND>(s[u] + 3, dx, dy, dz, ri, ax[k], ay[k], az[k]);
      }
    }
  });
}

// Far flags of the 64 tiles from tile t0 on (bit l: tile t0 + l; tiles
// below t_end only), wave-uniform.  Lane l tests tile t0 + l against the
// wave's target box [bl, bh].
__device__ __forceinline__ unsigned long long far_mask(const TileBox* __restrict__ tbox, int t0,
                                                       int t_end, const double (&bl)[3],
                                                       const double (&bh)[3], double far_ratio,
                                                       double eps) {
  const int t = t0 + (int)(threadIdx.x & 63);
  bool far = false;
  if (t < t_end) {
    const TileBox* __restrict__ tb = tbox + t;
    const double c0 = tb->c[0], c1 = tb->c[1], c2 = tb->c[2], rr = tb->r;
    const double e0 = fmax(0.0, fmax(bl[0] - c0, c0 - bh[0]));
    const double e1 = fmax(0.0, fmax(bl[1] - c1, c1 - bh[1]));
    const double e2 = fmax(0.0, fmax(bl[2] - c2, c2 - bh[2]));
    const double need = fmax(far_ratio * rr, rr + 2.0 * eps);
    far = rr >= 0.0 && e0 * e0 + e1 * e1 + e2 * e2 > need * need;
  }
  return __ballot(far);
}

// partial layout: part[(p*3 + c)*nt + i]; for T=float a second array holds
// the negated Kahan compensation.
template <typename T, int KIND, int K, bool CLAMP>
__global__ void __launch_bounds__(kBlock)
pairsum_kernel(const T* __restrict__ rec, const T* __restrict__ rec_far,
               const int* __restrict__ tperm, int chunk, int ns,
               int ns_pad, const double* __restrict__ tpos, long t_sp, long t_sc, int nt,
               const int* __restrict__ excl_lo, const int* __restrict__ excl_hi, int self_excl,
               double reg, double eps, const Geo* __restrict__ geo,
               const TileBox* __restrict__ tbox, double far_ratio, T* __restrict__ part_hi,
               T* __restrict__ part_lo) {
  constexpr bool kQuant = std::is_same<T, float>::value;
  // regularisation constants in the units of the coordinates used
  const double hs = kQuant ? geo->h : 1.0;
  const T reg2 = (T)((reg / hs) * (reg / hs));
  const T eps2 = (T)((eps / hs) * (eps / hs));
  // (FP64 clamps w = 2/r)
  const T rmax = (T)((TwoOverR<T, KIND>::value ? 2.0 : 1.0) * hs / eps);
  // The G = blockDim.x / 64 waves of a block share their 64 K targets and
  // take consecutive source chunks; their sums are combined in LDS, so the
  // block writes one partial per target (G times fewer partial bytes).
  // (readfirstlane: the wave index is wave-uniform, so the source records
  // stay scalar loads into SGPRs)
  const int lane = threadIdx.x & 63, G = blockDim.x >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int base = blockIdx.x * (64 * K) + lane;
  TCoord<T> tx[K], ty[K], tz[K];
  int lo[K], hi[K];
  int wlo = 0x7fffffff, whi = -0x7fffffff;
  double bl[3] = {1e300, 1e300, 1e300}, bh[3] = {-1e300, -1e300, -1e300};
#pragma unroll
  for (int k = 0; k < K; ++k) {
    // slot base + k 64 (the partial's index) holds target i (in spatial
    // order only without exclusion ranges, and for the self flag only when
    // sources and targets share one order: slot sl then excludes source sl)
    const int sl = base + k * 64;
    const int sc = sl < nt ? sl : nt - 1;
    const int i = tperm ? tperm[sc] : sc;
    const long b = (long)i * t_sp;
    const double px = tpos[b], py = tpos[b + t_sc], pz = tpos[b + 2 * t_sc];
    if (sl < nt) {
      bl[0] = fmin(bl[0], px); bh[0] = fmax(bh[0], px);
      bl[1] = fmin(bl[1], py); bh[1] = fmax(bh[1], py);
      bl[2] = fmin(bl[2], pz); bh[2] = fmax(bh[2], pz);
    }
    if constexpr (kQuant) {
      tx[k] = quantise(px, geo->origin[0], geo->inv_h);
      ty[k] = quantise(py, geo->origin[1], geo->inv_h);
      tz[k] = quantise(pz, geo->origin[2], geo->inv_h);
    } else {
      tx[k] = px;
      ty[k] = py;
      tz[k] = pz;
    }
    lo[k] = 0;
    hi[k] = 0;
    if ((excl_lo != nullptr || self_excl) && sl < nt) {
      lo[k] = self_excl ? sl : excl_lo[i];
      hi[k] = self_excl ? sl + 1 : excl_hi[i];
      if (hi[k] > lo[k]) {
        wlo = min(wlo, lo[k]);
        whi = max(whi, hi[k]);
      }
    }
  }
  // Wave-uniform union of the exclusion ranges (kept in SGPRs).
  wlo = __builtin_amdgcn_readfirstlane(wave_min_i32(wlo));
  whi = __builtin_amdgcn_readfirstlane(wave_max_i32(whi));
  // Wave-uniform box of the wave's targets (FP64).
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    bl[a] = readfirst_f64(wave_min_f64(bl[a]));
    bh[a] = readfirst_f64(wave_max_f64(bh[a]));
  }

  T sx[K], sy[K], sz[K], cx[K], cy[K], cz[K];
#pragma unroll
  for (int k = 0; k < K; ++k) sx[k] = sy[k] = sz[k] = cx[k] = cy[k] = cz[k] = 0;

  // this wave's chunk; past the last chunk (the final block's spare waves)
  // the loop is empty and the wave adds zeros
  const int j_begin = min((int)(blockIdx.y * G + wv) * chunk, ns_pad);
  const int j_end = min(j_begin + chunk, ns_pad);
  unsigned long long fmask = 0;
  for (int j0 = j_begin; j0 < j_end; j0 += kTile) {
    T ax[K], ay[K], az[K];
#pragma unroll
    for (int k = 0; k < K; ++k) ax[k] = ay[k] = az[k] = 0;
    // far flags, 64 tiles at a time
    if (((j0 - j_begin) & (64 * kTile - 1)) == 0)
      fmask = far_mask(tbox, j0 / kTile, j_end / kTile, bl, bh, far_ratio, eps);
    const bool masked = (j0 < whi && j0 + kTile > wlo) || j0 + kTile > ns;
    const bool far = !masked && (fmask & 1ull);
    fmask >>= 1;
    if (far) {
      T lx[K], ly[K], lz[K];
      if constexpr (kQuant) {
        const TileBox* __restrict__ tb = tbox + (j0 / kTile);
        const int q0 = tb->q[0], q1 = tb->q[1], q2 = tb->q[2];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          lx[k] = (T)(tx[k] - q0);
          ly[k] = (T)(ty[k] - q1);
          lz[k] = (T)(tz[k] - q2);
        }
        tile_far<T, KIND, K>(rec_far, j0, lx, ly, lz, ax, ay, az);
      } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          lx[k] = tx[k];
          ly[k] = ty[k];
          lz[k] = tz[k];
        }
        tile_far<T, KIND, K>(rec, j0, lx, ly, lz, ax, ay, az);
      }
    } else if (masked) {
      if (CLAMP)
        tile<T, KIND, K, MODE_CLAMP, true>(rec, j0, ns, tx, ty, tz, lo, hi, reg2, eps2,
                                           rmax, ax, ay, az);
      else
        tile<T, KIND, K, MODE_EXACT, true>(rec, j0, ns, tx, ty, tz, lo, hi, reg2, eps2,
                                           rmax, ax, ay, az);
    } else if (CLAMP) {
      tile<T, KIND, K, MODE_CLAMP, false>(rec, j0, ns, tx, ty, tz, lo, hi, reg2, eps2,
                                          rmax, ax, ay, az);
    } else if (tile<T, KIND, K, MODE_VOTE, false>(rec, j0, ns, tx, ty, tz, lo, hi, reg2,
                                                  eps2, rmax, ax, ay, az)) {
      // some lane met a pair with r < eps: redo the tile with the regularised form
#pragma unroll
      for (int k = 0; k < K; ++k) ax[k] = ay[k] = az[k] = 0;
      tile<T, KIND, K, MODE_EXACT, false>(rec, j0, ns, tx, ty, tz, lo, hi, reg2, eps2,
                                          rmax, ax, ay, az);
    }
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
  // each wave's sums in FP64 (FP32: the Kahan pair s - c), the block's
  // waves added in a fixed order
  double v[3][K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    v[0][k] = (double)sx[k] - (double)cx[k];
    v[1][k] = (double)sy[k] - (double)cy[k];
    v[2][k] = (double)sz[k] - (double)cz[k];
  }
  __shared__ double red[3][3][K][64];
  if (wv > 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 0; k < K; ++k) red[wv - 1][c][k][lane] = v[c][k];
  }
  __syncthreads();
  if (wv != 0) return;
  // (one wave's values at a time: the epilogue must not set the kernel's
  // register count)
#pragma unroll 1
  for (int w = 1; w < G; ++w) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 0; k < K; ++k) v[c][k] += red[w - 1][c][k][lane];
  }
  const size_t p = blockIdx.y;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = base + k * 64;
    if (i < nt) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (sizeof(T) == 4) {
          // the FP64 sum as an unevaluated FP32 pair (hi + lo)
          const T h = (T)v[c][k];
          part_hi[(p * 3 + c) * nt + i] = h;
          part_lo[(p * 3 + c) * nt + i] = (T)(v[c][k] - (double)h);
        } else {
          part_hi[(p * 3 + c) * nt + i] = (T)v[c][k];
        }
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

// Targets per block of reduce_kernel: its four waves split the source
// chunks p = w, w + 4, ... of the same 64 targets (4x the waves of one
// target per lane, so enough loads are in flight to cover HBM latency at
// 40k targets), then wave 0 adds the four sums in a fixed order.
constexpr int kRedTargets = 64;

template <typename T>
__global__ void __launch_bounds__(256)
reduce_kernel(const T* __restrict__ part_hi, const T* __restrict__ part_lo, int P, int nt,
              const int* __restrict__ tperm, double factor, const Geo* __restrict__ geo,
              int degree, double* __restrict__ out, long o_sp, long o_sc, int accumulate,
              ShellEpi epi) {
  if (geo) factor *= degree == 1 ? geo->inv_h : geo->inv_h * geo->inv_h;
  double sigma = 0.0;
  if (epi.sigma_part) {
    __shared__ double red[4];
    double v = epi.sigma_part[threadIdx.x];  // blockDim.x == kPackBlocks
    v = wave_sum_f64(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    sigma = red[0] + red[1] + red[2] + red[3];
  }
  __shared__ double sums[3][3][kRedTargets];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int sl = blockIdx.x * kRedTargets + lane;  // partial slot
  double a[3] = {0.0, 0.0, 0.0};
  if (sl < nt) {
#pragma unroll 2
    for (int p = wv; p < P; p += 4) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const size_t idx = ((size_t)p * 3 + c) * nt + sl;
        a[c] += (double)part_hi[idx];
        if (sizeof(T) == 4) a[c] += (double)part_lo[idx];
      }
    }
  }
  if (wv > 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) sums[wv - 1][c][lane] = a[c];
  }
  __syncthreads();
  if (wv != 0 || sl >= nt) return;
  const int i = tperm ? tperm[sl] : sl;  // the slot's target
  double u[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    u[c] = factor * (((a[c] + sums[0][c][lane]) + sums[1][c][lane]) + sums[2][c][lane]);
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
  int P;      // source chunks (one wave each per target group)
  int chunk;  // sources per chunk
  int G;      // chunks per block (waves of a block, combined in LDS)
  int Pg;     // partial sums per target: ceil(P / G)
  int sort;   // sources and targets in spatial order
};

static void set_groups(Plan& pl) {
  pl.G = pl.P >= 4 ? 4 : (pl.P >= 2 ? 2 : 1);
  pl.Pg = (int)ceil_div(pl.P, pl.G);
}

static int env_int(const char* name, int dflt) {
  const char* s = getenv(name);
  if (!s || !*s) return dflt;
  return atoi(s);
}

// Launch shape per (kernel, precision): targets per lane and the wave count
// the source split aims for.  Set by the autotuner through
// f3d_pairsum_set_launch; F3D_PAIRSUM_K32 / _K64 / _WAVES override.
static int g_launch_k[3][2] = {{2, 2}, {2, 2}, {2, 2}};
static int g_launch_waves[3][2] = {{8192, 8192}, {8192, 8192}, {8192, 8192}};
// Per size bucket: (ceil(log2 nt), ceil(log2 ns)), each clamped to
// [kLgMin, kLgMax]; k == 0 leaves the bucket on the (kernel, precision) shape.
constexpr int kLgMin = 8, kLgMax = 22, kLgN = kLgMax - kLgMin + 1;
static int g_bucket_k[3][2][kLgN][kLgN];
static int g_bucket_waves[3][2][kLgN][kLgN];

static int lg_bucket(int64_t n) {
  int l = 0;
  while (l < 62 && (1ll << l) < n) ++l;
  return (l < kLgMin ? kLgMin : (l > kLgMax ? kLgMax : l)) - kLgMin;
}

// Event pairs around the pair kernel, recorded while f3d_pairsum_timer is on.
constexpr int kTimerSlots = 4096;
struct KernelTimer {
  bool on = false;
  bool made = false;
  int n = 0;
  hipEvent_t ev[2 * kTimerSlots];
};
static KernelTimer g_timer;

static Plan make_plan(int kernel, int precision, int64_t ns, int64_t nt) {
  Plan pl;
  const int kk = (kernel >= 0 && kernel < 3) ? kernel : 0;
  const int pp = precision == F3D_F64 ? 0 : 1;
  const int bt = lg_bucket(nt), bs = lg_bucket(ns);
  const bool tuned = g_bucket_k[kk][pp][bt][bs] != 0;
  const int k0 = tuned ? g_bucket_k[kk][pp][bt][bs] : g_launch_k[kk][pp];
  const int w0 = tuned ? g_bucket_waves[kk][pp][bt][bs] : g_launch_waves[kk][pp];
  pl.K = env_int(precision == F3D_F64 ? "F3D_PAIRSUM_K64" : "F3D_PAIRSUM_K32", k0);
  if (pl.K != 1 && pl.K != 2 && pl.K != 4) pl.K = 2;
  pl.ns_pad = (int)align_up((size_t)(ns > 0 ? ns : 1), kTile);
  const int64_t waves_per_split = ceil_div(nt > 0 ? nt : 1, 64 * pl.K);
  int target_waves = env_int("F3D_PAIRSUM_WAVES", w0);
  if (target_waves < 64) target_waves = 64;
  const int64_t tiles = pl.ns_pad / kTile;
  int64_t P = ceil_div(target_waves, waves_per_split);
  if (P < 1) P = 1;
  if (P > tiles) P = tiles;
  const int64_t tiles_per = ceil_div(tiles, P);
  pl.chunk = (int)(tiles_per * kTile);
  pl.P = (int)ceil_div(pl.ns_pad, pl.chunk);
  set_groups(pl);
  // Oseen and rotlet calls with at least kSortMin sources and targets run in
  // spatial order (F3D_PAIRSUM_SORT=0 never, 1 at every size).  Not the
  // stresslet: its FP32 terms cancel (d.n d.q), and the larger tile sums
  // of spatially grouped sources raise its error by about 15%, past 1e-6 on
  // clustered points.
  const int sort_env = env_int("F3D_PAIRSUM_SORT", -1);
  pl.sort = kernel != F3D_STRESSLET &&
            (sort_env >= 0 ? sort_env != 0 : (ns >= kSortMin && nt >= kSortMin));
  return pl;
}

// source records; FP32C adds the far records (same size)
static size_t rec_bytes(int kernel, int precision, int ns_pad) {
  const size_t r = kernel == F3D_STRESSLET ? RecLen<double, F3D_STRESSLET>::value
                   : kernel == F3D_OSEEN && precision == F3D_F64 ? RecLen<double, F3D_OSEEN>::value
                                                                 : RecLen<float, F3D_OSEEN>::value;
  const size_t t = precision == F3D_F64 ? 8 : 4;
  const size_t one = align_up((size_t)ns_pad * r * t, 256);
  return precision == F3D_F64 ? one : 2 * one;
}

static size_t tile_bytes(int ns_pad) {
  return align_up((size_t)(ns_pad / kTile) * sizeof(TileBox), 256);
}

static size_t part_bytes(int precision, int P, int64_t nt) {
  const size_t per = precision == F3D_F64 ? 8 : 8;  // f64, or 2 x f32
  return align_up((size_t)P * 3 * (size_t)(nt > 0 ? nt : 1) * per, 256);
}

static size_t sort_bytes(const Plan& pl, int64_t nt) {
  return pl.sort ? sort_set_bytes(pl.ns_pad) + sort_set_bytes(nt) + sort_count_bytes() : 0;
}

static size_t plan_ws(int kernel, int precision, const Plan& pl, int64_t nt) {
  return rec_bytes(kernel, precision, pl.ns_pad) + tile_bytes(pl.ns_pad) +
         part_bytes(precision, pl.Pg, nt) +
         align_up(kPackBlocks * sizeof(double), 256) + align_up(sizeof(Geo), 256) +
         align_up(kBoxBlocks * 6 * sizeof(double), 256) + sort_bytes(pl, nt);
}

static size_t pairsum_ws(int kernel, int precision, int64_t ns, int64_t nt) {
  return plan_ws(kernel, precision, make_plan(kernel, precision, ns, nt), nt);
}

// The default plan with fewer source chunks (fewer partial sums) where its
// workspace exceeds cap bytes; P = 0 when even one chunk does not fit.
static Plan make_plan_within(int kernel, int precision, int64_t ns, int64_t nt, size_t cap) {
  Plan pl = make_plan(kernel, precision, ns, nt);
  if (plan_ws(kernel, precision, pl, nt) <= cap) return pl;
  Plan one = pl;
  one.Pg = 1;
  const size_t fixed = plan_ws(kernel, precision, one, nt) - part_bytes(precision, 1, nt);
  const size_t per = 3 * (size_t)(nt > 0 ? nt : 1) * 8;
  // partial groups that fit, each of up to 4 chunks
  int64_t gmax = cap > fixed + 256 ? (int64_t)((cap - fixed - 256) / per) : 0;
  if (gmax < 1) {
    pl.P = 0;
    return pl;
  }
  const int64_t tiles = pl.ns_pad / kTile;
  int64_t pmax = 4 * gmax;
  if (pmax > pl.P) pmax = pl.P;
  const int64_t tiles_per = ceil_div(tiles, pmax < tiles ? pmax : tiles);
  pl.chunk = (int)(tiles_per * kTile);
  pl.P = (int)ceil_div(pl.ns_pad, pl.chunk);
  set_groups(pl);
  return pl;
}

template <typename T, int KIND>
static int launch_pairsum(const Plan& pl, uint32_t flags, int64_t ns, const double* pos,
                          const double* nrm, const double* den, const double* w, int64_t s_sp,
                          int64_t s_sc, int64_t nt, const double* tpos, int64_t t_sp,
                          int64_t t_sc, const int32_t* lo, const int32_t* hi, double eta,
                          double reg, double eps, double* out, int64_t o_sp, int64_t o_sc,
                          char* ws, const ShellEpi* epi_in, hipStream_t st) {
  constexpr bool kQuant = std::is_same<T, float>::value;
  const int prec = kQuant ? F3D_F32C : F3D_F64;
  T* rec = reinterpret_cast<T*>(ws);
  T* rec_far = kQuant ? reinterpret_cast<T*>(ws + rec_bytes(KIND, prec, pl.ns_pad) / 2) : nullptr;
  char* p = ws + rec_bytes(KIND, prec, pl.ns_pad);
  TileBox* tbox = reinterpret_cast<TileBox*>(p);
  p += tile_bytes(pl.ns_pad);
  T* part_hi = reinterpret_cast<T*>(p);
  T* part_lo = part_hi + (size_t)pl.Pg * 3 * nt;
  p += part_bytes(prec, pl.Pg, nt);
  double* sigma_part = reinterpret_cast<double*>(p);
  p += align_up(kPackBlocks * sizeof(double), 256);
  // the quantisation geometry: FP32C's coordinates, and the Morton codes of
  // a sorted call in either precision (FP64 kernels take no geometry)
  Geo* geo_buf = reinterpret_cast<Geo*>(p);
  Geo* geo = kQuant ? geo_buf : nullptr;
  p += align_up(sizeof(Geo), 256);
  double* box_part = reinterpret_cast<double*>(p);
  p += align_up(kBoxBlocks * 6 * sizeof(double), 256);
  const bool same = pos == tpos && s_sp == t_sp && s_sc == t_sc && ns == nt;
  const bool self_excl_flag = (flags & F3D_EXCLUDE_SELF) != 0;
  const bool sorted = pl.sort && lo == nullptr && (!self_excl_flag || same);

  if (kQuant || sorted) {
    // the quantum must resolve eps; an extent below 64 (eps + reg) is raised
    const double min_extent = 64.0 * (eps + reg) + 1e-300;
    F3D_LAUNCH(bbox_partial_kernel, dim3(kBoxBlocks), dim3(256), 0, st, (int)ns, pos, (long)s_sp,
               (long)s_sc, (int)nt, tpos, (long)t_sp, (long)t_sc, box_part);
    F3D_LAUNCH(bbox_final_kernel, dim3(1), dim3(kBoxBlocks), 0, st, (const double*)box_part,
               min_extent, geo_buf);
  }
  // spatial order of sources and targets (one order when they are the same points)
  const int* sperm = nullptr;
  const int* tperm = nullptr;
  // (exclusion ranges keep the given order; the self flag needs one shared
  // order, in which slot i's own source is source i)
  if (sorted) {
    const SortBufs bs = sort_set(p, pl.ns_pad);
    p += sort_set_bytes(pl.ns_pad);
    const SortBufs bt = sort_set(p, nt);
    p += sort_set_bytes(nt);
    int* counts = reinterpret_cast<int*>(p);
    int rc = spatial_order(ns, pos, s_sp, s_sc, geo_buf, bs, counts, st, &sperm);
    if (rc != F3D_OK) return rc;
    tperm = sperm;
    if (!same) {
      rc = spatial_order(nt, tpos, t_sp, t_sc, geo_buf, bt, counts, st, &tperm);
      if (rc != F3D_OK) return rc;
    }
  }
  const bool want_sigma = epi_in != nullptr;
  const int tiles = pl.ns_pad / kTile;
  F3D_LAUNCH(tile_box_kernel, dim3((unsigned)tiles), dim3(64), 0, st, (int)ns, tiles, pos,
             (long)s_sp, (long)s_sc, (const Geo*)geo, sperm, tbox);
  F3D_LAUNCH((pack_kernel<T, KIND>), dim3(kPackBlocks), dim3(256), 0, st, (int)ns, pl.ns_pad,
             pos, nrm, den, w, (long)s_sp, (long)s_sc, (const Geo*)geo,
             (const TileBox*)tbox, sperm, rec, rec_far, want_sigma ? sigma_part : nullptr);

  const dim3 grid((unsigned)ceil_div(nt, (int64_t)64 * pl.K), (unsigned)pl.Pg);
  const dim3 block((unsigned)(64 * pl.G));
  const bool clamp = kQuant && (flags & F3D_CLAMP_NEAR);
  const int self_excl = (flags & F3D_EXCLUDE_SELF) ? 1 : 0;
  // F3D_PAIRSUM_FAR_RATIO overrides the FP32C far ratio (a very large value
  // turns the near/far split off)
  const char* fr = getenv("F3D_PAIRSUM_FAR_RATIO");
  const double far_ratio = kQuant ? ((fr && *fr) ? atof(fr) : kFarRatio32) : 0.0;
  const bool timed = g_timer.on && g_timer.n < kTimerSlots;
  if (timed) F3D_HIP_CHECK(hipEventRecord(g_timer.ev[2 * g_timer.n], st));
#define F3D_PS(KK, CL)                                                                        \
  F3D_LAUNCH((pairsum_kernel<T, KIND, KK, CL>), grid, block, 0, st, rec, (const T*)rec_far,  \
             tperm, pl.chunk, (int)ns, pl.ns_pad, tpos, (long)t_sp, (long)t_sc,               \
             (int)nt, lo, hi,                                                                 \
             self_excl, reg, eps, (const Geo*)geo, (const TileBox*)tbox, far_ratio, part_hi,  \
             part_lo)
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
  if (timed) {
    F3D_HIP_CHECK(hipEventRecord(g_timer.ev[2 * g_timer.n + 1], st));
    ++g_timer.n;
  }

  double factor;
  if (KIND == F3D_STRESSLET) factor = -3.0 / (4.0 * kPi * eta);
  else factor = 1.0 / (8.0 * kPi * eta);
  factor *= TwoOverR<T, KIND>::scale;
  ShellEpi epi{};
  if (epi_in) {
    epi = *epi_in;
    epi.sigma_part = sigma_part;
  }
  F3D_LAUNCH((reduce_kernel<T>), dim3((unsigned)ceil_div(nt, (int64_t)kRedTargets)), dim3(256),
             0, st, part_hi, part_lo, pl.Pg, (int)nt, tperm, factor, (const Geo*)geo,
             Degree<KIND>::value, out, (long)o_sp, (long)o_sc, (flags & F3D_ACCUMULATE) ? 1 : 0,
             epi);
  return F3D_OK;
}

// fit_ws: shrink the source split until the workspace fits ws_bytes (the
// allocation-free entry point, whose scratch has a fixed size) instead of
// requiring the default plan's size.
static int pairsum_impl(int kernel, int precision, uint32_t flags, int64_t ns, const double* pos,
                        const double* nrm, const double* den, const double* w, int64_t s_sp,
                        int64_t s_sc, int64_t nt, const double* tpos, int64_t t_sp, int64_t t_sc,
                        const int32_t* lo, const int32_t* hi, double eta, double reg, double eps,
                        double* out, int64_t o_sp, int64_t o_sc, void* ws, size_t ws_bytes,
                        const ShellEpi* epi, void* stream, bool fit_ws = false) {
  F3D_REQUIRE(kernel == F3D_OSEEN || kernel == F3D_ROTLET || kernel == F3D_STRESSLET, F3D_EINVAL,
              "pairsum: unknown kernel %d", kernel);
  F3D_REQUIRE(precision == F3D_F64 || precision == F3D_F32C, F3D_EINVAL,
              "pairsum: unknown precision %d", precision);
  F3D_REQUIRE(ns >= 0 && nt >= 0, F3D_EINVAL, "pairsum: negative size");
  F3D_REQUIRE(ns < (1ll << 30) && nt < (1ll << 30), F3D_EINVAL,
              "pairsum: ns=%lld nt=%lld exceed 2^30", (long long)ns, (long long)nt);
  F3D_REQUIRE((lo == nullptr) == (hi == nullptr), F3D_EINVAL,
              "pairsum: excl_lo and excl_hi must both be set or both NULL");
  F3D_REQUIRE(!((flags & F3D_EXCLUDE_SELF) && lo != nullptr), F3D_EINVAL,
              "pairsum: F3D_EXCLUDE_SELF and exclusion ranges are exclusive");
  F3D_REQUIRE((flags & ~(F3D_ACCUMULATE | F3D_CLAMP_NEAR | F3D_EXCLUDE_SELF)) == 0u, F3D_EINVAL,
              "pairsum: unknown flags 0x%x", flags);
  F3D_REQUIRE(eta > 0.0, F3D_EINVAL, "pairsum: eta must be positive");
  F3D_REQUIRE(eps >= 0.0 && reg >= 0.0, F3D_EINVAL, "pairsum: eps and reg must be >= 0");
  if (nt == 0) return F3D_OK;
  F3D_REQUIRE(tpos != nullptr && out != nullptr, F3D_EINVAL, "pairsum: NULL target/output");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (ns == 0 && epi == nullptr) {
    if (flags & F3D_ACCUMULATE) return F3D_OK;
    // out = 0 on the caller's layout: reuse the reduce kernel with P = 0.
    F3D_LAUNCH((reduce_kernel<double>), dim3((unsigned)ceil_div(nt, (int64_t)kRedTargets)),
               dim3(256), 0, st, (const double*)nullptr, (const double*)nullptr, 0, (int)nt,
               (const int*)nullptr, 0.0, (const Geo*)nullptr, 1, out, (long)o_sp, (long)o_sc, 0,
               ShellEpi{});
    return F3D_OK;
  }
  F3D_REQUIRE(pos != nullptr && den != nullptr, F3D_EINVAL, "pairsum: NULL source arrays");
  F3D_REQUIRE(kernel != F3D_STRESSLET || nrm != nullptr, F3D_EINVAL,
              "pairsum: stresslet needs normals");
  Plan pl;
  if (fit_ws) {
    pl = make_plan_within(kernel, precision, ns, nt, ws_bytes);
    F3D_REQUIRE(pl.P > 0 && plan_ws(kernel, precision, pl, nt) <= ws_bytes, F3D_EWORKSPACE,
                "pairsum: ns=%lld nt=%lld need more than the %zu bytes of static scratch",
                (long long)ns, (long long)nt, ws_bytes);
  } else {
    pl = make_plan(kernel, precision, ns, nt);
  }
  const size_t need = plan_ws(kernel, precision, pl, nt);
  F3D_REQUIRE(ws != nullptr && ws_bytes >= need, F3D_EWORKSPACE,
              "pairsum: workspace %zu bytes < required %zu", ws_bytes, need);
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

extern "C" int f3d_pairsum_set_launch(int kernel, int precision, int32_t k,
                                      int32_t target_waves) {
  F3D_REQUIRE(kernel == F3D_OSEEN || kernel == F3D_ROTLET || kernel == F3D_STRESSLET, F3D_EINVAL,
              "set_launch: unknown kernel %d", kernel);
  F3D_REQUIRE(precision == F3D_F64 || precision == F3D_F32C, F3D_EINVAL,
              "set_launch: unknown precision %d", precision);
  F3D_REQUIRE(k == 0 || k == 1 || k == 2 || k == 4, F3D_EINVAL, "set_launch: K must be 1, 2 or 4");
  F3D_REQUIRE(target_waves == 0 || (target_waves >= 64 && target_waves <= (1 << 20)), F3D_EINVAL,
              "set_launch: target_waves outside [64, 2^20]");
  const int pp = precision == F3D_F64 ? 0 : 1;
  g_launch_k[kernel][pp] = k ? k : 2;
  g_launch_waves[kernel][pp] = target_waves ? target_waves : 8192;
  return F3D_OK;
}

extern "C" int f3d_pairsum_set_launch_bucket(int kernel, int precision, int32_t lg_nt,
                                             int32_t lg_ns, int32_t k, int32_t target_waves) {
  F3D_REQUIRE(kernel == F3D_OSEEN || kernel == F3D_ROTLET || kernel == F3D_STRESSLET, F3D_EINVAL,
              "set_launch_bucket: unknown kernel %d", kernel);
  F3D_REQUIRE(precision == F3D_F64 || precision == F3D_F32C, F3D_EINVAL,
              "set_launch_bucket: unknown precision %d", precision);
  F3D_REQUIRE(lg_nt >= kLgMin && lg_nt <= kLgMax && lg_ns >= kLgMin && lg_ns <= kLgMax,
              F3D_EINVAL, "set_launch_bucket: log2 sizes (%d, %d) outside [%d, %d]", lg_nt, lg_ns,
              kLgMin, kLgMax);
  F3D_REQUIRE(k == 0 || k == 1 || k == 2 || k == 4, F3D_EINVAL,
              "set_launch_bucket: K must be 0, 1, 2 or 4");
  F3D_REQUIRE(k == 0 || (target_waves >= 64 && target_waves <= (1 << 20)), F3D_EINVAL,
              "set_launch_bucket: target_waves outside [64, 2^20]");
  const int pp = precision == F3D_F64 ? 0 : 1;
  g_bucket_k[kernel][pp][lg_nt - kLgMin][lg_ns - kLgMin] = k;
  g_bucket_waves[kernel][pp][lg_nt - kLgMin][lg_ns - kLgMin] = k ? target_waves : 0;
  return F3D_OK;
}

extern "C" int f3d_pairsum_plan(int kernel, int precision, int64_t ns, int64_t nt,
                                int32_t* out) {
  F3D_REQUIRE(kernel == F3D_OSEEN || kernel == F3D_ROTLET || kernel == F3D_STRESSLET, F3D_EINVAL,
              "pairsum_plan: unknown kernel %d", kernel);
  F3D_REQUIRE(precision == F3D_F64 || precision == F3D_F32C, F3D_EINVAL,
              "pairsum_plan: unknown precision %d", precision);
  F3D_REQUIRE(ns >= 0 && nt >= 0 && ns < (1ll << 30) && nt < (1ll << 30), F3D_EINVAL,
              "pairsum_plan: bad sizes");
  F3D_REQUIRE(out != nullptr, F3D_EINVAL, "pairsum_plan: NULL output");
  const Plan pl = make_plan(kernel, precision, ns, nt);
  out[0] = pl.K;
  out[1] = pl.P;
  out[2] = pl.chunk;
  return F3D_OK;
}

extern "C" int f3d_pairsum_timer(int enable) {
  if (enable && !g_timer.made) {
    for (int i = 0; i < 2 * kTimerSlots; ++i) F3D_HIP_CHECK(hipEventCreate(&g_timer.ev[i]));
    g_timer.made = true;
  }
  g_timer.on = enable != 0;
  if (enable) g_timer.n = 0;
  return F3D_OK;
}

extern "C" int f3d_pairsum_timer_read(double* total_ms, int32_t* launches) {
  F3D_REQUIRE(total_ms != nullptr && launches != nullptr, F3D_EINVAL,
              "timer_read: NULL output");
  double sum = 0.0;
  if (g_timer.made && g_timer.n > 0) {
    F3D_HIP_CHECK(hipEventSynchronize(g_timer.ev[2 * g_timer.n - 1]));
    for (int i = 0; i < g_timer.n; ++i) {
      float ms = 0.f;
      F3D_HIP_CHECK(hipEventElapsedTime(&ms, g_timer.ev[2 * i], g_timer.ev[2 * i + 1]));
      sum += ms;
    }
  }
  *total_ms = sum;
  *launches = g_timer.made ? g_timer.n : 0;
  return F3D_OK;
}

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

// ---------------------------------------------------------------------------
// Allocation-free pair sum: the workspace is a static device array, one copy
// per GPU.  Its address on the current device is taken once with
// hipGetSymbolAddress (the array's name is only a host-side stand-in) and
// kept per device.  Nothing in it has to start at zero: every call writes
// the records, partial sums and bounding box it reads.
namespace f3d {
constexpr size_t kScratchBytes = size_t(F3D_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_for_current_device(void** out) {
  int dev = -1;
  F3D_HIP_CHECK(hipGetDevice(&dev));
  F3D_REQUIRE(dev >= 0 && dev < kMaxDevices, F3D_EINVAL, "pairsum_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    F3D_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    F3D_REQUIRE(p != nullptr, F3D_EHIP, "pairsum_noalloc: no scratch address on device %d", dev);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return F3D_OK;
}
}  // namespace f3d

extern "C" int f3d_pairsum_noalloc(int kernel, int precision, uint32_t flags, int64_t ns,
                                   const double* src_pos, const double* src_nrm,
                                   const double* src_den, const double* src_w, int64_t s_sp,
                                   int64_t s_sc, int64_t nt, const double* trg_pos, int64_t t_sp,
                                   int64_t t_sc, const int32_t* excl_lo, const int32_t* excl_hi,
                                   double eta, double reg, double eps, double* out, int64_t o_sp,
                                   int64_t o_sc, void* stream) {
  void* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc != F3D_OK) return rc;
  return pairsum_impl(kernel, precision, flags, ns, src_pos, src_nrm, src_den, src_w, s_sp, s_sc,
                      nt, trg_pos, t_sp, t_sc, excl_lo, excl_hi, eta, reg, eps, out, o_sp, o_sc,
                      ws, kScratchBytes, nullptr, stream, true);
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
  return pairsum_impl(F3D_STRESSLET, precision, F3D_EXCLUDE_SELF, n, pos, nrm, q, nullptr, sp, sc, n, pos, sp,
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
// Dense Oseen tensor between nt targets and ns sources: block (i, j) =
// (I / r + d d^T / r^3) / (8 pi eta), d = t_i - s_j, with r replaced by
// sqrt(r^2 + reg^2) where r <= eps (the dense builders regularise at r <= eps,
// the sums at r < eps).  layout 0 (AoS, 'r'): entry (3 i + a, 3 j + b);
// layout 1 ('xyz'): entry (a nt + i, b ns + j).  One thread per (i, j).
__global__ void __launch_bounds__(256)
oseen_matrix_kernel(int ns, const double* __restrict__ src, long ssp, long ssc, int nt,
                    const double* __restrict__ trg, long tsp, long tsc, double c8, double reg2,
                    double eps, int layout, double* __restrict__ M, long ld) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= ns) return;
  const long bi = (long)i * tsp, bj = (long)j * ssp;
  const double d[3] = {trg[bi] - src[bj], trg[bi + tsc] - src[bj + ssc],
                       trg[bi + 2 * tsc] - src[bj + 2 * ssc]};
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double r = sqrt(r2);
  if (!(r > eps)) {
    r2 += reg2;
    r = sqrt(r2);
  }
  const double fr = c8 / r, gr = c8 / (r2 * r);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double v = (a == b ? fr : 0.0) + gr * d[a] * d[b];
      if (layout == 0)
        M[(3L * i + a) * ld + 3L * j + b] = v;
      else
        M[((long)a * nt + i) * ld + (long)b * ns + j] = v;
    }
  }
}
}  // namespace f3d

extern "C" int f3d_oseen_matrix(int64_t ns, const double* src, int64_t s_sp, int64_t s_sc,
                                int64_t nt, const double* trg, int64_t t_sp, int64_t t_sc,
                                double eta, double reg, double eps, int32_t layout, double* M,
                                int64_t ld, void* stream) {
  F3D_REQUIRE(ns >= 0 && ns < (1ll << 30) && nt >= 0 && nt < (1ll << 30), F3D_EINVAL,
              "oseen_matrix: bad sizes ns=%lld nt=%lld", (long long)ns, (long long)nt);
  F3D_REQUIRE(layout == 0 || layout == 1, F3D_EINVAL, "oseen_matrix: layout is 0 or 1");
  F3D_REQUIRE(ld >= 3 * ns, F3D_EINVAL, "oseen_matrix: ld < 3 ns");
  F3D_REQUIRE(eta > 0.0, F3D_EINVAL, "oseen_matrix: eta must be positive");
  if (ns == 0 || nt == 0) return F3D_OK;
  F3D_REQUIRE(src && trg && M, F3D_EINVAL, "oseen_matrix: NULL argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  F3D_LAUNCH(oseen_matrix_kernel, dim3((unsigned)ceil_div(ns, 256), (unsigned)nt), dim3(256), 0,
             st, (int)ns, src, (long)s_sp, (long)s_sc, (int)nt, trg, (long)t_sp, (long)t_sc,
             1.0 / (8.0 * kPi * eta), reg * reg, eps, (int)layout, M, (long)ld);
  return F3D_OK;
}

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
