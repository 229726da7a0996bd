// BF16 GEMM on MFMA with FP32 accumulation, with the epilogues the per-node
// MLP of the learned GMRES initial guess needs for training and inference.
// (No reference counterpart: the learned guess is new to this path.)
//
//   C[m, n] = epi( sum_k A[m, k] B[k, n] )
//   A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn] (FP32 in memory,
//   rounded to BF16 -- round to nearest even -- when staged into LDS);
//   C(m, n) = C[m*scm + n*scn] (FP32).
//
// General tiling for CDNA4 (any shape, K splits): a 256-thread workgroup owns
// a 64 x 64 tile of C; each wave owns a 32 x 32 quarter and issues
// v_mfma_f32_32x32x16_bf16 on 16-deep
// K slices staged through LDS (A row-major, B transposed so both fragments
// are 16-byte contiguous LDS reads).  grid.z splits K for the weight-gradient
// products (K = number of nodes), whose tiles are then added atomically.
// The per-node layers (N = 128 or 256, K <= 256) take the streaming kernel
// below.
#include "common.h"

namespace f3d {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kMT = 64, kNT = 64, kKT = 16;

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// epilogue modes
enum { EPI_STORE = 0, EPI_BIAS_RELU = 1, EPI_RELU_GRAD = 2, EPI_ATOMIC = 3, EPI_BIAS = 4 };

template <int EPI>
__global__ void __launch_bounds__(256)
gemm_bf16_kernel(int M, int N, int K, int kchunk, const float* __restrict__ A, long sam,
                 long sak, const float* __restrict__ B, long sbk, long sbn,
                 float* __restrict__ C, long scm, long scn, const float* __restrict__ bias,
                 float* __restrict__ Z, const float* __restrict__ Zin, float alpha) {
  __shared__ __attribute__((aligned(16))) unsigned short As[kMT][kKT + 8];
  __shared__ __attribute__((aligned(16))) unsigned short Bt[kNT][kKT + 8];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * kMT, n0 = blockIdx.x * kNT;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  f32x16_t acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

  for (int k0 = kbeg; k0 < kend; k0 += kKT) {
    // stage A (64 x 16) and B^T (64 x 16): 4 elements per thread each
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;   // 0..1023
      const int r = idx >> 4, kk = idx & 15;
      const int m = m0 + r, k = k0 + kk;
      float va = (m < M && k < kend) ? A[(long)m * sam + (long)k * sak] : 0.0f;
      As[r][kk] = f2bf(va);
      const int n = n0 + r;
      float vb = (n < N && k < kend) ? B[(long)k * sbk + (long)n * sbn] : 0.0f;
      Bt[r][kk] = f2bf(vb);
    }
    __syncthreads();
    const int row = 32 * wm + (lane & 31);
    const int col = 32 * wn + (lane & 31);
    const int kofs = 8 * (lane >> 5);
    // 16-byte LDS reads (row pitch 48 B, offsets 0 / 16 B: aligned)
    const bf16x8_t a = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4*>(&As[row][kofs]));
    const bf16x8_t b = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4*>(&Bt[col][kofs]));
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    __syncthreads();
  }

  // C layout of the 32 x 32 MFMA: register i of lane l holds
  // row 8*(i/4) + 4*(l/32) + i%4, column l%32.
  const int n = n0 + 32 * wn + (lane & 31);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int m = m0 + 32 * wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
    if (m >= M || n >= N) continue;
    float v = alpha * acc[i];
    const long ci = (long)m * scm + (long)n * scn;
    if (EPI == EPI_ATOMIC) {
      atomicAdd(&C[ci], v);
    } else if (EPI == EPI_BIAS_RELU || EPI == EPI_BIAS) {
      if (bias) v += bias[n];
      if (EPI == EPI_BIAS_RELU) {
        if (Z) Z[ci] = v;
        v = v > 0.0f ? v : 0.0f;
      }
      C[ci] = v;
    } else if (EPI == EPI_RELU_GRAD) {
      C[ci] = Zin[ci] > 0.0f ? v : 0.0f;
    } else {
      C[ci] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Streaming path for the per-node layers (M = nodes, large; N = 64 NT up to
// 256; K <= 256): a workgroup owns 128 rows and ALL N columns, so every
// activation row is read from HBM once and written once.  2 x NT waves, each
// owning a 64 x 64 block as 2 x 2 32x32 MFMA tiles (64 accumulator
// registers), so a 256-wide layer runs 8 waves per workgroup.  K is staged
// through LDS in 32-deep chunks, double
// buffered: the next chunk's global loads (float4) are issued before the
// current chunk's MFMAs, one barrier per chunk; each 16-deep step reuses
// every A fragment across the wave's two column tiles and every B fragment
// across its two row tiles.
constexpr int kSM = 128;  // rows per workgroup
constexpr int kSK = 32;   // K chunk
constexpr int kSP = kSK + 8;  // LDS row pitch (bf16 elements): 80 B, 16-byte aligned

template <int EPI, int NT>  // NT = N / 64 (2 or 4); 128 NT threads
__global__ void __launch_bounds__(512)
gemm_bf16_stream_kernel(int M, int N, int K, const float* __restrict__ A, long sam,
                        const float* __restrict__ B, long sbk, long sbn,
                        float* __restrict__ C, long scm, const float* __restrict__ bias,
                        float* __restrict__ Z, const float* __restrict__ Zin, float alpha) {
  constexpr int NW = 64 * NT;      // columns of the workgroup = N
  constexpr int NTH = 128 * NT;    // threads
  constexpr int AK = kSK / NT;     // A floats per thread and chunk
  constexpr int BI = kSK / 16;     // B items (k pair, column quad) per thread
  __shared__ __attribute__((aligned(16))) unsigned short As[2][kSM][kSP];
  __shared__ __attribute__((aligned(16))) unsigned short Bt[2][NW][kSP];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / NT, wn = wave % NT;
  const int m0 = blockIdx.x * kSM;
  f32x16_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.0f;

  // staging: A chunk 128 x kSK (row tid / NT, AK consecutive k per thread,
  // float4 pieces), B chunk kSK x NW as BI items per thread of (k pair, 4
  // columns): two float4 rows, stored transposed as four (k, k+1) bf16
  // pairs.  Vector loads where the row is 16-byte aligned and in range.
  float4 ra[AK / 4];
  float4 rb[BI][2];
  const bool bvec = sbn == 1 && ((reinterpret_cast<uintptr_t>(B) | (uintptr_t)(sbk * 4)) & 15) == 0;
  auto ld4 = [&](const float* __restrict__ p, int avail, bool vec) __attribute__((always_inline)) {
    if (vec && avail >= 4) return *reinterpret_cast<const float4*>(p);
    return make_float4(avail > 0 ? p[0] : 0.0f, avail > 1 ? p[1] : 0.0f,
                       avail > 2 ? p[2] : 0.0f, avail > 3 ? p[3] : 0.0f);
  };
  auto load = [&](int k0) __attribute__((always_inline)) {
    const int r = tid / NT, kh = (tid % NT) * AK;
    const int m = m0 + r;
    const float* __restrict__ arow = A + (long)(m < M ? m : 0) * sam;
#pragma unroll
    for (int q = 0; q < AK / 4; ++q) {
      const int k = k0 + kh + 4 * q;
      const float* __restrict__ pa = arow + k;
      const bool al = (reinterpret_cast<uintptr_t>(pa) & 15) == 0;
      ra[q] = m < M ? ld4(pa, K - k, al) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int e = 0; e < BI; ++e) {
      // over (k pair, column quad), k pairs fastest: consecutive lanes store
      // consecutive words of one LDS row (no bank conflicts)
      const int idx = tid + NTH * e;
      const int kp = idx % (kSK / 2), nq = idx / (kSK / 2);
      const int k = k0 + 2 * kp, n = 4 * nq;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int kk = k + h;
        if (kk < K && bvec) {
          rb[e][h] = ld4(B + (long)kk * sbk + n, N - n, true);
        } else if (kk < K) {
          const float* __restrict__ pb = B + (long)kk * sbk + (long)n * sbn;
          rb[e][h] = make_float4(n < N ? pb[0] : 0.0f, n + 1 < N ? pb[sbn] : 0.0f,
                                 n + 2 < N ? pb[2 * sbn] : 0.0f, n + 3 < N ? pb[3 * sbn] : 0.0f);
        } else {
          rb[e][h] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
      }
    }
  };
  auto pack2 = [](float lo, float hi) __attribute__((always_inline)) {
    return (unsigned)f2bf(lo) | ((unsigned)f2bf(hi) << 16);
  };
  auto store = [&](int buf) __attribute__((always_inline)) {
    const int r = tid / NT, kh = (tid % NT) * AK;
#pragma unroll
    for (int q = 0; q < AK / 8; ++q) {
      uint4 u;
      u.x = pack2(ra[2 * q].x, ra[2 * q].y);
      u.y = pack2(ra[2 * q].z, ra[2 * q].w);
      u.z = pack2(ra[2 * q + 1].x, ra[2 * q + 1].y);
      u.w = pack2(ra[2 * q + 1].z, ra[2 * q + 1].w);
      *reinterpret_cast<uint4*>(&As[buf][r][kh + 8 * q]) = u;
    }
#pragma unroll
    for (int e = 0; e < BI; ++e) {
      const int idx = tid + NTH * e;
      const int kp = idx % (kSK / 2), nq = idx / (kSK / 2);
      const int n = 4 * nq;
      *reinterpret_cast<unsigned*>(&Bt[buf][n + 0][2 * kp]) = pack2(rb[e][0].x, rb[e][1].x);
      *reinterpret_cast<unsigned*>(&Bt[buf][n + 1][2 * kp]) = pack2(rb[e][0].y, rb[e][1].y);
      *reinterpret_cast<unsigned*>(&Bt[buf][n + 2][2 * kp]) = pack2(rb[e][0].z, rb[e][1].z);
      *reinterpret_cast<unsigned*>(&Bt[buf][n + 3][2 * kp]) = pack2(rb[e][0].w, rb[e][1].w);
    }
  };
  load(0);
  store(0);
  __syncthreads();
  const int nchunks = (K + kSK - 1) / kSK;
  for (int c = 0; c < nchunks; ++c) {
    const int buf = c & 1;
    if (c + 1 < nchunks) load((c + 1) * kSK);  // in flight during the MFMAs
#pragma unroll
    for (int ks = 0; ks < kSK; ks += 16) {
      const int kofs = ks + 8 * (lane >> 5);
      bf16x8_t fa[2], fb[2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
        fa[a] = __builtin_bit_cast(
            bf16x8_t, *reinterpret_cast<const uint4*>(&As[buf][64 * wm + 32 * a + (lane & 31)][kofs]));
#pragma unroll
      for (int b = 0; b < 2; ++b)
        fb[b] = __builtin_bit_cast(
            bf16x8_t, *reinterpret_cast<const uint4*>(&Bt[buf][64 * wn + 32 * b + (lane & 31)][kofs]));
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    if (c + 1 < nchunks) store(buf ^ 1);
    __syncthreads();
  }
  // epilogue: register i of lane l holds row 8(i/4) + 4(l/32) + i%4, column l%32
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int n = 64 * wn + 32 * b + (lane & 31);
      const float bn = (bias && n < N) ? bias[n] : 0.0f;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int m = m0 + 64 * wm + 32 * a + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        if (m >= M || n >= N) continue;
        float v = alpha * acc[a][b][i];
        const long ci = (long)m * scm + n;
        if (EPI == EPI_BIAS_RELU || EPI == EPI_BIAS) {
          v += bn;
          if (EPI == EPI_BIAS_RELU) {
            if (Z) Z[ci] = v;
            v = v > 0.0f ? v : 0.0f;
          }
          C[ci] = v;
        } else if (EPI == EPI_RELU_GRAD) {
          C[ci] = Zin[ci] > 0.0f ? v : 0.0f;
        } else {
          C[ci] = v;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Resident-weight path for the per-node layers: one persistent 512-thread
// workgroup per CU holds all of B (K x N, BF16, transposed, 135 KB at
// N = K = 256) in LDS for the whole launch, so B is read once per CU instead
// of once per row block, and streams blocks of RM rows of A (32 rows at
// N = 256, 64 at N = 128).  The FP32 A rows of the next kRD blocks are in
// flight in registers while a block multiplies: at N = K = 256 that is 64 KB
// of HBM reads outstanding per CU.  Each of the 8 waves owns one 32 x 32
// tile of the block (v_mfma_f32_32x32x16_bf16 over K in 16-deep steps, A
// fragments from the block's LDS rows, B fragments from the resident
// columns).  K must be a multiple of 16 and at most 256, A row-major with
// 16-byte aligned rows.
constexpr int kRKmax = 256;
constexpr int kRP = kRKmax + 8;  // LDS pitch (bf16): 528 B, 16-byte aligned
constexpr int kRD = 2;           // A blocks in flight (2, 3, 4, 6 measured alike)

template <int EPI, int NT>  // N = 64 NT (2 or 4)
__global__ void __launch_bounds__(512)
gemm_bf16_resident_kernel(int M, int N, int K, const float* __restrict__ A, long sam,
                          const float* __restrict__ B, long sbk, long sbn,
                          float* __restrict__ C, long scm, const float* __restrict__ bias,
                          float* __restrict__ Z, const float* __restrict__ Zin, float alpha) {
  constexpr int NW = 64 * NT;        // columns
  constexpr int CT = NW / 32;        // column tiles
  constexpr int RT = 8 / CT;         // row tiles per block
  constexpr int RM = 32 * RT;        // rows per block
  constexpr int PP = RM * kRKmax / 4 / 512;  // float4 pieces per thread (K = 256)
  __shared__ __attribute__((aligned(16))) unsigned short Bs[NW][kRP];
  __shared__ __attribute__((aligned(16))) unsigned short As[RM][kRP];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int ct = wave % CT, rt = wave / CT;
  const int K4 = K / 4;                       // float4 pieces per row
  const int pieces = RM * K4;                 // per block (<= 512 PP)
  const int nblocks = (M + RM - 1) / RM;

  auto pack2 = [](float lo, float hi) __attribute__((always_inline)) {
    return (unsigned)f2bf(lo) | ((unsigned)f2bf(hi) << 16);
  };
  // A pieces of block bl into registers: piece q = tid + 512 j is row q / K4,
  // columns 4 (q % K4) ..; rows past M and pieces past the block read row 0
  // (in bounds) and are zeroed when staged.
  float4 ra[kRD][PP];
  auto load = [&](int bl, float4 (&r)[PP]) __attribute__((always_inline)) {
#pragma unroll
    for (int j = 0; j < PP; ++j) {
      const int q = tid + 512 * j;
      const int row = q / K4, c4 = q - row * K4;
      const long m = (long)bl * RM + row;
      const bool ok = q < pieces && m < M;
      r[j] = *reinterpret_cast<const float4*>(A + (ok ? m : 0) * sam + 4 * (ok ? c4 : 0));
    }
  };
  auto stage = [&](int bl, const float4 (&r)[PP]) __attribute__((always_inline)) {
#pragma unroll
    for (int j = 0; j < PP; ++j) {
      const int q = tid + 512 * j;
      if (q < pieces) {
        const int row = q / K4, c4 = q - row * K4;
        const bool ok = (long)bl * RM + row < M;
        uint2 u;
        u.x = ok ? pack2(r[j].x, r[j].y) : 0u;
        u.y = ok ? pack2(r[j].z, r[j].w) : 0u;
        *reinterpret_cast<uint2*>(&As[row][4 * c4]) = u;
      }
    }
  };

  // the first kRD blocks' A loads, then B into LDS (transposed, BF16)
  int bl = blockIdx.x;
#pragma unroll
  for (int d = 0; d < kRD; ++d) {
    const int b = bl + d * (int)gridDim.x;
    load(b < nblocks ? b : nblocks - 1, ra[d]);
  }
  for (int idx = tid; idx < NW * (K / 2); idx += 512) {
    // (k pair, column), k pairs fastest: consecutive lanes store
    // consecutive words of one LDS row
    const int kp = idx % (K / 2), n = idx / (K / 2);
    const int k = 2 * kp;
    const float b0 = n < N ? B[(long)k * sbk + (long)n * sbn] : 0.0f;
    const float b1 = n < N ? B[(long)(k + 1) * sbk + (long)n * sbn] : 0.0f;
    *reinterpret_cast<unsigned*>(&Bs[n][k]) = pack2(b0, b1);
  }

  const int n = 32 * ct + (lane & 31);
  const float bn = (bias && n < N) ? bias[n] : 0.0f;
  // block i of this workgroup uses register set i % kRD (unrolled by kRD)
  for (; bl < nblocks; bl += kRD * (int)gridDim.x) {
#pragma unroll
    for (int d = 0; d < kRD; ++d) {
      const int b = bl + d * (int)gridDim.x;
      if (b >= nblocks) continue;  // (uniform; the last round's spare set)
      __syncthreads();  // the previous block's MFMAs are done with As
      stage(b, ra[d]);
      __syncthreads();
      // this register set's next block (an in-bounds reload past the end)
      const int nb = b + kRD * (int)gridDim.x;
      load(nb < nblocks ? nb : nblocks - 1, ra[d]);
      f32x16_t acc;
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
      const int kofs = 8 * (lane >> 5);
      // (the compiler reports that it cannot honour this unroll hint in
      // full, but the loop it emits with it measured 4-8% faster than
      // without it or fully unrolled to K = 256)
#pragma unroll 4
      for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8_t fa = __builtin_bit_cast(
            bf16x8_t, *reinterpret_cast<const uint4*>(&As[32 * rt + (lane & 31)][k0 + kofs]));
        const bf16x8_t fb = __builtin_bit_cast(
            bf16x8_t, *reinterpret_cast<const uint4*>(&Bs[32 * ct + (lane & 31)][k0 + kofs]));
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
      }
      // register i of lane l holds row 8(i/4) + 4(l/32) + i%4, column l%32.
      // All 16 results are formed before any store, in registers of their
      // own: a store whose data register is overwritten before it has
      // issued makes the next write wait for it.
      const long m0 = (long)b * RM + 32 * rt + 4 * (lane >> 5);
      float v[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        v[i] = alpha * acc[i];
        if (EPI == EPI_BIAS_RELU || EPI == EPI_BIAS) v[i] += bn;
      }
      auto row_of = [](int i) { return 8 * (i >> 2) + (i & 3); };
      if ((long)(b + 1) * RM <= M) {
        // a whole block: no row checks
        const long base = m0 * scm + n;
        if (EPI == EPI_RELU_GRAD) {
          float zin[16];
#pragma unroll
          for (int i = 0; i < 16; ++i) zin[i] = Zin[base + row_of(i) * scm];
#pragma unroll
          for (int i = 0; i < 16; ++i) C[base + row_of(i) * scm] = zin[i] > 0.0f ? v[i] : 0.0f;
        } else {
          if (EPI == EPI_BIAS_RELU && Z) {
#pragma unroll
            for (int i = 0; i < 16; ++i) Z[base + row_of(i) * scm] = v[i];
          }
#pragma unroll
          for (int i = 0; i < 16; ++i)
            C[base + row_of(i) * scm] = EPI == EPI_BIAS_RELU ? (v[i] > 0.0f ? v[i] : 0.0f) : v[i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const long m = m0 + row_of(i);
          if (m >= M) continue;
          const long ci = m * scm + n;
          if (EPI == EPI_RELU_GRAD) {
            C[ci] = Zin[ci] > 0.0f ? v[i] : 0.0f;
          } else {
            if (EPI == EPI_BIAS_RELU && Z) Z[ci] = v[i];
            C[ci] = EPI == EPI_BIAS_RELU ? (v[i] > 0.0f ? v[i] : 0.0f) : v[i];
          }
        }
      }
    }
  }
}

}  // namespace f3d

using namespace f3d;

// Kernel for the per-node layers N = 128 / N = 256 (set by the autotuner
// through f3d_gemm_set_stream): 0 tiled, 1 streaming, 2 resident-weight.
static int g_gemm_stream[2] = {2, 2};

extern "C" int f3d_gemm_set_stream(int32_t n, int32_t enable) {
  F3D_REQUIRE(n == 128 || n == 256, F3D_EINVAL, "gemm_set_stream: N must be 128 or 256");
  F3D_REQUIRE(enable >= 0 && enable <= 2, F3D_EINVAL,
              "gemm_set_stream: kernel must be 0 (tiled), 1 (streaming) or 2 (resident)");
  g_gemm_stream[n == 256 ? 1 : 0] = enable;
  return F3D_OK;
}

static int gemm_num_cus() {
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

extern "C" int f3d_gemm_bf16(int epi, int64_t M, int64_t N, int64_t K, int32_t ksplit,
                             const float* A, int64_t sam, int64_t sak, const float* B,
                             int64_t sbk, int64_t sbn, float* C, int64_t scm, int64_t scn,
                             const float* bias, float* Z, const float* Zin, float alpha,
                             void* stream) {
  F3D_REQUIRE(epi >= 0 && epi <= 4, F3D_EINVAL, "gemm_bf16: unknown epilogue %d", epi);
  F3D_REQUIRE(M >= 0 && N >= 0 && K >= 0 && M < (1ll << 31) && N < (1ll << 31) &&
                  K < (1ll << 31),
              F3D_EINVAL, "gemm_bf16: bad sizes");
  F3D_REQUIRE(ksplit >= 1, F3D_EINVAL, "gemm_bf16: ksplit must be >= 1");
  F3D_REQUIRE(epi == EPI_ATOMIC || ksplit == 1, F3D_EINVAL,
              "gemm_bf16: K splitting needs the atomic epilogue");
  F3D_REQUIRE(epi != EPI_RELU_GRAD || Zin != nullptr, F3D_EINVAL,
              "gemm_bf16: relu-grad epilogue needs Zin");
  if (M == 0 || N == 0) return F3D_OK;
  F3D_REQUIRE(A && B && C, F3D_EINVAL, "gemm_bf16: NULL operand");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // the streaming path: row-major A (unit K stride) and C (unit N stride),
  // N in {128, 256}, K <= 256, no K split
  const int mode = (N == 128 || N == 256) ? g_gemm_stream[N == 256 ? 1 : 0] : 0;
  const bool stream_ok = epi != EPI_ATOMIC && ksplit == 1 && sak == 1 && scn == 1 && mode > 0 &&
                         K <= 256 && M >= kSM;
  // the resident-weight path also needs K a multiple of 16 and 16-byte rows of A
  const bool resident_ok = stream_ok && mode == 2 && K > 0 && K % 16 == 0 && sam % 4 == 0 &&
                           (reinterpret_cast<uintptr_t>(A) & 15) == 0;
  if (resident_ok) {
    const int RM = N == 256 ? 32 : 64;
    const int64_t nblocks = ceil_div(M, (int64_t)RM);
    const dim3 g((unsigned)std::min<int64_t>(nblocks, gemm_num_cus()));
#define F3D_R(E, T)                                                                          \
  F3D_LAUNCH((gemm_bf16_resident_kernel<E, T>), g, dim3(512), 0, st, (int)M, (int)N, (int)K, A, \
             (long)sam, B, (long)sbk, (long)sbn, C, (long)scm, bias, Z, Zin, alpha)
#define F3D_RE(E) if (N == 256) F3D_R(E, 4); else F3D_R(E, 2)
    switch (epi) {
      case EPI_STORE: F3D_RE(EPI_STORE); break;
      case EPI_BIAS_RELU: F3D_RE(EPI_BIAS_RELU); break;
      case EPI_RELU_GRAD: F3D_RE(EPI_RELU_GRAD); break;
      default: F3D_RE(EPI_BIAS); break;
    }
#undef F3D_RE
#undef F3D_R
    return F3D_OK;
  }
  if (stream_ok) {
    const dim3 g((unsigned)ceil_div(M, kSM));
#define F3D_S(E, T)                                                                            F3D_LAUNCH((gemm_bf16_stream_kernel<E, T>), g, dim3(128 * T), 0, st, (int)M, (int)N, (int)K, A,              (long)sam, B, (long)sbk, (long)sbn, C, (long)scm, bias, Z, Zin, alpha)
#define F3D_SE(E) if (N == 256) F3D_S(E, 4); else F3D_S(E, 2)
    switch (epi) {
      case EPI_STORE: F3D_SE(EPI_STORE); break;
      case EPI_BIAS_RELU: F3D_SE(EPI_BIAS_RELU); break;
      case EPI_RELU_GRAD: F3D_SE(EPI_RELU_GRAD); break;
      default: F3D_SE(EPI_BIAS); break;
    }
#undef F3D_SE
#undef F3D_S
    return F3D_OK;
  }
  int64_t kchunk = ceil_div(K > 0 ? K : 1, ksplit);
  kchunk = (int64_t)align_up((size_t)kchunk, kKT);
  const unsigned gz = (unsigned)ceil_div(K > 0 ? K : 1, kchunk);
  const dim3 grid((unsigned)ceil_div(N, kNT), (unsigned)ceil_div(M, kMT), gz);
#define F3D_G(E)                                                                            \
  F3D_LAUNCH((gemm_bf16_kernel<E>), grid, dim3(256), 0, st, (int)M, (int)N, (int)K,         \
             (int)kchunk, A, (long)sam, (long)sak, B, (long)sbk, (long)sbn, C, (long)scm,   \
             (long)scn, bias, Z, Zin, alpha)
  switch (epi) {
    case EPI_STORE: F3D_G(EPI_STORE); break;
    case EPI_BIAS_RELU: F3D_G(EPI_BIAS_RELU); break;
    case EPI_RELU_GRAD: F3D_G(EPI_RELU_GRAD); break;
    case EPI_ATOMIC: F3D_G(EPI_ATOMIC); break;
    default: F3D_G(EPI_BIAS); break;
  }
#undef F3D_G
  return F3D_OK;
}
