// BF16 GEMM on MFMA with FP32 accumulation, with the epilogues the per-node
// MLP of the learned GMRES initial guess needs for training and inference.
// (No reference counterpart: the learned guess is new to this path.)
//
//   C[m, n] = epi( sum_k A[m, k] B[k, n] )
//   A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn] (FP32 in memory,
//   rounded to BF16 -- round to nearest even -- when staged into LDS);
//   C(m, n) = C[m*scm + n*scn] (FP32).
//
// Tiling for CDNA4: a 256-thread workgroup owns a 64 x 64 tile of C; each
// wave owns a 32 x 32 quarter and issues v_mfma_f32_32x32x16_bf16 on 16-deep
// K slices staged through LDS (A row-major, B transposed so both fragments
// are 16-byte contiguous LDS reads).  grid.z splits K for the weight-gradient
// products (K = number of nodes), whose tiles are then added atomically.
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

}  // namespace f3d

using namespace f3d;

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
