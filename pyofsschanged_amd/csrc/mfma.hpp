// BF16 matrix-core GEMM for gfx950 with FP32 accumulation.
//
//   C[b] = op(A[b]) * op(B[b]),  op(A): M x K,  op(B): K x N
//
// AT = false: A is row-major M x K (A[m*lda + k]); AT = true: A is stored as
// K x M (A[k*lda + m]). BT = false: B is row-major K x N (B[k*ldb + n]);
// BT = true: B is stored as N x K (B[n*ldb + k]). Operands are bf16 (raw
// 16-bit patterns), accumulation is FP32 in the MFMA accumulators, and every
// output element is handed to an epilogue functor epi(batch, row, col, value)
// that stores it (any layout, any type, fused bias/activation/residual).
//
// Tiling: a workgroup of 256 threads (4 wavefronts, 2 x 2) computes a BM x BN
// tile; each wavefront owns (BM/2) x (BN/2) as 16 x 16 blocks of
// v_mfma_f32_16x16x32_bf16. K advances in slices of 32 staged through LDS in
// k-contiguous rows (A_s[m][k], B_s[n][k]), so each lane's MFMA operand is one
// 16-byte LDS read. Rows are padded by 8 bf16 (16 bytes) so that the 16 lanes
// reading one k-chunk hit distinct banks. The next slice is fetched from
// global memory into registers while the current one is multiplied.
//
// Split-K: blockIdx.z = batch * ksplit + s; split s covers K range
// [s*K/ksplit, (s+1)*K/ksplit) and the epilogue receives batch index
// batch * ksplit + s (partials are reduced by the caller).
//
// Requirements (checked by the host wrappers): K/ksplit a multiple of 32;
// lda, ldb multiples of 8 and base pointers 16-byte aligned; M and N are
// arbitrary (edges are zero-filled on load and masked on store).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gn {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

// float -> bf16 with round-to-nearest-even (NaN kept quiet)
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

struct GemmShape {
  long long M, N, K;
  long long lda, ldb;
  long long sA, sB;  // batch strides (elements)
  int ksplit;
};

template <int BM, int BN, bool AT, bool BT, class Epi>
__global__ void __launch_bounds__(256) k_gemm_bf16(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                                  GemmShape s, Epi epi) {
  constexpr int BK = 32;
  constexpr int LD = BK + 8;  // padded LDS row (bf16)
  constexpr int WM = BM / 2, WN = BN / 2;  // per-wavefront tile
  constexpr int TM = WM / 16, TN = WN / 16;
  static_assert(BM % 32 == 0 && BN % 32 == 0, "tile");
  __shared__ __attribute__((aligned(16))) uint16_t As[BM * LD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * LD];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
  const int z = blockIdx.z;
  const int batch = z / s.ksplit, split = z % s.ksplit;
  const long long Ks = s.K / s.ksplit, k_begin = (long long)split * Ks;
  const uint16_t* Ab = A + batch * s.sA;
  const uint16_t* Bb = B + batch * s.sB;

  // Global -> register staging. Non-transposed A (and transposed B) rows are
  // k-contiguous: 8 bf16 per 16-byte load. Transposed A (and non-transposed
  // B) are m- (n-) contiguous: 8 consecutive rows of one k per load, written
  // to LDS element by element.
  constexpr int A_LOADS = BM * BK / 8 / 256;
  constexpr int B_LOADS = BN * BK / 8 / 256;
  static_assert(A_LOADS >= 1 && B_LOADS >= 1, "tile too small for 256 threads");
  uint4 ra[A_LOADS], rb[B_LOADS];

  auto fetch = [&](long long k0) {
#pragma unroll
    for (int r = 0; r < A_LOADS; ++r) {
      const int i = tid + 256 * r;
      uint4 v = make_uint4(0, 0, 0, 0);
      if constexpr (!AT) {
        const int row = i / (BK / 8), kc = (i % (BK / 8)) * 8;
        const long long m = m0 + row;
        if (m < s.M) v = *reinterpret_cast<const uint4*>(Ab + m * s.lda + k0 + kc);
      } else {
        const int k = i / (BM / 8), mc = (i % (BM / 8)) * 8;
        const long long m = m0 + mc;
        if (m + 8 <= s.M) v = *reinterpret_cast<const uint4*>(Ab + (k0 + k) * s.lda + m);
        else if (m < s.M) {
          uint16_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          for (int j = 0; j < 8 && m + j < s.M; ++j) t[j] = Ab[(k0 + k) * s.lda + m + j];
          v = *reinterpret_cast<const uint4*>(t);
        }
      }
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < B_LOADS; ++r) {
      const int i = tid + 256 * r;
      uint4 v = make_uint4(0, 0, 0, 0);
      if constexpr (BT) {
        const int row = i / (BK / 8), kc = (i % (BK / 8)) * 8;
        const long long n = n0 + row;
        if (n < s.N) v = *reinterpret_cast<const uint4*>(Bb + n * s.ldb + k0 + kc);
      } else {
        const int k = i / (BN / 8), nc = (i % (BN / 8)) * 8;
        const long long n = n0 + nc;
        if (n + 8 <= s.N) v = *reinterpret_cast<const uint4*>(Bb + (k0 + k) * s.ldb + n);
        else if (n < s.N) {
          uint16_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          for (int j = 0; j < 8 && n + j < s.N; ++j) t[j] = Bb[(k0 + k) * s.ldb + n + j];
          v = *reinterpret_cast<const uint4*>(t);
        }
      }
      rb[r] = v;
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int r = 0; r < A_LOADS; ++r) {
      const int i = tid + 256 * r;
      if constexpr (!AT) {
        const int row = i / (BK / 8), kc = (i % (BK / 8)) * 8;
        *reinterpret_cast<uint4*>(&As[row * LD + kc]) = ra[r];
      } else {
        const int k = i / (BM / 8), mc = (i % (BM / 8)) * 8;
        const uint16_t* t = reinterpret_cast<const uint16_t*>(&ra[r]);
#pragma unroll
        for (int j = 0; j < 8; ++j) As[(mc + j) * LD + k] = t[j];
      }
    }
#pragma unroll
    for (int r = 0; r < B_LOADS; ++r) {
      const int i = tid + 256 * r;
      if constexpr (BT) {
        const int row = i / (BK / 8), kc = (i % (BK / 8)) * 8;
        *reinterpret_cast<uint4*>(&Bs[row * LD + kc]) = rb[r];
      } else {
        const int k = i / (BN / 8), nc = (i % (BN / 8)) * 8;
        const uint16_t* t = reinterpret_cast<const uint16_t*>(&rb[r]);
#pragma unroll
        for (int j = 0; j < 8; ++j) Bs[(nc + j) * LD + k] = t[j];
      }
    }
  };

  floatx4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};

  // lane's operand rows/cols and k-chunk inside a 16 x 16 x 32 block
  const int lr = lane & 15, kq = (lane >> 4) * 8;
  fetch(k_begin);
  for (long long k0 = k_begin; k0 < k_begin + Ks; k0 += BK) {
    __syncthreads();
    stash();
    __syncthreads();
    if (k0 + BK < k_begin + Ks) fetch(k0 + BK);
    bf16x8_t a[TM], b[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
      a[i] = *reinterpret_cast<const bf16x8_t*>(&As[(wm * WM + i * 16 + lr) * LD + kq]);
#pragma unroll
    for (int j = 0; j < TN; ++j)
      b[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[(wn * WN + j * 16 + lr) * LD + kq]);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }

  // accumulator layout of a 16 x 16 block: column = lane % 16, rows
  // 4 * (lane / 16) + v for v = 0..3
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const long long col = n0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long long row = m0 + wm * WM + i * 16 + (lane >> 4) * 4 + v;
        if (row < s.M && col < s.N) epi(z, row, col, acc[i][j][v]);
      }
    }
}

// Launch helper: grid (ceil(N/BN), ceil(M/BM), batch * ksplit).
template <int BM, int BN, bool AT, bool BT, class Epi>
hipError_t gemm_bf16(hipStream_t st, const uint16_t* A, const uint16_t* B, const GemmShape& s, long long batch,
                     const Epi& epi) {
  if (s.ksplit < 1 || s.K % ((long long)s.ksplit * 32) != 0 || s.lda % 8 || s.ldb % 8 ||
      ((uintptr_t)A & 15) || ((uintptr_t)B & 15) || s.sA % 8 || s.sB % 8)
    return hipErrorInvalidValue;
  const dim3 grid((unsigned)((s.N + BN - 1) / BN), (unsigned)((s.M + BM - 1) / BM), (unsigned)(batch * s.ksplit));
  hipLaunchKernelGGL((k_gemm_bf16<BM, BN, AT, BT, Epi>), grid, dim3(256), 0, st, A, B, s, epi);
  return hipGetLastError();
}

}  // namespace gn
