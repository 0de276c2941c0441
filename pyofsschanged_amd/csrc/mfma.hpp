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
#include <type_traits>

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

template <int BM, int BN>
struct MmaTile {
  static constexpr int BK = 32;
  static constexpr int LD = BK + 8;  // padded LDS row (bf16)
  static constexpr int WM = BM / 2, WN = BN / 2;  // per-wavefront tile
  static constexpr int TM = WM / 16, TN = WN / 16;
  static_assert(BM % 32 == 0 && BN % 32 == 0, "tile");
};

// Operand loaders. Each stages a ROWS x 32 slice of op(X) (rows = M or N
// index, columns = k) into LDS as rows of 32 k-contiguous bf16 (row pitch
// LD). A slice is fetched into registers (fetch) while the previous one is
// multiplied, then written to LDS (stash).
//
// RowSrc: the source is k-contiguous, X[r * ld + k] (16-byte loads, 16-byte
// LDS writes). Rows r >= rows are zero.
template <int ROWS>
struct RowSrc {
  static constexpr int LOADS = ROWS * 32 / 8 / 256;
  static_assert(LOADS >= 1, "tile too small for 256 threads");
  const uint16_t* X;
  long long ld, rows, r0;
  uint4 reg[LOADS];
  __device__ __forceinline__ void fetch(long long k0) {
#pragma unroll
    for (int q = 0; q < LOADS; ++q) {
      const int i = threadIdx.x + 256 * q;
      const int row = i / 4, kc = (i % 4) * 8;
      const long long r = r0 + row;
      reg[q] = (r < rows) ? *reinterpret_cast<const uint4*>(X + r * ld + k0 + kc) : make_uint4(0, 0, 0, 0);
    }
  }
  __device__ __forceinline__ void stash(uint16_t* lds) const {
    constexpr int LD = 40;
#pragma unroll
    for (int q = 0; q < LOADS; ++q) {
      const int i = threadIdx.x + 256 * q;
      const int row = i / 4, kc = (i % 4) * 8;
      *reinterpret_cast<uint4*>(&lds[row * LD + kc]) = reg[q];
    }
  }
};

// ColSrc: the source is row-contiguous, X[k * ld + r] (16-byte loads of 8
// consecutive rows of one k, transposed into LDS element by element).
template <int ROWS>
struct ColSrc {
  static constexpr int LOADS = ROWS * 32 / 8 / 256;
  static_assert(LOADS >= 1, "tile too small for 256 threads");
  const uint16_t* X;
  long long ld, rows, r0;
  uint4 reg[LOADS];
  __device__ __forceinline__ void fetch(long long k0) {
#pragma unroll
    for (int q = 0; q < LOADS; ++q) {
      const int i = threadIdx.x + 256 * q;
      const int k = i / (ROWS / 8), rc = (i % (ROWS / 8)) * 8;
      const long long r = r0 + rc;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (r + 8 <= rows) v = *reinterpret_cast<const uint4*>(X + (k0 + k) * ld + r);
      else if (r < rows) {
        uint16_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = 0; j < 8 && r + j < rows; ++j) t[j] = X[(k0 + k) * ld + r + j];
        v = *reinterpret_cast<const uint4*>(t);
      }
      reg[q] = v;
    }
  }
  __device__ __forceinline__ void stash(uint16_t* lds) const {
    constexpr int LD = 40;
#pragma unroll
    for (int q = 0; q < LOADS; ++q) {
      const int i = threadIdx.x + 256 * q;
      const int k = i / (ROWS / 8), rc = (i % (ROWS / 8)) * 8;
      const uint16_t* t = reinterpret_cast<const uint16_t*>(&reg[q]);
#pragma unroll
      for (int j = 0; j < 8; ++j) lds[(rc + j) * LD + k] = t[j];
    }
  }
};

// Row2Src: two k-contiguous sources concatenated along k: k < K0 reads
// X0[r * ld0 + k], k >= K0 reads X1[r * ld1 + k - K0] (K0 a multiple of 32,
// so a slice never straddles the two).
template <int ROWS>
struct Row2Src {
  static constexpr int LOADS = ROWS * 32 / 8 / 256;
  static_assert(LOADS >= 1, "tile too small for 256 threads");
  const uint16_t* X0;
  const uint16_t* X1;
  long long ld0, ld1, K0, rows, r0;
  uint4 reg[LOADS];
  __device__ __forceinline__ void fetch(long long k0) {
    const bool first = k0 < K0;
    const uint16_t* X = first ? X0 : X1;
    const long long ld = first ? ld0 : ld1;
    const long long kk = first ? k0 : k0 - K0;
#pragma unroll
    for (int q = 0; q < LOADS; ++q) {
      const int i = threadIdx.x + 256 * q;
      const int row = i / 4, kc = (i % 4) * 8;
      const long long r = r0 + row;
      reg[q] = (r < rows) ? *reinterpret_cast<const uint4*>(X + r * ld + kk + kc) : make_uint4(0, 0, 0, 0);
    }
  }
  __device__ __forceinline__ void stash(uint16_t* lds) const {
    constexpr int LD = 40;
#pragma unroll
    for (int q = 0; q < LOADS; ++q) {
      const int i = threadIdx.x + 256 * q;
      const int row = i / 4, kc = (i % 4) * 8;
      *reinterpret_cast<uint4*>(&lds[row * LD + kc]) = reg[q];
    }
  }
};

// The MFMA main loop over k in [k_begin, k_end) for one BM x BN tile:
// 4 wavefronts (2 x 2), each accumulating (BM/2) x (BN/2) as 16 x 16 blocks
// of v_mfma_f32_16x16x32_bf16. Accumulator block (i, j) of wavefront
// (wm, wn) holds rows wm*BM/2 + 16 i + 4 (lane / 16) + v, column
// wn*BN/2 + 16 j + lane % 16 in element v.
template <int BM, int BN, class LA, class LB>
__device__ __forceinline__ void mma_loop(LA& la, LB& lb, long long k_begin, long long k_end, uint16_t* As,
                                         uint16_t* Bs, floatx4_t (&acc)[MmaTile<BM, BN>::TM][MmaTile<BM, BN>::TN]) {
  using G = MmaTile<BM, BN>;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
  for (int i = 0; i < G::TM; ++i)
#pragma unroll
    for (int j = 0; j < G::TN; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
  const int lr = lane & 15, kq = (lane >> 4) * 8;
  la.fetch(k_begin);
  lb.fetch(k_begin);
  for (long long k0 = k_begin; k0 < k_end; k0 += G::BK) {
    __syncthreads();
    la.stash(As);
    lb.stash(Bs);
    __syncthreads();
    if (k0 + G::BK < k_end) {
      la.fetch(k0 + G::BK);
      lb.fetch(k0 + G::BK);
    }
    bf16x8_t a[G::TM], b[G::TN];
#pragma unroll
    for (int i = 0; i < G::TM; ++i)
      a[i] = *reinterpret_cast<const bf16x8_t*>(&As[(wm * G::WM + i * 16 + lr) * G::LD + kq]);
#pragma unroll
    for (int j = 0; j < G::TN; ++j)
      b[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[(wn * G::WN + j * 16 + lr) * G::LD + kq]);
#pragma unroll
    for (int i = 0; i < G::TM; ++i)
#pragma unroll
      for (int j = 0; j < G::TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
}

template <int BM, int BN, bool AT, bool BT, class Epi>
__global__ void __launch_bounds__(256) k_gemm_bf16(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                                  GemmShape s, Epi epi) {
  using G = MmaTile<BM, BN>;
  __shared__ __attribute__((aligned(16))) uint16_t As[BM * G::LD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * G::LD];
  const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
  const int z = blockIdx.z;
  const int batch = z / s.ksplit, split = z % s.ksplit;
  const long long Ks = s.K / s.ksplit, k_begin = (long long)split * Ks;
  const uint16_t* Ab = A + batch * s.sA;
  const uint16_t* Bb = B + batch * s.sB;
  using LA = typename std::conditional<AT, ColSrc<BM>, RowSrc<BM>>::type;
  using LB = typename std::conditional<BT, RowSrc<BN>, ColSrc<BN>>::type;
  LA la{Ab, s.lda, s.M, m0};
  LB lb{Bb, s.ldb, s.N, n0};
  floatx4_t acc[G::TM][G::TN];
  mma_loop<BM, BN>(la, lb, k_begin, k_begin + Ks, As, Bs, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
  for (int i = 0; i < G::TM; ++i)
#pragma unroll
    for (int j = 0; j < G::TN; ++j) {
      const long long col = n0 + wn * G::WN + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long long row = m0 + wm * G::WM + i * 16 + (lane >> 4) * 4 + v;
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
