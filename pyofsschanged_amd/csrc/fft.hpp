// Register-resident, LDS-exchanged Stockham FFT for one complex128 field per
// workgroup (gfx950).
//
// A workgroup of T threads owns an N-point field; thread `tid` holds the P =
// N/T points with indices tid + T*k (k = 0..P-1) in registers. Every pass of
// the mixed-radix Stockham recursion reads exactly those owned points (its
// inputs are j + r*N/R for butterfly j = tid + T*q), so a pass needs the LDS
// only to deliver its outputs to their owners: one exchange per pass
// boundary, none before the first pass or after the last one. The final pass
// writes its outputs to the owned positions again, so input and output are
// both in natural order at owned indices and pointwise operators fuse
// directly before and after a transform.
//
// Only the +i sign is compiled (X_k = sum_n x_n exp(+2 pi i k n / N)); the
// -i transform is conj(F+(conj(x))), which is exact. One transform body per
// kernel keeps code size and register pressure down.
//
// LDS layout: element e is stored at e + (e >> log2(R)), R the largest
// radix. The first pass writes with a stride of R elements per lane; the
// padding makes that stride R + 1 (odd), so consecutive lanes land in
// distinct banks. Reads of owned positions are unit-stride across lanes.
// When the padded field does not fit the 160 KiB LDS of a gfx950 CU, the
// exchange moves real parts and then imaginary parts through a half-size
// buffer (SPLIT mode).
//
// Twiddles: each thread keeps a few exact bases per pass (from a table tw[m] =
// exp(+2 pi i m / N) built on the host in extended precision) and forms the
// powers it needs in registers, so a pass issues no memory loads for them.
#pragma once
#include <hip/hip_runtime.h>

namespace gn {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 conjg(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 mul_i(double2 a) { return make_double2(-a.y, a.x); }

constexpr double kR2 = 0.7071067811865476;   // sqrt(2)/2
constexpr double kC8 = 0.9238795325112867;   // cos(pi/8)
constexpr double kS8 = 0.3826834323650898;   // sin(pi/8)

// a * exp(i pi/4)
__device__ __forceinline__ double2 mul_w8_1(double2 a) { return make_double2(kR2 * (a.x - a.y), kR2 * (a.y + a.x)); }
// a * exp(3 i pi/4)
__device__ __forceinline__ double2 mul_w8_3(double2 a) { return make_double2(kR2 * (-a.x - a.y), kR2 * (a.x - a.y)); }
__device__ __forceinline__ double2 mul_cs(double2 a, double c, double s) {
  return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// In-register DFTs, X_k = sum_n v_n exp(+2 pi i n k / R).
__device__ __forceinline__ void dft2(double2* v) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
__device__ __forceinline__ void dft4(double2* v) {
  double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  double2 t2 = cadd(v[1], v[3]), t3 = mul_i(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}
__device__ __forceinline__ void dft8(double2* v) {
  double2 e[4] = {v[0], v[2], v[4], v[6]};
  double2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e);
  dft4(o);
  o[1] = mul_w8_1(o[1]);
  o[2] = mul_i(o[2]);
  o[3] = mul_w8_3(o[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}
__device__ __forceinline__ void dft16(double2* v) {
  double2 e[8], o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  dft8(e);
  dft8(o);
  o[1] = mul_cs(o[1], kC8, kS8);
  o[2] = mul_w8_1(o[2]);
  o[3] = mul_cs(o[3], kS8, kC8);
  o[4] = mul_i(o[4]);
  o[5] = mul_cs(o[5], -kS8, kC8);
  o[6] = mul_w8_3(o[6]);
  o[7] = mul_cs(o[7], -kC8, kS8);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 8] = csub(e[k], o[k]);
  }
}
template <int R> __device__ __forceinline__ void dft(double2* v) {
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 4) dft4(v);
  else if constexpr (R == 8) dft8(v);
  else if constexpr (R == 16) dft16(v);
  else static_assert(R == 2 || R == 4 || R == 8 || R == 16, "unsupported radix");
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

// Optimisation barrier (empty asm, no instruction). The transform is called
// inside long-running loops; everything it derives from the thread index and
// the twiddle bases is loop-invariant, and hoisting it (dozens of twiddle
// powers and LDS addresses) would keep it live across the whole kernel.
// Passing those inputs through an opaque copy at each call keeps that work
// local to the transform.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
__device__ __forceinline__ double2 opaque(double2 w) {
  asm volatile("" : "+v"(w.x), "+v"(w.y));
  return w;
}

// Pass plan: radix RM = min(P, 16) for every pass except one remainder pass,
// placed second to last (first when there are only two passes). Every radix
// divides P, so a thread runs P/R butterflies per pass.
struct Plan {
  int npass;
  int radix[8];
};
constexpr int max_radix(int P) { return P < 16 ? P : 16; }
constexpr Plan make_plan(int N, int P) {
  const int RM = max_radix(P);
  Plan p{0, {0, 0, 0, 0, 0, 0, 0, 0}};
  int m = 0, n = N;
  while (n % RM == 0 && n > 1) {
    n /= RM;
    ++m;
  }
  if (n == 1) {
    p.npass = m;
    for (int i = 0; i < m; ++i) p.radix[i] = RM;
  } else {
    p.npass = m + 1;
    const int at = m > 1 ? m - 1 : 0;
    for (int i = 0; i < m + 1; ++i) p.radix[i] = (i == at) ? n : RM;
  }
  return p;
}
constexpr int plan_ns(int N, int P, int pass) {
  Plan p = make_plan(N, P);
  int ns = 1;
  for (int i = 0; i < pass; ++i) ns *= p.radix[i];
  return ns;
}

template <int N, int T>
struct FFT {
  static constexpr int P = N / T;
  // padding period: the largest radix, i.e. the write stride of a pass
  static constexpr int LP = ilog2(max_radix(P));
  static constexpr int PADDED = N + (N >> LP);
  static constexpr bool SPLIT = (size_t)PADDED * 16 > 160 * 1024;
  static constexpr int LDS_DOUBLES = SPLIT ? PADDED : 2 * PADDED;
  static constexpr Plan plan = make_plan(N, P);
  static constexpr int NPASS = plan.npass;
  static_assert(N % T == 0 && P >= 2 && P <= 32 && (T % 64) == 0, "bad shape");

  // Per-thread twiddle bases. In pass p (NS = product of earlier radices)
  // butterfly j = tid + T*q needs the powers w^r of w = tw[(j mod NS) *
  // N/(NS*R)]. When NS divides T, j mod NS = tid mod NS for every q and one
  // base serves all of the thread's butterflies; otherwise (only the last
  // pass, NS = N/R > T) there are NS/T distinct bases, indexed by q mod
  // (NS/T). The bases are loaded once per kernel and kept in registers.
  static constexpr int nbase(int p) {
    const int ns = plan_ns(N, P, p);
    return ns <= T ? 1 : ns / T;
  }
  static constexpr int base_off(int p) {
    int o = 0;
    for (int i = 0; i < p; ++i) o += nbase(i);
    return o;
  }
  static constexpr int NBASE = base_off(NPASS);
  struct Twiddles {
    double2 w[NBASE];
  };
  // tw = exp(+2 pi i m / (S N)): S > 1 reads the table of a longer transform.
  template <int S = 1>
  __device__ static __forceinline__ void load_twiddles(Twiddles& t, const double2* __restrict__ tw) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
      const int ns = plan_ns(N, P, p);
      const int R = plan.radix[p];
#pragma unroll
      for (int b = 0; b < nbase(p); ++b) t.w[base_off(p) + b] = tw[S * (((tid + T * b) & (ns - 1)) * (N / (ns * R)))];
    }
  }

  __device__ static __forceinline__ int pad(int e) { return e + (e >> LP); }

  // Deliver the pass outputs held in x to their owners through the LDS.
  template <int R, int NS>
  __device__ static __forceinline__ void exchange(double2 (&x)[P], double* lds, int tid) {
    constexpr int NB = P / R;
    auto dst = [&](int k) {
      const int q = k % NB, r = k / NB;
      const int j = tid + T * q;
      return (j / NS) * NS * R + (j & (NS - 1)) + r * NS;
    };
    if constexpr (!SPLIT) {
      double2* l2 = reinterpret_cast<double2*>(lds);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < P; ++k) l2[pad(dst(k))] = x[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < P; ++k) x[k] = l2[pad(tid + T * k)];
    } else {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < P; ++k) lds[pad(dst(k))] = x[k].x;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < P; ++k) x[k].x = lds[pad(tid + T * k)];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < P; ++k) lds[pad(dst(k))] = x[k].y;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < P; ++k) x[k].y = lds[pad(tid + T * k)];
    }
  }

  // v[r] *= w^r, r = 1..R-1, with powers from a product tree of depth <= 4
  // (w^2, w^4, w^8 by squaring; w^(a+b) = w^a w^b), about 3 ulp at worst.
  template <int R>
  __device__ static __forceinline__ void twiddle(double2* v, double2 w) {
    if constexpr (R == 2) {
      v[1] = cmul(v[1], w);
    } else {
      const double2 w2 = cmul(w, w);
      const double2 w3 = cmul(w2, w);
      v[1] = cmul(v[1], w);
      v[2] = cmul(v[2], w2);
      v[3] = cmul(v[3], w3);
      if constexpr (R >= 8) {
        const double2 w4 = cmul(w2, w2);
        v[4] = cmul(v[4], w4);
        v[5] = cmul(v[5], cmul(w4, w));
        v[6] = cmul(v[6], cmul(w4, w2));
        v[7] = cmul(v[7], cmul(w4, w3));
        if constexpr (R == 16) {
          const double2 w8 = cmul(w4, w4);
          v[8] = cmul(v[8], w8);
          v[9] = cmul(v[9], cmul(w8, w));
          v[10] = cmul(v[10], cmul(w8, w2));
          v[11] = cmul(v[11], cmul(w8, w3));
          const double2 w12 = cmul(w8, w4);
          v[12] = cmul(v[12], w12);
          v[13] = cmul(v[13], cmul(w12, w));
          v[14] = cmul(v[14], cmul(w12, w2));
          v[15] = cmul(v[15], cmul(w12, w3));
        }
      }
    }
  }

  template <int PASS>
  __device__ static __forceinline__ void pass(double2 (&x)[P], double* lds, const Twiddles& tw, int tid) {
    constexpr int R = plan.radix[PASS];
    constexpr int NS = plan_ns(N, P, PASS);
    constexpr int NB = P / R;
    constexpr bool LAST = (PASS == NPASS - 1);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      double2 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = x[q + r * NB];
      if constexpr (NS > 1) twiddle<R>(v, opaque(tw.w[base_off(PASS) + q % nbase(PASS)]));
      dft<R>(v);
#pragma unroll
      for (int r = 0; r < R; ++r) x[q + r * NB] = v[r];
    }
    if constexpr (!LAST) {
      exchange<R, NS>(x, lds, tid);
      pass<PASS + 1>(x, lds, tw, opaque(tid));
    }
  }

  // X_k = sum_n x_n exp(+2 pi i k n / N), unscaled, in place.
  __device__ static __forceinline__ void plus(double2 (&x)[P], double* lds, const Twiddles& tw) {
    pass<0>(x, lds, tw, opaque((int)threadIdx.x));
  }
};

}  // namespace gn
