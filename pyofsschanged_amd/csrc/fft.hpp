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
constexpr double kC16 = 0.9807852804032304;  // cos(pi/16)
constexpr double kS16 = 0.19509032201612828; // sin(pi/16)
constexpr double kC316 = 0.8314696123025452; // cos(3 pi/16)
constexpr double kS316 = 0.5555702330196022; // sin(3 pi/16)

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
// Radix 32 as two radix-16 halves and the twiddles exp(+2 pi i k/32).
__device__ __forceinline__ void dft32(double2* v) {
  double2 e[16], o[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  dft16(e);
  dft16(o);
  o[1] = mul_cs(o[1], kC16, kS16);
  o[2] = mul_cs(o[2], kC8, kS8);
  o[3] = mul_cs(o[3], kC316, kS316);
  o[4] = mul_w8_1(o[4]);
  o[5] = mul_cs(o[5], kS316, kC316);
  o[6] = mul_cs(o[6], kS8, kC8);
  o[7] = mul_cs(o[7], kS16, kC16);
  o[8] = mul_i(o[8]);
  o[9] = mul_cs(o[9], -kS16, kC16);
  o[10] = mul_cs(o[10], -kS8, kC8);
  o[11] = mul_cs(o[11], -kS316, kC316);
  o[12] = mul_w8_3(o[12]);
  o[13] = mul_cs(o[13], -kC316, kS316);
  o[14] = mul_cs(o[14], -kC8, kS8);
  o[15] = mul_cs(o[15], -kC16, kS16);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 16] = csub(e[k], o[k]);
  }
}
template <int R> __device__ __forceinline__ void dft(double2* v) {
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 4) dft4(v);
  else if constexpr (R == 8) dft8(v);
  else if constexpr (R == 16) dft16(v);
  else if constexpr (R == 32) dft32(v);
  else static_assert(R == 2 || R == 4 || R == 8 || R == 16 || R == 32, "unsupported radix");
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

// Pass plan. PL selects the policy (the autotuner's choice among compiled
// plans):
//   bit 0: radix up to 32 in a thread (RM = min(P, 32)), else RM = min(P, 16);
//   bits 1-2: where the one remainder pass goes (when N is not a power of the
//     radix): 0 second to last (first when there are only two passes),
//     1 first, 2 last;
//   bit 3: lane-pair passes. A pass of radix 2P runs across the two lanes l
//     and l + 32 of a wavefront, each holding P of its inputs (threads are
//     renumbered so that owned indices v and v + T/2 sit in those lanes, see
//     FFT::tid): each lane transforms its P inputs, the halves of the two
//     results are swapped between the lanes by v_permlane32_swap (no LDS, no
//     barrier) and a radix-2 step combines them. Such a pass covers one more
//     index bit than a P-point pass, so the plan needs fewer passes and LDS
//     exchanges: the fewest passes, then the fewest lane-pair passes, lane-pair
//     passes first;
//   bit 4: the renumbered threads of bit 3 without lane-pair passes (for a
//     second transform length in a kernel whose other transform has them:
//     every transform and pointwise pass of a kernel follows one ownership
//     rule).
// In-thread radices divide P, so a thread runs P/R butterflies per pass.
struct Plan {
  int npass;
  int radix[16];
};
constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }
constexpr int max_radix(int P, int PL = 0) { return (PL & 1) ? (P < 32 ? P : 32) : (P < 16 ? P : 16); }
constexpr Plan make_plan_lanepair(int N, int P, int PL) {
  const int L = ilog2c(N), ip = ilog2c(P), lp = ip + 1;
  int best_b = 0, best_passes = 1000;
  for (int b = 0; b * lp <= L; ++b) {
    const int rest = L - b * lp;
    const int passes = b + rest / ip + (rest % ip ? 1 : 0);
    if (passes < best_passes) {
      best_passes = passes;
      best_b = b;
    }
  }
  const int rest = L - best_b * lp, a = rest / ip, rem = rest % ip;
  Plan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int full[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int m = 0;
  for (int i = 0; i < best_b; ++i) full[m++] = 2 * P;
  for (int i = 0; i < a; ++i) full[m++] = P;
  if (!rem) {
    p.npass = m;
    for (int i = 0; i < m; ++i) p.radix[i] = full[i];
    return p;
  }
  const int where = (PL >> 1) & 3;
  const int at = where == 1 ? 0 : where == 2 ? m : (m > 1 ? m - 1 : 0);
  p.npass = m + 1;
  for (int i = 0, f = 0; i < m + 1; ++i) p.radix[i] = (i == at) ? (1 << rem) : full[f++];
  return p;
}
constexpr Plan make_plan(int N, int P, int PL = 0) {
  if (PL & 8) return make_plan_lanepair(N, P, PL);
  const int RM = max_radix(P, PL);
  Plan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
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
    const int where = (PL >> 1) & 3;
    const int at = where == 1 ? 0 : where == 2 ? m : (m > 1 ? m - 1 : 0);
    for (int i = 0; i < m + 1; ++i) p.radix[i] = (i == at) ? n : RM;
  }
  return p;
}
constexpr int plan_ns(int N, int P, int pass, int PL = 0) {
  Plan p = make_plan(N, P, PL);
  int ns = 1;
  for (int i = 0; i < pass; ++i) ns *= p.radix[i];
  return ns;
}
constexpr int plan_rmax(int N, int P, int PL) {
  Plan p = make_plan(N, P, PL);
  int r = 0;
  for (int i = 0; i < p.npass; ++i) r = p.radix[i] > r ? p.radix[i] : r;
  return r;
}

// exp(+2 pi i r / 32), r = 0..15: the radix-2 step of a lane-pair pass of
// radix 2P reads entry r 32/(2P).
__device__ constexpr double kW32[16][2] = {
    {1.0, 0.0},
    {0.9807852804032304, 0.19509032201612828},
    {0.9238795325112867, 0.3826834323650898},
    {0.8314696123025452, 0.5555702330196022},
    {0.7071067811865476, 0.7071067811865476},
    {0.5555702330196022, 0.8314696123025452},
    {0.3826834323650898, 0.9238795325112867},
    {0.19509032201612828, 0.9807852804032304},
    {0.0, 1.0},
    {-0.19509032201612828, 0.9807852804032304},
    {-0.3826834323650898, 0.9238795325112867},
    {-0.5555702330196022, 0.8314696123025452},
    {-0.7071067811865476, 0.7071067811865476},
    {-0.8314696123025452, 0.5555702330196022},
    {-0.9238795325112867, 0.3826834323650898},
    {-0.9807852804032304, 0.19509032201612828},
};

// Lanes 32-63 of a trade places with lanes 0-31 of b (v_permlane32_swap on
// both 32-bit halves).
__device__ __forceinline__ void swap_lane_halves(double& a, double& b) {
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
  a = __hiloint2double((int)hi[0], (int)lo[0]);
  b = __hiloint2double((int)hi[1], (int)lo[1]);
}

template <int N, int T, int PL = 0>
struct FFT {
  static constexpr int P = N / T;
  static constexpr bool LANEPAIR = (PL & 8) != 0;
  static constexpr bool RENUMBER = (PL & 24) != 0;
  static constexpr Plan plan = make_plan(N, P, PL);
  static constexpr int NPASS = plan.npass;
  // padding period: the largest radix, i.e. the largest write stride of a pass
  static constexpr int LP = ilog2(plan_rmax(N, P, PL));
  static constexpr int PADDED = N + (N >> LP);
  static constexpr bool SPLIT = (size_t)PADDED * 16 > 160 * 1024;
  static constexpr int LDS_DOUBLES = SPLIT ? PADDED : 2 * PADDED;
  static_assert(N % T == 0 && P >= 2 && P <= 32 && (T % 64) == 0, "bad shape");
  static_assert(NPASS >= 2, "at least two passes");
  static_assert(!LANEPAIR || (P >= 2 && P <= 16), "lane-pair passes: 2 <= P <= 16");

  // Index v of the calling thread in the ownership rule (v owns v + T k).
  // Renumbered (PL bits 3, 4), lane l of wavefront w is v = 32 w + (l mod 32) +
  // (T/2) [l >= 32], so v and v + T/2 share a wavefront; each half-wavefront
  // still owns 32 consecutive indices (coalesced 512-byte segments).
  __device__ static __forceinline__ int tid() {
    const int p = threadIdx.x;
    if constexpr (RENUMBER) return (p & 31) + ((p >> 6) << 5) + ((p >> 5) & 1) * (T / 2);
    else return p;
  }
  static constexpr bool lane_pass(int p) { return plan.radix[p] == 2 * P; }

  // Per-thread twiddle bases. In pass p (NS = product of earlier radices)
  // butterfly j needs the powers w^r of w = tw[(j mod NS) N/(NS R)]. In a
  // thread's own butterflies j = tid + T q: when NS divides T, j mod NS = tid
  // mod NS for every q and one base serves all of them; otherwise (only the
  // last pass, NS = N/R > T) there are NS/T distinct bases, indexed by q mod
  // (NS/T). A lane-pair pass has one butterfly j = v mod T/2 per lane pair
  // and keeps two bases, w and w^2. The bases are loaded once per kernel and
  // kept in registers.
  static constexpr int nbase(int p) {
    if (lane_pass(p)) return 2;
    const int ns = plan_ns(N, P, p, PL);
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
    const int v = tid();
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
      const int ns = plan_ns(N, P, p, PL);
      const int R = plan.radix[p];
      if (lane_pass(p)) {
        const int m = (v & (T / 2 - 1) & (ns - 1)) * (N / (ns * R));
        t.w[base_off(p)] = tw[S * m];
        t.w[base_off(p) + 1] = tw[S * 2 * m];
      } else {
#pragma unroll
        for (int b = 0; b < nbase(p); ++b) t.w[base_off(p) + b] = tw[S * (((v + T * b) & (ns - 1)) * (N / (ns * R)))];
      }
    }
  }

  __device__ static __forceinline__ int pad(int e) { return e + (e >> LP); }

  // Deliver the pass outputs held in x to their owners through the LDS.
  // After an in-thread pass register k holds output r = k / NB of butterfly
  // tid + T (k mod NB); after a lane-pair pass it holds output 2k + s of
  // butterfly tid mod T/2, s = [tid >= T/2].
  template <int R, int NS, bool LANE>
  __device__ static __forceinline__ void exchange(double2 (&x)[P], double* lds, int tid) {
    constexpr int NB = LANE ? 1 : P / R;
    auto dst = [&](int k) {
      if constexpr (LANE) {
        const int j = tid & (T / 2 - 1), s = tid >= T / 2 ? 1 : 0;
        return (j / NS) * NS * R + (j & (NS - 1)) + (2 * k + s) * NS;
      } else {
        const int q = k % NB, r = k / NB;
        const int j = tid + T * q;
        return (j / NS) * NS * R + (j & (NS - 1)) + r * NS;
      }
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
  // (w^2, w^4, w^8 by squaring; w^(a+b) = w^a w^b), about 3 ulp at worst; each
  // power is applied as soon as it is formed, so only the bases stay live.
  // Radix 32 reuses the tree for v[16 + r] with w^16 (depth 5).
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
        if constexpr (R >= 16) {
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
          if constexpr (R == 32) {
            const double2 w16 = cmul(w8, w8);
            v[16] = cmul(v[16], w16);
            twiddle<16>(v + 16, w);
#pragma unroll
            for (int r = 17; r < 32; ++r) v[r] = cmul(v[r], w16);
          }
        }
      }
    }
  }

  // One lane-pair butterfly of radix 2P: register m holds input r = 2m + s
  // of butterfly j = tid mod T/2 (s = [tid >= T/2]), i.e. the lane with s = 0
  // holds the even inputs and its partner (lane + 32) the odd ones. With
  // w^(2m+s) = w^s (w^2)^m applied, each lane forms the P-point transform of
  // its half (E in the even lane, O in the odd one); the lanes trade the
  // halves that the other needs (even lane keeps E[2i] and receives O[2i],
  // odd lane keeps O[2i+1] and receives E[2i+1]) and the radix-2 step
  // X[r] = E[r] + W^r O[r], X[r + P] = E[r] - W^r O[r] (W = exp(+2 pi i/2P))
  // leaves output 2k + s in register k.
  template <int NS, int B>
  __device__ static __forceinline__ void lane_butterfly(double2 (&x)[P], const Twiddles& tw, int tid) {
    const bool odd = tid >= T / 2;
    if constexpr (NS > 1) {
      twiddle<P>(x, opaque(tw.w[B + 1]));
      const double2 w = opaque(tw.w[B]);
      const double2 u = make_double2(odd ? w.x : 1.0, odd ? w.y : 0.0);
#pragma unroll
      for (int m = 0; m < P; ++m) x[m] = cmul(x[m], u);
    }
    dft<P>(x);
#pragma unroll
    for (int i = 0; i < P / 2; ++i) {
      swap_lane_halves(x[2 * i].x, x[2 * i + 1].x);
      swap_lane_halves(x[2 * i].y, x[2 * i + 1].y);
    }
    double2 y[P];
#pragma unroll
    for (int i = 0; i < P / 2; ++i) {
      constexpr int STEP = 32 / (2 * P);
      const double* c0 = kW32[(2 * i) * STEP];
      const double* c1 = kW32[(2 * i + 1) * STEP];
      const double2 c = make_double2(odd ? c1[0] : c0[0], odd ? c1[1] : c0[1]);
      const double2 t = cmul(c, x[2 * i + 1]);
      y[i] = cadd(x[2 * i], t);
      y[i + P / 2] = csub(x[2 * i], t);
    }
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = y[k];
  }

  template <int PASS>
  __device__ static __forceinline__ void pass(double2 (&x)[P], double* lds, const Twiddles& tw, int tid) {
    constexpr int R = plan.radix[PASS];
    constexpr int NS = plan_ns(N, P, PASS, PL);
    constexpr bool LAST = (PASS == NPASS - 1);
    constexpr bool LANE = lane_pass(PASS);
    if constexpr (LANE) {
      lane_butterfly<NS, base_off(PASS)>(x, tw, tid);
    } else {
      constexpr int NB = P / R;
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
    }
    if constexpr (!LAST) {
      exchange<R, NS, LANE>(x, lds, tid);
      pass<PASS + 1>(x, lds, tw, opaque(tid));
    }
  }

  // X_k = sum_n x_n exp(+2 pi i k n / N), unscaled, in place (x[k] at the
  // owned index tid() + T k before and after).
  __device__ static __forceinline__ void plus(double2 (&x)[P], double* lds, const Twiddles& tw) {
    pass<0>(x, lds, tw, opaque(tid()));
  }
};

}  // namespace gn
