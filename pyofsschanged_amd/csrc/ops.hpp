// Device-side GNLSE operators for one field per workgroup.
//
// Each thread owns the points tid + T*k of the field (k < P = N/T): pointwise
// work and the per-instance global slots are touched only at owned indices,
// so they need no barrier; the FFTs exchange data through LDS.
//
// The operators restate, in this register layout, the arithmetic of
//   pyofss/modules/linearity.py:241-272    exp_lin (with amplifier coupling)
//   pyofss/modules/nonlinearity.py:201-270 non / exp_non variants
//   pyofss/modules/amplifier.py:149-158    saturable gain factor
//   pyofss/modules/amplifier.py:368-402    two-level Yb gain factor and pump
//   pyofss/modules/solver.py:266-275, 339-354  ss_symmetric, rk4ip
// with the pyofss transform convention of pyofss/field.py:109-136:
//   fft(A)  = (1/N) sum_n A_n exp(+2 pi i k n / N)   = F+(A)/N
//   ifft(A) =       sum_k A_k exp(-2 pi i k n / N)   = conj(F+(conj(A)))
// Every spectral operator ifft(m * fft(y)) is one routine with a single
// F+ body run twice, so a kernel holds few copies of the transform.
#pragma once
#include "fft.hpp"

namespace gn {

enum NlKind { NL_GENERIC = -1, NL_KERR = 0, NL_ALL = 1, NL_SS = 2, NL_RS = 3, NL_SS_RS = 4 };
enum Method { M_RK4IP = 0, M_SS_SYMMETRIC = 1, M_SS_SYM_RK4 = 2 };
// spectral multiplier kinds
enum MulKind { MUL_ARRAY = 0, MUL_OMEGA = 1, MUL_IOMEGA = 2, MUL_GAIN = 3, MUL_YB = 4 };
// gain models coupled into the linear operator
enum GainKind { GAIN_NONE = 0, GAIN_SATURABLE = 1, GAIN_YB = 2 };

struct Ops {
  int nl_kind;
  int gain;
  double f_R, ss, rs;
  double2 fac;              // nonlinear factor i*gamma (nonlinearity.py:187-192)
  const double2* lin;       // linear factor i*D(omega) - alpha/2, FFT order
  const double2* efix;      // exp(lin * h/2) of the fixed-step path
  const double2* nlmul;     // fac * (1 + omega*ss), FFT order
  const double2* hr;        // fft(h_R) * window_t, FFT order
  const double* omega;      // omega - omega_c, FFT order
  const double2* tw;        // exp(+2 pi i m / N), m < N
  double gain_log;          // ln(10^(0.1 G))
  double gain_len2;         // 2 * amplifier length
  double esat;              // saturation energy
  const double* gprof;      // spectral gain profile, FFT order
  const double* simpw;      // Simpson weights of the energy integral
  // simpw = simp_c everywhere but at points 0, 1, n-2, n-1, where it is
  // simp_c + simp_d[0..3] (simp_flat), so the energy of a field follows from
  // its spectrum (Parseval) and four time samples
  int simp_flat;
  double simp_c, simp_d[4];
  // two-level Yb model (amplifier.py:178-402), arrays in FFT order; terms of
  // zero cross section (NaN or inf in the reference) hold 0
  const double* yb_rs;      // ratio_s / Psat_s
  const double* yb_ip;      // 1 / Psat_s
  const double* yb_as;      // alpha_s
  const double* yb_es;      // eta_s
  double yb_pfac;           // len(A) * dt / Tr: spectral power -> watts
  double yb_rp, yb_psp;     // ratio_p, Psat_p
  double yb_n;              // total ion number N
  double yb_ap, yb_ep;      // alpha_p, eta_p
  double* yb_pp;            // pump power per instance (read at start, written at end)
  // two-channel WDM (linearity.py:188-211, 306-320; nonlinearity.py:197-200,
  // 287-304): channel 0 uses lin/efix/fac, channel 1 these
  int wdm;
  const double2* lin2;      // channel 1 linear factor, FFT order
  const double2* efix2;     // exp(lin2 * h/2) of the fixed-step path
  double2 fac2;             // i*gamma_1
};

__device__ __forceinline__ double2 cexp(double2 z) {
  double s, c;
  sincos(z.y, &s, &c);
  const double e = exp(z.x);
  return make_double2(e * c, e * s);
}

// numpy's complex division (Smith's method), used for ss/B.
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double ac = fabs(b.x), ad = fabs(b.y);
  if (ac >= ad) {
    if (ac == 0.0 && ad == 0.0) return make_double2(a.x / ac, a.y / ad);
    const double r = b.y / b.x, den = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
  }
  const double r = b.x / b.y, den = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}

__device__ __forceinline__ double abs2(double2 a) { return a.x * a.x + a.y * a.y; }
// The 1/6 of the RK4 combinations as a product with the rounded constant
// (within one unit in the last place of the quotient; an FP64 division is
// an instruction sequence ten times longer).
constexpr double kSixth = 1.0 / 6.0;

// The dispersion factor of a linear call: E as stored, or its square (the
// coarse step's exp(factor h/2) from the stored exp(factor h/4): the same
// product the per-attempt set-up would store).
__device__ __forceinline__ double2 efac(double2 e, bool esq) { return esq ? cmul(e, e) : e; }

// Component-wise select. A ternary between two double2 array elements can be
// lowered to a select of their addresses, which pins both arrays in scratch
// memory; selecting the scalars keeps them in registers.
__device__ __forceinline__ double2 sel(bool c, double2 a, double2 b) {
  return make_double2(c ? a.x : b.x, c ? a.y : b.y);
}

// Energy of a field (field.py:161-172, the Simpson sum of |A|^2 over t) from
// its spectrum X = fft(A) when the weights are simp_c except at points 0,
// 1, n-2, n-1 (Ops::simp_flat): E = simp_c n sum |X_k|^2 + sum_j simp_d[j]
// |A_j|^2, with A_j = sum_k X_k exp(-2 pi i k j / n). energy_acc adds point
// i's terms to the partial sums a = [sum |X|^2, A_0, A_1, A_n-1, A_n-2]
// (complex samples as two doubles); energy_of forms E from the reduced sums.
template <int N>
__device__ __forceinline__ void energy_acc(double* a, double2 X, unsigned i, const double2* __restrict__ tw) {
  a[0] += abs2(X);
  a[1] += X.x;
  a[2] += X.y;
  const double2 w1 = tw[i];
  const double2 p1 = cmul(X, conjg(w1));
  a[3] += p1.x;
  a[4] += p1.y;
  const double2 m1 = cmul(X, w1);
  a[5] += m1.x;
  a[6] += m1.y;
  const double2 m2 = cmul(X, tw[(2u * i) & (unsigned)(N - 1)]);
  a[7] += m2.x;
  a[8] += m2.y;
}
template <int N>
__device__ __forceinline__ double energy_of(const Ops& op, const double* a) {
  return op.simp_c * ((double)N * a[0]) + op.simp_d[0] * (a[1] * a[1] + a[2] * a[2]) +
         op.simp_d[1] * (a[3] * a[3] + a[4] * a[4]) + op.simp_d[2] * (a[7] * a[7] + a[8] * a[8]) +
         op.simp_d[3] * (a[5] * a[5] + a[6] * a[6]);
}
// amplifier.py:149-158: the gain exponent g of a linear call of step hl on a
// field of energy En (as Dev::linear forms it).
__device__ __forceinline__ double sat_gain(const Ops& op, double En, double hl) {
  const double G0 = (op.gain_log * hl) / op.gain_len2;
  return G0 / (1.0 + En / op.esat);
}

// Pointwise passes of rk4ip_sp that load and store different arrays. Each
// pointer is offset to the calling thread's first owned point and the points
// follow at strides of T. The compiler keeps loads below earlier stores
// unless it can prove they differ, which it does not here, so a pass written
// point by point waits for one memory round trip per point; these passes
// load a block of C points of every input first, then compute and store it.
// Entry of rk4ip_sp: A_I~ = E A~ into AIs (x is the load buffer).
template <int P, int T>
__device__ __forceinline__ void pw_sp_entry(double2 (&x)[P], const double2* E, bool esq, const double2* As,
                                            double2* AIs) {
  double2 e[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    e[k] = efac(E[T * k], esq);
    x[k] = As[T * k];
  }
#pragma unroll
  for (int k = 0; k < P; ++k) AIs[T * k] = cmul(e[k], x[k]);
}
// The same with A~ already in x (registers).
template <int P, int T>
__device__ __forceinline__ void pw_sp_entry_reg(const double2 (&x)[P], const double2* E, bool esq, double2* AIs) {
  double2 e[P];
#pragma unroll
  for (int k = 0; k < P; ++k) e[k] = efac(E[T * k], esq);
#pragma unroll
  for (int k = 0; k < P; ++k) AIs[T * k] = cmul(e[k], x[k]);
}
// Stage ST of rk4ip_sp once x holds F+(x R(x)) of the stage's field (or,
// when !fresh, Nt itself): Nt = M F+(.)/N (stored to nA when given); k~ = h
// Nt (times E at stage 1) enters the bracket S of A'~; x <- conj of the
// spectrum of the next stage's field (A_I~ + k~/2, E (A_I~ + k2~), or A'~).
// Blocks of C points.
template <int P, int T, int ST, int C>
__device__ __forceinline__ void pw_sp_stage(double2 (&x)[P], bool fresh, double2* nA, double h, double invn,
                                            const double2* M, const double2* E, bool esq, double2* S,
                                            const double2* AIs, double2* out) {
  static_assert(P % C == 0, "block size");
#pragma unroll
  for (int c = 0; c < P; c += C) {
    double2 m[C], a[C], b[C], e[C];
#pragma unroll
    for (int j = 0; j < C; ++j) {
      if (fresh) m[j] = M[T * (c + j)];
      if (ST != 2) e[j] = efac(E[T * (c + j)], esq);
      if (ST != 1) b[j] = S[T * (c + j)];
      if (ST != 4) a[j] = AIs[T * (c + j)];
    }
#pragma unroll
    for (int j = 0; j < C; ++j) {
      double2 n = x[c + j];
      if (fresh) n = cmul(m[j], cscale(n, invn));
      if (nA) nA[T * (c + j)] = n;
      const double2 v = cscale(n, h);
      double2 w;
      if (ST == 1) {  // k0~ = E h Nt(A); next input A_I~ + k0~/2
        const double2 k0 = cmul(e[j], v);
        S[T * (c + j)] = k0;
        w = cadd(a[j], cscale(k0, 0.5));
      } else if (ST == 2) {  // k1~; next input A_I~ + k1~/2
        S[T * (c + j)] = cadd(b[j], cscale(v, 2.0));
        w = cadd(a[j], cscale(v, 0.5));
      } else if (ST == 3) {  // k2~: S <- A_I~ + S/6; next input E (A_I~ + k2~)
        const double2 q = cadd(b[j], cscale(v, 2.0));
        S[T * (c + j)] = cadd(a[j], make_double2(q.x * kSixth, q.y * kSixth));
        w = cmul(e[j], cadd(a[j], v));
      } else {  // k3~: A'~ = k3~/6 + E S
        w = cadd(make_double2(v.x * kSixth, v.y * kSixth), cmul(e[j], b[j]));
        if (out) out[T * (c + j)] = w;
      }
      x[c + j] = conjg(w);
    }
  }
}

// Sum of v over a workgroup of NW wavefronts through red (NW doubles of
// LDS); every thread gets the same value: a butterfly over the 64 lanes, then
// the wavefront partial sums in wavefront order.
template <int NW>
__device__ __forceinline__ double wg_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) s += red[w];
  return s;
}
// M sums at once through buf (M NW doubles of LDS), each in the order of
// wg_sum, so the results are those of M wg_sum calls, bit for bit, with two
// barriers instead of 2 M.
template <int NW, int M>
__device__ __forceinline__ void wg_sum_n(double (&v)[M], double* buf) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] += __shfl_xor(v[m], o, 64);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int m = 0; m < M; ++m) buf[m * NW + (threadIdx.x >> 6)] = v[m];
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += buf[m * NW + w];
    v[m] = s;
  }
}

// KIND fixes the nonlinear operator at compile time (NL_KERR, NL_ALL) or
// leaves it to ops.nl_kind (NL_GENERIC); a fixed kind lets the compiler drop
// the other operators and their register live ranges.
// YB compiles the two-level Yb gain into the linear operator (its branch
// costs registers, so kernels without it do not carry it).
// CFG is the compiled schedule (the autotuner picks among compiled ones):
// bits 0-3 the plan policy of the N-point transform (FFT's PL), bits 4-7
// that of the N/2-point transform of the Raman convolution, bits 8-9 the
// block of the stage passes of rk4ip_sp (0: 4 points, 1: 2, 2: 8). A
// lane-pair plan in either transform renumbers the threads of the whole
// kernel (every transform and pass follows one ownership rule). Bit 10
// (kCfgNoGain): the launch has checked that the operators carry no gain, so
// the linear calls compile without the saturable-gain branch; bit 12
// (kCfgLine): rk4ip runs written out stage by stage (rk4ip_line). The Kerr
// launches without gain set both (kCfgKerrFree): the runtime gain branch and
// the merging paths of the 7-stage loop cost the 16384-point Kerr kernels
// half their speed in register spills. The gain-free launches of the other
// time-domain operators (self-steepening, T_R Raman) set kCfgNoGain: 16384
// points, 35-39 % less time per attempted step. Bit 13 (kCfgSpUnroll): the twelve
// passes of rk4ip_sp unrolled (each pass compiled for its stage and phase);
// bit 14 (kCfgRamanUnroll): the two transforms of raman_real at two call
// sites. Both trade code size for fewer register copies where the loops'
// paths merge: 8192 points at 512 threads spill 100 B per thread instead of
// 264 B, and an attempted step of the 8192-point Raman + self-steepening
// kernel (256 instances) drops from 501 to 443 us (with 8-point stage
// blocks).
constexpr int kCfgNoGain = 0x400, kCfgLine = 0x1000, kCfgKerrFree = kCfgNoGain | kCfgLine;
constexpr int kCfgSpUnroll = 0x2000, kCfgRamanUnroll = 0x4000, kCfgUnrolled = kCfgSpUnroll | kCfgRamanUnroll;
template <int N, int T, int KIND, bool YB = false, int CFG = 0>
struct Dev {
  static constexpr int RENUM = ((CFG & 8) || (CFG & 0x80)) ? 16 : 0;
  using F = FFT<N, T, (CFG & 15) | RENUM>;
  using FH = FFT<N / 2, T, ((CFG >> 4) & 15) | RENUM>;
  static constexpr int TILE = ((CFG >> 8) & 3) == 1 ? 2 : ((CFG >> 8) & 3) == 2 ? 8 : 4;
  static constexpr int P = N / T;
  static constexpr double INVN = 1.0 / N;  // exact: N is a power of two
  static constexpr int NW = T / 64;
  static constexpr bool NOGAIN = (CFG & kCfgNoGain) != 0;
  static constexpr bool LINE = (CFG & kCfgLine) != 0;
  static constexpr int SP_UNROLL = (CFG & kCfgSpUnroll) ? 12 : 1;
  static constexpr int RAMAN_UNROLL = (CFG & kCfgRamanUnroll) ? 2 : 1;
  double* lds;
  double* red;
  const Ops& op;
  typename F::Twiddles tw;
  double pp = 0.0;  // Yb pump power of this instance (identical in every thread)

  __device__ Dev(double* l, double* r, const Ops& o) : lds(l), red(r), op(o) { F::load_twiddles(tw, o.tw); }

  // Owned index of register k. The thread index goes through opaque() so
  // that slot addresses are formed at each use instead of being hoisted out
  // of the step loops and kept live for the whole kernel.
  __device__ __forceinline__ unsigned tid() const { return (unsigned)opaque(F::tid()); }
  __device__ __forceinline__ unsigned idx(int k) const { return tid() + T * k; }
  // The thread's first owned point of an array; its other points follow at
  // strides of T (the pointwise helpers below index it so).
  template <class V>
  __device__ __forceinline__ V* mine(V* p) const { return p + tid(); }
  __device__ __forceinline__ int kind() const { return KIND == NL_GENERIC ? op.nl_kind : KIND; }

  __device__ __forceinline__ void load(double2 (&x)[P], const double2* __restrict__ g) const {
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = g[idx(k)];
  }
  __device__ __forceinline__ void store(double2* __restrict__ g, const double2 (&x)[P]) const {
#pragma unroll
    for (int k = 0; k < P; ++k) g[idx(k)] = x[k];
  }

  // pyofss fft / ifft of a register field (standalone transforms)
  __device__ __forceinline__ void fft(double2 (&x)[P]) {
    F::plus(x, lds, tw);
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = cscale(x[k], INVN);
  }
  __device__ __forceinline__ void ifft(double2 (&x)[P]) {
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = conjg(x[k]);
    F::plus(x, lds, tw);
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = conjg(x[k]);
  }

  // y <- ifft(m * fft(y)); m selected by `mul` (MUL_GAIN: m = exp(g*prof)*arr).
  __device__ __forceinline__ void spectral(double2 (&y)[P], int mul, const double2* __restrict__ arr, double g) {
#pragma unroll 1
    for (int it = 0; it < 2; ++it) {
      F::plus(y, lds, tw);
      if (it == 0) {
        if (mul == MUL_ARRAY) {
#pragma unroll
          for (int k = 0; k < P; ++k) y[k] = conjg(cmul(arr[idx(k)], cscale(y[k], INVN)));
        } else if (mul == MUL_OMEGA) {
#pragma unroll
          for (int k = 0; k < P; ++k) y[k] = conjg(cscale(cscale(y[k], INVN), op.omega[idx(k)]));
        } else if (mul == MUL_IOMEGA) {
#pragma unroll
          for (int k = 0; k < P; ++k) y[k] = conjg(cmul(make_double2(0.0, op.omega[idx(k)]), cscale(y[k], INVN)));
        } else if (mul == MUL_GAIN) {
#pragma unroll
          for (int k = 0; k < P; ++k)
            y[k] = conjg(cscale(cmul(arr[idx(k)], cscale(y[k], INVN)), exp(g * op.gprof[idx(k)])));
        } else if constexpr (YB) {
          // amplifier.py:368-402 with g = h: inversion N2 from the spectral
          // power of the array being propagated (its fft is y), the signal
          // gain exp(g_s h 1e3 / 2) and the pump update exp(g_p h 1e3)
          double s1 = 0.0, s2 = 0.0;
#pragma unroll
          for (int k = 0; k < P; ++k) {
            y[k] = cscale(y[k], INVN);
            const double ps = abs2(y[k]) * op.yb_pfac;
            s1 += op.yb_rs[idx(k)] * ps;
            s2 += op.yb_ip[idx(k)] * ps;
          }
          s1 = block_sum(s1);
          s2 = block_sum(s2);
          const double n2 = (((op.yb_rp * pp) / op.yb_psp + s1) / (1.0 + pp / op.yb_psp + s2)) * op.yb_n;
          const double hg = g * 1e3;
#pragma unroll
          for (int k = 0; k < P; ++k) {
            const double gs = op.yb_as[idx(k)] * n2 - op.yb_es[idx(k)];
            y[k] = conjg(cscale(cmul(arr[idx(k)], y[k]), exp(gs * hg * 0.5)));
          }
          pp = pp * exp((op.yb_ap * n2 - op.yb_ep) * hg);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) y[k] = conjg(y[k]);
  }

  // Sum over the workgroup; every thread gets the same value (fixed order).
  __device__ __forceinline__ double block_sum(double v) { return wg_sum<NW>(v, red); }

  // M block sums at once through buf (M * NW doubles of LDS, not in use);
  // each value is reduced in the order of block_sum.
  template <int M>
  __device__ __forceinline__ void block_sum_n(double (&v)[M], double* buf) {
    wg_sum_n<NW, M>(v, buf);
  }

  // Pump power of instance b (the Yb model only): read before the first
  // linear call, written back after the last.
  __device__ __forceinline__ void pump_begin(size_t b) {
    if constexpr (YB) pp = op.yb_pp[b];
  }
  __device__ __forceinline__ void pump_end(size_t b) const {
    if constexpr (YB) {
      if (threadIdx.x == 0) op.yb_pp[b] = pp;
    }
  }

  // linearity.py:241-272 with amplifier.py:149-158 (saturable gain):
  // x <- ifft(exp(G0/(1+E(x)/E_sat) * profile) * E * fft(x)); with the Yb
  // model (amplifier.py:390-402) the gain comes from the spectrum of x and
  // each call advances the pump power pp.
  __device__ __forceinline__ void linear(double2 (&x)[P], const double2* __restrict__ E, double h) {
    if constexpr (YB) {
      spectral(x, MUL_YB, E, h);
    } else if (NOGAIN) {
      spectral(x, MUL_ARRAY, E, 0.0);
    } else if (op.gain == GAIN_SATURABLE) {
      double e = 0.0;
#pragma unroll
      for (int k = 0; k < P; ++k) e += op.simpw[idx(k)] * abs2(x[k]);
      const double En = block_sum(e);
      const double G0 = (op.gain_log * h) / op.gain_len2;
      spectral(x, MUL_GAIN, E, G0 / (1.0 + En / op.esat));
    } else {
      spectral(x, MUL_ARRAY, E, 0.0);
    }
  }

  // nonlinearity.py non(A): x <- N(x). The input is parked in slot sX while
  // the transforms run, so only one field is live in registers; sT holds the
  // self-steepening term of NL_SS_RS.
  __device__ __forceinline__ void nonlin(double2 (&x)[P], double2* __restrict__ sX, double2* __restrict__ sT) {
    const int kind = this->kind();
    const double2 fac = op.fac;
    if (kind == NL_KERR) {  // :266-267  factor * |A|^2 * A
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double p = abs2(x[k]);
        x[k] = cmul(make_double2(fac.x * p, fac.y * p), x[k]);
      }
      return;
    }
    const int npass = (kind == NL_ALL || kind == NL_SS_RS) ? 2 : 1;
    double2 y[P];
    if (kind == NL_ALL || kind == NL_RS) {
#pragma unroll
      for (int k = 0; k < P; ++k) {
        y[k] = make_double2(abs2(x[k]), 0.0);
        sX[idx(k)] = x[k];
      }
    } else if (kind == NL_SS) {
#pragma unroll
      for (int k = 0; k < P; ++k) {
        y[k] = cscale(x[k], abs2(x[k]));  // term_spm of :225
        sX[idx(k)] = y[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < P; ++k) {
        y[k] = cscale(x[k], abs2(x[k]));  // |A|^2 A of :253
        sX[idx(k)] = x[k];
      }
    }
#pragma unroll 1
    for (int pass = 0; pass < npass; ++pass) {
      int mul = MUL_ARRAY;
      const double2* arr = op.hr;
      if (kind == NL_ALL) arr = pass == 0 ? op.hr : op.nlmul;
      else if (kind == NL_SS) mul = MUL_OMEGA;
      else if (kind == NL_RS) mul = MUL_IOMEGA;
      else mul = pass == 0 ? MUL_OMEGA : MUL_IOMEGA;
      spectral(y, mul, arr, 0.0);
      if (pass == 0 && npass == 2) {
        if (kind == NL_ALL) {  // :205
          const double a1 = 1.0 - op.f_R;
#pragma unroll
          for (int k = 0; k < P; ++k) {
            const double2 a = sX[idx(k)];
            y[k] = cadd(cscale(cscale(a, a1), abs2(a)), cmul(cscale(a, op.f_R), y[k]));
          }
        } else {  // NL_SS_RS: keep term_ss, start term_rs
#pragma unroll
          for (int k = 0; k < P; ++k) {
            sT[idx(k)] = cscale(y[k], op.ss);
            y[k] = make_double2(abs2(sX[idx(k)]), 0.0);
          }
        }
      }
    }
    if (kind == NL_ALL) {  // :207
#pragma unroll
      for (int k = 0; k < P; ++k) x[k] = y[k];
    } else if (kind == NL_SS) {  // :228
#pragma unroll
      for (int k = 0; k < P; ++k) x[k] = cmul(fac, cadd(sX[idx(k)], cscale(y[k], op.ss)));
    } else {  // :241 (RS), :256 (SS_RS adds term_ss outside the factor)
      const bool with_ss = (kind == NL_SS_RS);
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 a = sX[idx(k)];
        const double2 tr = cscale(y[k], op.rs);
        const double2 v = cmul(cmul(fac, make_double2(abs2(a) + tr.x, tr.y)), a);
        double2 ts = make_double2(0.0, 0.0);
        if (with_ss) ts = sT[idx(k)];
        x[k] = cadd(v, ts);
      }
    }
  }

  // nonlinearity.py exp_non(A, h, B): B <- exp(h*factor*(...)) * B, with A
  // read from the global slot gA; B is parked in sX during the transforms.
  __device__ __forceinline__ void exp_nonlin(double2 (&B)[P], const double2* __restrict__ gA, double h, double2* __restrict__ sX,
                             double2* __restrict__ sT) {
    const int kind = this->kind();
    const double2 hf = make_double2(h * op.fac.x, h * op.fac.y);
    if (kind == NL_KERR) {  // :269-270
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double p = abs2(gA[idx(k)]);
        B[k] = cmul(cexp(make_double2(hf.x * p, hf.y * p)), B[k]);
      }
      return;
    }
    const int npass = (kind == NL_ALL || kind == NL_SS_RS) ? 2 : 1;
    double2 y[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const double2 a = gA[idx(k)];
      const double p = abs2(a);
      y[k] = sel(kind == NL_ALL || kind == NL_RS, make_double2(p, 0.0), cscale(a, p));
    }
    store(sX, B);
#pragma unroll 1
    for (int pass = 0; pass < npass; ++pass) {
      int mul;
      if (kind == NL_ALL) mul = pass == 0 ? MUL_ARRAY : MUL_OMEGA;
      else if (kind == NL_SS) mul = MUL_OMEGA;
      else if (kind == NL_RS) mul = MUL_IOMEGA;
      else mul = pass == 0 ? MUL_OMEGA : MUL_IOMEGA;
      spectral(y, mul, op.hr, 0.0);
      if (pass == 0 && npass == 2) {
        if (kind == NL_ALL) {  // :212-214 term_all, then term_all * A
          const double a1 = 1.0 - op.f_R;
#pragma unroll
          for (int k = 0; k < P; ++k) {
            const double2 a = gA[idx(k)];
            const double2 ta = cadd(make_double2(a1 * abs2(a), 0.0), cscale(y[k], op.f_R));
            sT[idx(k)] = ta;
            y[k] = cmul(ta, a);
          }
        } else {  // NL_SS_RS :261 term_ss, then |A|^2 for term_rs
#pragma unroll
          for (int k = 0; k < P; ++k) {
            sT[idx(k)] = cmul(cdiv(make_double2(op.ss, 0.0), sX[idx(k)]), y[k]);
            y[k] = make_double2(abs2(gA[idx(k)]), 0.0);
          }
        }
      }
    }
    if (kind == NL_ALL) {  // :216, :221
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 b = sX[idx(k)];
        const double2 q = sel(hypot(b.x, b.y) > 1e-15, cdiv(make_double2(op.ss, 0.0), b), make_double2(0.0, 0.0));
        B[k] = cmul(cexp(cmul(hf, cadd(sT[idx(k)], cmul(q, y[k])))), b);
      }
    } else if (kind == NL_SS) {  // :233-234
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 b = sX[idx(k)];
        const double2 tss = cmul(cdiv(make_double2(op.ss, 0.0), b), y[k]);
        B[k] = cmul(cexp(cmul(hf, make_double2(abs2(gA[idx(k)]) + tss.x, tss.y))), b);
      }
    } else {  // :246-248 (RS), :263-264 (SS_RS)
      const bool with_ss = (kind == NL_SS_RS);
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 b = sX[idx(k)];
        double2 ts = make_double2(0.0, 0.0);
        if (with_ss) ts = sT[idx(k)];
        const double2 tr = cscale(y[k], op.rs);
        const double p = abs2(gA[idx(k)]);
        const double2 s = make_double2(p + ts.x + tr.x, ts.y + tr.y);  // ts is zero for RS
        B[k] = cmul(cexp(cmul(hf, s)), b);
      }
    }
  }

  // E <- exp(lin * h) at owned indices (linearity.py:242, 246)
  __device__ __forceinline__ void make_exp(double2* __restrict__ E, double h) const { make_exp_of(E, op.lin, h); }
  // exp(factor * h) of a given linear factor (the WDM channels have two)
  __device__ __forceinline__ void make_exp_of(double2* __restrict__ E, const double2* __restrict__ fac, double h) const {
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const double2 f = fac[idx(k)];
      E[idx(k)] = cexp(make_double2(f.x * h, f.y * h));
    }
  }
};

// Per-instance scratch slots (N complex each).
enum Slot { S_A = 0, S_NA, S_H, S_AI, S_K0, S_K1, S_EQ, S_EH, S_T, S_U, S_HS, NSLOT };

// solver.py:339-354 rk4ip as a 7-stage loop (one call site per operator),
// with the linear calls in the reference's order (L(A), L(h N(A)),
// L(A_I + k2), L(acc)), which the Yb gain's pump state depends on.
// Input field in slot `in`; E = exp(factor*h/2). Result in x. N(in) is read
// from nA instead of being computed when nA_ready, and stored to nA when
// nA_store (the step-doubling controller evaluates N of the same field for
// its half step and its coarse step).
template <int N, int T, int K, bool YB, int CFG>
__device__ __forceinline__ void rk4ip(Dev<N, T, K, YB, CFG>& d, double2 (&x)[N / T], const double2* in, const double2* E, double h,
                      double2* ws, double2* nA = nullptr, bool nA_ready = false, bool nA_store = false) {
  constexpr int P = N / T;
  double2* sAI = ws + S_AI * N;
  double2* sK0 = ws + S_K0 * N;
  double2* sK1 = ws + S_K1 * N;
  double2* sT = ws + S_T * N;
  double2* sU = ws + S_U * N;
  const double hh = 0.5 * h;
#pragma unroll 1
  for (int st = 0; st < 7; ++st) {
    // pre
    if (st == 0 || (st == 1 && !nA_ready)) d.load(x, in);
    else if (st == 3 || st == 4) {  // A_I + 0.5*k0, A_I + 0.5*k1
#pragma unroll
      for (int k = 0; k < P; ++k) x[k] = cadd(sAI[d.idx(k)], cscale(x[k], 0.5));
    } else if (st == 6) d.load(x, sK0);
    if (st == 0 || st == 2 || st == 5 || st == 6) d.linear(x, E, hh);
    if (st == 1 || st == 3 || st == 4 || st == 5) {
      if (st == 1 && nA_ready) {
        d.load(x, nA);
      } else {
        d.nonlin(x, sU, sT);
        if (st == 1 && nA_store) d.store(nA, x);
      }
#pragma unroll
      for (int k = 0; k < P; ++k) x[k] = cscale(x[k], h);
    }
    // post
    if (st == 0) d.store(sAI, x);         // A_I = L(A)
    else if (st == 2) d.store(sK0, x);    // k0 = L(h N(A))
    else if (st == 3) d.store(sK1, x);    // k1
    else if (st == 4) {                   // k2 in x
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const int i = d.idx(k);
        const double2 ai = sAI[i];
        const double2 s = cadd(sK0[i], cscale(cadd(sK1[i], x[k]), 2.0));
        sK0[i] = cadd(ai, make_double2(s.x * kSixth, s.y * kSixth));  // A_I + (k0 + 2(k1+k2))/6
        x[k] = cadd(ai, x[k]);                                  // A_I + k2
      }
    } else if (st == 5) d.store(sK1, x);  // k3
    else if (st == 6) {                   // k3/6 + L(acc)
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 k3 = sK1[d.idx(k)];
        x[k] = cadd(make_double2(k3.x * kSixth, k3.y * kSixth), x[k]);
      }
    }
  }
}

// The steps of rk4ip above written out in order, one call site per
// operator call (four linear calls, four N(x)), for kernels whose N(x) is
// pointwise (Kerr). The same operations on the same slots; the only
// difference from the loop is how the compiler schedules and allocates it:
// 16384 points at 1024 threads spill 36 B per thread in this form against
// 436 B in the loop, whose merging paths keep extra copies of the field.
template <int N, int T, int K, bool YB, int CFG>
__device__ __forceinline__ void rk4ip_line(Dev<N, T, K, YB, CFG>& d, double2 (&x)[N / T], const double2* in,
                                           const double2* E, double h, double2* ws, double2* nA, bool nA_ready,
                                           bool nA_store) {
  constexpr int P = N / T;
  double2* sAI = ws + S_AI * N;
  double2* sK0 = ws + S_K0 * N;
  double2* sK1 = ws + S_K1 * N;
  double2* sT = ws + S_T * N;
  double2* sU = ws + S_U * N;
  const double hh = 0.5 * h;
  d.load(x, in);  // A_I = L(A)
  d.linear(x, E, hh);
  d.store(sAI, x);
  if (nA_ready) {  // k0 = L(h N(A))
    d.load(x, nA);
  } else {
    d.load(x, in);
    d.nonlin(x, sU, sT);
    if (nA_store) d.store(nA, x);
  }
#pragma unroll
  for (int k = 0; k < P; ++k) x[k] = cscale(x[k], h);
  d.linear(x, E, hh);
  d.store(sK0, x);
#pragma unroll
  for (int k = 0; k < P; ++k) x[k] = cadd(d.mine(sAI)[T * k], cscale(x[k], 0.5));  // k1 = h N(A_I + k0/2)
  d.nonlin(x, sU, sT);
#pragma unroll
  for (int k = 0; k < P; ++k) x[k] = cscale(x[k], h);
  d.store(sK1, x);
#pragma unroll
  for (int k = 0; k < P; ++k) x[k] = cadd(d.mine(sAI)[T * k], cscale(x[k], 0.5));  // k2 = h N(A_I + k1/2)
  d.nonlin(x, sU, sT);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int i = d.idx(k);
    const double2 k2 = cscale(x[k], h);
    const double2 ai = sAI[i];
    const double2 s = cadd(sK0[i], cscale(cadd(sK1[i], k2), 2.0));
    sK0[i] = cadd(ai, make_double2(s.x * kSixth, s.y * kSixth));  // A_I + (k0 + 2(k1+k2))/6
    x[k] = cadd(ai, k2);                                    // A_I + k2
  }
  d.linear(x, E, hh);  // k3 = h N(L(A_I + k2))
  d.nonlin(x, sU, sT);
#pragma unroll
  for (int k = 0; k < P; ++k) x[k] = cscale(x[k], h);
  d.store(sK1, x);
  d.load(x, sK0);  // A' = L(A_I + (k0 + 2(k1+k2))/6) + k3/6
  d.linear(x, E, hh);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const double2 k3 = sK1[d.idx(k)];
    x[k] = cadd(make_double2(k3.x * kSixth, k3.y * kSixth), x[k]);
  }
}

// Gain models of rk4ip_sp's linear calls.
enum SpGain { SPG_NONE = 0, SPG_SAT = 1, SPG_YB = 2 };
// Partial sums that fix a linear call's gain, per point of the spectrum Y
// of the array passed to it: the energy terms (SPG_SAT, 9 sums) or the
// two spectral-power sums of amplifier.py:376-381 (SPG_YB).
template <int GK>
struct SpSums {
  static constexpr int M = GK == SPG_SAT ? 9 : 2;
};
template <int N, int GK>
__device__ __forceinline__ void gain_acc(const Ops& op, double* a, double2 Y, unsigned i) {
  if constexpr (GK == SPG_SAT) {
    energy_acc<N>(a, Y, i, op.tw);
  } else {
    const double ps = abs2(Y) * op.yb_pfac;
    a[0] += op.yb_rs[i] * ps;
    a[1] += op.yb_ip[i] * ps;
  }
}
// The gain parameter of a linear call of step hl from the reduced sums: the
// exponent g (SPG_SAT) or the inversion N2, which also advances the pump
// power (SPG_YB, amplifier.py:382-388).
template <int N, int GK, class D>
__device__ __forceinline__ double gain_param(D& d, const double* a, double hl) {
  const Ops& op = d.op;
  if constexpr (GK == SPG_SAT) {
    return sat_gain(op, energy_of<N>(op, a), hl);
  } else {
    const double pp = d.pp;
    const double n2 = (((op.yb_rp * pp) / op.yb_psp + a[0]) / (1.0 + pp / op.yb_psp + a[1])) * op.yb_n;
    d.pp = pp * exp((op.yb_ap * n2 - op.yb_ep) * (hl * 1e3));
    return n2;
  }
}
// Per-point operands of the gain factor and the factor itself (as
// Dev::spectral forms it: the spectrum times E, then times the factor).
template <int GK>
__device__ __forceinline__ double2 gain_load(const Ops& op, unsigned i) {
  if constexpr (GK == SPG_SAT) return make_double2(op.gprof[i], 0.0);
  else return make_double2(op.yb_as[i], op.yb_es[i]);
}
template <int GK>
__device__ __forceinline__ double gain_factor(double2 q, double gp, double hl) {
  if constexpr (GK == SPG_SAT) {
    return exp(gp * q.x);
  } else {
    const double gs = q.x * gp - q.y;
    return exp(gs * (hl * 1e3) * 0.5);
  }
}

// Entry of rk4ip_sp with a gain: A_I~ = factor E A~ into AIs. For SPG_SAT
// the exponent g0 comes from A's time samples (the caller); for SPG_YB the
// inversion comes from the spectrum A~ itself. x is the load buffer.
template <int N, int T, int GK, int C, int KIND, bool YB, int CFG>
__device__ __forceinline__ void pw_sp_entry_g(Dev<N, T, KIND, YB, CFG>& d, double2 (&x)[N / T], const double2* E, bool esq,
                                              const double2* As, double2* AIs, double g0, double hl, double* red) {
  constexpr int P = N / T;
  const Ops& op = d.op;
  const unsigned t = d.tid();
#pragma unroll
  for (int k = 0; k < P; ++k) x[k] = As[T * k];
  double gp = g0;
  if constexpr (GK == SPG_YB) {
    double a[SpSums<GK>::M] = {};
#pragma unroll
    for (int k = 0; k < P; ++k) gain_acc<N, GK>(op, a, x[k], t + T * k);
    d.block_sum_n(a, red);
    gp = gain_param<N, GK>(d, a, hl);
  }
#pragma unroll
  for (int c = 0; c < P; c += C) {
    double2 e[C], q[C];
#pragma unroll
    for (int j = 0; j < C; ++j) {
      e[j] = efac(E[T * (c + j)], esq);
      q[j] = gain_load<GK>(op, t + T * (c + j));
    }
#pragma unroll
    for (int j = 0; j < C; ++j) AIs[T * (c + j)] = cscale(cmul(e[j], x[c + j]), gain_factor<GK>(q[j], gp, hl));
  }
}

// The stages of rk4ip_sp whose linear call carries a gain: the gain of L(y)
// follows from y (its energy, or its spectral power), so the spectrum y is
// formed and its sums reduced over the workgroup before it is multiplied by
// the gain factor and E.
//   ST = 1: y = h Nt(A) (x holds F+(x R(x)), or Nt(A) when !fresh);
//           k0~ = factor E y into S, x <- conj(A_I~ + k0~/2).
//   ST = 3: k2~ = h Nt; S <- A_I~ + (S + 2 k2~)/6; y = A_I~ + k2~;
//           x <- conj(factor E y); gp4 <- the gain parameter of L(acc) (in
//           the reference's call order: after L(A_I + k2)).
//   ST = 4: x <- conj(k3~/6 + factor(gp4) E S), A'~ also to out.
// red holds 2 SpSums<GK>::M NW doubles.
template <int N, int T, int ST, int C, int GK, int KIND, bool YB, int CFG>
__device__ __forceinline__ void pw_sp_stage_g(Dev<N, T, KIND, YB, CFG>& d, double2 (&x)[N / T], bool fresh, double2* nA,
                                              double h, double hl, const double2* E, bool esq, double2* S,
                                              const double2* AIs, double2* out, double& gp4, double* red) {
  constexpr int P = N / T;
  constexpr int M = SpSums<GK>::M;
  constexpr double INVN = 1.0 / N;
  const Ops& op = d.op;
  const unsigned t = d.tid();
  if constexpr (ST == 1) {
    double a[M] = {};
#pragma unroll
    for (int k = 0; k < P; ++k) {
      double2 n = x[k];
      if (fresh) n = cmul(op.nlmul[t + T * k], cscale(n, INVN));
      if (nA) nA[t + T * k] = n;
      x[k] = cscale(n, h);
      gain_acc<N, GK>(op, a, x[k], t + T * k);
    }
    d.block_sum_n(a, red);
    const double gp = gain_param<N, GK>(d, a, hl);
#pragma unroll
    for (int c = 0; c < P; c += C) {
      double2 e[C], ai[C], q[C];
#pragma unroll
      for (int j = 0; j < C; ++j) {
        e[j] = efac(E[T * (c + j)], esq);
        ai[j] = AIs[T * (c + j)];
        q[j] = gain_load<GK>(op, t + T * (c + j));
      }
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const double2 k0 = cscale(cmul(e[j], x[c + j]), gain_factor<GK>(q[j], gp, hl));
        S[T * (c + j)] = k0;
        x[c + j] = conjg(cadd(ai[j], cscale(k0, 0.5)));
      }
    }
  } else if constexpr (ST == 3) {
    double a[2 * M] = {};
#pragma unroll
    for (int c = 0; c < P; c += C) {
      double2 m[C], ai[C], b[C];
#pragma unroll
      for (int j = 0; j < C; ++j) {
        m[j] = op.nlmul[t + T * (c + j)];
        ai[j] = AIs[T * (c + j)];
        b[j] = S[T * (c + j)];
      }
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const double2 v = cscale(cmul(m[j], cscale(x[c + j], INVN)), h);
        const double2 q = cadd(b[j], cscale(v, 2.0));
        const double2 sn = cadd(ai[j], make_double2(q.x * kSixth, q.y * kSixth));
        S[T * (c + j)] = sn;
        x[c + j] = cadd(ai[j], v);
        gain_acc<N, GK>(op, a, x[c + j], t + T * (c + j));
        gain_acc<N, GK>(op, a + M, sn, t + T * (c + j));
      }
    }
    d.block_sum_n(a, red);
    const double gp = gain_param<N, GK>(d, a, hl);
    gp4 = gain_param<N, GK>(d, a + M, hl);
#pragma unroll
    for (int c = 0; c < P; c += C) {
      double2 e[C], q[C];
#pragma unroll
      for (int j = 0; j < C; ++j) {
        e[j] = efac(E[T * (c + j)], esq);
        q[j] = gain_load<GK>(op, t + T * (c + j));
      }
#pragma unroll
      for (int j = 0; j < C; ++j)
        x[c + j] = conjg(cscale(cmul(e[j], x[c + j]), gain_factor<GK>(q[j], gp, hl)));
    }
  } else {
    static_assert(ST == 4, "stages 1, 3 and 4 carry a linear call");
#pragma unroll
    for (int c = 0; c < P; c += C) {
      double2 m[C], b[C], e[C], q[C];
#pragma unroll
      for (int j = 0; j < C; ++j) {
        m[j] = op.nlmul[t + T * (c + j)];
        b[j] = S[T * (c + j)];
        e[j] = efac(E[T * (c + j)], esq);
        q[j] = gain_load<GK>(op, t + T * (c + j));
      }
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const double2 v = cscale(cmul(m[j], cscale(x[c + j], INVN)), h);
        const double2 w =
            cadd(make_double2(v.x * kSixth, v.y * kSixth), cscale(cmul(e[j], b[j]), gain_factor<GK>(q[j], gp4, hl)));
        if (out) out[T * (c + j)] = w;
        x[c + j] = conjg(w);
      }
    }
  }
}

// LDS (doubles) of a kernel that runs rk4ip_sp: the N-point transform, the
// N/2-point one and the N real values of the Raman exchanges.
template <int N, int T, int CFG = 0>
constexpr int sp_lds_doubles() {
  constexpr int a = Dev<N, T, NL_ALL, false, CFG>::F::LDS_DOUBLES, b = Dev<N, T, NL_ALL, false, CFG>::FH::LDS_DOUBLES;
  constexpr int m = a > b ? a : b;
  return m > N ? m : N;
}

// The Raman convolution of nonlinearity.py:204, ifft(h_R fft(|x|^2)), of a
// real signal p = |x|^2 through N/2-point transforms (the standard
// real-signal method): z_n = p_2n + i p_2n+1 and Z = F+(z) give
//   P_k = E_k + W^k O_k,  P_k+M = E_k - W^k O_k   (M = N/2, W = e^(2 pi i/N)),
//   E_k = (Z_k + conj Z_M-k)/2,  O_k = (Z_k - conj Z_M-k)/(2i);
// with Q = h_R P/N the inverse is y = F-(Q_k + Q_k+M + i (Q_k - Q_k+M) W^-k)
// over M points, y_m = q_2m + i q_2m+1. The imaginary part of the
// reference's convolution is rounding of a real quantity (h_R is the
// transform of a real response) and is dropped. On entry the LDS holds p
// (raman_stage); on exit x_k holds the convolution at owned point k (real
// part; imaginary part 0).
template <int N, int T, int K, bool YB, int CFG>
__device__ __forceinline__ void raman_stage(Dev<N, T, K, YB, CFG>& d, const double2 (&x)[N / T]) {
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N / T; ++k) d.lds[d.idx(k)] = abs2(x[k]);
}
template <int N, int T, int K, bool YB, int CFG>
__device__ __forceinline__ void raman_real(Dev<N, T, K, YB, CFG>& d, double2 (&x)[N / T]) {
  constexpr int M = N / 2, P = N / T, P2 = M / T;
  using F2 = typename Dev<N, T, K, YB, CFG>::FH;
  constexpr double INVN = 1.0 / N;
  const Ops& op = d.op;
  double* ld = d.lds;
  double2* l2 = reinterpret_cast<double2*>(d.lds);
  const unsigned t = d.tid();
  typename F2::Twiddles tw2;
  F2::template load_twiddles<2>(tw2, op.tw);
  __syncthreads();
  double2 z[P2];
#pragma unroll
  for (int j = 0; j < P2; ++j) z[j] = l2[t + T * j];
#pragma unroll Dev<N, T, K, YB, CFG>::RAMAN_UNROLL
  for (int it = 0; it < 2; ++it) {
    F2::plus(z, d.lds, tw2);
    if (it == 0) {  // pair Z_k with Z_M-k, multiply, fold back to M points
      const unsigned u = d.tid();
      __syncthreads();
#pragma unroll
      for (int j = 0; j < P2; ++j) l2[u + T * j] = z[j];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < P2; ++j) {
        const double2 w = op.tw[u + T * j];
        const double2 zr = conjg(l2[(M - (u + T * j)) & (M - 1)]);
        const double2 e = cscale(cadd(z[j], zr), 0.5);
        const double2 dz = csub(z[j], zr);
        const double2 o = make_double2(0.5 * dz.y, -0.5 * dz.x);  // dz / 2i
        const double2 wo = cmul(w, o);
        const double2 q0 = cmul(op.hr[u + T * j], cscale(cadd(e, wo), INVN));
        const double2 q1 = cmul(op.hr[u + T * j + M], cscale(csub(e, wo), INVN));
        const double2 go = cmul(csub(q0, q1), conjg(w));
        const double2 y = cadd(cadd(q0, q1), make_double2(-go.y, go.x));  // + i go
        z[j] = conjg(y);  // F- = conj . F+ . conj
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < P2; ++j) l2[t + T * j] = conjg(z[j]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < P; ++k) x[k] = make_double2(ld[t + T * k], 0.0);
}

// solver.py:339-354 rk4ip for the full Raman + self-steepening operator
// (nonlinearity.py:201-207), carried in the spectral domain between
// operators (with a gain in the linear calls: GK below, where E stands for
// the gain factor times E). The reference evaluates each linear call as
// ifft(E fft(x)) and each N(x) as ifft(m fft(x R(x))), m = i gamma (1 +
// omega ss); wherever a forward transform follows an inverse one the pair
// cancels, so with Nt(x) = m fft(x R(x)) = fft(N(x)):
//   A_I~ = E A~,          k0~ = E h Nt(A),
//   k1~ = h Nt(ifft(A_I~ + k0~/2)),  k2~ = h Nt(ifft(A_I~ + k1~/2)),
//   k3~ = h Nt(ifft(E (A_I~ + k2~))),
//   A'~ = E (A_I~ + (k0~ + 2 (k1~ + k2~))/6) + k3~/6,  A' = ifft(A'~)
// (the reference forms A_I + k/2 in the time domain; the sum of spectra is
// the same field). That is 8 N-point transforms per call plus 4 Raman
// convolutions of N/2-point transform pairs (raman_real), instead of the
// reference's 24 N-point transforms. Passes p = 0..11 are, per stage
// s = 1..4, the Raman convolution of the stage's field, the forward
// transform of x R(x) and the inverse transform of the next stage's field;
// the N-point transform runs at one call site.
// Input field: time domain in `in`, spectral in `inS` (without a gain, inS
// may be null when Nt(in) is not stored: the spectrum is then formed from
// `in` with one more transform). Result: time domain
// in x and, when outS is given, spectral in outS (neither of them one of
// the slots S_K0, S_K1, S_U used here; `in` may be S_U when Nt(in) is not
// ready, since it is read before the first field is parked there). The
// kernel's LDS holds
// sp_lds_doubles<N, T>() doubles.
// GK: the gain in every linear call (SpGain): SPG_SAT, the saturable gain
// of amplifier.py:149-158 (energies of the arrays the reference passes: A
// from its time samples, the others from their spectra; Ops::simp_flat must
// hold); SPG_YB, the two-level Yb gain of amplifier.py:368-402 (from the
// spectra of those arrays; the pump power advances in the reference's call
// order). The workgroup's red buffer holds 18 NW doubles.
template <int GK, int N, int T, int K, bool YB, int CFG>
__device__ __forceinline__ void rk4ip_sp(Dev<N, T, K, YB, CFG>& d, double2 (&x)[N / T], const double2* in, const double2* inS,
                                         double2* outS, const double2* E, bool esq, double h, double2* ws, double2* nA,
                                         bool nA_ready, bool nA_store) {
  static_assert(K == NL_ALL && YB == (GK == SPG_YB), "spectral-domain rk4ip: full nonlinear operator");
  using F = typename Dev<N, T, K, YB, CFG>::F;
  constexpr int P = N / T;
  constexpr double INVN = Dev<N, T, K, YB, CFG>::INVN;
  double2* sAIs = ws + S_K0 * N;  // A_I~
  double2* sS = ws + S_K1 * N;    // k0~, then k0~ + 2 k1~, then the bracket of A'~
  double2* sU = ws + S_U * N;     // the field x of the running N(x)
  const Ops& op = d.op;
  const double a1 = 1.0 - op.f_R;
  constexpr int TL = Dev<N, T, K, YB, CFG>::TILE;
  constexpr int CW = P < TL ? P : TL;
  int p0 = 0;
  double gp4 = 0.0;  // gain parameter of L(acc), from stage 3
  if constexpr (GK != SPG_NONE) {
    double g0 = 0.0;
    if (GK == SPG_SAT || !nA_ready) d.load(x, in);
    if constexpr (GK == SPG_SAT) {  // L(A): energy of A from its time samples (field.py:161-172)
      double e = 0.0;
#pragma unroll
      for (int k = 0; k < P; ++k) e += op.simpw[d.idx(k)] * abs2(x[k]);
      g0 = sat_gain(op, d.block_sum(e), 0.5 * h);
    }
    if (!nA_ready) raman_stage(d, x);
    pw_sp_entry_g<N, T, GK, CW>(d, x, d.mine(E), esq, d.mine(inS), d.mine(sAIs), g0, 0.5 * h, d.red);
    if (nA_ready) {
      d.load(x, nA);
      p0 = 1;
    }
  } else if (nA_ready) {  // stage 1 starts from the stored Nt(A)
    pw_sp_entry<P, T>(x, d.mine(E), esq, d.mine(inS), d.mine(sAIs));
    d.load(x, nA);
    p0 = 1;
  } else {  // the product after the convolution reads A from `in`
    if (!inS) {  // no stored spectrum: A~ = fft(A) formed here, A_I~ = E A~
      d.load(x, in);
      d.fft(x);
      pw_sp_entry_reg<P, T>(x, d.mine(E), esq, d.mine(sAIs));
    }
    d.load(x, in);
    raman_stage(d, x);
    if (inS) pw_sp_entry<P, T>(x, d.mine(E), esq, d.mine(inS), d.mine(sAIs));
  }
  bool skip = nA_ready;  // the next pass has no transform (Nt(A) was loaded)
#pragma unroll Dev<N, T, K, YB, CFG>::SP_UNROLL
  for (int p = 0; p < 12; ++p) {
    if (p < p0) continue;
    const int st = p / 3 + 1;
    const int ph = p % 3;
    if (ph == 0) {  // x R(x) = (1-f_R) x |x|^2 + f_R x conv, x parked in sU
      raman_real(d, x);
      const double2* xs = st == 1 ? in : sU;
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 a = xs[d.idx(k)];
        x[k] = cadd(cscale(cscale(a, a1), abs2(a)), cscale(cscale(a, op.f_R), x[k].x));
      }
      continue;
    }
    if (!skip) F::plus(x, d.lds, d.tw);
    skip = false;
    if (ph == 1) {  // fft(x R(x)) done: Nt, then the stage's combination
      const bool fresh = !(st == 1 && nA_ready);
      double2* nAo = (st == 1 && nA_store) ? d.mine(nA) : nullptr;
      const double2* M = d.mine(op.nlmul);
      if constexpr (GK != SPG_NONE) {
        const double hl = 0.5 * h;
        double2* out = outS ? d.mine(outS) : nullptr;
        if (st == 1)
          pw_sp_stage_g<N, T, 1, CW, GK>(d, x, fresh, nA_store ? nA : nullptr, h, hl, d.mine(E), esq, d.mine(sS),
                                         d.mine(sAIs), nullptr, gp4, d.red);
        else if (st == 2) pw_sp_stage<P, T, 2, CW>(x, true, nullptr, h, INVN, M, nullptr, false, d.mine(sS), d.mine(sAIs), nullptr);
        else if (st == 3)
          pw_sp_stage_g<N, T, 3, CW, GK>(d, x, true, nullptr, h, hl, d.mine(E), esq, d.mine(sS), d.mine(sAIs), nullptr, gp4,
                                         d.red);
        else pw_sp_stage_g<N, T, 4, CW, GK>(d, x, true, nullptr, h, hl, d.mine(E), esq, d.mine(sS), nullptr, out, gp4, d.red);
        continue;
      }
      if (st == 1) pw_sp_stage<P, T, 1, CW>(x, fresh, nAo, h, INVN, M, d.mine(E), esq, d.mine(sS), d.mine(sAIs), nullptr);
      else if (st == 2) pw_sp_stage<P, T, 2, CW>(x, true, nullptr, h, INVN, M, nullptr, false, d.mine(sS), d.mine(sAIs), nullptr);
      else if (st == 3) pw_sp_stage<P, T, 3, CW>(x, true, nullptr, h, INVN, M, d.mine(E), esq, d.mine(sS), d.mine(sAIs), nullptr);
      else pw_sp_stage<P, T, 4, CW>(x, true, nullptr, h, INVN, M, d.mine(E), esq, d.mine(sS), nullptr, outS ? d.mine(outS) : nullptr);
    } else {  // inverse transform done: the next stage's field (or A')
#pragma unroll
      for (int k = 0; k < P; ++k) x[k] = conjg(x[k]);
      if (st == 4) break;
      d.store(sU, x);  // parked for the product after the convolution
      raman_stage(d, x);
    }
  }
}

// solver.py:266-275 ss_symmetric: L(h/2), exp_non with the phase of the
// un-propagated input, L(h/2).
template <int N, int T, int K, bool YB, int CFG>
__device__ __forceinline__ void ss_symmetric(Dev<N, T, K, YB, CFG>& d, double2 (&x)[N / T], const double2* in, const double2* E, double h,
                             double2* ws) {
  double2* sT = ws + S_T * N;
  double2* sU = ws + S_U * N;
  const double hh = 0.5 * h;
  d.load(x, in);
#pragma unroll 1
  for (int st = 0; st < 3; ++st) {
    if (st == 1) d.exp_nonlin(x, in, h, sU, sT);
    else d.linear(x, E, hh);
  }
}

// solver.py:330-337 ss_sym_rk4: L(h/2), the classical rk4 of solver.py
// (k_j = h N(A_L + c_j k_(j-1)), A_N = A_L + (k0 + 2 (k1 + k2) + k3)/6) on the
// nonlinear operator alone, L(h/2); opencl_fibre.py:929-970 cl_ss_sym_rk4 is
// the same scheme. One nonlinear call site in a 4-stage loop.
template <int N, int T, int K, bool YB, int CFG>
__device__ __forceinline__ void ss_sym_rk4(Dev<N, T, K, YB, CFG>& d, double2 (&x)[N / T], const double2* in,
                                           const double2* E, double h, double2* ws) {
  constexpr int P = N / T;
  double2* sAL = ws + S_AI * N;
  double2* sK0 = ws + S_K0 * N;
  double2* sK1 = ws + S_K1 * N;
  double2* sT = ws + S_T * N;
  double2* sU = ws + S_U * N;
  const double hh = 0.5 * h;
  d.load(x, in);
  d.linear(x, E, hh);
  d.store(sAL, x);
#pragma unroll 1
  for (int st = 0; st < 4; ++st) {
    if (st > 0) {  // A_L + c k_(st-1), c = 1/2, 1/2, 1
      const double c = st == 3 ? 1.0 : 0.5;
#pragma unroll
      for (int k = 0; k < P; ++k) x[k] = cadd(sAL[d.idx(k)], cscale(x[k], c));
    }
    d.nonlin(x, sU, sT);
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = cscale(x[k], h);
    if (st == 0) d.store(sK0, x);
    else if (st == 1) d.store(sK1, x);
    else if (st == 2) {  // k1 + k2 kept; x stays k2 for the next stage
#pragma unroll
      for (int k = 0; k < P; ++k) sK1[d.idx(k)] = cadd(sK1[d.idx(k)], x[k]);
    } else {  // A_N = A_L + (k0 + 2 (k1 + k2) + k3) / 6
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const int i = d.idx(k);
        const double2 s = cadd(cadd(sK0[i], cscale(sK1[i], 2.0)), x[k]);
        x[k] = cadd(sAL[i], make_double2(s.x * kSixth, s.y * kSixth));
      }
    }
  }
  d.linear(x, E, hh);
}

template <int METHOD, int N, int T, int K, bool YB, int CFG>
__device__ __forceinline__ void integrator(Dev<N, T, K, YB, CFG>& d, double2 (&x)[N / T], const double2* in, const double2* E,
                                           double h, double2* ws, double2* nA = nullptr, bool nA_ready = false,
                                           bool nA_store = false) {
  if constexpr (METHOD == M_RK4IP && Dev<N, T, K, YB, CFG>::LINE)
    rk4ip_line(d, x, in, E, h, ws, nA, nA_ready, nA_store);
  else if constexpr (METHOD == M_RK4IP) rk4ip(d, x, in, E, h, ws, nA, nA_ready, nA_store);
  else if constexpr (METHOD == M_SS_SYM_RK4) ss_sym_rk4(d, x, in, E, h, ws);
  else ss_symmetric(d, x, in, E, h, ws);
}

}  // namespace gn
