// Propagation on grids too large for one workgroup's registers and LDS:
// N = 32768 and 65536 (pyofss allows up to 262143 samples, domain.py:176;
// its OpenCL examples run 2^15 and 2^16, mamyshev_oscillator_cl.ipynb and
// fopo_cars/FOPO_model_ocl.py:193).
//
// Transform. N = R M with M = 8192 (R = 4 or 8). Index time as
// n = n1 + M n2 and frequency as k = R k1 + k2 (n1, k1 < M; n2, k2 < R):
//   X[R k1 + k2] = sum_n1 e^(2 pi i k1 n1 / M) [ e^(2 pi i k2 n1 / N)
//                    sum_n2 x[n1 + M n2] e^(2 pi i k2 n2 / R) ],
// i.e. R-point DFTs down the columns n1 (a "column pass", with the twiddles
// e^(2 pi i k2 n1 / N)), then one M-point transform per k2 (a "block pass",
// the register-resident Stockham transform of fft.hpp over 1024 threads).
// The spectrum is kept in block layout: position k2 M + k1 holds X[R k1 +
// k2]. Every operator array indexed by frequency (linear factor, Raman
// response, self-steepening and gain arrays) is uploaded in that layout, so
// spectral multiplies stay pointwise; only gnlse_fft's inputs and outputs are
// in natural order. The inverse runs the two passes backwards.
//
// Ownership. Thread t of the 1024 owns positions t + 1024 j + M c (j < 8,
// c < R) of every field, in time and in block-layout frequency alike: the
// column pass of n1 = t + 1024 j touches exactly those positions, and the
// block pass of block c reads and writes positions t + 1024 j of the block
// at owned indices (the Stockham transform's natural-order property). So
// every operator, transform included, reads and writes global memory only at
// the calling thread's own positions; only the block transform's LDS
// exchanges cross threads. Fields live in per-instance workspace slots
// (global memory), not in registers.
//
// The operators restate ops.hpp's reference-order arithmetic on slots:
// rk4ip and ss_symmetric (solver.py:266-275, 339-354), the linear call with
// the saturable or two-level Yb gain (linearity.py:241-272, amplifier.py:
// 149-158, 368-402) and the five nonlinear variants (nonlinearity.py:
// 201-270). The transform and the reductions are non-inlined functions with
// one LDS allocation each, so a kernel holds one copy of the transform body.
#pragma once
#include "kernels.hpp"

namespace gn {

constexpr int kBigT = 1024;
constexpr int kBigM = 8192;

// Workspace slots of an instance on a large grid (N complex each).
enum BigSlot { B_A = 0, B_NA, B_H, B_C, B_AI, B_K0, B_K1, B_U, B_T, B_V, B_EQ, B_EH, NSLOT_BIG };

template <int N>
struct BigGeo {
  static constexpr int T = kBigT, M = kBigM, R = N / M, PB = M / T, PT = N / T;
  static_assert(N % M == 0 && (R == 4 || R == 8), "large grids: 4 or 8 blocks of 8192");
};

// Owned position q (q < PT) of the calling thread.
template <int N>
__device__ __forceinline__ unsigned big_pos(int q) {
  using G = BigGeo<N>;
  return (unsigned)opaque((int)threadIdx.x) + G::T * (q % G::PB) + G::M * (q / G::PB);
}

// Workgroup sum over 1024 threads, every thread gets it (fixed order).
__device__ __noinline__ double big_block_sum(double v) {
  __shared__ double red[kBigT / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kBigT / 64; ++w) s += red[w];
  return s;
}

// Column passes. Forward: y_k2[n1] = W^(k2 n1) sum_n2 x[n1 + M n2]
// W_R^(k2 n2) into out[k2 M + n1] (W = e^(2 pi i / N)). Inverse: the
// adjoint, out[n1 + M n2] = post * sum_k2 W_R^(-k2 n2) W^(-k2 n1)
// in[k2 M + n1]. Both may run in place (a column is read before it is
// written).
template <int N>
__device__ __forceinline__ void big_col_fwd(const Ops& op, const double2* in, double2* out) {
  using G = BigGeo<N>;
  const unsigned t = (unsigned)opaque((int)threadIdx.x);
#pragma unroll 1
  for (int j = 0; j < G::PB; ++j) {
    const unsigned n1 = t + G::T * j;
    double2 v[G::R];
#pragma unroll
    for (int c = 0; c < G::R; ++c) v[c] = in[n1 + G::M * c];
    dft<G::R>(v);
#pragma unroll
    for (int k = 1; k < G::R; ++k) v[k] = cmul(v[k], op.tw[k * n1]);
#pragma unroll
    for (int k = 0; k < G::R; ++k) out[k * G::M + n1] = v[k];
  }
}
template <int N>
__device__ __forceinline__ void big_col_inv(const Ops& op, const double2* in, double2* out, double post) {
  using G = BigGeo<N>;
  const unsigned t = (unsigned)opaque((int)threadIdx.x);
#pragma unroll 1
  for (int j = 0; j < G::PB; ++j) {
    const unsigned n1 = t + G::T * j;
    double2 v[G::R];
#pragma unroll
    for (int k = 0; k < G::R; ++k) v[k] = in[k * G::M + n1];
#pragma unroll
    for (int k = 1; k < G::R; ++k) v[k] = cmul(v[k], conjg(op.tw[k * n1]));
#pragma unroll
    for (int k = 0; k < G::R; ++k) v[k] = conjg(v[k]);
    dft<G::R>(v);
#pragma unroll
    for (int c = 0; c < G::R; ++c) out[n1 + G::M * c] = cscale(conjg(v[c]), post);
  }
}

// Block-pass modes of big_blocks.
enum BigBlockMode {
  BB_SPECTRAL = 0,  // per block of dst: F+, times the multiplier / N, F-; in place
  BB_FWD_NAT = 1,   // per block of src: F+ / N, stored to dst in natural order
  BB_INV_NAT = 2,   // per block: F- of the natural-order spectrum in src, into dst's blocks
  BB_YB = 3         // two-level Yb gain on dst: all blocks F+ / N, sums, pump update, multiply, F-
};

// Arguments of one spectral multiply (ops.hpp Dev::spectral): kind MUL_*,
// the array (E, h_R or the nonlinear factor), g (MUL_GAIN: the exponent;
// MUL_YB: the step h of the linear call) and the pump power (MUL_YB).
struct BigMul {
  int kind;
  const double2* arr;
  double g;
};

// The block pass (the only M-point transform call site of a kernel).
template <int N>
__device__ __noinline__ void big_blocks(const Ops* opp, int mode, BigMul m, double* pp, const double2* src,
                                        double2* dst) {
  using G = BigGeo<N>;
  using F = FFT<G::M, G::T>;
  constexpr double INVN = 1.0 / N;
  __shared__ double lds[F::LDS_DOUBLES];
  const Ops& op = *opp;
  const unsigned t = (unsigned)opaque((int)threadIdx.x);
  typename F::Twiddles tw;
  F::template load_twiddles<G::R>(tw, op.tw);
  double2 x[G::PB];
  if (mode == BB_YB) {
    // amplifier.py:368-402 with g = h: inversion from the spectral power of
    // the array propagated, signal gain exp(g_s h 1e3 / 2), pump update
    double s1 = 0.0, s2 = 0.0;
#pragma unroll 1
    for (int c = 0; c < G::R; ++c) {
      double2* blk = dst + c * G::M;
#pragma unroll
      for (int j = 0; j < G::PB; ++j) x[j] = blk[t + G::T * j];
      F::plus(x, lds, tw);
#pragma unroll
      for (int j = 0; j < G::PB; ++j) {
        const unsigned p = c * G::M + t + G::T * j;
        x[j] = cscale(x[j], INVN);
        const double ps = abs2(x[j]) * op.yb_pfac;
        s1 += op.yb_rs[p] * ps;
        s2 += op.yb_ip[p] * ps;
        blk[t + G::T * j] = x[j];
      }
    }
    s1 = big_block_sum(s1);
    s2 = big_block_sum(s2);
    const double p0 = *pp;
    const double n2 = (((op.yb_rp * p0) / op.yb_psp + s1) / (1.0 + p0 / op.yb_psp + s2)) * op.yb_n;
    const double hg = m.g * 1e3;
#pragma unroll 1
    for (int c = 0; c < G::R; ++c) {
      double2* blk = dst + c * G::M;
#pragma unroll
      for (int j = 0; j < G::PB; ++j) {
        const unsigned p = c * G::M + t + G::T * j;
        const double gs = op.yb_as[p] * n2 - op.yb_es[p];
        x[j] = conjg(cscale(cmul(m.arr[p], blk[t + G::T * j]), exp(gs * hg * 0.5)));
      }
      F::plus(x, lds, tw);
#pragma unroll
      for (int j = 0; j < G::PB; ++j) blk[t + G::T * j] = conjg(x[j]);
    }
    *pp = p0 * exp((op.yb_ap * n2 - op.yb_ep) * hg);
    return;
  }
#pragma unroll 1
  for (int c = 0; c < G::R; ++c) {
    double2* blk = dst + c * G::M;
    if (mode == BB_INV_NAT) {
#pragma unroll
      for (int j = 0; j < G::PB; ++j) x[j] = conjg(src[G::R * (t + G::T * j) + c]);
    } else if (mode == BB_FWD_NAT) {
#pragma unroll
      for (int j = 0; j < G::PB; ++j) x[j] = src[c * G::M + t + G::T * j];
    } else {
#pragma unroll
      for (int j = 0; j < G::PB; ++j) x[j] = blk[t + G::T * j];
    }
    const int nit = mode == BB_SPECTRAL ? 2 : 1;
#pragma unroll 1
    for (int it = 0; it < nit; ++it) {
      F::plus(x, lds, tw);
      if (mode == BB_SPECTRAL && it == 0) {
#pragma unroll
        for (int j = 0; j < G::PB; ++j) {
          const unsigned p = c * G::M + t + G::T * j;
          const double2 y = cscale(x[j], INVN);
          double2 r;
          if (m.kind == MUL_ARRAY) r = cmul(m.arr[p], y);
          else if (m.kind == MUL_OMEGA) r = cscale(y, op.omega[p]);
          else if (m.kind == MUL_IOMEGA) r = cmul(make_double2(0.0, op.omega[p]), y);
          else r = cscale(cmul(m.arr[p], y), exp(m.g * op.gprof[p]));  // MUL_GAIN
          x[j] = conjg(r);
        }
      }
    }
    if (mode == BB_FWD_NAT) {
#pragma unroll
      for (int j = 0; j < G::PB; ++j) dst[G::R * (t + G::T * j) + c] = cscale(x[j], INVN);
    } else {
#pragma unroll
      for (int j = 0; j < G::PB; ++j) blk[t + G::T * j] = conjg(x[j]);
    }
  }
}

// out <- post * ifft(m * fft(in)) in pyofss's convention (ops.hpp
// Dev::spectral); out may be in. MUL_GAIN: m.g is the step hl of the linear
// call and the exponent follows from the energy of `in` (amplifier.py:
// 149-158 with field.py:161-172's Simpson weights).
template <int N>
__device__ __noinline__ void big_spectral(const Ops* opp, const double2* in, double2* out, BigMul m, double post,
                                          double* pp) {
  using G = BigGeo<N>;
  const Ops& op = *opp;
  if (m.kind == MUL_GAIN) {
    double e = 0.0;
#pragma unroll 4
    for (int q = 0; q < G::PT; ++q) {
      const unsigned p = big_pos<N>(q);
      e += op.simpw[p] * abs2(in[p]);
    }
    m.g = sat_gain(op, big_block_sum(e), m.g);
  }
  big_col_fwd<N>(op, in, out);
  big_blocks<N>(opp, m.kind == MUL_YB ? BB_YB : BB_SPECTRAL, m, pp, nullptr, out);
  big_col_inv<N>(op, out, out, post);
}

// pyofss field.fft (sign > 0) / field.ifft of a natural-order field, with
// one workspace slot `w`.
template <int N>
__device__ __noinline__ void big_fft_nat(const Ops* opp, int sign, const double2* in, double2* out, double2* w) {
  const BigMul none{MUL_ARRAY, nullptr, 0.0};
  if (sign > 0) {
    big_col_fwd<N>(*opp, in, w);
    big_blocks<N>(opp, BB_FWD_NAT, none, nullptr, w, out);
  } else {
    big_blocks<N>(opp, BB_INV_NAT, none, nullptr, in, w);
    big_col_inv<N>(*opp, w, out, 1.0);
  }
}

// Operators on workspace slots of one instance (the large-grid Dev).
template <int N, bool YB>
struct BigDev {
  using G = BigGeo<N>;
  const Ops& op;
  double2* ws;
  double pp = 0.0;  // Yb pump power of this instance (identical in every thread)

  __device__ BigDev(const Ops& o, double2* w) : op(o), ws(w) {}
  __device__ __forceinline__ double2* slot(int s) const { return ws + (size_t)s * N; }
  __device__ __forceinline__ unsigned pos(int q) const { return big_pos<N>(q); }

  __device__ __forceinline__ void copy(double2* dst, const double2* src) const {
#pragma unroll 4
    for (int q = 0; q < G::PT; ++q) dst[pos(q)] = src[pos(q)];
  }
  __device__ __forceinline__ void pump_begin(size_t b) {
    if constexpr (YB) pp = op.yb_pp[b];
  }
  __device__ __forceinline__ void pump_end(size_t b) const {
    if constexpr (YB) {
      if (threadIdx.x == 0) op.yb_pp[b] = pp;
    }
  }
  __device__ __forceinline__ void spectral(const double2* in, double2* out, int kind, const double2* arr, double g,
                                           double post) {
    big_spectral<N>(&op, in, out, BigMul{kind, arr, g}, post, &pp);
  }

  // E <- exp(lin * h) (linearity.py:242, 246)
  __device__ __forceinline__ void make_exp(double2* E, double h) const {
#pragma unroll 4
    for (int q = 0; q < G::PT; ++q) {
      const unsigned p = pos(q);
      const double2 f = op.lin[p];
      E[p] = cexp(make_double2(f.x * h, f.y * h));
    }
  }

  // linearity.py:241-272 with the amplifier coupling: out <- L(in)
  __device__ __forceinline__ void linear(const double2* in, double2* out, const double2* E, double hl) {
    if constexpr (YB) spectral(in, out, MUL_YB, E, hl, 1.0);
    else if (op.gain == GAIN_SATURABLE) spectral(in, out, MUL_GAIN, E, hl, 1.0);
    else spectral(in, out, MUL_ARRAY, E, 0.0, 1.0);
  }

  // nonlinearity.py non(A): out <- post * N(x) (out may be x; x is neither
  // B_T nor B_V, which hold intermediates)
  __device__ __forceinline__ void nonlin(const double2* x, double2* out, double post) {
    const int kind = op.nl_kind;
    const double2 fac = op.fac;
    double2* sT = slot(B_T);
    double2* sV = slot(B_V);
    if (kind == NL_KERR) {  // :266-267
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = pos(q);
        const double2 a = x[p];
        const double pw = abs2(a);
        out[p] = cscale(cmul(make_double2(fac.x * pw, fac.y * pw), a), post);
      }
      return;
    }
#pragma unroll 4
    for (int q = 0; q < G::PT; ++q) {
      const unsigned p = pos(q);
      const double2 a = x[p];
      sV[p] = (kind == NL_ALL || kind == NL_RS) ? make_double2(abs2(a), 0.0) : cscale(a, abs2(a));
    }
    if (kind == NL_ALL) {  // :201-207
      spectral(sV, sV, MUL_ARRAY, op.hr, 0.0, 1.0);
      const double a1 = 1.0 - op.f_R;
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = pos(q);
        const double2 a = x[p];
        sV[p] = cadd(cscale(cscale(a, a1), abs2(a)), cmul(cscale(a, op.f_R), sV[p]));
      }
      spectral(sV, out, MUL_ARRAY, op.nlmul, 0.0, post);
      return;
    }
    if (kind == NL_SS) {  // :223-228
      spectral(sV, sT, MUL_OMEGA, nullptr, 0.0, 1.0);
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = pos(q);
        out[p] = cscale(cmul(fac, cadd(sV[p], cscale(sT[p], op.ss))), post);
      }
      return;
    }
    const bool with_ss = (kind == NL_SS_RS);
    if (with_ss) {  // :250-256: term_ss (outside the factor), then |A|^2 for term_rs
      spectral(sV, sT, MUL_OMEGA, nullptr, 0.0, op.ss);
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = pos(q);
        sV[p] = make_double2(abs2(x[p]), 0.0);
      }
    }
    spectral(sV, sV, MUL_IOMEGA, nullptr, 0.0, 1.0);  // :236-241
#pragma unroll 4
    for (int q = 0; q < G::PT; ++q) {
      const unsigned p = pos(q);
      const double2 a = x[p];
      const double2 tr = cscale(sV[p], op.rs);
      const double2 v = cmul(cmul(fac, make_double2(abs2(a) + tr.x, tr.y)), a);
      const double2 ts = with_ss ? sT[p] : make_double2(0.0, 0.0);
      out[p] = cscale(cadd(v, ts), post);
    }
  }

  // nonlinearity.py exp_non(A, h, B): out <- exp(h factor (...)) B (out may
  // be B; A and B are neither B_T nor B_V)
  __device__ __forceinline__ void exp_nonlin(const double2* B, const double2* A, double h, double2* out) {
    const int kind = op.nl_kind;
    const double2 hf = make_double2(h * op.fac.x, h * op.fac.y);
    double2* sT = slot(B_T);
    double2* sV = slot(B_V);
    if (kind == NL_KERR) {  // :269-270
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = pos(q);
        const double pw = abs2(A[p]);
        out[p] = cmul(cexp(make_double2(hf.x * pw, hf.y * pw)), B[p]);
      }
      return;
    }
#pragma unroll 4
    for (int q = 0; q < G::PT; ++q) {
      const unsigned p = pos(q);
      const double2 a = A[p];
      sV[p] = (kind == NL_ALL || kind == NL_RS) ? make_double2(abs2(a), 0.0) : cscale(a, abs2(a));
    }
    if (kind == NL_ALL) {  // :209-221
      spectral(sV, sV, MUL_ARRAY, op.hr, 0.0, 1.0);
      const double a1 = 1.0 - op.f_R;
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = pos(q);
        const double2 a = A[p];
        const double2 ta = cadd(make_double2(a1 * abs2(a), 0.0), cscale(sV[p], op.f_R));
        sT[p] = ta;
        sV[p] = cmul(ta, a);
      }
      spectral(sV, sV, MUL_OMEGA, nullptr, 0.0, 1.0);
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = pos(q);
        const double2 b = B[p];
        const double2 qq = sel(hypot(b.x, b.y) > 1e-15, cdiv(make_double2(op.ss, 0.0), b), make_double2(0.0, 0.0));
        out[p] = cmul(cexp(cmul(hf, cadd(sT[p], cmul(qq, sV[p])))), b);
      }
      return;
    }
    if (kind == NL_SS) {  // :230-234
      spectral(sV, sV, MUL_OMEGA, nullptr, 0.0, 1.0);
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = pos(q);
        const double2 b = B[p];
        const double2 tss = cmul(cdiv(make_double2(op.ss, 0.0), b), sV[p]);
        out[p] = cmul(cexp(cmul(hf, make_double2(abs2(A[p]) + tss.x, tss.y))), b);
      }
      return;
    }
    const bool with_ss = (kind == NL_SS_RS);
    if (with_ss) {  // :258-264: term_ss, then |A|^2 for term_rs
      spectral(sV, sV, MUL_OMEGA, nullptr, 0.0, 1.0);
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = pos(q);
        sT[p] = cmul(cdiv(make_double2(op.ss, 0.0), B[p]), sV[p]);
        sV[p] = make_double2(abs2(A[p]), 0.0);
      }
    }
    spectral(sV, sV, MUL_IOMEGA, nullptr, 0.0, 1.0);  // :243-248
#pragma unroll 4
    for (int q = 0; q < G::PT; ++q) {
      const unsigned p = pos(q);
      const double2 b = B[p];
      const double2 ts = with_ss ? sT[p] : make_double2(0.0, 0.0);
      const double2 tr = cscale(sV[p], op.rs);
      const double pw = abs2(A[p]);
      out[p] = cmul(cexp(cmul(hf, make_double2(pw + ts.x + tr.x, ts.y + tr.y))), b);
    }
  }
};

// solver.py:339-354 rk4ip with the linear calls in the reference's order
// (L(A), L(h N(A)), L(A_I + k2), L(acc); the Yb pump state depends on it).
// in -> out (out may be in). N(in) is read from nA instead of computed when
// nA_ready, and stored to nA when nA_store.
template <int N, bool YB>
__device__ __forceinline__ void big_rk4ip(BigDev<N, YB>& d, const double2* in, double2* out, const double2* E,
                                          double h, double2* nA = nullptr, bool nA_ready = false,
                                          bool nA_store = false) {
  using G = BigGeo<N>;
  double2* sAI = d.slot(B_AI);
  double2* sK0 = d.slot(B_K0);
  double2* sK1 = d.slot(B_K1);
  double2* sU = d.slot(B_U);
  const double hh = 0.5 * h;
  d.linear(in, sAI, E, hh);  // A_I = L(A)
  if (nA_ready || nA_store) {
    if (!nA_ready) d.nonlin(in, nA, 1.0);
#pragma unroll 4
    for (int q = 0; q < G::PT; ++q) {
      const unsigned p = d.pos(q);
      sK0[p] = cscale(nA[p], h);
    }
  } else {
    d.nonlin(in, sK0, h);
  }
  d.linear(sK0, sK0, E, hh);  // k0 = L(h N(A))
#pragma unroll 4
  for (int q = 0; q < G::PT; ++q) {
    const unsigned p = d.pos(q);
    sU[p] = cadd(sAI[p], cscale(sK0[p], 0.5));
  }
  d.nonlin(sU, sK1, h);  // k1
#pragma unroll 4
  for (int q = 0; q < G::PT; ++q) {
    const unsigned p = d.pos(q);
    sU[p] = cadd(sAI[p], cscale(sK1[p], 0.5));
  }
  d.nonlin(sU, sU, h);  // k2
#pragma unroll 4
  for (int q = 0; q < G::PT; ++q) {
    const unsigned p = d.pos(q);
    const double2 ai = sAI[p];
    const double2 k2 = sU[p];
    const double2 s = cadd(sK0[p], cscale(cadd(sK1[p], k2), 2.0));
    sK0[p] = cadd(ai, make_double2(s.x * kSixth, s.y * kSixth));  // A_I + (k0 + 2(k1+k2))/6
    sU[p] = cadd(ai, k2);                                   // A_I + k2
  }
  d.linear(sU, sU, E, hh);
  d.nonlin(sU, sK1, h);  // k3
  d.linear(sK0, out, E, hh);
#pragma unroll 4
  for (int q = 0; q < G::PT; ++q) {
    const unsigned p = d.pos(q);
    const double2 k3 = sK1[p];
    out[p] = cadd(make_double2(k3.x * kSixth, k3.y * kSixth), out[p]);
  }
}

// solver.py:266-275 ss_symmetric: L(h/2), exp_non with the phase of the
// un-propagated input, L(h/2). in -> out (out may be in).
template <int N, bool YB>
__device__ __forceinline__ void big_ss_symmetric(BigDev<N, YB>& d, const double2* in, double2* out,
                                                 const double2* E, double h) {
  double2* sU = d.slot(B_U);
  d.linear(in, sU, E, 0.5 * h);
  d.exp_nonlin(sU, in, h, sU);
  d.linear(sU, out, E, 0.5 * h);
}

// solver.py:330-337 ss_sym_rk4 (opencl_fibre.py:929-970 cl_ss_sym_rk4): L(h/2),
// the classical rk4 on the nonlinear operator, L(h/2). in -> out (out may
// be in).
template <int N, bool YB>
__device__ __forceinline__ void big_ss_sym_rk4(BigDev<N, YB>& d, const double2* in, double2* out, const double2* E,
                                               double h) {
  using G = BigGeo<N>;
  double2* sAL = d.slot(B_AI);
  double2* sK0 = d.slot(B_K0);
  double2* sK1 = d.slot(B_K1);
  double2* sU = d.slot(B_U);
  d.linear(in, sAL, E, 0.5 * h);
  d.nonlin(sAL, sK0, h);  // k0
#pragma unroll 4
  for (int q = 0; q < G::PT; ++q) {
    const unsigned p = d.pos(q);
    sU[p] = cadd(sAL[p], cscale(sK0[p], 0.5));
  }
  d.nonlin(sU, sK1, h);  // k1
#pragma unroll 4
  for (int q = 0; q < G::PT; ++q) {
    const unsigned p = d.pos(q);
    sU[p] = cadd(sAL[p], cscale(sK1[p], 0.5));
  }
  d.nonlin(sU, sU, h);  // k2
#pragma unroll 4
  for (int q = 0; q < G::PT; ++q) {
    const unsigned p = d.pos(q);
    const double2 k2 = sU[p];
    sK1[p] = cadd(sK1[p], k2);  // k1 + k2
    sU[p] = cadd(sAL[p], k2);   // A_L + k2
  }
  d.nonlin(sU, sU, h);  // k3
#pragma unroll 4
  for (int q = 0; q < G::PT; ++q) {
    const unsigned p = d.pos(q);
    const double2 s = cadd(cadd(sK0[p], cscale(sK1[p], 2.0)), sU[p]);
    sU[p] = cadd(sAL[p], make_double2(s.x * kSixth, s.y * kSixth));
  }
  d.linear(sU, out, E, 0.5 * h);
}

template <int METHOD, int N, bool YB>
__device__ __forceinline__ void big_integrator(BigDev<N, YB>& d, const double2* in, double2* out, const double2* E,
                                               double h, double2* nA = nullptr, bool nA_ready = false,
                                               bool nA_store = false) {
  if constexpr (METHOD == M_RK4IP) big_rk4ip(d, in, out, E, h, nA, nA_ready, nA_store);
  else if constexpr (METHOD == M_SS_SYM_RK4) big_ss_sym_rk4(d, in, out, E, h);
  else big_ss_symmetric(d, in, out, E, h);
}

template <int N>
__global__ void __launch_bounds__(kBigT) k_fft_big(Ops ops, int sign, const double2* __restrict__ in,
                                                   double2* __restrict__ out, double2* ws_all) {
  const size_t b = blockIdx.x;
  big_fft_nat<N>(&ops, sign, in + b * N, out + b * N, ws_all + b * N);
}

template <int N>
__global__ void __launch_bounds__(kBigT) k_make_exp_big(Ops ops, double h, double2* __restrict__ out) {
  BigDev<N, false> d(ops, nullptr);
  d.make_exp(out, h);
}

template <int N, bool YB>
__global__ void __launch_bounds__(kBigT) k_linear_big(Ops ops, double h, const double2* __restrict__ in,
                                                      double2* __restrict__ out, double2* ws_all) {
  const size_t b = blockIdx.x;
  BigDev<N, YB> d(ops, ws_all + b * NSLOT_BIG * N);
  d.make_exp(d.slot(B_EH), h);
  d.pump_begin(b);
  d.linear(in + b * N, out + b * N, d.slot(B_EH), h);
  d.pump_end(b);
}

// linearity.py:238-239 default_f: ifft(factor * fft(A))
template <int N>
__global__ void __launch_bounds__(kBigT) k_linfac_big(Ops ops, const double2* __restrict__ in,
                                                      double2* __restrict__ out) {
  const size_t b = blockIdx.x;
  BigDev<N, false> d(ops, nullptr);
  d.spectral(in + b * N, out + b * N, MUL_ARRAY, ops.lin, 0.0, 1.0);
}

template <int N>
__global__ void __launch_bounds__(kBigT) k_nonlin_big(Ops ops, const double2* __restrict__ in,
                                                      double2* __restrict__ out, double2* ws_all) {
  const size_t b = blockIdx.x;
  BigDev<N, false> d(ops, ws_all + b * NSLOT_BIG * N);
  d.nonlin(in + b * N, out + b * N, 1.0);
}

template <int N>
__global__ void __launch_bounds__(kBigT) k_exp_nonlin_big(Ops ops, double h, const double2* __restrict__ A,
                                                          const double2* __restrict__ B, double2* __restrict__ out,
                                                          double2* ws_all) {
  const size_t b = blockIdx.x;
  BigDev<N, false> d(ops, ws_all + b * NSLOT_BIG * N);
  d.exp_nonlin(B + b * N, A + b * N, h, out + b * N);
}

template <int N, int METHOD, bool YB>
__global__ void __launch_bounds__(kBigT) k_step_big(Ops ops, double h, double h_exp, const double2* __restrict__ in,
                                                    double2* __restrict__ out, double2* ws_all) {
  const size_t b = blockIdx.x;
  BigDev<N, YB> d(ops, ws_all + b * NSLOT_BIG * N);
  d.make_exp(d.slot(B_EH), h_exp);
  d.pump_begin(b);
  big_integrator<METHOD>(d, in + b * N, out + b * N, d.slot(B_EH), h);
  d.pump_end(b);
}

// stepper.py:164-222 with the cached exp(factor h/2) in ops.efix.
template <int N, int METHOD, bool YB>
__global__ void __launch_bounds__(kBigT) k_fixed_big(Ops ops, FixedParams prm, const double2* __restrict__ in,
                                                     double2* __restrict__ out, double2* ws_all) {
  const size_t b = blockIdx.x;
  BigDev<N, YB> d(ops, ws_all + b * NSLOT_BIG * N);
  double2* sA = d.slot(B_A);
  d.copy(sA, in + b * N);
  d.pump_begin(b);
  long long n_tr = 0;
  for (long long i = 0; i < prm.nsteps; ++i) {
    big_integrator<METHOD>(d, sA, sA, ops.efix, prm.h);
    if (prm.trace_every > 0 && (i % prm.trace_every) == 0) {  // stepper.py:206-209
      if (n_tr < prm.max_traces) d.copy(prm.tr_A + (b * prm.max_traces + n_tr) * N, sA);
      ++n_tr;
    }
  }
  d.copy(out + b * N, sA);
  d.pump_end(b);
  if (threadIdx.x == 0 && prm.n_traces) prm.n_traces[b] = n_tr;
}

// stepper.py:247-376 on slots: the controller of kernels.hpp adaptive_body
// step for step (half step, fine step, coarse step in the reference's order;
// error, controller, acceptance, local extrapolation, trace schedule,
// replay).
template <int N, int METHOD, bool YB>
__global__ void __launch_bounds__(kBigT) k_adaptive_big(Ops ops, AdaptiveParams prm, const double2* __restrict__ in,
                                                        double2* __restrict__ out, double2* ws_all) {
  using G = BigGeo<N>;
  const size_t b = blockIdx.x;
  BigDev<N, YB> d(ops, ws_all + b * NSLOT_BIG * N);
  double2* sA = d.slot(B_A);
  double2* sH = d.slot(B_H);
  double2* sC = d.slot(B_C);
  double2* sNA = d.slot(B_NA);
  double2* sEQ = d.slot(B_EQ);
  double2* sEH = d.slot(B_EH);
  d.copy(sA, in + b * N);
  d.pump_begin(b);
  bool nA_valid = false;
  double z = 0.0, h = prm.h0, z_ignore = prm.step_storage;
  long long n_tr = 0, n_att = 0, n_acc = 0;
  int status = GNLSE_STATUS_OK;
  bool finished = false;

  auto store_trace = [&](double zt, double ht) {
    if (n_tr < prm.max_traces) {
      d.copy(prm.tr_A + (b * prm.max_traces + n_tr) * N, sA);
      if (threadIdx.x == 0) {
        prm.tr_z[b * prm.max_traces + n_tr] = zt;
        prm.st_z[b * prm.max_traces + n_tr] = zt;
        prm.st_h[b * prm.max_traces + n_tr] = ht;
      }
      ++n_tr;
      return true;
    }
    return false;
  };

  if (prm.traces != 1 && prm.max_traces > 0) store_trace(0.0, 0.0);  // stepper.py:284-285

  for (long long s = 1; s < prm.steps_max; ++s) {
    if ((z + h) > prm.length) h = prm.length - z;  // stepper.py:291-292
    bool accepted = false;
    for (long long ta = 0; ta < prm.total_attempts; ++ta) {
      if (prm.replay_h) {
        if (n_att >= prm.replay_len) {
          status = GNLSE_STATUS_REPLAY_EXHAUSTED;
          break;
        }
        h = prm.replay_h[n_att];
      }
      const double h_half = 0.5 * h;
      d.make_exp(sEH, 0.5 * h);
      d.make_exp(sEQ, 0.5 * h_half);
      // stepper.py:303-308: half step, fine step, coarse step
      big_integrator<METHOD>(d, sA, sH, sEQ, h_half, sNA, nA_valid, !nA_valid);
      nA_valid = true;
      big_integrator<METHOD>(d, sH, sH, sEQ, h_half);
      big_integrator<METHOD>(d, sA, sC, sEH, h, sNA, true, false);
      // stepper.py:230-245 relative_local_error
      double sd = 0.0, sf = 0.0, bad = 0.0;
#pragma unroll 4
      for (int q = 0; q < G::PT; ++q) {
        const unsigned p = d.pos(q);
        const double2 f = sH[p];
        sd += abs2(csub(f, sC[p]));
        sf += abs2(f);
        if (!(isfinite(f.x) && isfinite(f.y))) bad = 1.0;
      }
      bad = big_block_sum(bad);
      sd = big_block_sum(sd);
      sf = big_block_sum(sf);
      double delta;
      if (bad > 0.0) delta = 1.0;
      else if (sf != 0.0) delta = sqrt(sd) / sqrt(sf);
      else delta = sqrt(sd);
      // stepper.py:313-325 controller
      const double h_temp = h;
      if (delta > 0.0) {
        const double ratio = prm.tol / delta;
        const double factor = prm.safety * pow(ratio, prm.inv_eta);
        h = h_temp * fmin(prm.max_factor, fmax(prm.min_factor, factor));
      } else {
        h = h_temp * prm.max_factor;
      }
      const bool acc = prm.replay_h ? (prm.replay_accept[n_att] != 0) : (delta < 2.0 * prm.tol);
      if (threadIdx.x == 0 && n_att < prm.log_cap) {
        prm.log_h[b * prm.log_cap + n_att] = h_temp;
        prm.log_delta[b * prm.log_cap + n_att] = delta;
        prm.log_acc[b * prm.log_cap + n_att] = acc ? 1 : 0;
      }
      ++n_att;
      if (acc) {
        z += h_temp;
        ++n_acc;
        // local extrapolation, stepper.py:336-338
#pragma unroll 4
        for (int q = 0; q < G::PT; ++q) {
          const unsigned p = d.pos(q);
          sA[p] = csub(cscale(sH[p], prm.f_alpha), cscale(sC[p], prm.f_beta));
        }
        nA_valid = false;
        if (z > z_ignore) {  // stepper.py:342-345
          if (prm.max_traces > 0 && !store_trace(z, h_temp)) status = GNLSE_STATUS_TRACE_OVERFLOW;
          z_ignore += prm.step_storage;
        }
        accepted = true;
        break;
      } else if (h < prm.hmin) {  // stepper.py:349-351
        status = GNLSE_STATUS_SMALL_STEP;
        break;
      }
    }
    if (status != GNLSE_STATUS_OK) break;
    if (!accepted) {  // stepper.py:353-354
      status = GNLSE_STATUS_SUITABLE_STEP;
      break;
    }
    if (z >= prm.length) {  // stepper.py:357
      finished = true;
      break;
    }
  }
  if (!finished && status == GNLSE_STATUS_OK) status = GNLSE_STATUS_MAX_STEPS;
  d.pump_end(b);
  d.copy(out + b * N, sA);
  if (threadIdx.x == 0) {
    gnlse_instance_result r;
    r.z = z;
    r.h = h;
    r.n_attempts = n_att;
    r.n_accepted = n_acc;
    r.n_traces = n_tr;
    r.status = status;
    r.reserved = 0;
    prm.res[b] = r;
  }
}

template <int N>
struct BigLauncher {
  static hipError_t fft(hipStream_t s, const Ops& o, int sign, long long batch, const double2* in, double2* out,
                        double2* ws) {
    if (!ws) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_fft_big<N>), dim3(batch), dim3(kBigT), 0, s, o, sign, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t make_exp(hipStream_t s, const Ops& o, double h, double2* out) {
    hipLaunchKernelGGL((k_make_exp_big<N>), dim3(1), dim3(kBigT), 0, s, o, h, out);
    return hipGetLastError();
  }
  static hipError_t linear(hipStream_t s, const Ops& o, double h, long long batch, const double2* in, double2* out,
                           double2* ws) {
    if (o.wdm) return hipErrorNotSupported;
    if (o.gain == GAIN_YB) hipLaunchKernelGGL((k_linear_big<N, true>), dim3(batch), dim3(kBigT), 0, s, o, h, in, out, ws);
    else hipLaunchKernelGGL((k_linear_big<N, false>), dim3(batch), dim3(kBigT), 0, s, o, h, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t linfac(hipStream_t s, const Ops& o, long long batch, const double2* in, double2* out) {
    if (o.wdm) return hipErrorNotSupported;
    hipLaunchKernelGGL((k_linfac_big<N>), dim3(batch), dim3(kBigT), 0, s, o, in, out);
    return hipGetLastError();
  }
  static hipError_t nonlin(hipStream_t s, const Ops& o, long long batch, const double2* in, double2* out,
                           double2* ws) {
    if (o.wdm) return hipErrorNotSupported;
    hipLaunchKernelGGL((k_nonlin_big<N>), dim3(batch), dim3(kBigT), 0, s, o, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t exp_nonlin(hipStream_t s, const Ops& o, double h, long long batch, const double2* A,
                               const double2* B, double2* out, double2* ws) {
    if (o.wdm) return hipErrorNotSupported;
    hipLaunchKernelGGL((k_exp_nonlin_big<N>), dim3(batch), dim3(kBigT), 0, s, o, h, A, B, out, ws);
    return hipGetLastError();
  }
#define GN_BIG_BY(KERNEL, ...)                                                                     \
  do {                                                                                            \
    if (o.wdm) return hipErrorNotSupported;                                                       \
    if (method == M_RK4IP && o.gain == GAIN_YB) hipLaunchKernelGGL((KERNEL<N, M_RK4IP, true>), __VA_ARGS__);  \
    else if (method == M_RK4IP) hipLaunchKernelGGL((KERNEL<N, M_RK4IP, false>), __VA_ARGS__);     \
    else if (method == M_SS_SYM_RK4 && o.gain == GAIN_YB)                                          \
      hipLaunchKernelGGL((KERNEL<N, M_SS_SYM_RK4, true>), __VA_ARGS__);                            \
    else if (method == M_SS_SYM_RK4) hipLaunchKernelGGL((KERNEL<N, M_SS_SYM_RK4, false>), __VA_ARGS__); \
    else if (o.gain == GAIN_YB) hipLaunchKernelGGL((KERNEL<N, M_SS_SYMMETRIC, true>), __VA_ARGS__); \
    else hipLaunchKernelGGL((KERNEL<N, M_SS_SYMMETRIC, false>), __VA_ARGS__);                    \
  } while (0)
  static hipError_t step(hipStream_t s, const Ops& o, int method, double h, double h_exp, long long batch,
                         const double2* in, double2* out, double2* ws) {
    GN_BIG_BY(k_step_big, dim3(batch), dim3(kBigT), 0, s, o, h, h_exp, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t fixed(hipStream_t s, const Ops& o, int method, const FixedParams& p, long long batch,
                          const double2* in, double2* out, double2* ws) {
    GN_BIG_BY(k_fixed_big, dim3(batch), dim3(kBigT), 0, s, o, p, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t adaptive(hipStream_t s, const Ops& o, int method, const AdaptiveParams& p, long long batch,
                             const double2* in, double2* out, double2* ws) {
    GN_BIG_BY(k_adaptive_big, dim3(batch), dim3(kBigT), 0, s, o, p, in, out, ws);
    return hipGetLastError();
  }
#undef GN_BIG_BY
  static hipError_t characts(hipStream_t s, long long count, const double2* A, const double2* S, const double* w,
                             double* out, int* npk, int* pk, int max_pk, double2* scratch) {
    if (!scratch) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_characts_global<N>), dim3((unsigned)count), dim3(kCharT), 0, s, A, S, w, out, npk, pk,
                       max_pk, scratch);
    return hipGetLastError();
  }
  static constexpr LaunchTable table = {N,      kBigT,       NSLOT_BIG, BigGeo<N>::R, &fft,     &make_exp,
                                        &linear, &linfac,    &nonlin,   &exp_nonlin,  &step,    &fixed,
                                        &adaptive, &characts};
};

}  // namespace gn

#define GN_DEFINE_BIG_GRID(NN) \
  namespace gn {               \
  const LaunchTable* table_##NN() { return &BigLauncher<NN>::table; } \
  }
