// Two-channel WDM propagation (pyofss sim_type="wdm").
//
// An instance holds two channels A_0, A_1 on the same grid, each with its
// own linear factor (linearity.py:188-211) and nonlinear coefficient
// (nonlinearity.py:197-200), coupled by cross-phase modulation
// (nonlinearity.py:287-304):
//   N_c(A) = i gamma_c (|A_c|^2 + 2 |A_o|^2) A_c,   o = 1 - c.
// The coupling is pointwise, so a workgroup keeps one channel in registers
// at a time and reads the other channel's field of the same stage from its
// slot. Each channel has its own NSLOT workspace slots; instance b's slots
// start at ws + 2 b NSLOT N, channel c's at + c NSLOT N. Fields and traces are
// laid out [instance][channel][N].
#pragma once
#include "ops.hpp"

namespace gn {

template <int N>
struct WdmSlots {
  double2* base;
  __device__ __forceinline__ double2* at(int c, int slot) const { return base + ((size_t)c * NSLOT + slot) * N; }
};

// nonlinearity.py:287-293 wdm_f, channel c in x
template <int N, int T, int K, bool YB>
__device__ __forceinline__ void wdm_nonlin(Dev<N, T, K, YB>& d, double2 (&x)[N / T], int c,
                                           const double2* __restrict__ other) {
  constexpr int P = N / T;
  const double2 fac = c ? d.op.fac2 : d.op.fac;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const double p = abs2(x[k]) + 2.0 * abs2(other[d.idx(k)]);
    x[k] = cmul(make_double2(fac.x * p, fac.y * p), x[k]);
  }
}

// nonlinearity.py:295-304 wdm_exp_f: B_c <- exp(h i gamma_c (|A_c|^2 +
// 2 |A_o|^2)) B_c, B_c in x, A read from its slots
template <int N, int T, int K, bool YB>
__device__ __forceinline__ void wdm_exp_nonlin(Dev<N, T, K, YB>& d, double2 (&x)[N / T], int c,
                                               const double2* __restrict__ a_c, const double2* __restrict__ a_o,
                                               double h) {
  constexpr int P = N / T;
  const double2 fac = c ? d.op.fac2 : d.op.fac;
  const double2 hf = make_double2(h * fac.x, h * fac.y);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const double p = abs2(a_c[d.idx(k)]) + 2.0 * abs2(a_o[d.idx(k)]);
    x[k] = cmul(cexp(cscale(hf, p)), x[k]);
  }
}

// solver.py:339-354 rk4ip on both channels: from slot `in` to slot `out`
// (out must differ from in and from the stage slots). E_c = exp(factor_c h/2).
template <int N, int T, int K, bool YB>
__device__ __forceinline__ void rk4ip_wdm(Dev<N, T, K, YB>& d, const WdmSlots<N>& W, int in, const double2* E0,
                                          const double2* E1, double h, int out) {
  constexpr int P = N / T;
  const double hh = 0.5 * h;
  double2 x[P];
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // A_I = L(A)
    d.load(x, W.at(c, in));
    d.linear(x, c ? E1 : E0, hh);
    d.store(W.at(c, S_AI), x);
  }
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // k0 = L(h N(A))
    d.load(x, W.at(c, in));
    wdm_nonlin(d, x, c, W.at(1 - c, in));
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = cscale(x[k], h);
    d.linear(x, c ? E1 : E0, hh);
    d.store(W.at(c, S_K0), x);
  }
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // A_I + 0.5 k0
    const double2* ai = W.at(c, S_AI);
    const double2* k0 = W.at(c, S_K0);
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = cadd(ai[d.idx(k)], cscale(k0[d.idx(k)], 0.5));
    d.store(W.at(c, S_T), x);
  }
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // k1 = h N(A_I + 0.5 k0)
    d.load(x, W.at(c, S_T));
    wdm_nonlin(d, x, c, W.at(1 - c, S_T));
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = cscale(x[k], h);
    d.store(W.at(c, S_K1), x);
  }
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // A_I + 0.5 k1
    const double2* ai = W.at(c, S_AI);
    const double2* k1 = W.at(c, S_K1);
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = cadd(ai[d.idx(k)], cscale(k1[d.idx(k)], 0.5));
    d.store(W.at(c, S_U), x);
  }
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // k2 = h N(A_I + 0.5 k1); acc; L(A_I + k2)
    d.load(x, W.at(c, S_U));
    wdm_nonlin(d, x, c, W.at(1 - c, S_U));
    double2* ai = W.at(c, S_AI);
    double2* k0 = W.at(c, S_K0);
    const double2* k1 = W.at(c, S_K1);
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int i = d.idx(k);
      const double2 k2 = cscale(x[k], h);
      const double2 a = ai[i];
      const double2 s = cadd(k0[i], cscale(cadd(k1[i], k2), 2.0));
      k0[i] = cadd(a, make_double2(s.x * kSixth, s.y * kSixth));  // A_I + (k0 + 2(k1+k2))/6
      x[k] = cadd(a, k2);                                    // A_I + k2
    }
    d.linear(x, c ? E1 : E0, hh);
    d.store(W.at(c, S_T), x);
  }
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // k3 = h N(L(A_I + k2))
    d.load(x, W.at(c, S_T));
    wdm_nonlin(d, x, c, W.at(1 - c, S_T));
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = cscale(x[k], h);
    d.store(W.at(c, S_K1), x);
  }
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // k3/6 + L(acc)
    d.load(x, W.at(c, S_K0));
    d.linear(x, c ? E1 : E0, hh);
    const double2* k3 = W.at(c, S_K1);
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const double2 v = k3[d.idx(k)];
      x[k] = cadd(make_double2(v.x * kSixth, v.y * kSixth), x[k]);
    }
    d.store(W.at(c, out), x);
  }
}

// solver.py:266-275 ss_symmetric on both channels (the phase of exp_non from
// the un-propagated fields), from slot `in` to slot `out`.
template <int N, int T, int K, bool YB>
__device__ __forceinline__ void ss_symmetric_wdm(Dev<N, T, K, YB>& d, const WdmSlots<N>& W, int in, const double2* E0,
                                                 const double2* E1, double h, int out) {
  constexpr int P = N / T;
  const double hh = 0.5 * h;
  double2 x[P];
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // A_L = L(A)
    d.load(x, W.at(c, in));
    d.linear(x, c ? E1 : E0, hh);
    d.store(W.at(c, S_T), x);
  }
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {  // L(exp_non(A, h, A_L))
    d.load(x, W.at(c, S_T));
    wdm_exp_nonlin(d, x, c, W.at(c, in), W.at(1 - c, in), h);
    d.linear(x, c ? E1 : E0, hh);
    d.store(W.at(c, out), x);
  }
}

template <int METHOD, int N, int T, int K, bool YB>
__device__ __forceinline__ void integrator_wdm(Dev<N, T, K, YB>& d, const WdmSlots<N>& W, int in, const double2* E0,
                                               const double2* E1, double h, int out) {
  if constexpr (METHOD == M_RK4IP) rk4ip_wdm(d, W, in, E0, E1, h, out);
  else ss_symmetric_wdm(d, W, in, E0, E1, h, out);
}

template <int N, int T>
__device__ __forceinline__ void wdm_copy(const Dev<N, T, NL_KERR, false>& d, double2* __restrict__ dst,
                                         const double2* __restrict__ src) {
  double2 x[N / T];
  d.load(x, src);
  d.store(dst, x);
}

// linearity.py:306-320 wdm_exp_f(_cached): channel c of instance b is row
// 2b + c; ws holds NSLOT slots per row
template <int N, int T>
__global__ void __launch_bounds__(T)
    k_linear_wdm(Ops ops, double h, const double2* __restrict__ in, double2* __restrict__ out, double2* ws_all) {
  GN_SHARED(N, T, NL_KERR);
  constexpr int P = N / T;
  const size_t r = blockIdx.x;
  const int c = (int)(r & 1);
  double2* E = ws_all + (r * NSLOT + S_EH) * N;
  d.make_exp_of(E, c ? ops.lin2 : ops.lin, h);
  double2 x[P];
  d.load(x, in + r * N);
  d.linear(x, E, h);
  d.store(out + r * N, x);
}

// linearity.py:303-304 wdm_f: ifft(factor_c * fft(A_c))
template <int N, int T>
__global__ void __launch_bounds__(T) k_linfac_wdm(Ops ops, const double2* __restrict__ in, double2* __restrict__ out) {
  GN_SHARED(N, T, NL_KERR);
  constexpr int P = N / T;
  const size_t r = blockIdx.x;
  double2 x[P];
  d.load(x, in + r * N);
  d.spectral(x, MUL_ARRAY, (r & 1) ? ops.lin2 : ops.lin, 0.0);
  d.store(out + r * N, x);
}

// nonlinearity.py:287-304 on whole instances: non (B == nullptr) or exp_non
template <int N, int T>
__global__ void __launch_bounds__(T) k_nonlin_wdm(Ops ops, double h, const double2* __restrict__ A,
                                                  const double2* __restrict__ B, double2* __restrict__ out) {
  GN_SHARED(N, T, NL_KERR);
  constexpr int P = N / T;
  const size_t b = blockIdx.x;
  const double2* a = A + 2 * b * N;
  double2 x[P];
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {
    if (B) {
      d.load(x, B + (2 * b + c) * N);
      wdm_exp_nonlin(d, x, c, a + c * N, a + (1 - c) * N, h);
    } else {
      d.load(x, a + c * N);
      wdm_nonlin(d, x, c, a + (1 - c) * N);
    }
    d.store(out + (2 * b + c) * N, x);
  }
}

// one integrator call per instance (Solver.rk4ip / ss_symmetric with WDM arrays)
template <int N, int T, int METHOD>
__global__ void __launch_bounds__(T) k_step_wdm(Ops ops, double h, double h_exp, const double2* __restrict__ in,
                                                double2* __restrict__ out, double2* ws_all) {
  GN_SHARED(N, T, NL_KERR);
  const size_t b = blockIdx.x;
  const WdmSlots<N> W{ws_all + 2 * b * NSLOT * N};
#pragma unroll 1
  for (int c = 0; c < 2; ++c) {
    d.make_exp_of(W.at(c, S_EH), c ? ops.lin2 : ops.lin, h_exp);
    wdm_copy<N, T>(d, W.at(c, S_A), in + (2 * b + c) * N);
  }
  integrator_wdm<METHOD>(d, W, S_A, W.at(0, S_EH), W.at(1, S_EH), h, S_NA);
#pragma unroll 1
  for (int c = 0; c < 2; ++c) wdm_copy<N, T>(d, out + (2 * b + c) * N, W.at(c, S_NA));
}

// stepper.py:164-222 with WDM fields: state ping-pongs between two slots
template <int N, int T, int METHOD>
__global__ void __launch_bounds__(T) k_fixed_wdm(Ops ops, FixedParams prm, const double2* __restrict__ in,
                                                 double2* __restrict__ out, double2* ws_all) {
  GN_SHARED(N, T, NL_KERR);
  const size_t b = blockIdx.x;
  const WdmSlots<N> W{ws_all + 2 * b * NSLOT * N};
#pragma unroll 1
  for (int c = 0; c < 2; ++c) wdm_copy<N, T>(d, W.at(c, S_A), in + (2 * b + c) * N);
  int cur = S_A, nxt = S_NA;
  long long n_tr = 0;
  for (long long i = 0; i < prm.nsteps; ++i) {
    integrator_wdm<METHOD>(d, W, cur, ops.efix, ops.efix2, prm.h, nxt);
    const int t = cur;
    cur = nxt;
    nxt = t;
    if (prm.trace_every > 0 && (i % prm.trace_every) == 0) {  // stepper.py:206-209
      if (n_tr < prm.max_traces)
#pragma unroll 1
        for (int c = 0; c < 2; ++c) wdm_copy<N, T>(d, prm.tr_A + ((b * prm.max_traces + n_tr) * 2 + c) * N, W.at(c, cur));
      ++n_tr;
    }
  }
#pragma unroll 1
  for (int c = 0; c < 2; ++c) wdm_copy<N, T>(d, out + (2 * b + c) * N, W.at(c, cur));
  if (threadIdx.x == 0 && prm.n_traces) prm.n_traces[b] = n_tr;
}

// stepper.py:247-376 with WDM fields: half step (A -> H), fine step
// (H -> C), coarse step (A -> H); the relative error is taken over both
// channels (numpy's norm of the 2 x N arrays).
template <int N, int T, int METHOD>
__global__ void __launch_bounds__(T) k_adaptive_wdm(Ops ops, AdaptiveParams prm, const double2* __restrict__ in,
                                                    double2* __restrict__ out, double2* ws_all) {
  GN_SHARED(N, T, NL_KERR);
  constexpr int P = N / T;
  constexpr int S_C = S_NA;  // the fine result
  const size_t b = blockIdx.x;
  const WdmSlots<N> W{ws_all + 2 * b * NSLOT * N};
#pragma unroll 1
  for (int c = 0; c < 2; ++c) wdm_copy<N, T>(d, W.at(c, S_A), in + (2 * b + c) * N);
  double z = 0.0, h = prm.h0, z_ignore = prm.step_storage;
  long long n_tr = 0, n_att = 0, n_acc = 0;
  int status = GNLSE_STATUS_OK;
  bool finished = false;
  auto store_trace = [&](double zt, double ht) {
    if (n_tr < prm.max_traces) {
#pragma unroll 1
      for (int c = 0; c < 2; ++c) wdm_copy<N, T>(d, prm.tr_A + ((b * prm.max_traces + n_tr) * 2 + c) * N, W.at(c, S_A));
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
  double2 x[P];
  for (long long s = 1; s < prm.steps_max; ++s) {
    if ((z + h) > prm.length) h = prm.length - z;  // stepper.py:291-292
    bool accepted = false;
    for (long long ta = 0; ta < prm.total_attempts; ++ta) {
      const double h_half = 0.5 * h;
#pragma unroll 1
      for (int c = 0; c < 2; ++c) {
        const double2* fac = c ? ops.lin2 : ops.lin;
        d.make_exp_of(W.at(c, S_EH), fac, 0.5 * h);
        d.make_exp_of(W.at(c, S_EQ), fac, 0.5 * h_half);
      }
      // stepper.py:303-308: half, fine, coarse
      integrator_wdm<METHOD>(d, W, S_A, W.at(0, S_EQ), W.at(1, S_EQ), h_half, S_H);
      integrator_wdm<METHOD>(d, W, S_H, W.at(0, S_EQ), W.at(1, S_EQ), h_half, S_C);
      integrator_wdm<METHOD>(d, W, S_A, W.at(0, S_EH), W.at(1, S_EH), h, S_H);
      // stepper.py:230-245 relative_local_error over both channels
      double sd = 0.0, sf = 0.0, bad = 0.0;
#pragma unroll 1
      for (int c = 0; c < 2; ++c) {
        const double2* fn = W.at(c, S_C);
        const double2* co = W.at(c, S_H);
#pragma unroll
        for (int k = 0; k < P; ++k) {
          const double2 f = fn[d.idx(k)];
          const double2 df = csub(f, co[d.idx(k)]);
          sd += abs2(df);
          sf += abs2(f);
          if (!(isfinite(f.x) && isfinite(f.y))) bad = 1.0;
        }
      }
      bad = d.block_sum(bad);
      sd = d.block_sum(sd);
      sf = d.block_sum(sf);
      double delta;
      if (bad > 0.0) delta = 1.0;
      else if (sf != 0.0) delta = sqrt(sd) / sqrt(sf);
      else delta = sqrt(sd);
      const double h_temp = h;  // stepper.py:313-325
      if (delta > 0.0) {
        const double ratio = prm.tol / delta;
        const double factor = prm.safety * pow(ratio, prm.inv_eta);
        h = h_temp * fmin(prm.max_factor, fmax(prm.min_factor, factor));
      } else {
        h = h_temp * prm.max_factor;
      }
      const bool acc = delta < 2.0 * prm.tol;
      if (threadIdx.x == 0 && n_att < prm.log_cap) {
        prm.log_h[b * prm.log_cap + n_att] = h_temp;
        prm.log_delta[b * prm.log_cap + n_att] = delta;
        prm.log_acc[b * prm.log_cap + n_att] = acc ? 1 : 0;
      }
      ++n_att;
      if (acc) {
        z += h_temp;
        ++n_acc;
#pragma unroll 1
        for (int c = 0; c < 2; ++c) {  // stepper.py:336-338
          const double2* fn = W.at(c, S_C);
          const double2* co = W.at(c, S_H);
#pragma unroll
          for (int k = 0; k < P; ++k)
            x[k] = csub(cscale(fn[d.idx(k)], prm.f_alpha), cscale(co[d.idx(k)], prm.f_beta));
          d.store(W.at(c, S_A), x);
        }
        if (z > z_ignore) {  // stepper.py:342-345
          if (prm.max_traces > 0 && !store_trace(z, h_temp)) status = GNLSE_STATUS_TRACE_OVERFLOW;
          z_ignore += prm.step_storage;
        }
        accepted = true;
        break;
      } else if (h < prm.hmin) {
        status = GNLSE_STATUS_SMALL_STEP;
        break;
      }
    }
    if (status != GNLSE_STATUS_OK) break;
    if (!accepted) {
      status = GNLSE_STATUS_SUITABLE_STEP;
      break;
    }
    if (z >= prm.length) {
      finished = true;
      break;
    }
  }
  if (!finished && status == GNLSE_STATUS_OK) status = GNLSE_STATUS_MAX_STEPS;
#pragma unroll 1
  for (int c = 0; c < 2; ++c) wdm_copy<N, T>(d, out + (2 * b + c) * N, W.at(c, S_A));
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

}  // namespace gn
