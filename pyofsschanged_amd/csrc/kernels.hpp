// Kernels of the GNLSE engine and their per-grid launch tables.
//
// Execution model: one workgroup owns one propagation ("instance"). The
// propagation kernels are persistent: the whole fixed-step or adaptive run of
// an instance, including the step-doubling controller, its reductions, the
// accept/reject decision and trace capture, happens inside one launch. A
// batch of instances is one launch with one workgroup per instance.
//
// Each grid size is compiled in its own translation unit (grid_<N>.hip),
// which instantiates Launcher<N, T, C> and exports its table.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/gnlse.h"
#include "characts.hpp"
#include "ops.hpp"

namespace gn {

struct FixedParams {
  double h;
  long long nsteps, trace_every, max_traces;
  double2* tr_A;
  long long* n_traces;
};

struct AdaptiveParams {
  double length, h0, tol, safety, max_factor, min_factor, hmin, inv_eta, f_alpha, f_beta, step_storage;
  long long traces, total_attempts, steps_max, max_traces, log_cap, replay_len;
  const double* replay_h;
  const unsigned char* replay_accept;
  double2* tr_A;
  double* tr_z;
  double* st_z;
  double* st_h;
  double* log_h;
  double* log_delta;
  unsigned char* log_acc;
  gnlse_instance_result* res;
};

struct LaunchTable {
  int n, t;
  int nslot;  // workspace slots (N complex) per field of an instance
  // spectral layout of the operator arrays and of every spectrum inside the
  // kernels: 1 = FFT order; R > 1 = R blocks, position k2 (N/R) + k1 holding
  // frequency index R k1 + k2 (the large grids of big.hpp)
  int blocks;
  // gnlse_fft: natural (FFT) order in and out; ws holds one field per
  // instance (used by the large grids only)
  hipError_t (*fft)(hipStream_t, const Ops&, int sign, long long batch, const double2* in, double2* out,
                    double2* ws);
  hipError_t (*make_exp)(hipStream_t, const Ops&, double h, double2* out);
  hipError_t (*linear)(hipStream_t, const Ops&, double h, long long batch, const double2* in, double2* out,
                       double2* ws);
  hipError_t (*linfac)(hipStream_t, const Ops&, long long batch, const double2* in, double2* out);
  hipError_t (*nonlin)(hipStream_t, const Ops&, long long batch, const double2* in, double2* out, double2* ws);
  hipError_t (*exp_nonlin)(hipStream_t, const Ops&, double h, long long batch, const double2* A, const double2* B,
                           double2* out, double2* ws);
  hipError_t (*step)(hipStream_t, const Ops&, int method, double h, double h_exp, long long batch, const double2* in,
                     double2* out, double2* ws);
  hipError_t (*fixed)(hipStream_t, const Ops&, int method, const FixedParams&, long long batch, const double2* in,
                      double2* out, double2* ws);
  hipError_t (*adaptive)(hipStream_t, const Ops&, int method, const AdaptiveParams&, long long batch,
                         const double2* in, double2* out, double2* ws);
  // scratch: count x (N doubles + N bytes), the large grids' powers and
  // peak flags (the other grids hold them in LDS)
  hipError_t (*characts)(hipStream_t, long long count, const double2* A, const double2* S, const double* w,
                         double* out, int* npk, int* pk, int max_pk, double2* scratch);
  // schedule (Dev's CFG) of the spectral-domain RK4IP kernels in this table
  int cfg;
  // adaptive propagation on two workgroups per instance (k_adaptive_pair;
  // workspace 2 kPairSlots N per instance, two zeroed flags per instance):
  // hipErrorNotSupported for operators without a two-workgroup kernel,
  // hipErrorCooperativeLaunchTooLarge when the batch does not fit; null on
  // grids without one
  hipError_t (*adaptive_pair)(hipStream_t, const Ops&, int method, const AdaptiveParams&, long long batch,
                              const double2* in, double2* out, double2* ws, unsigned* flags);
};

const LaunchTable* find_table(long long n);

// A compiled schedule of the spectral-domain adaptive RK4IP kernels of a grid
// (full Raman + self-steepening operator, without gain or with the saturable
// gain): threads per instance and Dev's CFG (transform plans, stage-pass
// block). A grid's default schedule is its LaunchTable's; the autotuner
// (gnlse_tune_adaptive) times the alternatives on the caller's inputs.
struct Schedule {
  int t, cfg;
  hipError_t (*adaptive)(hipStream_t, const Ops&, const AdaptiveParams&, long long batch, const double2* in,
                         double2* out, double2* ws);
};
// The launches a schedule covers: rk4ip with the full nonlinear operator and
// no gain (SP 1) or the saturable gain with the flat interior Simpson weight
// (SP 2), single-channel fields.
inline bool schedule_applies(const Ops& o, int method) {
  return method == M_RK4IP && !o.wdm && o.nl_kind == NL_ALL &&
         (o.gain == GAIN_NONE || (o.gain == GAIN_SATURABLE && o.simp_flat));
}

#define COMMA ,
#define GN_SHARED_G(N, T, K, YB)                 \
  __shared__ double lds[FFT<N, T>::LDS_DOUBLES]; \
  __shared__ double red[(T) / 64 + 2];           \
  Dev<N, T, K, YB> d(lds, red, ops)
#define GN_SHARED(N, T, K) GN_SHARED_G(N, T, K, false)

template <int N, int T>
__global__ void __launch_bounds__(T) k_fft(Ops ops, int sign, const double2* __restrict__ in, double2* __restrict__ out) {
  GN_SHARED(N, T, NL_KERR);
  constexpr int P = N / T;
  const size_t b = blockIdx.x;
  double2 x[P];
  d.load(x, in + b * N);
  if (sign > 0) d.fft(x);
  else d.ifft(x);
  d.store(out + b * N, x);
}

template <int N, int T>
__global__ void __launch_bounds__(T) k_make_exp(Ops ops, double h, double2* __restrict__ out) {
  GN_SHARED(N, T, NL_KERR);
  d.make_exp(out, h);
}

template <int N, int T, bool YB>
__global__ void __launch_bounds__(T)
    k_linear(Ops ops, double h, const double2* __restrict__ in, double2* __restrict__ out, double2* ws_all) {
  GN_SHARED_G(N, T, NL_KERR, YB);
  constexpr int P = N / T;
  const size_t b = blockIdx.x;
  double2* ws = ws_all + b * NSLOT * N;
  d.make_exp(ws + S_EH * N, h);
  double2 x[P];
  d.load(x, in + b * N);
  d.pump_begin(b);
  d.linear(x, ws + S_EH * N, h);
  d.pump_end(b);
  d.store(out + b * N, x);
}

// linearity.py:238-239 default_f: ifft(factor * fft(A))
template <int N, int T>
__global__ void __launch_bounds__(T) k_linfac(Ops ops, const double2* __restrict__ in, double2* __restrict__ out) {
  GN_SHARED(N, T, NL_KERR);
  constexpr int P = N / T;
  const size_t b = blockIdx.x;
  double2 x[P];
  d.load(x, in + b * N);
  d.spectral(x, MUL_ARRAY, ops.lin, 0.0);
  d.store(out + b * N, x);
}

template <int N, int T, int K>
__global__ void __launch_bounds__(T)
    k_nonlin(Ops ops, const double2* __restrict__ in, double2* __restrict__ out, double2* ws_all) {
  GN_SHARED(N, T, K);
  constexpr int P = N / T;
  const size_t b = blockIdx.x;
  double2* ws = ws_all + b * NSLOT * N;
  double2 x[P];
  d.load(x, in + b * N);
  d.nonlin(x, ws + S_U * N, ws + S_T * N);
  d.store(out + b * N, x);
}

template <int N, int T, int K>
__global__ void __launch_bounds__(T) k_exp_nonlin(Ops ops, double h, const double2* __restrict__ A,
                                                  const double2* __restrict__ B, double2* __restrict__ out,
                                                  double2* ws_all) {
  GN_SHARED(N, T, K);
  constexpr int P = N / T;
  const size_t b = blockIdx.x;
  double2* ws = ws_all + b * NSLOT * N;
  double2 x[P];
  d.load(x, A + b * N);
  d.store(ws + S_A * N, x);
  d.load(x, B + b * N);
  d.exp_nonlin(x, ws + S_A * N, h, ws + S_U * N, ws + S_T * N);
  d.store(out + b * N, x);
}

template <int N, int T, int METHOD, int K, bool YB>
__global__ void __launch_bounds__(T)
    k_step(Ops ops, double h, double h_exp, const double2* __restrict__ in, double2* __restrict__ out, double2* ws_all) {
  GN_SHARED_G(N, T, K, YB);
  constexpr int P = N / T;
  const size_t b = blockIdx.x;
  double2* ws = ws_all + b * NSLOT * N;
  double2 x[P];
  d.make_exp(ws + S_EH * N, h_exp);
  d.load(x, in + b * N);
  d.store(ws + S_A * N, x);
  d.pump_begin(b);
  integrator<METHOD>(d, x, ws + S_A * N, ws + S_EH * N, h, ws);
  d.pump_end(b);
  d.store(out + b * N, x);
}

// stepper.py:164-222 with the cached exp(factor*h/2) in ops.efix. SP as for
// k_adaptive: rk4ip_sp with gain SP - 1, the field's spectrum carried from
// step to step in S_T.
template <int N, int T, int METHOD, int K, bool YB, int SP = 0, int CFG = 0>
__global__ void __launch_bounds__(T) k_fixed(Ops ops, FixedParams prm, const double2* __restrict__ in,
                                             double2* __restrict__ out, double2* ws_all) {
  static_assert(!SP || (METHOD == M_RK4IP && K == NL_ALL && YB == (SP == 1 + SPG_YB)),
                "spectral-domain stepping: rk4ip, NL_ALL");
  __shared__ double lds[SP ? sp_lds_doubles<N, T, CFG>() : Dev<N, T, K, YB, CFG>::F::LDS_DOUBLES];
  __shared__ double red[SP ? 18 * (T / 64) + 2 : T / 64 + 2];
  Dev<N, T, K, YB, CFG> d(lds, red, ops);
  constexpr int P = N / T;
  const size_t b = blockIdx.x;
  double2* ws = ws_all + b * NSLOT * N;
  double2* sA = ws + S_A * N;
  double2* sAs = ws + S_T * N;
  double2 x[P];
  d.load(x, in + b * N);
  d.store(sA, x);
  if constexpr (SP != 0) {
    d.fft(x);
    d.store(sAs, x);
  }
  d.pump_begin(b);
  long long n_tr = 0;
  for (long long i = 0; i < prm.nsteps; ++i) {
    if constexpr (SP != 0) rk4ip_sp<SP - 1>(d, x, sA, sAs, sAs, ops.efix, false, prm.h, ws, nullptr, false, false);
    else integrator<METHOD>(d, x, sA, ops.efix, prm.h, ws);
    d.store(sA, x);
    if (prm.trace_every > 0 && (i % prm.trace_every) == 0) {  // stepper.py:206-209
      if (n_tr < prm.max_traces) d.store(prm.tr_A + (b * prm.max_traces + n_tr) * N, x);
      ++n_tr;
    }
  }
  d.store(out + b * N, x);
  d.pump_end(b);
  if (threadIdx.x == 0 && prm.n_traces) prm.n_traces[b] = n_tr;
}

// stepper.py:247-376: step doubling, local extrapolation and the controller,
// all per instance on the device.
// One adaptive propagation (stepper.py:247-376) by the calling workgroup:
// instance b of a batch whose inputs, outputs and workspaces start at in,
// out and ws_all. With spec != nullptr the pyofss spectrum fft(A_out) is
// also written to spec + b*N (input of the ensemble coherence).
// SP selects rk4ip_sp (RK4IP, full nonlinear operator) with gain SP - 1
// (SpGain: 1 none, 2 saturable, 3 two-level Yb): the accepted field is then
// also kept as its spectrum in S_T and the half step's spectrum, left in
// S_HS, starts the fine step. Without a gain (SP = 1) the workspace an
// attempt touches is cut to 7 slots (so that a batch of 256 at 8192 samples,
// 224 MiB, fits the 256 MiB Infinity Cache): the accepted field lives in S_U
// (read only by the first call after an acceptance, before that call parks
// its first stage field there) and the fine step forms the half step's
// spectrum from its time samples instead of reading it from S_HS.
// Two workgroups per instance (PAIR): the half and fine steps (role 0) and
// the coarse step (role 1) of an attempt are independent up to the error
// norm, so they run side by side and an attempt waits for two integrator
// calls instead of three. Each role has its own workspace of kPairSlots
// slots (the NSLOT of a single workgroup and two result slots) and computes
// N(A) for its own step; the fine and coarse results go to result slot
// n_att mod 2 of their role (so an attempt never overwrites what the
// partner may still read), the roles meet
// at pair_sync, and both then form the same error norm, controller decision
// and accepted field from both results (the same operations in the same
// order: identical values in both roles and equal, bit for bit, to the
// single-workgroup kernel). Role 0 alone writes traces, the attempt log, the
// output and the result. The launch must keep both workgroups of a pair
// resident (a cooperative launch); the wait is bounded
// (GNLSE_STATUS_PAIR_TIMEOUT).
constexpr long long kPairSpin = 1ll << 26;  // s_sleep(10) spins, about 20 s
constexpr int kPairSlots = NSLOT + 2;
__device__ __forceinline__ bool pair_sync(unsigned* flags, int role, unsigned epoch) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();  // this role's results visible to the partner
    __hip_atomic_store(&flags[role], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    long long it = 0;
    while (__hip_atomic_load(&flags[1 - role], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch && it < kPairSpin) {
      __builtin_amdgcn_s_sleep(10);
      ++it;
    }
    ok = it < kPairSpin;
    __threadfence();  // the partner's results visible here
  }
  __syncthreads();
  return ok != 0;
}

template <int N, int T, int METHOD, int K, bool YB = false, int SP = 0, int CFG = 0, bool PAIR = false>
__device__ __forceinline__ void adaptive_body(double* lds, double* red, const Ops& ops, const AdaptiveParams& prm,
                                              size_t b, const double2* __restrict__ in, double2* __restrict__ out,
                                              double2* ws_all, double2* __restrict__ spec, int role = 0,
                                              unsigned* pflags = nullptr) {
  static_assert(!SP || (METHOD == M_RK4IP && K == NL_ALL && YB == (SP == 1 + SPG_YB)),
                "spectral-domain stepping: rk4ip, NL_ALL");
  static_assert(!PAIR || !YB, "two-workgroup instances: not with the Yb gain (its pump follows the call order)");
  Dev<N, T, K, YB, CFG> d(lds, red, ops);
  constexpr int P = N / T;
  double2* ws = ws_all + (PAIR ? (2 * b + role) * kPairSlots : b * NSLOT) * N;
  double2* ws_f = ws_all + (2 * b) * kPairSlots * N;      // PAIR: role 0 (fine result)
  double2* ws_c = ws_all + (2 * b + 1) * kPairSlots * N;  // PAIR: role 1 (coarse result)
  const bool lead = !PAIR || role == 0;
  double2* sA = ws + S_A * N;
  double2* sH = ws + S_H * N;
  double2* sEQ = ws + S_EQ * N;
  double2* sEH = ws + S_EH * N;
  double2* sAs = ws + S_T * N;   // SP: fft(A)
  double2* sHs = SP == 1 ? nullptr : ws + S_HS * N;  // SP > 1: fft of the half step's result
  double2* sAt = SP == 1 ? ws + S_U * N : sA;         // the accepted field (time domain)
  double2 x[P];
  d.load(x, in + b * N);
  d.store(sAt, x);
  d.pump_begin(b);

  double2* sNA = ws + S_NA * N;
  bool nA_valid = false;  // sNA holds N(sA)
  double z = 0.0, h = prm.h0, z_ignore = prm.step_storage;
  long long n_tr = 0, n_att = 0, n_acc = 0;
  int status = GNLSE_STATUS_OK;
  bool finished = false;

  auto store_trace = [&](double zt, double ht) {  // both roles count, role 0 stores
    if (n_tr < prm.max_traces) {
      if (lead) {
        d.store(prm.tr_A + (b * prm.max_traces + n_tr) * N, x);
        if (threadIdx.x == 0) {
          prm.tr_z[b * prm.max_traces + n_tr] = zt;
          prm.st_z[b * prm.max_traces + n_tr] = zt;
          prm.st_h[b * prm.max_traces + n_tr] = ht;
        }
      }
      ++n_tr;
      return true;
    }
    return false;
  };

  if (prm.traces != 1 && prm.max_traces > 0) store_trace(0.0, 0.0);  // stepper.py:284-285
  if constexpr (SP != 0) {
    d.fft(x);
    d.store(sAs, x);
  }

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
      if constexpr (SP != 0) {
        d.make_exp(sEQ, 0.5 * h_half);  // the coarse step squares it where it is read
      } else {
        d.make_exp(sEH, 0.5 * h);
        d.make_exp(sEQ, 0.5 * h_half);
      }
      // stepper.py:303-308 in its order (the Yb pump state depends on it):
      // half step, fine step (both h/2), then the coarse step h. The fine
      // result replaces the half-step field in sH; the coarse one stays in x.
      // PAIR: the result slot of this attempt in each role's workspace
      const int rs = NSLOT + (int)(n_att & 1);
#pragma unroll 1
      for (int c = 0; c < 3; ++c) {
        if (PAIR && ((c == 2) != (role == 1))) continue;
        const double2* src = (c == 1) ? sH : sAt;
        const double2* E = (c == 2) ? sEH : sEQ;
        // N(A) is shared by the half and the coarse step, and by every
        // attempt from the same accepted field
        const bool from_a = (c != 1);
        if constexpr (SP != 0)
          rk4ip_sp<SP - 1>(d, x, src, c == 1 ? sHs : sAs, c == 0 ? sHs : nullptr, sEQ, c == 2, c == 2 ? h : h_half, ws,
                           sNA, from_a && nA_valid, from_a && !nA_valid);
        else
          integrator<METHOD>(d, x, src, E, c == 2 ? h : h_half, ws, sNA, from_a && nA_valid, from_a && !nA_valid);
        if (from_a) nA_valid = true;
        if (PAIR && c >= 1) d.store(ws + rs * N, x);  // fine (role 0) or coarse (role 1) result
        else if (c < 2) d.store(sH, x);
      }
      const double2* sF = sH;  // the fine result; x holds the coarse one
      if constexpr (PAIR) {
        if (!pair_sync(pflags, role, (unsigned)(n_att + 1))) {
          status = GNLSE_STATUS_PAIR_TIMEOUT;
          break;
        }
        sF = ws_f + rs * N;
        d.load(x, ws_c + rs * N);
      }
      // stepper.py:230-245 relative_local_error
      double sd = 0.0, sf = 0.0, bad = 0.0;
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 f = sF[d.idx(k)];
        const double2 df = csub(f, x[k]);
        sd += abs2(df);
        sf += abs2(f);
        if (!(isfinite(f.x) && isfinite(f.y))) bad = 1.0;
        // Gain and Kerr kernels: the local extrapolation of stepper.py:336-338
        // formed here (a rejected attempt drops x, the next one starts from
        // sA); it saves a pass over sH there and costs the gain-free Raman
        // kernels registers
        if constexpr (SP >= 2 || K == NL_KERR) x[k] = csub(cscale(f, prm.f_alpha), cscale(x[k], prm.f_beta));
      }
      bad = d.block_sum(bad);
      sd = d.block_sum(sd);
      sf = d.block_sum(sf);
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
      if (lead && threadIdx.x == 0 && n_att < prm.log_cap) {
        prm.log_h[b * prm.log_cap + n_att] = h_temp;
        prm.log_delta[b * prm.log_cap + n_att] = delta;
        prm.log_acc[b * prm.log_cap + n_att] = acc ? 1 : 0;
      }
      ++n_att;
      if (acc) {
        z += h_temp;
        ++n_acc;
        // local extrapolation, stepper.py:336-338
        if constexpr (!(SP >= 2 || K == NL_KERR)) {
#pragma unroll
          for (int k = 0; k < P; ++k)
            x[k] = csub(cscale(sF[d.idx(k)], prm.f_alpha), cscale(x[k], prm.f_beta));
        }
        d.store(sAt, x);
        nA_valid = false;
        if (z > z_ignore) {  // stepper.py:342-345
          if (prm.max_traces > 0 && !store_trace(z, h_temp)) status = GNLSE_STATUS_TRACE_OVERFLOW;
          z_ignore += prm.step_storage;
        }
        if constexpr (SP != 0) {
          d.fft(x);
          d.store(sAs, x);
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
  if (!lead && status != GNLSE_STATUS_PAIR_TIMEOUT) return;
  if (SP == 1 && !finished) {  // S_U may hold a stage field: the accepted field from its spectrum
    d.load(x, sAs);
    d.ifft(x);
  } else {
    d.load(x, sAt);
  }
  if (lead) d.store(out + b * N, x);
  if (spec) {
    d.fft(x);
    d.store(spec + b * N, x);
  }
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

template <int N, int T, int METHOD, int K, bool YB, int SP = 0, int CFG = 0>
__global__ void __launch_bounds__(T) k_adaptive(Ops ops, AdaptiveParams prm, const double2* __restrict__ in,
                                                double2* __restrict__ out, double2* ws_all) {
  __shared__ double lds[SP ? sp_lds_doubles<N, T, CFG>() : Dev<N, T, K, YB, CFG>::F::LDS_DOUBLES];
  __shared__ double red[SP ? 18 * (T / 64) + 2 : T / 64 + 2];
  adaptive_body<N, T, METHOD, K, YB, SP, CFG>(lds, red, ops, prm, blockIdx.x, in, out, ws_all, nullptr);
}

// Two workgroups per instance (see pair_sync): workgroup w serves instance
// 8 (w / 16) + w % 8 in role (w / 8) % 2, so both roles of an instance sit
// in the same position of two consecutive groups of eight (the same XCD
// when workgroups are dealt to the eight XCDs in turn). Workgroups past the
// batch return at once (they are nobody's partner). flags: two epoch
// counters per instance, zero at launch.
template <int N, int T, int METHOD, int K, int SP, int C>
__global__ void __launch_bounds__(T) k_adaptive_pair(Ops ops, AdaptiveParams prm, const double2* __restrict__ in,
                                                     double2* __restrict__ out, double2* ws_all, unsigned* flags,
                                                     long long batch) {
  __shared__ double lds[SP ? sp_lds_doubles<N, T, C>() : Dev<N, T, K, false, C>::F::LDS_DOUBLES];
  __shared__ double red[SP ? 18 * (T / 64) + 2 : T / 64 + 2];
  const long long w = blockIdx.x;
  const long long b = (w / 16) * 8 + (w % 8);
  const int role = (int)((w / 8) % 2);
  if (b >= batch) return;
  adaptive_body<N, T, METHOD, K, false, SP, C, true>(lds, red, ops, prm, b, in, out, ws_all, nullptr, role,
                                                    flags + 2 * b);
}

// Cooperative launch of the two-workgroup kernel when every workgroup fits
// at once (hipErrorCooperativeLaunchTooLarge otherwise: the caller runs the
// one-workgroup kernel).
template <int N, int T, int METHOD, int K, int SP, int C>
hipError_t launch_pair(hipStream_t s, const Ops& o, const AdaptiveParams& p, long long batch, const double2* in,
                       double2* out, double2* ws, unsigned* flags) {
  const void* k = reinterpret_cast<const void*>(&k_adaptive_pair<N, T, METHOD, K, SP, C>);
  const long long grid = 16 * ((batch + 7) / 8);
  int per_cu = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, T, 0);
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  if ((long long)per_cu * cus < grid) return hipErrorCooperativeLaunchTooLarge;
  Ops ops = o;
  AdaptiveParams prm = p;
  void* args[] = {&ops, &prm, &in, &out, &ws, &flags, &batch};
  return hipLaunchCooperativeKernel(k, dim3((unsigned)grid), dim3(T), args, 0, s);
}

// Dispatch on the nonlinearity: Kerr and the full Raman + self-steepening
// operator get their own instantiations; the other variants share one with
// runtime selection.
#define GN_BY_KIND(ops, KERNEL, ...)                                                          \
  do {                                                                                       \
    if ((ops).nl_kind == NL_KERR) hipLaunchKernelGGL((KERNEL, NL_KERR>), __VA_ARGS__);       \
    else if ((ops).nl_kind == NL_ALL) hipLaunchKernelGGL((KERNEL, NL_ALL>), __VA_ARGS__);    \
    else hipLaunchKernelGGL((KERNEL, NL_GENERIC>), __VA_ARGS__);                             \
  } while (0)
// Propagating kernels also dispatch on the gain model: the Yb gain has its
// own instantiations (full Raman + self-steepening operator, or runtime kind).
#define GN_BY_KIND_GAIN(ops, KERNEL, ...)                                                                 \
  do {                                                                                                  \
    if ((ops).gain == GAIN_YB) {                                                                        \
      if ((ops).nl_kind == NL_ALL) hipLaunchKernelGGL((KERNEL, NL_ALL, true>), __VA_ARGS__);            \
      else hipLaunchKernelGGL((KERNEL, NL_GENERIC, true>), __VA_ARGS__);                                \
    } else if ((ops).nl_kind == NL_KERR) hipLaunchKernelGGL((KERNEL, NL_KERR, false>), __VA_ARGS__);    \
    else if ((ops).nl_kind == NL_ALL) hipLaunchKernelGGL((KERNEL, NL_ALL, false>), __VA_ARGS__);        \
    else hipLaunchKernelGGL((KERNEL, NL_GENERIC, false>), __VA_ARGS__);                                 \
  } while (0)

// The same with the schedule CC (Dev's CFG) of the grid: the propagating
// kernels (fixed and adaptive steppers) run their transforms on the grid's
// plan.
#define GN_BY_KIND_GAIN_C(ops, KERNEL, CC, ...)                                                          \
  do {                                                                                                  \
    if ((ops).gain == GAIN_YB) {                                                                        \
      if ((ops).nl_kind == NL_ALL) hipLaunchKernelGGL((KERNEL, NL_ALL, true, 0, CC>), __VA_ARGS__);     \
      else hipLaunchKernelGGL((KERNEL, NL_GENERIC, true, 0, CC>), __VA_ARGS__);                         \
    } else if ((ops).nl_kind == NL_KERR && (ops).gain == GAIN_NONE)                                     \
      hipLaunchKernelGGL((KERNEL, NL_KERR, false, 0, (CC) | kCfgKerrFree>), __VA_ARGS__);                   \
    else if ((ops).nl_kind == NL_KERR) hipLaunchKernelGGL((KERNEL, NL_KERR, false, 0, CC>), __VA_ARGS__); \
    else if ((ops).nl_kind == NL_ALL) hipLaunchKernelGGL((KERNEL, NL_ALL, false, 0, CC>), __VA_ARGS__); \
    else if ((ops).gain == GAIN_NONE)                                                                   \
      hipLaunchKernelGGL((KERNEL, NL_GENERIC, false, 0, (CC) | kCfgNoGain>), __VA_ARGS__);                  \
    else hipLaunchKernelGGL((KERNEL, NL_GENERIC, false, 0, CC>), __VA_ARGS__);                          \
  } while (0)

}  // namespace gn

#include "wdm.hpp"

namespace gn {

// C is the schedule (Dev's CFG) of the grid's propagating kernels (fixed and
// adaptive steppers, every method and operator); the single-operator
// kernels and the WDM kernels use the plain one.
template <int N, int T, int C = 0>
struct Launcher {
  static hipError_t fft(hipStream_t s, const Ops& o, int sign, long long batch, const double2* in, double2* out,
                        double2*) {
    hipLaunchKernelGGL((k_fft<N, T>), dim3(batch), dim3(T), 0, s, o, sign, in, out);
    return hipGetLastError();
  }
  static hipError_t make_exp(hipStream_t s, const Ops& o, double h, double2* out) {
    hipLaunchKernelGGL((k_make_exp<N, T>), dim3(1), dim3(T), 0, s, o, h, out);
    return hipGetLastError();
  }
  static hipError_t linear(hipStream_t s, const Ops& o, double h, long long batch, const double2* in, double2* out,
                           double2* ws) {
    if (o.wdm) hipLaunchKernelGGL((k_linear_wdm<N, T>), dim3(2 * batch), dim3(T), 0, s, o, h, in, out, ws);
    else if (o.gain == GAIN_YB) hipLaunchKernelGGL((k_linear<N, T, true>), dim3(batch), dim3(T), 0, s, o, h, in, out, ws);
    else hipLaunchKernelGGL((k_linear<N, T, false>), dim3(batch), dim3(T), 0, s, o, h, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t linfac(hipStream_t s, const Ops& o, long long batch, const double2* in, double2* out) {
    if (o.wdm) hipLaunchKernelGGL((k_linfac_wdm<N, T>), dim3(2 * batch), dim3(T), 0, s, o, in, out);
    else hipLaunchKernelGGL((k_linfac<N, T>), dim3(batch), dim3(T), 0, s, o, in, out);
    return hipGetLastError();
  }
  static hipError_t nonlin(hipStream_t s, const Ops& o, long long batch, const double2* in, double2* out,
                           double2* ws) {
    if (o.wdm) hipLaunchKernelGGL((k_nonlin_wdm<N, T>), dim3(batch), dim3(T), 0, s, o, 0.0, in, nullptr, out);
    else GN_BY_KIND(o, k_nonlin<N COMMA T, dim3(batch), dim3(T), 0, s, o, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t exp_nonlin(hipStream_t s, const Ops& o, double h, long long batch, const double2* A,
                               const double2* B, double2* out, double2* ws) {
    if (o.wdm) hipLaunchKernelGGL((k_nonlin_wdm<N, T>), dim3(batch), dim3(T), 0, s, o, h, A, B, out);
    else GN_BY_KIND(o, k_exp_nonlin<N COMMA T, dim3(batch), dim3(T), 0, s, o, h, A, B, out, ws);
    return hipGetLastError();
  }
  static hipError_t step(hipStream_t s, const Ops& o, int method, double h, double h_exp, long long batch, const double2* in,
                         double2* out, double2* ws) {
    if (method == M_SS_SYM_RK4) {
      if (o.wdm) return hipErrorNotSupported;
      if (o.gain == GAIN_YB)
        hipLaunchKernelGGL((k_step<N, T, M_SS_SYM_RK4, NL_GENERIC, true>), dim3(batch), dim3(T), 0, s, o, h, h_exp, in,
                           out, ws);
      else
        hipLaunchKernelGGL((k_step<N, T, M_SS_SYM_RK4, NL_GENERIC, false>), dim3(batch), dim3(T), 0, s, o, h, h_exp, in,
                           out, ws);
      return hipGetLastError();
    }
    if (o.wdm && method == M_RK4IP)
      hipLaunchKernelGGL((k_step_wdm<N, T, M_RK4IP>), dim3(batch), dim3(T), 0, s, o, h, h_exp, in, out, ws);
    else if (o.wdm)
      hipLaunchKernelGGL((k_step_wdm<N, T, M_SS_SYMMETRIC>), dim3(batch), dim3(T), 0, s, o, h, h_exp, in, out, ws);
    else if (method == M_RK4IP) GN_BY_KIND_GAIN(o, k_step<N COMMA T COMMA M_RK4IP, dim3(batch), dim3(T), 0, s, o, h, h_exp, in, out, ws);
    else GN_BY_KIND_GAIN(o, k_step<N COMMA T COMMA M_SS_SYMMETRIC, dim3(batch), dim3(T), 0, s, o, h, h_exp, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t fixed(hipStream_t s, const Ops& o, int method, const FixedParams& p, long long batch,
                          const double2* in, double2* out, double2* ws) {
    if (method == M_SS_SYM_RK4) {
      if (o.wdm) return hipErrorNotSupported;
      if (o.gain == GAIN_YB)
        hipLaunchKernelGGL((k_fixed<N, T, M_SS_SYM_RK4, NL_GENERIC, true, 0, C>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
      else
        hipLaunchKernelGGL((k_fixed<N, T, M_SS_SYM_RK4, NL_GENERIC, false, 0, C>), dim3(batch), dim3(T), 0, s, o, p, in, out,
                           ws);
      return hipGetLastError();
    }
    if (o.wdm && method == M_RK4IP)
      hipLaunchKernelGGL((k_fixed_wdm<N, T, M_RK4IP>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else if (o.wdm)
      hipLaunchKernelGGL((k_fixed_wdm<N, T, M_SS_SYMMETRIC>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else if (method == M_RK4IP && o.nl_kind == NL_ALL && o.gain == GAIN_YB)
      hipLaunchKernelGGL((k_fixed<N, T, M_RK4IP, NL_ALL, true, 1 + SPG_YB, C>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else if (method == M_RK4IP && o.nl_kind == NL_ALL && o.gain == GAIN_NONE)
      hipLaunchKernelGGL((k_fixed<N, T, M_RK4IP, NL_ALL, false, 1, C>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else if (method == M_RK4IP && o.nl_kind == NL_ALL && o.gain == GAIN_SATURABLE && o.simp_flat)
      hipLaunchKernelGGL((k_fixed<N, T, M_RK4IP, NL_ALL, false, 1 + SPG_SAT, C>), dim3(batch), dim3(T), 0, s, o, p, in, out,
                         ws);
    else if (method == M_RK4IP) GN_BY_KIND_GAIN_C(o, k_fixed<N COMMA T COMMA M_RK4IP, C, dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else GN_BY_KIND_GAIN_C(o, k_fixed<N COMMA T COMMA M_SS_SYMMETRIC, C, dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t adaptive(hipStream_t s, const Ops& o, int method, const AdaptiveParams& p, long long batch,
                             const double2* in, double2* out, double2* ws) {
    if (method == M_SS_SYM_RK4) {
      if (o.wdm) return hipErrorNotSupported;
      if (o.gain == GAIN_YB)
        hipLaunchKernelGGL((k_adaptive<N, T, M_SS_SYM_RK4, NL_GENERIC, true, 0, C>), dim3(batch), dim3(T), 0, s, o, p, in, out,
                           ws);
      else
        hipLaunchKernelGGL((k_adaptive<N, T, M_SS_SYM_RK4, NL_GENERIC, false, 0, C>), dim3(batch), dim3(T), 0, s, o, p, in,
                           out, ws);
      return hipGetLastError();
    }
    if (o.wdm && method == M_RK4IP)
      hipLaunchKernelGGL((k_adaptive_wdm<N, T, M_RK4IP>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else if (o.wdm)
      hipLaunchKernelGGL((k_adaptive_wdm<N, T, M_SS_SYMMETRIC>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else if (method == M_RK4IP && o.gain == GAIN_NONE && o.nl_kind == NL_ALL)
      hipLaunchKernelGGL((k_adaptive<N, T, M_RK4IP, NL_ALL, false, 1, C>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else if (method == M_RK4IP && o.gain == GAIN_SATURABLE && o.simp_flat && o.nl_kind == NL_ALL)
      hipLaunchKernelGGL((k_adaptive<N, T, M_RK4IP, NL_ALL, false, 2, C>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else if (method == M_RK4IP && o.gain == GAIN_YB && o.nl_kind == NL_ALL)
      hipLaunchKernelGGL((k_adaptive<N, T, M_RK4IP, NL_ALL, true, 1 + SPG_YB, C>), dim3(batch), dim3(T), 0, s, o, p, in, out,
                         ws);
    else if (method == M_RK4IP)
      GN_BY_KIND_GAIN_C(o, k_adaptive<N COMMA T COMMA M_RK4IP, C, dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else GN_BY_KIND_GAIN_C(o, k_adaptive<N COMMA T COMMA M_SS_SYMMETRIC, C, dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    return hipGetLastError();
  }
  static hipError_t characts(hipStream_t s, long long count, const double2* A, const double2* S, const double* w,
                             double* out, int* npk, int* pk, int max_pk, double2*) {
    hipLaunchKernelGGL((k_characts<N>), dim3((unsigned)count), dim3(kCharT), 0, s, A, S, w, out, npk, pk, max_pk);
    return hipGetLastError();
  }
  // The two-workgroup kernels exist on the 4096-16384 grids for rk4ip with
  // the full operator (no gain or the saturable gain) and for rk4ip and
  // ss_symmetric with the Kerr operator.
  static hipError_t adaptive_pair(hipStream_t s, const Ops& o, int method, const AdaptiveParams& p, long long batch,
                                  const double2* in, double2* out, double2* ws, unsigned* flags) {
    if constexpr (N < 4096) {
      return hipErrorNotSupported;
    } else {
      if (o.wdm || o.gain == GAIN_YB) return hipErrorNotSupported;
      if (method == M_RK4IP && o.nl_kind == NL_ALL && o.gain == GAIN_NONE)
        return launch_pair<N, T, M_RK4IP, NL_ALL, 1, C>(s, o, p, batch, in, out, ws, flags);
      if (method == M_RK4IP && o.nl_kind == NL_ALL && o.gain == GAIN_SATURABLE && o.simp_flat)
        return launch_pair<N, T, M_RK4IP, NL_ALL, 2, C>(s, o, p, batch, in, out, ws, flags);
      if (o.nl_kind == NL_KERR && o.gain == GAIN_NONE) {
        if (method == M_RK4IP)
          return launch_pair<N, T, M_RK4IP, NL_KERR, 0, C | kCfgKerrFree>(s, o, p, batch, in, out, ws, flags);
        if (method == M_SS_SYMMETRIC)
          return launch_pair<N, T, M_SS_SYMMETRIC, NL_KERR, 0, C | kCfgKerrFree>(s, o, p, batch, in, out, ws, flags);
      }
      return hipErrorNotSupported;
    }
  }
  static constexpr LaunchTable table = {N,       T,     NSLOT, 1,           &fft,  &make_exp, &linear,  &linfac,
                                        &nonlin, &exp_nonlin, &step, &fixed, &adaptive, &characts, C, &adaptive_pair};
};

template <int N, int T, int C>
struct SchedLauncher {
  static hipError_t adaptive(hipStream_t s, const Ops& o, const AdaptiveParams& p, long long batch, const double2* in,
                             double2* out, double2* ws) {
    if (o.gain == GAIN_NONE)
      hipLaunchKernelGGL((k_adaptive<N, T, M_RK4IP, NL_ALL, false, 1, C>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else if (o.gain == GAIN_SATURABLE && o.simp_flat)
      hipLaunchKernelGGL((k_adaptive<N, T, M_RK4IP, NL_ALL, false, 2, C>), dim3(batch), dim3(T), 0, s, o, p, in, out, ws);
    else return hipErrorNotSupported;
    return hipGetLastError();
  }
  static constexpr Schedule sched = {T, C, &adaptive};
};


}  // namespace gn

#define GN_DEFINE_SCHEDULE(NN, TT, CC, ID) \
  namespace gn {                            \
  const Schedule* schedule_##NN##_##ID() { return &SchedLauncher<NN, TT, CC>::sched; } \
  }


#define GN_DEFINE_GRID(NN, TT, CC) \
  namespace gn {                   \
  const LaunchTable* table_##NN() { return &Launcher<NN, TT, CC>::table; } \
  }
