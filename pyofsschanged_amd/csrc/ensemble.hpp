// Noise ensembles on mixed grids in one persistent launch, and their
// spectral coherence |g12|.
//
// An ensemble is a list of groups; a group is a set of realisations on one
// grid (N = 4096, 8192 or 16384) sharing the operators of one context. The
// launch has one workgroup per realisation over all groups; workgroup w runs
// the adaptive propagation of realisation w - first(g) of the group g whose
// block range contains w, with that group's operators, controller constants
// and buffers. Every realisation keeps its own step size, attempt count and
// status. All grids share the launch's thread count T, so the per-grid code
// paths differ only in points per thread (N/T) and pass plan; T is chosen per
// ensemble by a timed trial of the compiled candidates (512, 1024).
//
// Ensembles whose groups all carry the full Raman + self-steepening operator
// without gain (rk4ip) run the spectral-domain RK4IP of the single-grid
// launches (k_ensemble_sp), each grid with its default schedule.
//
// Groups are laid out in the order the host gives them (largest grid first),
// so the workgroups dispatched first are the longest-running ones.
//
// Coherence (per group of M realisations with pyofss spectra a_i = fft(A_i)):
//   g12 = | |sum_i a_i|^2 - sum_i |a_i|^2 | / ((M - 1) sum_i |a_i|^2),
// which is |sum_{i != j} a_i* a_j| / ((M - 1) sum_i |a_i|^2), the mean
// cross-correlation over ordered pairs normalised by the mean power; the
// sums run over i in a fixed order, so the result is reproducible bit for
// bit.
#pragma once
#include "kernels.hpp"

namespace gn {

constexpr int kEnsMaxGroups = 4;

struct EnsembleGroup {
  Ops ops;
  AdaptiveParams prm;
  const double2* in;
  double2* out;
  double2* ws;
  double2* spec;
  long long n, first, count;
};

// Schedule (Dev's CFG) of a grid inside an ensemble launch: the default
// schedule of the 8192- and 16384-sample grids' own launches (lane-pair
// N-point plan, unrolled spectral-domain passes; stage blocks of 8 points
// at 8192 and 2 at 16384) where the points per thread allow it.
template <int N, int T>
constexpr int ens_cfg() {
  return (N >= 8192 && N / T <= 16)
             ? (8 | kCfgUnrolled | (N == 8192 && T == 512 ? 0x200 : 0) | (N == 16384 && T == 1024 ? 0x100 : 0))
             : 0;
}

template <int T>
struct EnsembleShape {
  static constexpr int a = Dev<4096, T, NL_KERR, false, ens_cfg<4096, T>()>::F::LDS_DOUBLES;
  static constexpr int b = Dev<8192, T, NL_KERR, false, ens_cfg<8192, T>()>::F::LDS_DOUBLES;
  static constexpr int c = Dev<16384, T, NL_KERR, false, ens_cfg<16384, T>()>::F::LDS_DOUBLES;
  static constexpr int LDS_DOUBLES = a > b ? (a > c ? a : c) : (b > c ? b : c);
};

// X: further Dev CFG bits of every grid (kCfgKerrFree when the groups are
// Kerr without gain).
template <int T, int METHOD, int K, int X = 0>
__global__ void __launch_bounds__(T) k_ensemble(const EnsembleGroup* __restrict__ groups, int ngroups) {
  __shared__ double lds[EnsembleShape<T>::LDS_DOUBLES];
  __shared__ double red[T / 64 + 2];
  const long long w = blockIdx.x;
  int gi = 0;
#pragma unroll
  for (int i = 1; i < kEnsMaxGroups; ++i)
    if (i < ngroups && w >= groups[i].first) gi = i;
  const EnsembleGroup& g = groups[gi];
  const size_t b = (size_t)(w - g.first);
  if (g.n == 4096)
    adaptive_body<4096, T, METHOD, K, false, 0, ens_cfg<4096, T>() | X>(lds, red, g.ops, g.prm, b, g.in, g.out, g.ws,
                                                                     g.spec);
  else if (g.n == 8192)
    adaptive_body<8192, T, METHOD, K, false, 0, ens_cfg<8192, T>() | X>(lds, red, g.ops, g.prm, b, g.in, g.out, g.ws,
                                                                     g.spec);
  else
    adaptive_body<16384, T, METHOD, K, false, 0, ens_cfg<16384, T>() | X>(lds, red, g.ops, g.prm, b, g.in, g.out,
                                                                       g.ws, g.spec);
}

template <int T>
struct EnsembleShapeSP {
  static constexpr int a = sp_lds_doubles<4096, T, ens_cfg<4096, T>()>();
  static constexpr int b = sp_lds_doubles<8192, T, ens_cfg<8192, T>()>();
  static constexpr int c = sp_lds_doubles<16384, T, ens_cfg<16384, T>()>();
  static constexpr int LDS_DOUBLES = a > b ? (a > c ? a : c) : (b > c ? b : c);
};

template <int T>
__global__ void __launch_bounds__(T) k_ensemble_sp(const EnsembleGroup* __restrict__ groups, int ngroups) {
  __shared__ double lds[EnsembleShapeSP<T>::LDS_DOUBLES];
  __shared__ double red[18 * (T / 64) + 2];
  const long long w = blockIdx.x;
  int gi = 0;
#pragma unroll
  for (int i = 1; i < kEnsMaxGroups; ++i)
    if (i < ngroups && w >= groups[i].first) gi = i;
  const EnsembleGroup& g = groups[gi];
  const size_t b = (size_t)(w - g.first);
  if (g.n == 4096)
    adaptive_body<4096, T, M_RK4IP, NL_ALL, false, 1, ens_cfg<4096, T>()>(lds, red, g.ops, g.prm, b, g.in, g.out,
                                                                              g.ws, g.spec);
  else if (g.n == 8192)
    adaptive_body<8192, T, M_RK4IP, NL_ALL, false, 1, ens_cfg<8192, T>()>(lds, red, g.ops, g.prm, b, g.in, g.out,
                                                                              g.ws, g.spec);
  else
    adaptive_body<16384, T, M_RK4IP, NL_ALL, false, 1, ens_cfg<16384, T>()>(lds, red, g.ops, g.prm, b, g.in,
                                                                                g.out, g.ws, g.spec);
}

// g12 of one group: one thread per spectral sample, the realisations summed
// in index order (reads are unit-stride across the wavefront for each i).
template <int BS>
__global__ void __launch_bounds__(BS) k_coherence(const double2* __restrict__ spec, long long M, long long n,
                                                 double* __restrict__ g12) {
  const long long k = (long long)blockIdx.x * BS + threadIdx.x;
  if (k >= n) return;
  double sx = 0.0, sy = 0.0, p = 0.0;
  for (long long i = 0; i < M; ++i) {
    const double2 a = spec[i * n + k];
    sx += a.x;
    sy += a.y;
    p += a.x * a.x + a.y * a.y;
  }
  const double cross = (sx * sx + sy * sy) - p;
  g12[k] = (p > 0.0) ? fabs(cross) / ((double)(M - 1) * p) : 0.0;
}

struct EnsembleTable {
  int t;
  hipError_t (*run)(hipStream_t, int method, int kind, bool gain, const EnsembleGroup* groups, int ngroups,
                    long long blocks);
};

template <int T>
struct EnsembleLauncher {
  // gain: some group's operators carry a gain
  static hipError_t run(hipStream_t s, int method, int kind, bool gain, const EnsembleGroup* groups, int ngroups,
                        long long blocks) {
    if (method == M_RK4IP) {
      if (kind == NL_ALL) hipLaunchKernelGGL((k_ensemble_sp<T>), dim3(blocks), dim3(T), 0, s, groups, ngroups);
      else if (kind == NL_KERR && !gain)
        hipLaunchKernelGGL((k_ensemble<T, M_RK4IP, NL_KERR, kCfgKerrFree>), dim3(blocks), dim3(T), 0, s, groups, ngroups);
      else if (kind == NL_KERR) hipLaunchKernelGGL((k_ensemble<T, M_RK4IP, NL_KERR>), dim3(blocks), dim3(T), 0, s, groups, ngroups);
      else hipLaunchKernelGGL((k_ensemble<T, M_RK4IP, NL_GENERIC>), dim3(blocks), dim3(T), 0, s, groups, ngroups);
    } else {
      if (kind == NL_KERR && !gain)
        hipLaunchKernelGGL((k_ensemble<T, M_SS_SYMMETRIC, NL_KERR, kCfgKerrFree>), dim3(blocks), dim3(T), 0, s, groups,
                           ngroups);
      else if (kind == NL_KERR)
        hipLaunchKernelGGL((k_ensemble<T, M_SS_SYMMETRIC, NL_KERR>), dim3(blocks), dim3(T), 0, s, groups, ngroups);
      else
        hipLaunchKernelGGL((k_ensemble<T, M_SS_SYMMETRIC, NL_GENERIC>), dim3(blocks), dim3(T), 0, s, groups, ngroups);
    }
    return hipGetLastError();
  }
  static constexpr EnsembleTable table = {T, &run};
};

const EnsembleTable* ensemble_table_512();
const EnsembleTable* ensemble_table_1024();

}  // namespace gn

#define GN_DEFINE_ENSEMBLE(TT) \
  namespace gn {               \
  const EnsembleTable* ensemble_table_##TT() { return &EnsembleLauncher<TT>::table; } \
  }
