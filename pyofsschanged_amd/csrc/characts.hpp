// Trace characteristics on the device: pyofss's Storage.update_characts
// (pyofss/modules/storage.py:194-209) for every stored trace in one launch.
//
// Per trace (one workgroup of 256 threads, the power held in LDS):
//   energy      = sum_n w_n |A_n|^2 (w: old-Simpson weights of t * 1e-3,
//                 field.py:161-172)
//   max_power   = max |A_n|^2
//   duration    = width at half prominence of the highest peak of |A|^2, in
//                 samples (max_peak_params + get_duration, field.py:207-229,
//                 283-287)
//   spec_width  = spectrum_width_params + get_duration_spec on the spectral
//                 power in natural frequency order (field.py:60-75,
//                 232-280), in samples; NaN where the reference finds no
//                 qualifying peak
//   peaks       = indices of the peaks of |A|^2 with prominence >= max/10
//                 (get_peaks, field.py:272-274), in increasing order
// Peak semantics are scipy.signal's: a peak is a sample whose neighbours are
// smaller (a flat top counts once, at its middle sample rounded down; the
// first and last samples are never peaks); prominence extends a horizontal
// line from the peak to the window border or a higher sample (equal heights
// do not stop it), takes the minimum on each side (the base; ties nearest the
// peak) and subtracts the higher base from the height; the width at
// rel_height 0.5 is measured at h = height - prominence / 2 from the peak
// outwards, not past the bases, with linear interpolation where the sample
// falls below h.
//
// The highest peak's prominence and width are block reductions (nearest
// higher sample, minimum and its index, crossing indices). Prominence
// filtering screens candidates by height - global minimum >= threshold (an
// upper bound of the prominence) and scans the survivors serially in LDS.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace gn {

constexpr int kCharT = 256;

struct CharScratch {
  double v[kCharT];
  long long i[kCharT];
};

// Block reduction of (value, index) pairs; better(a, b) says a wins.
template <class Better>
__device__ __forceinline__ void char_reduce(CharScratch& s, double& v, long long& idx, Better better) {
  __syncthreads();
  s.v[threadIdx.x] = v;
  s.i[threadIdx.x] = idx;
  __syncthreads();
  for (int o = kCharT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const double v2 = s.v[threadIdx.x + o];
      const long long i2 = s.i[threadIdx.x + o];
      if (better(v2, i2, s.v[threadIdx.x], s.i[threadIdx.x])) {
        s.v[threadIdx.x] = v2;
        s.i[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  v = s.v[0];
  idx = s.i[0];
  __syncthreads();
}

// Peak starting at a rising edge i (x[i-1] < x[i]); returns the peak index or
// -1, and the index where the next rising edge search continues.
__device__ __forceinline__ int char_peak_at(const double* x, int n, int i) {
  if (i < 1 || i >= n - 1 || !(x[i - 1] < x[i])) return -1;
  int j = i + 1;
  while (j < n - 1 && x[j] == x[i]) ++j;
  return (x[j] < x[i]) ? (i + j - 1) / 2 : -1;
}

// Serial prominence of peak p (one thread).
__device__ __forceinline__ double char_prominence(const double* x, int n, int p, int& lb, int& rb) {
  const double h = x[p];
  double lmin = h, rmin = h;
  lb = rb = p;
  for (int i = p; i >= 0 && x[i] <= h; --i)
    if (x[i] < lmin) {
      lmin = x[i];
      lb = i;
    }
  for (int i = p; i < n && x[i] <= h; ++i)
    if (x[i] < rmin) {
      rmin = x[i];
      rb = i;
    }
  return h - fmax(lmin, rmin);
}

// Width at half prominence of peak p with known bases, by block reductions.
__device__ __forceinline__ double char_width(CharScratch& s, const double* x, int n, int p, double prom, int lb,
                                             int rb) {
  const double h = x[p] - 0.5 * prom;
  long long il = -1, ir = n;
  for (int i = threadIdx.x; i < n; i += kCharT) {
    if (i >= lb && i <= p && x[i] <= h && i > il) il = i;
    if (i >= p && i <= rb && x[i] <= h && i < ir) ir = i;
  }
  double dummy = 0.0;
  char_reduce(s, dummy, il, [](double, long long a, double, long long b) { return a > b; });
  dummy = 0.0;
  char_reduce(s, dummy, ir, [](double, long long a, double, long long b) { return a < b; });
  double left = (double)il, right = (double)ir;
  if (x[il] < h) left += (h - x[il]) / (x[il + 1] - x[il]);
  if (x[ir] < h) right -= (h - x[ir]) / (x[ir - 1] - x[ir]);
  return right - left;
}

// Prominence and bases of peak p by block reductions (no serial scan).
__device__ __forceinline__ double char_prominence_par(CharScratch& s, const double* x, int n, int p, int& lb,
                                                      int& rb) {
  const double h = x[p];
  long long hl = -1, hr = n;  // nearest higher samples
  for (int i = threadIdx.x; i < n; i += kCharT) {
    if (i < p && x[i] > h && i > hl) hl = i;
    if (i > p && x[i] > h && i < hr) hr = i;
  }
  double d = 0.0;
  char_reduce(s, d, hl, [](double, long long a, double, long long b) { return a > b; });
  d = 0.0;
  char_reduce(s, d, hr, [](double, long long a, double, long long b) { return a < b; });
  double lmin = INFINITY, rmin = INFINITY;
  long long il = p, ir = p;
  for (int i = threadIdx.x; i < n; i += kCharT) {
    if (i > hl && i <= p && (x[i] < lmin || (x[i] == lmin && i > il))) {
      lmin = x[i];
      il = i;
    }
    if (i >= p && i < hr && (x[i] < rmin || (x[i] == rmin && i < ir))) {
      rmin = x[i];
      ir = i;
    }
  }
  // ties nearest the peak: larger index on the left, smaller on the right
  char_reduce(s, lmin, il, [](double va, long long a, double vb, long long b) { return va < vb || (va == vb && a > b); });
  char_reduce(s, rmin, ir, [](double va, long long a, double vb, long long b) { return va < vb || (va == vb && a < b); });
  lb = (int)il;
  rb = (int)ir;
  return h - fmax(lmin, rmin);
}

// The characteristics of trace b, with its powers in x[N] and the peak flags
// in flag[N] (LDS, or global scratch on the large grids).
template <int N>
__device__ __forceinline__ void characts_body(const double2* __restrict__ A, const double2* __restrict__ S,
                                              const double* __restrict__ w, double* __restrict__ out,
                                              int* __restrict__ npk, int* __restrict__ pk, int max_pk, double* x,
                                              unsigned char* flag, CharScratch& s, int* cnt) {
  const long long b = blockIdx.x;
  const int tid = threadIdx.x;

  // ---- temporal power --------------------------------------------------
  double e = 0.0;
  for (int i = tid; i < N; i += kCharT) {
    const double2 a = A[b * N + i];
    x[i] = a.x * a.x + a.y * a.y;
    e += w[i] * x[i];
    flag[i] = 0;
  }
  // energy: fixed-order additive tree over the per-thread partial sums
  {
    __syncthreads();
    s.v[tid] = e;
    __syncthreads();
    for (int o = kCharT / 2; o > 0; o >>= 1) {
      if (tid < o) s.v[tid] += s.v[tid + o];
      __syncthreads();
    }
    e = s.v[0];
    __syncthreads();
  }
  double mx = -INFINITY, mn = INFINITY;
  long long dm = 0;
  for (int i = tid; i < N; i += kCharT) {
    mx = fmax(mx, x[i]);
    mn = fmin(mn, x[i]);
  }
  char_reduce(s, mx, dm, [](double va, long long, double vb, long long) { return va > vb; });
  char_reduce(s, mn, dm, [](double va, long long, double vb, long long) { return va < vb; });

  // highest peak (first in index order among equal heights)
  double hv = -INFINITY;
  long long hp = -1;
  for (int i = tid + 1; i < N - 1; i += kCharT) {
    const int p = char_peak_at(x, N, i);
    if (p >= 0 && (x[p] > hv || (x[p] == hv && p < hp))) {
      hv = x[p];
      hp = p;
    }
  }
  char_reduce(s, hv, hp, [](double va, long long a, double vb, long long b2) {
    return b2 < 0 || (a >= 0 && (va > vb || (va == vb && a < b2)));
  });
  double duration = NAN;
  if (hp >= 0) {
    int lb, rb;
    const double prom = char_prominence_par(s, x, N, (int)hp, lb, rb);
    duration = char_width(s, x, N, (int)hp, prom, lb, rb);
  }

  // peaks with prominence >= max/10: flag, then ordered compaction
  const double thr10 = mx / 10.0;
  for (int i = tid + 1; i < N - 1; i += kCharT) {
    const int p = char_peak_at(x, N, i);
    if (p >= 0 && x[p] - mn >= thr10) {
      int lb, rb;
      if (char_prominence(x, N, p, lb, rb) >= thr10) flag[p] = 1;
    }
  }
  __syncthreads();
  constexpr int CH = N / kCharT;
  int c = 0;
  for (int i = tid * CH; i < (tid + 1) * CH; ++i) c += flag[i];
  cnt[tid + 1] = c;
  __syncthreads();
  if (tid == 0) {
    cnt[0] = 0;
    for (int k = 1; k <= kCharT; ++k) cnt[k] += cnt[k - 1];
  }
  __syncthreads();
  {
    int o = cnt[tid];
    for (int i = tid * CH; i < (tid + 1) * CH; ++i)
      if (flag[i]) {
        if (o < max_pk) pk[b * max_pk + o] = i;
        ++o;
      }
  }
  if (tid == 0) npk[b] = cnt[kCharT];
  __syncthreads();

  // ---- spectral power in natural order ---------------------------------
  for (int i = tid; i < N; i += kCharT) {
    const double2 a = S[b * N + ((i + N / 2) % N)];
    x[i] = a.x * a.x + a.y * a.y;
  }
  double smx = -INFINITY, smn = INFINITY;
  for (int i = tid; i < N; i += kCharT) {
    smx = fmax(smx, x[i]);
    smn = fmin(smn, x[i]);
  }
  char_reduce(s, smx, dm, [](double va, long long, double vb, long long) { return va > vb; });
  char_reduce(s, smn, dm, [](double va, long long, double vb, long long) { return va < vb; });
  const double thr100 = smx / 100.0;
  // prominent peaks: count, first, last, and the bases of one of them
  long long first = N, last = -1, count = 0;
  double one_prom = 0.0;
  long long one_p = -1, one_lb = 0, one_rb = 0;
  for (int i = tid + 1; i < N - 1; i += kCharT) {
    const int p = char_peak_at(x, N, i);
    if (p >= 0 && x[p] - smn >= thr100) {
      int lb, rb;
      const double prom = char_prominence(x, N, p, lb, rb);
      if (prom >= thr100) {
        ++count;
        if (p < first) first = p;
        if (p > last) last = p;
        one_p = p;
        one_prom = prom;
        one_lb = lb;
        one_rb = rb;
      }
    }
  }
  double cd;
  {
    __syncthreads();
    s.v[tid] = (double)count;
    __syncthreads();
    for (int o = kCharT / 2; o > 0; o >>= 1) {
      if (tid < o) s.v[tid] += s.v[tid + o];
      __syncthreads();
    }
    cd = s.v[0];
    __syncthreads();
  }
  double fd = 0.0;
  char_reduce(s, fd, first, [](double, long long a, double, long long b2) { return a < b2; });
  fd = 0.0;
  char_reduce(s, fd, last, [](double, long long a, double, long long b2) { return a > b2; });
  double spec = NAN;
  if (cd > 1.5) {
    const double h10 = smx / 10.0;
    long long lft = -1, rgt = N;
    for (int i = tid; i < N; i += kCharT) {
      if (i < first && x[i] < h10 && i > lft) lft = i;
      if (i > last && x[i] < h10 && i < rgt) rgt = i;
    }
    fd = 0.0;
    char_reduce(s, fd, lft, [](double, long long a, double, long long b2) { return a > b2; });
    fd = 0.0;
    char_reduce(s, fd, rgt, [](double, long long a, double, long long b2) { return a < b2; });
    const long long left = lft >= 0 ? lft : 0, right = rgt < N ? rgt : 0;
    spec = fabs((double)(right - left));
  } else if (cd > 0.5) {
    // the single prominent peak: its owner publishes prominence and bases
    double pv = one_prom;
    long long owner = one_p;
    char_reduce(s, pv, owner, [](double, long long a, double, long long b2) { return a > b2; });
    __syncthreads();
    if (one_p == owner && one_p >= 0) {
      s.i[0] = one_lb;
      s.i[1] = one_rb;
      s.v[0] = one_prom;
    }
    __syncthreads();
    const int lb = (int)s.i[0], rb = (int)s.i[1];
    const double prom = s.v[0];
    __syncthreads();
    spec = fabs(char_width(s, x, N, (int)owner, prom, lb, rb));
  }
  if (tid == 0) {
    out[b * 4 + 0] = e;
    out[b * 4 + 1] = mx;
    out[b * 4 + 2] = duration;
    out[b * 4 + 3] = spec;
  }
}

template <int N>
__global__ void __launch_bounds__(kCharT) k_characts(const double2* __restrict__ A, const double2* __restrict__ S,
                                                    const double* __restrict__ w, double* __restrict__ out,
                                                    int* __restrict__ npk, int* __restrict__ pk, int max_pk) {
  __shared__ double x[N];
  __shared__ unsigned char flag[N];
  __shared__ CharScratch s;
  __shared__ int cnt[kCharT + 1];
  characts_body<N>(A, S, w, out, npk, pk, max_pk, x, flag, s, cnt);
}

// Large grids: the powers and flags of trace b in scratch + b (N + N/8)
// doubles (written and read by this workgroup only; its barriers order them).
template <int N>
__global__ void __launch_bounds__(kCharT) k_characts_global(const double2* __restrict__ A,
                                                           const double2* __restrict__ S, const double* __restrict__ w,
                                                           double* __restrict__ out, int* __restrict__ npk,
                                                           int* __restrict__ pk, int max_pk, double2* scratch) {
  __shared__ CharScratch s;
  __shared__ int cnt[kCharT + 1];
  double* x = reinterpret_cast<double*>(scratch) + (size_t)blockIdx.x * (N + N / 8);
  characts_body<N>(A, S, w, out, npk, pk, max_pk, x, reinterpret_cast<unsigned char*>(x + N), s, cnt);
}

}  // namespace gn
