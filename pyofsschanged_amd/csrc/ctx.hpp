// Engine context internals shared by the C ABI translation units: device
// buffers, the per-grid context, and error reporting.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ensemble.hpp"

namespace gn {
const LaunchTable* find_table(long long n);
}  // namespace gn

using namespace gn;

extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(GNLSE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define RET_IF(x)                  \
  do {                             \
    int r_ = (x);                  \
    if (r_ != GNLSE_OK) return r_; \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct gnlse_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int64_t n = 0;
  const LaunchTable* tab = nullptr;
  DevBuf lin, efix, nlmul, hr, omega, gprof, simpw, tw;
  double efix_h = NAN;
  bool have_lin = false, have_nl = false, have_lin2 = false;
  Ops ops{};
  DevBuf ws, io_in, io_out, io_b, tr_A, tr_z, st_z, st_h, lg_h, lg_d, lg_a, rp_h, rp_a, res, ntr;
  DevBuf ch_w, ch_out, ch_npk, ch_pk;  // trace characteristics
  DevBuf yb_rs, yb_ip, yb_as, yb_es;   // two-level Yb gain arrays
  DevBuf yb_pp;                        // pump power per instance
  int64_t yb_count = 0;                // instances with a pump power set
  DevBuf lin2, efix2;                  // WDM channel 1 linear factor and its fixed-step exponential
  DevBuf rows;                         // row index table of gnlse_copy_rows
  void* scratch = nullptr;             // this device's no-allocation workspace (device address)
  int sched = 0;                       // schedule of the spectral-domain RK4IP launches (0: the grid's default)
  DevBuf pair_flags;                   // two-workgroup launches: two epoch counters per instance
  int pairing = 1;                     // two workgroups per instance when the batch leaves room
};

// Fields per instance: 2 with the WDM channels, else 1.
inline int64_t fields_per_instance(const gnlse_ctx* c) { return c->ops.wdm ? 2 : 1; }

inline int ensure(DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return GNLSE_OK;
  if (b.p) HIPCHK(hipFree(b.p));
  b.p = nullptr;
  b.bytes = 0;
  HIPCHK(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  return GNLSE_OK;
}
inline void release(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

inline int upload(gnlse_ctx* c, DevBuf& b, const void* src, size_t bytes) {
  RET_IF(ensure(b, bytes));
  HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

inline int upload_raw(int device, hipStream_t s, DevBuf& b, const void* src, size_t bytes) {
  HIPCHK(hipSetDevice(device));
  RET_IF(ensure(b, bytes));
  HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  return GNLSE_OK;
}

// exp(+2 pi i m / N) in extended precision, rounded to double.
inline int build_twiddles(gnlse_ctx* c) {
  const int64_t n = c->n;
  std::vector<double> tw(2 * n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int64_t m = 0; m < n; ++m) {
    const long double a = two_pi * (long double)m / (long double)n;
    tw[2 * m] = (double)cosl(a);
    tw[2 * m + 1] = (double)sinl(a);
  }
  return upload(c, c->tw, tw.data(), tw.size() * sizeof(double));
}

inline void refresh_ops(gnlse_ctx* c) {
  c->ops.lin = (const double2*)c->lin.p;
  c->ops.efix = (const double2*)c->efix.p;
  c->ops.nlmul = (const double2*)c->nlmul.p;
  c->ops.hr = (const double2*)c->hr.p;
  c->ops.omega = (const double*)c->omega.p;
  c->ops.tw = (const double2*)c->tw.p;
  c->ops.gprof = (const double*)c->gprof.p;
  c->ops.simpw = (const double*)c->simpw.p;
  c->ops.yb_rs = (const double*)c->yb_rs.p;
  c->ops.yb_ip = (const double*)c->yb_ip.p;
  c->ops.yb_as = (const double*)c->yb_as.p;
  c->ops.yb_es = (const double*)c->yb_es.p;
  c->ops.yb_pp = (double*)c->yb_pp.p;
  c->ops.lin2 = (const double2*)c->lin2.p;
  c->ops.efix2 = (const double2*)c->efix2.p;
}

// With the Yb gain every instance of a launch carries a pump power.
inline int check_pump(gnlse_ctx* c, int64_t batch) {
  if (c->ops.gain == GAIN_YB && c->yb_count < batch)
    return fail(GNLSE_ERR_STATE, "the Yb gain needs a pump power per instance (gnlse_set_pump with count >= batch)");
  return GNLSE_OK;
}

inline int check_ready(gnlse_ctx* c, bool need_lin, bool need_nl) {
  if (!c) return fail(GNLSE_ERR_ARG, "null context");
  if (c->n == 0 || !c->tab) return fail(GNLSE_ERR_STATE, "gnlse_set_grid has not been called");
  if (need_lin && !c->have_lin) return fail(GNLSE_ERR_STATE, "gnlse_set_linear has not been called");
  if (need_nl && !c->have_nl) return fail(GNLSE_ERR_STATE, "gnlse_set_nonlinear has not been called");
  if (need_lin && c->ops.wdm && !c->have_lin2)
    return fail(GNLSE_ERR_STATE, "WDM fields need both channels' linear factors (gnlse_set_linear_wdm)");
  return GNLSE_OK;
}

inline int ensure_ws(gnlse_ctx* c, int64_t batch) {
  return ensure(c->ws, (size_t)batch * fields_per_instance(c) * c->tab->nslot * c->n * sizeof(double2));
}
// Workspace of the standalone transform (one field per instance; the large
// grids use it, and their trace characteristics reuse it afterwards).
inline int ensure_fft_ws(gnlse_ctx* c, int64_t batch) {
  return ensure(c->ws, (size_t)batch * c->n * sizeof(double2));
}

// Upload an array indexed by frequency (FFT order) with `comp` doubles per
// element, in the kernels' spectral layout: FFT order, or on the large grids
// (tab->blocks = R > 1) block layout, position k2 (n/R) + k1 holding index
// R k1 + k2.
inline int upload_spec(gnlse_ctx* c, DevBuf& b, const double* src, int comp) {
  const int64_t n = c->n, R = c->tab->blocks;
  if (R <= 1) return upload(c, b, src, (size_t)n * comp * sizeof(double));
  const int64_t M = n / R;
  std::vector<double> perm((size_t)n * comp);
  for (int64_t k2 = 0; k2 < R; ++k2)
    for (int64_t k1 = 0; k1 < M; ++k1)
      for (int q = 0; q < comp; ++q) perm[(size_t)(k2 * M + k1) * comp + q] = src[(size_t)(R * k1 + k2) * comp + q];
  return upload(c, b, perm.data(), perm.size() * sizeof(double));
}

