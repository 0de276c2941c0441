// C ABI of the GNLSE engine (include/gnlse.h): contexts, operator upload,
// device buffers, and dispatch to the per-grid launch tables.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"

namespace gn {
const LaunchTable* table_256();
const LaunchTable* table_512();
const LaunchTable* table_1024();
const LaunchTable* table_2048();
const LaunchTable* table_4096();
const LaunchTable* table_8192();
const LaunchTable* table_16384();
const LaunchTable* table_32768();
const LaunchTable* table_65536();

const LaunchTable* find_table(long long n) {
  const LaunchTable* all[] = {table_256(),  table_512(),   table_1024(),  table_2048(), table_4096(),
                               table_8192(), table_16384(), table_32768(), table_65536()};
  for (const LaunchTable* t : all)
    if (t->n == n) return t;
  return nullptr;
}

// Alternative schedules of the spectral-domain adaptive RK4IP kernels
// (sched_<N>.hip); schedule 0 of a grid is its launch table's own.
#define GN_SCHEDULES(X) \
  X(4096, 1) X(4096, 2) X(4096, 3) X(8192, 1) X(8192, 2) X(8192, 3) X(8192, 4) X(16384, 1) X(16384, 2)
#define GN_DECLARE_SCHEDULE(NN, ID) const Schedule* schedule_##NN##_##ID();
GN_SCHEDULES(GN_DECLARE_SCHEDULE)
#undef GN_DECLARE_SCHEDULE

struct ScheduleEntry {
  long long n;
  const Schedule* (*get)();
};
static std::vector<const Schedule*> grid_schedules(long long n) {
#define GN_SCHEDULE_ENTRY(NN, ID) {NN, &schedule_##NN##_##ID},
  static const ScheduleEntry all[] = {GN_SCHEDULES(GN_SCHEDULE_ENTRY)};
#undef GN_SCHEDULE_ENTRY
  std::vector<const Schedule*> v;
  for (const ScheduleEntry& e : all)
    if (e.n == n) v.push_back(e.get());
  return v;
}

// "T512 N:32L.16.16 R:8.8.8.8 tile4": threads, the pass radices of the
// N-point and N/2-point transforms (L: lane-pair pass) and the stage block.
static std::string schedule_desc(long long n, int t, int cfg) {
  const int renum = ((cfg & 8) || (cfg & 0x80)) ? 16 : 0;
  auto plan = [&](long long len, int pl) {
    const int p = (int)(len / t);
    const Plan pn = make_plan((int)len, p, pl);
    std::string s;
    for (int i = 0; i < pn.npass; ++i) {
      if (i) s += ".";
      s += std::to_string(pn.radix[i]);
      if (pn.radix[i] == 2 * p && (pl & 8)) s += "L";
    }
    return s;
  };
  const int tb = (cfg >> 8) & 3;
  const int tile = tb == 1 ? 2 : tb == 2 ? 8 : 4;
  const bool sp = cfg & kCfgSpUnroll, ra = cfg & kCfgRamanUnroll;
  return "T" + std::to_string(t) + " N:" + plan(n, (cfg & 15) | renum) + " R:" + plan(n / 2, ((cfg >> 4) & 15) | renum) +
         " tile" + std::to_string(tile) + (sp && ra ? " unrolled" : sp ? " passes unrolled" : ra ? " raman unrolled" : "");
}
}  // namespace gn

using namespace gn;

thread_local std::string g_err;

// Workspace of gnlse_propagate_adaptive_device_noalloc: a library-owned
// device array, one copy per GPU, held until the process ends. In host code
// the name `scratch` is not a device address; gnlse_ctx_create looks up the
// address of the current device's copy once (noalloc_scratch) and every
// launch takes that pointer.
static __device__ __align__(256) unsigned char scratch[GNLSE_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // per device

// Address of this device's `scratch` (the device must be current).
static int noalloc_scratch(int device, void** ptr) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GNLSE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return GNLSE_OK;
}

static int timed_begin(gnlse_ctx* c) {
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  return GNLSE_OK;
}
static int timed_end(gnlse_ctx* c, hipError_t launch) {
  if (launch != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(launch));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  return GNLSE_OK;
}

static int io_prepare(gnlse_ctx* c, int64_t batch, const double* in, DevBuf& dst) {
  if (batch <= 0) return fail(GNLSE_ERR_ARG, "batch must be positive");
  if (batch > 0x7fffffff) return fail(GNLSE_ERR_ARG, "batch too large");
  if (!in) return fail(GNLSE_ERR_ARG, "null input");
  const size_t bytes = (size_t)batch * c->n * sizeof(double2);
  RET_IF(ensure(dst, bytes));
  HIPCHK(hipMemcpyAsync(dst.p, in, bytes, hipMemcpyHostToDevice, c->stream));
  return GNLSE_OK;
}
static int io_finish(gnlse_ctx* c, int64_t batch, double* out) {
  if (!out) return fail(GNLSE_ERR_ARG, "null output");
  const size_t bytes = (size_t)batch * c->n * sizeof(double2);
  HIPCHK(hipMemcpyAsync(out, c->io_out.p, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

static bool method_ok(int m) {
  return m == GNLSE_METHOD_RK4IP || m == GNLSE_METHOD_SS_SYMMETRIC || m == GNLSE_METHOD_SS_SYM_RK4;
}

// Down-sampled powers of traces (one workgroup row per trace, a thread per
// output point). A: fields; S: their pyofss spectra fft(A) in FFT order.
// PYOFSS: field.get_downsampled (field.py:300-303), i.e. scipy's interp1d
//   at x_i = i (N-1)/(n_out-1) (numpy linspace), lo = clip(searchsorted(x),
//   1, N-1) - 1, y = (y[lo+1] - y[lo]) (x - lo) + y[lo], of temporal_power
//   and of spectral_power (natural order), storage.py:304-350.
// OPENCL: cl_power, cl_fftshift and cl_interpolate of opencl_fibre.py:199-216,
//   272-281: t = i/n_out in single precision, j = (int)(t (N-1)),
//   y = mix(y[j], y[j+1], t (N-1) - j).
// The products and sums are rounded separately (no contraction), as in numpy.
constexpr int kPowT = 256;

__device__ __forceinline__ double pow_at(const double2* a, long long k) {
  const double2 v = a[k];
  return __dadd_rn(__dmul_rn(v.x, v.x), __dmul_rn(v.y, v.y));
}

__global__ void __launch_bounds__(kPowT) k_trace_powers(const double2* __restrict__ A, const double2* __restrict__ S,
                                                        long long n, long long n_out, int mode,
                                                        double* __restrict__ out_t, double* __restrict__ out_s) {
  const long long b = blockIdx.x;
  const long long i = (long long)blockIdx.y * kPowT + threadIdx.x;
  if (i >= n_out) return;
  const double2* a = A + b * n;
  const double2* s = S + b * n;
  const long long half = n / 2;
  long long lo;
  double u;
  if (mode == GNLSE_DOWNSAMPLE_PYOFSS) {
    const double x = (i == n_out - 1) ? (double)(n - 1) : __dmul_rn((double)i, (double)(n - 1) / (double)(n_out - 1));
    long long hi = (long long)ceil(x);
    hi = hi < 1 ? 1 : (hi > n - 1 ? n - 1 : hi);
    lo = hi - 1;
    u = __dsub_rn(x, (double)lo);
  } else {
    const float t = (float)i / (float)n_out;
    const double tt = __dmul_rn((double)t, (double)(n - 1));
    lo = (long long)tt;
    u = __dsub_rn(tt, (double)lo);
  }
  const double t0 = pow_at(a, lo), t1 = pow_at(a, lo + 1);
  // natural (shifted) order: index k holds the FFT-order bin (k + N/2) mod N
  const double s0 = pow_at(s, (lo + half) % n), s1 = pow_at(s, (lo + 1 + half) % n);
  out_t[b * n_out + i] = __dadd_rn(__dmul_rn(__dsub_rn(t1, t0), u), t0);
  out_s[b * n_out + i] = __dadd_rn(__dmul_rn(__dsub_rn(s1, s0), u), s0);
}

extern "C" {

const char* gnlse_last_error(void) { return g_err.c_str(); }
int gnlse_version(void) { return 1; }
int gnlse_grid_supported(int64_t n) { return find_table(n) ? 1 : 0; }

int gnlse_device_count(int* count) {
  if (!count) return fail(GNLSE_ERR_ARG, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return fail(GNLSE_ERR_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  *count = n;
  return GNLSE_OK;
}

int gnlse_ctx_create(int device, gnlse_ctx** out) {
  if (!out) return fail(GNLSE_ERR_ARG, "null out");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return fail(GNLSE_ERR_HIP, std::string("no HIP device available: ") + hipGetErrorString(e));
  if (device < 0 || device >= n) return fail(GNLSE_ERR_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  void* sp = nullptr;
  RET_IF(noalloc_scratch(device, &sp));
  gnlse_ctx* c = new gnlse_ctx();
  c->device = device;
  c->scratch = sp;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return fail(GNLSE_ERR_HIP, "stream/event creation failed");
  }
  *out = c;
  return GNLSE_OK;
}

int gnlse_ctx_destroy(gnlse_ctx* c) {
  if (!c) return GNLSE_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  DevBuf* bufs[] = {&c->lin,  &c->efix,  &c->nlmul, &c->hr,   &c->omega, &c->gprof, &c->simpw, &c->tw,
                    &c->ws,   &c->io_in, &c->io_out, &c->io_b, &c->tr_A, &c->tr_z,  &c->st_z,  &c->st_h,
                    &c->lg_h, &c->lg_d,  &c->lg_a,  &c->rp_h, &c->rp_a,  &c->res,   &c->ntr,
                    &c->ch_w, &c->ch_out, &c->ch_npk, &c->ch_pk, &c->yb_rs, &c->yb_ip, &c->yb_as, &c->yb_es,
                    &c->yb_pp, &c->lin2, &c->efix2, &c->rows, &c->pair_flags};
  for (DevBuf* b : bufs) release(*b);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GNLSE_OK;
}

int gnlse_set_grid(gnlse_ctx* c, int64_t n) {
  if (!c) return fail(GNLSE_ERR_ARG, "null context");
  const LaunchTable* t = find_table(n);
  if (!t) return fail(GNLSE_ERR_UNSUPPORTED, "grid size " + std::to_string(n) + " has no kernel plan");
  HIPCHK(hipSetDevice(c->device));
  if (c->n != n) {
    c->n = n;
    c->tab = t;
    c->sched = 0;
    c->have_lin = c->have_nl = c->have_lin2 = false;
    c->efix_h = NAN;
    RET_IF(build_twiddles(c));
    std::vector<double> zeros(2 * n, 0.0);  // every operator pointer stays valid
    RET_IF(upload(c, c->nlmul, zeros.data(), 2 * n * sizeof(double)));
    RET_IF(upload(c, c->hr, zeros.data(), 2 * n * sizeof(double)));
    RET_IF(upload(c, c->omega, zeros.data(), n * sizeof(double)));
    RET_IF(upload(c, c->gprof, zeros.data(), n * sizeof(double)));
    RET_IF(upload(c, c->simpw, zeros.data(), n * sizeof(double)));
    RET_IF(ensure(c->efix, 2 * n * sizeof(double)));
    c->ops.gain = GAIN_NONE;
    c->ops.wdm = 0;
    refresh_ops(c);
  }
  return GNLSE_OK;
}

int gnlse_set_linear(gnlse_ctx* c, const double* factor) {
  RET_IF(check_ready(c, false, false));
  if (!factor) return fail(GNLSE_ERR_ARG, "null factor");
  HIPCHK(hipSetDevice(c->device));
  RET_IF(upload_spec(c, c->lin, factor, 2));
  c->have_lin = true;
  c->have_lin2 = false;
  c->efix_h = NAN;
  refresh_ops(c);
  return GNLSE_OK;
}

int gnlse_set_nonlinear(gnlse_ctx* c, int kind, double gamma, double f_R, double ss, double rs, const double* omega,
                        const double* nl_mult, const double* h_R) {
  RET_IF(check_ready(c, false, false));
  if (kind < GNLSE_NL_KERR || kind > GNLSE_NL_SS_RS) return fail(GNLSE_ERR_ARG, "unknown nonlinearity kind");
  if (kind != GNLSE_NL_KERR && !omega) return fail(GNLSE_ERR_ARG, "omega is required for this nonlinearity");
  if (kind == GNLSE_NL_ALL && (!nl_mult || !h_R))
    return fail(GNLSE_ERR_ARG, "nl_mult and h_R are required for NL_ALL");
  HIPCHK(hipSetDevice(c->device));
  const size_t n = (size_t)c->n;
  if (omega) RET_IF(upload_spec(c, c->omega, omega, 1));
  if (nl_mult) RET_IF(upload_spec(c, c->nlmul, nl_mult, 2));
  if (h_R) RET_IF(upload_spec(c, c->hr, h_R, 2));
  c->ops.nl_kind = kind;
  c->ops.wdm = 0;  // single-channel fields again
  c->ops.fac = make_double2(0.0, gamma);
  c->ops.f_R = f_R;
  c->ops.ss = ss;
  c->ops.rs = rs;
  c->have_nl = true;
  refresh_ops(c);
  return GNLSE_OK;
}

int gnlse_set_gain(gnlse_ctx* c, int enabled, double gain_log, double length, double E_sat, const double* profile,
                   const double* simpson_w) {
  RET_IF(check_ready(c, false, false));
  HIPCHK(hipSetDevice(c->device));
  if (!enabled) {
    c->ops.gain = GAIN_NONE;
    return GNLSE_OK;
  }
  if (!profile || !simpson_w) return fail(GNLSE_ERR_ARG, "profile and simpson_w are required");
  if (c->ops.wdm) return fail(GNLSE_ERR_UNSUPPORTED, "WDM fields carry no gain model");
  if (!(E_sat > 0.0) || !(length > 0.0)) return fail(GNLSE_ERR_ARG, "E_sat and length must be positive");
  const size_t n = (size_t)c->n;
  RET_IF(upload_spec(c, c->gprof, profile, 1));
  RET_IF(upload(c, c->simpw, simpson_w, n * sizeof(double)));
  c->ops.gain = GAIN_SATURABLE;
  c->ops.gain_log = gain_log;
  c->ops.gain_len2 = 2.0 * length;
  c->ops.esat = E_sat;
  // the old Simpson rule of an even number of points weighs every interior
  // point alike, up to the rounding of the grid spacing (1.1e-12 relative on
  // a 16384-point grid): then the spectral-domain stepper forms energies from
  // spectra (kernels.hpp, k_adaptive with SP = 2), with a relative error
  // below the tolerance accepted here
  c->ops.simp_flat = 0;
  if (n >= 8) {
    const double w = simpson_w[2];
    bool flat = w > 0.0;
    for (size_t i = 2; i + 2 < n && flat; ++i) flat = fabs(simpson_w[i] - w) <= 1e-10 * w;
    if (flat) {
      c->ops.simp_flat = 1;
      c->ops.simp_c = w;
      c->ops.simp_d[0] = simpson_w[0] - w;
      c->ops.simp_d[1] = simpson_w[1] - w;
      c->ops.simp_d[2] = simpson_w[n - 2] - w;
      c->ops.simp_d[3] = simpson_w[n - 1] - w;
    }
  }
  refresh_ops(c);
  return GNLSE_OK;
}

int gnlse_set_linear_wdm(gnlse_ctx* c, const double* factor0, const double* factor1) {
  RET_IF(check_ready(c, false, false));
  if (!factor0 || !factor1) return fail(GNLSE_ERR_ARG, "null factor");
  if (c->tab->blocks > 1) return fail(GNLSE_ERR_UNSUPPORTED, "WDM fields run on grids of up to 16384 samples");
  HIPCHK(hipSetDevice(c->device));
  RET_IF(upload(c, c->lin, factor0, 2 * (size_t)c->n * sizeof(double)));
  RET_IF(upload(c, c->lin2, factor1, 2 * (size_t)c->n * sizeof(double)));
  RET_IF(ensure(c->efix2, 2 * (size_t)c->n * sizeof(double)));
  c->have_lin = true;
  c->have_lin2 = true;
  c->efix_h = NAN;
  refresh_ops(c);
  return GNLSE_OK;
}

int gnlse_set_nonlinear_wdm(gnlse_ctx* c, double gamma0, double gamma1) {
  RET_IF(check_ready(c, false, false));
  if (c->tab->blocks > 1) return fail(GNLSE_ERR_UNSUPPORTED, "WDM fields run on grids of up to 16384 samples");
  if (c->ops.gain != GAIN_NONE) return fail(GNLSE_ERR_UNSUPPORTED, "WDM fields carry no gain model");
  c->ops.wdm = 1;
  c->ops.nl_kind = NL_KERR;
  c->ops.fac = make_double2(0.0, gamma0);
  c->ops.fac2 = make_double2(0.0, gamma1);
  c->have_nl = true;
  c->efix_h = NAN;
  return GNLSE_OK;
}

int gnlse_set_yb_gain(gnlse_ctx* c, const double* ratio_over_psat, const double* inv_psat, const double* alpha_s,
                      const double* eta_s, double power_factor, double ratio_p, double psat_p, double n_total,
                      double alpha_p, double eta_p) {
  RET_IF(check_ready(c, false, false));
  if (!ratio_over_psat || !inv_psat || !alpha_s || !eta_s) return fail(GNLSE_ERR_ARG, "null Yb array");
  if (c->ops.wdm) return fail(GNLSE_ERR_UNSUPPORTED, "WDM fields carry no gain model");
  if (!(psat_p > 0.0) || !std::isfinite(power_factor) || !std::isfinite(n_total))
    return fail(GNLSE_ERR_ARG, "invalid Yb constants");
  const size_t n = (size_t)c->n;
  for (const double* a : {ratio_over_psat, inv_psat, alpha_s, eta_s})
    for (size_t i = 0; i < n; ++i)
      if (!std::isfinite(a[i])) return fail(GNLSE_ERR_ARG, "Yb arrays must be finite (zero where a cross section vanishes)");
  HIPCHK(hipSetDevice(c->device));
  RET_IF(upload_spec(c, c->yb_rs, ratio_over_psat, 1));
  RET_IF(upload_spec(c, c->yb_ip, inv_psat, 1));
  RET_IF(upload_spec(c, c->yb_as, alpha_s, 1));
  RET_IF(upload_spec(c, c->yb_es, eta_s, 1));
  c->ops.gain = GAIN_YB;
  c->ops.yb_pfac = power_factor;
  c->ops.yb_rp = ratio_p;
  c->ops.yb_psp = psat_p;
  c->ops.yb_n = n_total;
  c->ops.yb_ap = alpha_p;
  c->ops.yb_ep = eta_p;
  refresh_ops(c);
  return GNLSE_OK;
}

int gnlse_set_pump(gnlse_ctx* c, int64_t count, const double* pump) {
  if (!c) return fail(GNLSE_ERR_ARG, "null context");
  if (count < 1 || !pump) return fail(GNLSE_ERR_ARG, "count must be positive and pump non-null");
  HIPCHK(hipSetDevice(c->device));
  RET_IF(upload(c, c->yb_pp, pump, (size_t)count * sizeof(double)));
  c->yb_count = count;
  refresh_ops(c);
  return GNLSE_OK;
}

int gnlse_get_pump(gnlse_ctx* c, int64_t count, double* pump) {
  if (!c) return fail(GNLSE_ERR_ARG, "null context");
  if (count < 1 || !pump) return fail(GNLSE_ERR_ARG, "count must be positive and pump non-null");
  if (count > c->yb_count) return fail(GNLSE_ERR_ARG, "count exceeds the pump powers set");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(pump, c->yb_pp.p, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

int gnlse_fft(gnlse_ctx* c, int sign, int64_t batch, const double* in, double* out) {
  RET_IF(check_ready(c, false, false));
  if (sign != 1 && sign != -1) return fail(GNLSE_ERR_ARG, "sign must be +1 or -1");
  HIPCHK(hipSetDevice(c->device));
  RET_IF(io_prepare(c, batch, in, c->io_in));
  RET_IF(ensure(c->io_out, (size_t)batch * c->n * sizeof(double2)));
  RET_IF(ensure_fft_ws(c, batch));
  RET_IF(timed_begin(c));
  RET_IF(timed_end(c, c->tab->fft(c->stream, c->ops, sign, batch, (const double2*)c->io_in.p, (double2*)c->io_out.p,
                                  (double2*)c->ws.p)));
  return io_finish(c, batch, out);
}

int gnlse_apply_linear(gnlse_ctx* c, double h, int64_t batch, const double* in, double* out) {
  RET_IF(check_ready(c, true, false));
  RET_IF(check_pump(c, batch));
  HIPCHK(hipSetDevice(c->device));
  const int64_t rows = batch * fields_per_instance(c);
  RET_IF(io_prepare(c, rows, in, c->io_in));
  RET_IF(ensure(c->io_out, (size_t)rows * c->n * sizeof(double2)));
  RET_IF(ensure_ws(c, batch));
  RET_IF(timed_begin(c));
  RET_IF(timed_end(c, c->tab->linear(c->stream, c->ops, h, batch, (const double2*)c->io_in.p, (double2*)c->io_out.p,
                                     (double2*)c->ws.p)));
  return io_finish(c, rows, out);
}

int gnlse_apply_linear_factor(gnlse_ctx* c, int64_t batch, const double* in, double* out) {
  RET_IF(check_ready(c, true, false));
  HIPCHK(hipSetDevice(c->device));
  const int64_t rows = batch * fields_per_instance(c);
  RET_IF(io_prepare(c, rows, in, c->io_in));
  RET_IF(ensure(c->io_out, (size_t)rows * c->n * sizeof(double2)));
  RET_IF(timed_begin(c));
  RET_IF(timed_end(c, c->tab->linfac(c->stream, c->ops, batch, (const double2*)c->io_in.p, (double2*)c->io_out.p)));
  return io_finish(c, rows, out);
}

int gnlse_apply_nonlinear(gnlse_ctx* c, int64_t batch, const double* in, double* out) {
  RET_IF(check_ready(c, false, true));
  HIPCHK(hipSetDevice(c->device));
  const int64_t rows = batch * fields_per_instance(c);
  RET_IF(io_prepare(c, rows, in, c->io_in));
  RET_IF(ensure(c->io_out, (size_t)rows * c->n * sizeof(double2)));
  RET_IF(ensure_ws(c, batch));
  RET_IF(timed_begin(c));
  RET_IF(timed_end(c, c->tab->nonlin(c->stream, c->ops, batch, (const double2*)c->io_in.p, (double2*)c->io_out.p,
                                     (double2*)c->ws.p)));
  return io_finish(c, rows, out);
}

int gnlse_apply_exp_nonlinear(gnlse_ctx* c, double h, int64_t batch, const double* A, const double* B, double* out) {
  RET_IF(check_ready(c, false, true));
  HIPCHK(hipSetDevice(c->device));
  const int64_t rows = batch * fields_per_instance(c);
  RET_IF(io_prepare(c, rows, A, c->io_in));
  RET_IF(io_prepare(c, rows, B, c->io_b));
  RET_IF(ensure(c->io_out, (size_t)rows * c->n * sizeof(double2)));
  RET_IF(ensure_ws(c, batch));
  RET_IF(timed_begin(c));
  RET_IF(timed_end(c, c->tab->exp_nonlin(c->stream, c->ops, h, batch, (const double2*)c->io_in.p,
                                         (const double2*)c->io_b.p, (double2*)c->io_out.p, (double2*)c->ws.p)));
  return io_finish(c, rows, out);
}

int gnlse_step(gnlse_ctx* c, int method, double h, double h_exp, int64_t batch, const double* in, double* out) {
  RET_IF(check_ready(c, true, true));
  RET_IF(check_pump(c, batch));
  if (!method_ok(method)) return fail(GNLSE_ERR_ARG, "unknown method");
  HIPCHK(hipSetDevice(c->device));
  const int64_t rows = batch * fields_per_instance(c);
  RET_IF(io_prepare(c, rows, in, c->io_in));
  RET_IF(ensure(c->io_out, (size_t)rows * c->n * sizeof(double2)));
  RET_IF(ensure_ws(c, batch));
  RET_IF(timed_begin(c));
  RET_IF(timed_end(c, c->tab->step(c->stream, c->ops, method, h, h_exp > 0.0 ? h_exp : 0.5 * h, batch, (const double2*)c->io_in.p,
                                   (double2*)c->io_out.p, (double2*)c->ws.p)));
  return io_finish(c, rows, out);
}

int gnlse_propagate_fixed(gnlse_ctx* c, int method, int64_t batch, double h, double h_exp, int64_t nsteps,
                          int64_t trace_every, int64_t max_traces, const double* A_in, double* A_out, double* traces,
                          int64_t* n_traces) {
  RET_IF(check_ready(c, true, true));
  if (!method_ok(method)) return fail(GNLSE_ERR_ARG, "unknown method");
  if (nsteps < 0 || trace_every < 0 || max_traces < 0) return fail(GNLSE_ERR_ARG, "negative count");
  if (max_traces > 0 && (!traces || !n_traces)) return fail(GNLSE_ERR_ARG, "null trace outputs");
  RET_IF(check_pump(c, batch));
  HIPCHK(hipSetDevice(c->device));
  const int64_t n = c->n * fields_per_instance(c);  // field length of an instance
  // linearity.py:220-224, 258-262: the cached factor exp(factor*h_exp), where
  // h_exp is the step of the first linear call (h/2 for both integrators)
  const double hh = h_exp > 0.0 ? h_exp : 0.5 * h;
  if (!(c->efix_h == hh)) {
    hipError_t e = c->tab->make_exp(c->stream, c->ops, hh, (double2*)c->efix.p);
    if (e == hipSuccess && c->ops.wdm) {  // channel 1's factor (linearity.py:220-231 wdm_cache)
      Ops o1 = c->ops;
      o1.lin = c->ops.lin2;
      e = c->tab->make_exp(c->stream, o1, hh, (double2*)c->efix2.p);
    }
    if (e != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("make_exp: ") + hipGetErrorString(e));
    c->efix_h = hh;
  }
  RET_IF(io_prepare(c, batch * fields_per_instance(c), A_in, c->io_in));
  RET_IF(ensure(c->io_out, (size_t)batch * n * sizeof(double2)));
  RET_IF(ensure_ws(c, batch));
  RET_IF(ensure(c->tr_A, (size_t)batch * max_traces * n * sizeof(double2)));
  RET_IF(ensure(c->ntr, (size_t)batch * sizeof(int64_t)));
  FixedParams prm;
  prm.h = h;
  prm.nsteps = nsteps;
  prm.trace_every = trace_every;
  prm.max_traces = max_traces;
  prm.tr_A = (double2*)c->tr_A.p;
  prm.n_traces = (long long*)c->ntr.p;
  RET_IF(timed_begin(c));
  RET_IF(timed_end(c, c->tab->fixed(c->stream, c->ops, method, prm, batch, (const double2*)c->io_in.p,
                                    (double2*)c->io_out.p, (double2*)c->ws.p)));
  if (max_traces > 0) {
    HIPCHK(hipMemcpyAsync(traces, c->tr_A.p, (size_t)batch * max_traces * n * sizeof(double2), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(n_traces, c->ntr.p, (size_t)batch * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  }
  return io_finish(c, batch * fields_per_instance(c), A_out);
}

static int adaptive_params(int method, const gnlse_adaptive_params* p, AdaptiveParams& prm) {
  if (!p) return fail(GNLSE_ERR_ARG, "null params");
  if (!method_ok(method)) return fail(GNLSE_ERR_ARG, "unknown method");
  if (!(p->length > 0.0) || !(p->h0 > 0.0) || p->traces < 1 || p->eta < 2 || p->total_attempts < 1)
    return fail(GNLSE_ERR_ARG, "invalid adaptive parameters");
  std::memset(&prm, 0, sizeof(prm));
  prm.length = p->length;
  prm.h0 = p->h0;
  prm.tol = p->local_error;
  prm.safety = p->safety;
  prm.max_factor = p->max_factor;
  prm.min_factor = p->min_factor;
  prm.hmin = p->step_size_min;
  prm.inv_eta = 1.0 / (double)p->eta;
  const double f_eta = std::pow(2.0, (double)p->eta - 1.0);  // stepper.py:272-274
  prm.f_alpha = f_eta / (f_eta - 1.0);
  prm.f_beta = 1.0 / (f_eta - 1.0);
  prm.step_storage = p->length / (double)p->traces;
  prm.traces = p->traces;
  prm.total_attempts = p->total_attempts;
  prm.steps_max = p->steps_max;
  return GNLSE_OK;
}

// The context's schedule for this launch (null: the grid's table, for
// schedule 0 and for launches no schedule covers).
static const Schedule* active_schedule(gnlse_ctx* c, int method) {
  if (c->sched > 0 && schedule_applies(c->ops, method)) {
    const std::vector<const Schedule*> v = grid_schedules(c->n);
    if (c->sched <= (int)v.size()) return v[c->sched - 1];
  }
  return nullptr;
}

// The one-workgroup adaptive launch through the context's schedule or the
// grid's table; no allocation (the no-allocation entry point uses it).
static hipError_t adaptive_enqueue(gnlse_ctx* c, int method, int64_t batch, const AdaptiveParams& prm,
                                   const double2* in, double2* out, double2* ws) {
  const Schedule* sc = active_schedule(c, method);
  if (sc) return sc->adaptive(c->stream, c->ops, prm, batch, in, out, ws);
  return c->tab->adaptive(c->stream, c->ops, method, prm, batch, in, out, ws);
}

static int launch_adaptive(gnlse_ctx* c, int method, int64_t batch, const AdaptiveParams& prm, const double2* in,
                           double2* out) {
  RET_IF(check_pump(c, batch));
  // Two workgroups per instance (the grid's default schedule, pairing on)
  // when the operators have that kernel and every workgroup of the batch
  // fits on the device at once; else one workgroup per instance. The
  // 16384-point Kerr kernels share the device's memory bandwidth worst:
  // pairs there gain only while they fill at most three quarters of the CUs
  // (measured: 0.90x the time at 16 instances, 0.99x at 96, 1.06x at 128).
  bool pair = c->pairing && c->sched == 0 && c->tab->adaptive_pair && !c->ops.wdm;
  if (pair && c->ops.nl_kind == NL_KERR && c->n >= 16384) {
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    pair = 8 * batch <= 3 * (int64_t)cus;
  }
  if (pair) {
    RET_IF(ensure(c->ws, (size_t)2 * batch * kPairSlots * c->n * sizeof(double2)));
    RET_IF(ensure(c->pair_flags, (size_t)2 * batch * sizeof(unsigned)));
    HIPCHK(hipMemsetAsync(c->pair_flags.p, 0, (size_t)2 * batch * sizeof(unsigned), c->stream));
    RET_IF(timed_begin(c));
    const hipError_t e = c->tab->adaptive_pair(c->stream, c->ops, method, prm, batch, in, out, (double2*)c->ws.p,
                                               (unsigned*)c->pair_flags.p);
    if (e != hipErrorNotSupported && e != hipErrorCooperativeLaunchTooLarge) return timed_end(c, e);
  }
  RET_IF(ensure_ws(c, batch));
  RET_IF(timed_begin(c));
  return timed_end(c, adaptive_enqueue(c, method, batch, prm, in, out, (double2*)c->ws.p));
}

int gnlse_propagate_adaptive(gnlse_ctx* c, int method, int64_t batch, const gnlse_adaptive_params* p,
                             const double* replay_h, const uint8_t* replay_accept, int64_t replay_len,
                             const double* A_in, double* A_out, int64_t max_traces, double* trace_A,
                             double* trace_z, double* step_z, double* step_h, int64_t log_cap, double* log_h,
                             double* log_delta, uint8_t* log_accept, gnlse_instance_result* results) {
  RET_IF(check_ready(c, true, true));
  AdaptiveParams prm;
  RET_IF(adaptive_params(method, p, prm));
  if (!results) return fail(GNLSE_ERR_ARG, "null results");
  if (max_traces < 0 || log_cap < 0 || replay_len < 0) return fail(GNLSE_ERR_ARG, "negative count");
  if (max_traces > 0 && (!trace_A || !trace_z || !step_z || !step_h))
    return fail(GNLSE_ERR_ARG, "null trace outputs");
  if (log_cap > 0 && (!log_h || !log_delta || !log_accept)) return fail(GNLSE_ERR_ARG, "null log outputs");
  if (replay_h && (!replay_accept || replay_len <= 0)) return fail(GNLSE_ERR_ARG, "invalid replay");
  if (replay_h && c->ops.wdm) return fail(GNLSE_ERR_UNSUPPORTED, "step replay is not available for WDM fields");
  HIPCHK(hipSetDevice(c->device));
  const int64_t n = c->n * fields_per_instance(c);  // field length of an instance
  RET_IF(io_prepare(c, batch * fields_per_instance(c), A_in, c->io_in));
  RET_IF(ensure(c->io_out, (size_t)batch * n * sizeof(double2)));
  RET_IF(ensure(c->tr_A, (size_t)batch * max_traces * n * sizeof(double2)));
  RET_IF(ensure(c->tr_z, (size_t)batch * max_traces * sizeof(double)));
  RET_IF(ensure(c->st_z, (size_t)batch * max_traces * sizeof(double)));
  RET_IF(ensure(c->st_h, (size_t)batch * max_traces * sizeof(double)));
  RET_IF(ensure(c->lg_h, (size_t)batch * log_cap * sizeof(double)));
  RET_IF(ensure(c->lg_d, (size_t)batch * log_cap * sizeof(double)));
  RET_IF(ensure(c->lg_a, (size_t)batch * log_cap));
  RET_IF(ensure(c->res, (size_t)batch * sizeof(gnlse_instance_result)));
  prm.max_traces = max_traces;
  prm.log_cap = log_cap;
  prm.tr_A = (double2*)c->tr_A.p;
  prm.tr_z = (double*)c->tr_z.p;
  prm.st_z = (double*)c->st_z.p;
  prm.st_h = (double*)c->st_h.p;
  prm.log_h = (double*)c->lg_h.p;
  prm.log_delta = (double*)c->lg_d.p;
  prm.log_acc = (unsigned char*)c->lg_a.p;
  prm.res = (gnlse_instance_result*)c->res.p;
  if (replay_h) {
    RET_IF(upload(c, c->rp_h, replay_h, (size_t)replay_len * sizeof(double)));
    RET_IF(upload(c, c->rp_a, replay_accept, (size_t)replay_len));
    prm.replay_h = (const double*)c->rp_h.p;
    prm.replay_accept = (const unsigned char*)c->rp_a.p;
    prm.replay_len = replay_len;
  }
  RET_IF(launch_adaptive(c, method, batch, prm, (const double2*)c->io_in.p, (double2*)c->io_out.p));
  auto d2h = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes == 0) return GNLSE_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    return GNLSE_OK;
  };
  if (max_traces > 0) {
    RET_IF(d2h(trace_A, c->tr_A.p, (size_t)batch * max_traces * n * sizeof(double2)));
    RET_IF(d2h(trace_z, c->tr_z.p, (size_t)batch * max_traces * sizeof(double)));
    RET_IF(d2h(step_z, c->st_z.p, (size_t)batch * max_traces * sizeof(double)));
    RET_IF(d2h(step_h, c->st_h.p, (size_t)batch * max_traces * sizeof(double)));
  }
  if (log_cap > 0) {
    RET_IF(d2h(log_h, c->lg_h.p, (size_t)batch * log_cap * sizeof(double)));
    RET_IF(d2h(log_delta, c->lg_d.p, (size_t)batch * log_cap * sizeof(double)));
    RET_IF(d2h(log_accept, c->lg_a.p, (size_t)batch * log_cap));
  }
  RET_IF(d2h(results, c->res.p, (size_t)batch * sizeof(gnlse_instance_result)));
  return io_finish(c, batch * fields_per_instance(c), A_out);
}

int gnlse_malloc(gnlse_ctx* c, int64_t bytes, void** dptr) {
  if (!c || !dptr || bytes <= 0) return fail(GNLSE_ERR_ARG, "invalid malloc request");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMalloc(dptr, (size_t)bytes));
  return GNLSE_OK;
}
int gnlse_free(gnlse_ctx* c, void* dptr) {
  if (!c) return fail(GNLSE_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->device));
  if (dptr) HIPCHK(hipFree(dptr));
  return GNLSE_OK;
}
int gnlse_memcpy_htod(gnlse_ctx* c, void* dst, const void* src, int64_t bytes) {
  if (!c || !dst || !src || bytes < 0) return fail(GNLSE_ERR_ARG, "invalid copy");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}
int gnlse_memcpy_dtoh(gnlse_ctx* c, void* dst, const void* src, int64_t bytes) {
  if (!c || !dst || !src || bytes < 0) return fail(GNLSE_ERR_ARG, "invalid copy");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

int gnlse_memcpy_dtod(gnlse_ctx* c, void* dst, const void* src, int64_t bytes) {
  if (!c || !dst || !src || bytes < 0) return fail(GNLSE_ERR_ARG, "invalid copy arguments");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

// Row r of the copy: n2 double2 from src + srow[r] n2 to dst + drow[r] n2.
__global__ void __launch_bounds__(256) k_copy_rows(const double2* __restrict__ src, double2* __restrict__ dst,
                                                  const int64_t* __restrict__ rows, long long count, long long n2) {
  const long long r = blockIdx.x;
  if (r >= count) return;
  const double2* s = src + rows[r] * n2;
  double2* d = dst + rows[count + r] * n2;
  for (long long i = threadIdx.x; i < n2; i += 256) d[i] = s[i];
}

int gnlse_copy_rows(gnlse_ctx* c, int64_t count, int64_t row_bytes, const int64_t* src_rows, const int64_t* dst_rows,
                    const void* d_src, void* d_dst) {
  if (!c || count < 1 || count > 0x7fffffff || row_bytes < 16 || row_bytes % 16 || !src_rows || !dst_rows || !d_src ||
      !d_dst)
    return fail(GNLSE_ERR_ARG, "invalid row copy arguments");
  for (int64_t r = 0; r < count; ++r)
    if (src_rows[r] < 0 || dst_rows[r] < 0) return fail(GNLSE_ERR_ARG, "negative row index");
  HIPCHK(hipSetDevice(c->device));
  std::vector<int64_t> rows(2 * count);
  std::memcpy(rows.data(), src_rows, count * sizeof(int64_t));
  std::memcpy(rows.data() + count, dst_rows, count * sizeof(int64_t));
  RET_IF(upload(c, c->rows, rows.data(), rows.size() * sizeof(int64_t)));
  hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)count), dim3(256), 0, c->stream, (const double2*)d_src, (double2*)d_dst,
                     (const int64_t*)c->rows.p, (long long)count, (long long)(row_bytes / 16));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

// The propagation kernels take input and output as __restrict__ pointers
// (each workgroup reads its instance once and writes it once, after all of
// its steps): the two device ranges of a batch must not overlap.
static int check_disjoint(const gnlse_ctx* c, int64_t batch, const double* in, const double* out) {
  const uintptr_t bytes = (uintptr_t)batch * fields_per_instance(c) * c->n * sizeof(double2);
  const uintptr_t a = (uintptr_t)in, b = (uintptr_t)out;
  if (a < b + bytes && b < a + bytes)
    return fail(GNLSE_ERR_ARG, "d_A_in and d_A_out overlap; give the output its own device buffer");
  return GNLSE_OK;
}

int gnlse_propagate_adaptive_device(gnlse_ctx* c, int method, int64_t batch, const gnlse_adaptive_params* p,
                                    const double* d_A_in, double* d_A_out, gnlse_instance_result* results) {
  RET_IF(check_ready(c, true, true));
  AdaptiveParams prm;
  RET_IF(adaptive_params(method, p, prm));
  if (batch <= 0 || batch > 0x7fffffff || !d_A_in || !d_A_out || !results)
    return fail(GNLSE_ERR_ARG, "invalid arguments");
  RET_IF(check_disjoint(c, batch, d_A_in, d_A_out));
  HIPCHK(hipSetDevice(c->device));
  RET_IF(ensure(c->res, (size_t)batch * sizeof(gnlse_instance_result)));
  prm.res = (gnlse_instance_result*)c->res.p;
  RET_IF(launch_adaptive(c, method, batch, prm, (const double2*)d_A_in, (double2*)d_A_out));
  HIPCHK(hipMemcpyAsync(results, c->res.p, (size_t)batch * sizeof(gnlse_instance_result), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

int gnlse_propagate_adaptive_device_noalloc(gnlse_ctx* c, int method, int64_t batch, const gnlse_adaptive_params* p,
                                            const double* d_A_in, double* d_A_out, gnlse_instance_result* d_results) {
  RET_IF(check_ready(c, true, true));
  AdaptiveParams prm;
  RET_IF(adaptive_params(method, p, prm));
  if (batch <= 0 || batch > 0x7fffffff || !d_A_in || !d_A_out || !d_results)
    return fail(GNLSE_ERR_ARG, "invalid arguments");
  RET_IF(check_disjoint(c, batch, d_A_in, d_A_out));
  RET_IF(check_pump(c, batch));
  const size_t per = (size_t)fields_per_instance(c) * c->tab->nslot * c->n * sizeof(double2);
  if ((size_t)batch * per > (size_t)GNLSE_NOALLOC_SCRATCH_BYTES)
    return fail(GNLSE_ERR_ARG, "batch " + std::to_string(batch) + " needs " + std::to_string((size_t)batch * per) +
                                   " bytes of workspace; the no-allocation entry point holds " +
                                   std::to_string((size_t)GNLSE_NOALLOC_SCRATCH_BYTES) + " (at most " +
                                   std::to_string((size_t)GNLSE_NOALLOC_SCRATCH_BYTES / per) +
                                   " instances on this grid)");
  if (!c->scratch) return fail(GNLSE_ERR_STATE, "no workspace address for this device");
  HIPCHK(hipSetDevice(c->device));
  prm.res = d_results;
  const hipError_t e = adaptive_enqueue(c, method, batch, prm, (const double2*)d_A_in, (double2*)d_A_out,
                                        (double2*)c->scratch);
  if (e != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return GNLSE_OK;
}

int gnlse_schedule_count(int64_t n, int* count) {
  if (!count) return fail(GNLSE_ERR_ARG, "null count");
  if (!find_table(n)) return fail(GNLSE_ERR_UNSUPPORTED, "grid size " + std::to_string(n) + " has no kernel plan");
  *count = 1 + (int)grid_schedules(n).size();
  return GNLSE_OK;
}

int gnlse_schedule_info(int64_t n, int index, int* threads, int* cfg, char* desc, int desc_len) {
  const LaunchTable* t = find_table(n);
  if (!t) return fail(GNLSE_ERR_UNSUPPORTED, "grid size " + std::to_string(n) + " has no kernel plan");
  const std::vector<const Schedule*> v = grid_schedules(n);
  if (index < 0 || index > (int)v.size()) return fail(GNLSE_ERR_ARG, "schedule index out of range");
  const int th = index == 0 ? t->t : v[index - 1]->t;
  const int cf = index == 0 ? t->cfg : v[index - 1]->cfg;
  if (threads) *threads = th;
  if (cfg) *cfg = cf;
  if (desc && desc_len > 0) {
    const std::string d = t->blocks > 1 ? "T" + std::to_string(th) + " two-level" : schedule_desc(n, th, cf);
    std::strncpy(desc, d.c_str(), (size_t)desc_len - 1);
    desc[desc_len - 1] = 0;
  }
  return GNLSE_OK;
}

int gnlse_set_schedule(gnlse_ctx* c, int index) {
  RET_IF(check_ready(c, false, false));
  if (index < 0 || index > (int)grid_schedules(c->n).size()) return fail(GNLSE_ERR_ARG, "schedule index out of range");
  c->sched = index;
  return GNLSE_OK;
}

int gnlse_set_pairing(gnlse_ctx* c, int enable) {
  if (!c) return fail(GNLSE_ERR_ARG, "null context");
  c->pairing = enable ? 1 : 0;
  return GNLSE_OK;
}

int gnlse_get_schedule(gnlse_ctx* c, int* index) {
  if (!c || !index) return fail(GNLSE_ERR_ARG, "null argument");
  *index = c->sched;
  return GNLSE_OK;
}

int gnlse_tune_adaptive(gnlse_ctx* c, int64_t batch, const gnlse_adaptive_params* p, const double* A_in,
                        int64_t trial_attempts, int repeats, int* chosen, double* ms) {
  RET_IF(check_ready(c, true, true));
  if (!chosen || trial_attempts < 1 || repeats < 2) return fail(GNLSE_ERR_ARG, "invalid tuning arguments");
  if (!schedule_applies(c->ops, GNLSE_METHOD_RK4IP))
    return fail(GNLSE_ERR_UNSUPPORTED, "no schedules for these operators (rk4ip with the full Raman + "
                                       "self-steepening operator, no gain or the saturable gain)");
  AdaptiveParams prm;
  RET_IF(adaptive_params(GNLSE_METHOD_RK4IP, p, prm));
  HIPCHK(hipSetDevice(c->device));
  RET_IF(io_prepare(c, batch, A_in, c->io_in));
  RET_IF(ensure(c->io_out, (size_t)batch * c->n * sizeof(double2)));
  RET_IF(ensure(c->res, (size_t)batch * sizeof(gnlse_instance_result)));
  // every schedule runs the same attempts: steps of h0, all accepted
  const std::vector<double> rh((size_t)trial_attempts, p->h0);
  const std::vector<unsigned char> ra((size_t)trial_attempts, 1);
  RET_IF(upload(c, c->rp_h, rh.data(), rh.size() * sizeof(double)));
  RET_IF(upload(c, c->rp_a, ra.data(), ra.size()));
  prm.replay_h = (const double*)c->rp_h.p;
  prm.replay_accept = (const unsigned char*)c->rp_a.p;
  prm.replay_len = trial_attempts;
  prm.res = (gnlse_instance_result*)c->res.p;
  const int keep = c->sched;
  const int count = 1 + (int)grid_schedules(c->n).size();
  int best = 0;
  double best_ms = 0.0;
  for (int i = 0; i < count; ++i) {
    c->sched = i;
    double sum = 0.0;
    for (int r = 0; r < repeats; ++r) {  // the first run loads the code object
      const int rc = launch_adaptive(c, GNLSE_METHOD_RK4IP, batch, prm, (const double2*)c->io_in.p,
                                     (double2*)c->io_out.p);
      if (rc != GNLSE_OK) {
        c->sched = keep;
        return rc;
      }
      if (r > 0) sum += c->last_ms;
    }
    const double t = sum / (repeats - 1);
    if (ms) ms[i] = t;
    if (i == 0 || t < best_ms) {
      best = i;
      best_ms = t;
    }
  }
  c->sched = best;
  *chosen = best;
  return GNLSE_OK;
}

int gnlse_trace_characteristics(gnlse_ctx* c, int64_t count, const double* A, const double* simpson_w, double* out,
                                int32_t* n_peaks, int32_t* peaks, int max_peaks) {
  RET_IF(check_ready(c, false, false));
  if (count < 1 || count > 0x7fffffff || !A || !simpson_w || !out || !n_peaks || (max_peaks > 0 && !peaks) ||
      max_peaks < 0)
    return fail(GNLSE_ERR_ARG, "invalid characteristics arguments");
  HIPCHK(hipSetDevice(c->device));
  const size_t fb = (size_t)count * c->n * sizeof(double2);
  RET_IF(io_prepare(c, count, A, c->io_in));
  RET_IF(ensure(c->io_b, fb));
  RET_IF(ensure(c->ch_out, (size_t)count * 4 * sizeof(double)));
  RET_IF(ensure(c->ch_npk, (size_t)count * sizeof(int)));
  RET_IF(ensure(c->ch_pk, (size_t)count * (max_peaks > 0 ? max_peaks : 1) * sizeof(int)));
  RET_IF(ensure(c->ch_w, (size_t)c->n * sizeof(double)));
  HIPCHK(hipMemcpyAsync(c->ch_w.p, simpson_w, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // pyofss spectra (field.fft), then the characteristics of every trace
  // (the large grids' transform workspace then holds their powers and flags)
  RET_IF(ensure_fft_ws(c, count));
  hipError_t e = c->tab->fft(c->stream, c->ops, +1, count, (const double2*)c->io_in.p, (double2*)c->io_b.p,
                             (double2*)c->ws.p);
  if (e == hipSuccess)
    e = c->tab->characts(c->stream, count, (const double2*)c->io_in.p, (const double2*)c->io_b.p,
                         (const double*)c->ch_w.p, (double*)c->ch_out.p, (int*)c->ch_npk.p, (int*)c->ch_pk.p,
                         max_peaks, (double2*)c->ws.p);
  if (e != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("characteristics: ") + hipGetErrorString(e));
  HIPCHK(hipMemcpyAsync(out, c->ch_out.p, (size_t)count * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(n_peaks, c->ch_npk.p, (size_t)count * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (max_peaks > 0)
    HIPCHK(hipMemcpyAsync(peaks, c->ch_pk.p, (size_t)count * max_peaks * sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

int gnlse_trace_powers(gnlse_ctx* c, int64_t count, const double* A, int64_t n_out, int mode, double* temporal,
                       double* spectral) {
  RET_IF(check_ready(c, false, false));
  if (count < 1 || count > 0x7fffffff || !A || !temporal || !spectral || n_out < 2 || n_out > (1 << 24) ||
      (mode != GNLSE_DOWNSAMPLE_PYOFSS && mode != GNLSE_DOWNSAMPLE_OPENCL))
    return fail(GNLSE_ERR_ARG, "invalid trace power arguments");
  HIPCHK(hipSetDevice(c->device));
  const size_t fb = (size_t)count * c->n * sizeof(double2);
  const size_t ob = (size_t)count * n_out * sizeof(double);
  RET_IF(io_prepare(c, count, A, c->io_in));
  RET_IF(ensure(c->io_b, fb));
  RET_IF(ensure(c->ch_out, 2 * ob));
  RET_IF(ensure_fft_ws(c, count));
  hipError_t e = c->tab->fft(c->stream, c->ops, +1, count, (const double2*)c->io_in.p, (double2*)c->io_b.p,
                             (double2*)c->ws.p);
  if (e == hipSuccess) {
    const dim3 grid((unsigned)count, (unsigned)((n_out + kPowT - 1) / kPowT));
    double* out = (double*)c->ch_out.p;
    hipLaunchKernelGGL(k_trace_powers, grid, dim3(kPowT), 0, c->stream, (const double2*)c->io_in.p,
                       (const double2*)c->io_b.p, (long long)c->n, (long long)n_out, mode, out,
                       out + (size_t)count * n_out);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("trace powers: ") + hipGetErrorString(e));
  HIPCHK(hipMemcpyAsync(temporal, c->ch_out.p, ob, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(spectral, (const char*)c->ch_out.p + ob, ob, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

int gnlse_last_kernel_ms(gnlse_ctx* c, double* ms) {
  if (!c || !ms) return fail(GNLSE_ERR_ARG, "null argument");
  *ms = c->last_ms;
  return GNLSE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Ensembles: realisations on up to four grids (ensemble.hpp). The groups that
// share a thread count run in one launch; a second thread count runs in a
// second launch beside it, on a second stream (joined before the coherence
// kernels), so each grid can keep the thread count the tuner finds fastest
// for it.
// ---------------------------------------------------------------------------
// One launch of a run: its groups (launch order) with block offsets within it.
struct EnsLaunch {
  const EnsembleTable* tab = nullptr;
  int ng = 0;
  int64_t blocks = 0;
  EnsembleGroup g[kEnsMaxGroups];
};

struct gnlse_ensemble {
  int device = 0, method = 0, kind = NL_KERR, ngroups = 0;
  bool gain = false;  // some group's operators carry a gain
  hipStream_t stream = nullptr, stream2 = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_fork = nullptr, ev_join = nullptr;
  double last_ms = 0.0;
  gnlse_adaptive_params params{};
  // caller's group order; slot[i] = position of caller group i in launch order
  int64_t n[kEnsMaxGroups] = {}, batch[kEnsMaxGroups] = {};
  int slot[kEnsMaxGroups] = {};
  size_t field_off[kEnsMaxGroups] = {}, g12_off[kEnsMaxGroups] = {}, res_off[kEnsMaxGroups] = {};
  int64_t blocks = 0;
  std::vector<EnsembleGroup> groups;  // launch order (largest grid first)
  DevBuf in, out, ws, spec, g12, res, dgroups, rp_h, rp_a;
  const EnsembleTable* tabs[kEnsMaxGroups] = {};  // thread count of each group (launch order)
  EnsLaunch launch[2];                            // the launches of a run; their tables in dgroups
  int nlaunch = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static const EnsembleTable* ensemble_for_threads(int t) {
  if (t == 512) return ensemble_table_512();
  if (t == 1024) return ensemble_table_1024();
  return nullptr;
}

static void drop_graph(gnlse_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

// Split the groups by thread count into at most two launches (group order
// kept within each) and upload their tables to d (kEnsMaxGroups entries per
// launch). tabs gives each group's table; only groups with sel set take part.
static int ens_plan(gnlse_ensemble* e, const std::vector<EnsembleGroup>& groups, const EnsembleTable* const* tabs,
                    const bool* sel, DevBuf& d, EnsLaunch (&L)[2], int& nl) {
  nl = 0;
  for (int s = 0; s < e->ngroups; ++s) {
    if (!sel[s]) continue;
    int k = 0;
    while (k < nl && L[k].tab != tabs[s]) ++k;
    if (k == nl) {
      L[k] = EnsLaunch();
      L[k].tab = tabs[s];
      ++nl;
    }
    EnsembleGroup g = groups[s];
    g.first = L[k].blocks;
    L[k].g[L[k].ng++] = g;
    L[k].blocks += g.count;
  }
  RET_IF(ensure(d, 2 * kEnsMaxGroups * sizeof(EnsembleGroup)));
  for (int k = 0; k < nl; ++k)
    HIPCHK(hipMemcpyAsync((EnsembleGroup*)d.p + k * kEnsMaxGroups, L[k].g, L[k].ng * sizeof(EnsembleGroup),
                          hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return GNLSE_OK;
}

// The launches of a run from the groups' current thread counts.
static int ens_replan(gnlse_ensemble* e) {
  bool all[kEnsMaxGroups];
  for (int s = 0; s < kEnsMaxGroups; ++s) all[s] = true;
  return ens_plan(e, e->groups, e->tabs, all, e->dgroups, e->launch, e->nlaunch);
}

// The launch sequence of one ensemble run: the propagation launches (a
// second one forked onto stream2 and joined back), then one coherence
// reduction per group of two or more realisations.
static hipError_t ensemble_enqueue(gnlse_ensemble* e, const EnsLaunch* L, int nl, const EnsembleGroup* dbase) {
  hipError_t r = hipSuccess;
  if (nl > 1) {
    r = hipEventRecord(e->ev_fork, e->stream);
    if (r == hipSuccess) r = hipStreamWaitEvent(e->stream2, e->ev_fork, 0);
  }
  for (int k = 0; k < nl && r == hipSuccess; ++k)
    r = L[k].tab->run(k == 0 ? e->stream : e->stream2, e->method, e->kind, e->gain, dbase + k * kEnsMaxGroups, L[k].ng,
                      L[k].blocks);
  if (nl > 1 && r == hipSuccess) {
    r = hipEventRecord(e->ev_join, e->stream2);
    if (r == hipSuccess) r = hipStreamWaitEvent(e->stream, e->ev_join, 0);
  }
  if (r != hipSuccess) return r;
  for (int i = 0; i < e->ngroups; ++i) {
    const EnsembleGroup& g = e->groups[i];
    if (g.count < 2) continue;
    const unsigned nb = (unsigned)((g.n + 255) / 256);
    hipLaunchKernelGGL((k_coherence<256>), dim3(nb), dim3(256), 0, e->stream, (const double2*)g.spec, g.count, g.n,
                       (double*)((double*)e->g12.p + e->g12_off[i]));
    r = hipGetLastError();
    if (r != hipSuccess) return r;
  }
  return hipSuccess;
}

static int ensemble_time(gnlse_ensemble* e, const EnsLaunch* L, int nl, const EnsembleGroup* dbase, bool graph) {
  HIPCHK(hipEventRecord(e->ev0, e->stream));
  if (graph) HIPCHK(hipGraphLaunch(e->exec, e->stream));
  else {
    hipError_t r = ensemble_enqueue(e, L, nl, dbase);
    if (r != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("ensemble launch: ") + hipGetErrorString(r));
  }
  HIPCHK(hipEventRecord(e->ev1, e->stream));
  HIPCHK(hipEventSynchronize(e->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
  e->last_ms = ms;
  return GNLSE_OK;
}

extern "C" {

int gnlse_ensemble_destroy(gnlse_ensemble* e) {
  if (!e) return GNLSE_OK;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graph(e);
  DevBuf* bufs[] = {&e->in, &e->out, &e->ws, &e->spec, &e->g12, &e->res, &e->dgroups, &e->rp_h, &e->rp_a};
  for (DevBuf* b : bufs) release(*b);
  if (e->stream2) (void)hipStreamSynchronize(e->stream2);
  hipEvent_t evs[] = {e->ev0, e->ev1, e->ev_fork, e->ev_join};
  for (hipEvent_t v : evs)
    if (v) (void)hipEventDestroy(v);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  if (e->stream2) (void)hipStreamDestroy(e->stream2);
  delete e;
  return GNLSE_OK;
}

int gnlse_ensemble_create(int method, const gnlse_adaptive_params* p, int ngroups, gnlse_ctx* const* ctxs,
                          const int64_t* batch, gnlse_ensemble** out) {
  if (!out) return fail(GNLSE_ERR_ARG, "null out");
  *out = nullptr;
  if (ngroups < 1 || ngroups > kEnsMaxGroups || !ctxs || !batch)
    return fail(GNLSE_ERR_ARG, "an ensemble has 1 to 4 groups");
  AdaptiveParams prm;
  RET_IF(adaptive_params(method, p, prm));
  if (method == GNLSE_METHOD_SS_SYM_RK4)
    return fail(GNLSE_ERR_UNSUPPORTED, "ensembles run rk4ip or ss_symmetric");
  for (int i = 0; i < ngroups; ++i) {
    RET_IF(check_ready(ctxs[i], true, true));
    const int64_t n = ctxs[i]->n;
    if (n != 4096 && n != 8192 && n != 16384)
      return fail(GNLSE_ERR_UNSUPPORTED, "ensemble grids are 4096, 8192 or 16384 samples, got " + std::to_string(n));
    if (ctxs[i]->device != ctxs[0]->device) return fail(GNLSE_ERR_ARG, "ensemble groups must share one device");
    if (batch[i] < 1 || batch[i] > (1 << 24)) return fail(GNLSE_ERR_ARG, "group batch out of range");
    if (ctxs[i]->ops.wdm) return fail(GNLSE_ERR_UNSUPPORTED, "ensembles hold single-channel fields");
    if (ctxs[i]->ops.gain == GAIN_YB)
      return fail(GNLSE_ERR_UNSUPPORTED, "ensembles do not carry the Yb pump state; propagate such groups directly");
  }
  gnlse_ensemble* e = new gnlse_ensemble();
  e->device = ctxs[0]->device;
  e->method = method;
  e->ngroups = ngroups;
  e->params = *p;
  if (hipSetDevice(e->device) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming) != hipSuccess) {
    gnlse_ensemble_destroy(e);
    return fail(GNLSE_ERR_HIP, "stream/event creation failed");
  }
  // launch order: larger grids first (stable for equal sizes)
  std::vector<int> order(ngroups);
  for (int i = 0; i < ngroups; ++i) order[i] = i;
  for (int a = 1; a < ngroups; ++a)
    for (int b = a; b > 0 && ctxs[order[b]]->n > ctxs[order[b - 1]]->n; --b) std::swap(order[b], order[b - 1]);
  size_t fields = 0, g12s = 0, ress = 0;
  bool all_kerr = true, all_sp = method == GNLSE_METHOD_RK4IP;
  for (int s = 0; s < ngroups; ++s) {
    const int i = order[s];
    e->slot[i] = s;
    e->n[i] = ctxs[i]->n;
    e->batch[i] = batch[i];
    e->field_off[s] = fields;
    e->g12_off[s] = g12s;
    e->res_off[s] = ress;
    fields += (size_t)batch[i] * ctxs[i]->n;
    g12s += (size_t)ctxs[i]->n;
    ress += (size_t)batch[i];
    all_kerr = all_kerr && ctxs[i]->ops.nl_kind == NL_KERR;
    all_sp = all_sp && ctxs[i]->ops.nl_kind == NL_ALL && ctxs[i]->ops.gain == GAIN_NONE && !ctxs[i]->ops.wdm;
    e->gain = e->gain || ctxs[i]->ops.gain != GAIN_NONE;
  }
  // the spectral-domain RK4IP when every group has the full operator without gain
  e->kind = all_kerr ? NL_KERR : all_sp ? NL_ALL : NL_GENERIC;
  int rc = GNLSE_OK;
  auto alloc = [&](DevBuf& b, size_t bytes) {
    if (rc == GNLSE_OK) rc = ensure(b, bytes);
  };
  alloc(e->in, fields * sizeof(double2));
  alloc(e->out, fields * sizeof(double2));
  alloc(e->spec, fields * sizeof(double2));
  alloc(e->ws, fields * NSLOT * sizeof(double2));
  alloc(e->g12, g12s * sizeof(double));
  alloc(e->res, ress * sizeof(gnlse_instance_result));
  alloc(e->dgroups, 2 * kEnsMaxGroups * sizeof(EnsembleGroup));
  if (rc != GNLSE_OK) {
    gnlse_ensemble_destroy(e);
    return rc;
  }
  if (hipMemsetAsync(e->in.p, 0, e->in.bytes, e->stream) != hipSuccess ||
      hipMemsetAsync(e->g12.p, 0, e->g12.bytes, e->stream) != hipSuccess) {
    gnlse_ensemble_destroy(e);
    return fail(GNLSE_ERR_HIP, "ensemble buffer initialisation failed");
  }
  e->groups.resize(ngroups);
  int64_t first = 0;
  for (int s = 0; s < ngroups; ++s) {
    const int i = order[s];
    gnlse_ctx* c = ctxs[i];
    refresh_ops(c);
    EnsembleGroup& g = e->groups[s];
    std::memset(&g, 0, sizeof(g));
    g.ops = c->ops;
    g.prm = prm;
    g.prm.res = (gnlse_instance_result*)e->res.p + e->res_off[s];
    g.in = (const double2*)e->in.p + e->field_off[s];
    g.out = (double2*)e->out.p + e->field_off[s];
    g.spec = (double2*)e->spec.p + e->field_off[s];
    g.ws = (double2*)e->ws.p + e->field_off[s] * NSLOT;
    g.n = c->n;
    g.first = first;
    g.count = batch[i];
    first += batch[i];
  }
  e->blocks = first;
  for (int s = 0; s < ngroups; ++s) e->tabs[s] = ensemble_table_1024();
  if ((rc = ens_replan(e)) != GNLSE_OK) {
    gnlse_ensemble_destroy(e);
    return rc;
  }
  *out = e;
  return GNLSE_OK;
}

int gnlse_ensemble_tune(gnlse_ensemble* e, int threads, int trial_attempts, int* chosen) {
  if (!e) return fail(GNLSE_ERR_ARG, "null ensemble");
  HIPCHK(hipSetDevice(e->device));
  const EnsembleTable* pick[kEnsMaxGroups];
  if (threads != 0) {
    const EnsembleTable* t = ensemble_for_threads(threads);
    if (!t) return fail(GNLSE_ERR_ARG, "threads per realisation must be 0 (tune), 512 or 1024");
    for (int s = 0; s < e->ngroups; ++s) pick[s] = t;
  } else {
    if (trial_attempts < 1 || trial_attempts > 64) return fail(GNLSE_ERR_ARG, "trial_attempts must be 1..64");
    // Trial: each group alone, at each compiled thread count, replays
    // trial_attempts accepted steps of h0 in every realisation (a fixed
    // amount of work per candidate); outputs are overwritten by the next run.
    std::vector<double> rh(trial_attempts, e->params.h0);
    std::vector<unsigned char> ra(trial_attempts, 1);
    RET_IF(upload_raw(e->device, e->stream, e->rp_h, rh.data(), rh.size() * sizeof(double)));
    RET_IF(upload_raw(e->device, e->stream, e->rp_a, ra.data(), ra.size()));
    std::vector<EnsembleGroup> trial = e->groups;
    for (EnsembleGroup& g : trial) {
      g.prm.replay_h = (const double*)e->rp_h.p;
      g.prm.replay_accept = (const unsigned char*)e->rp_a.p;
      g.prm.replay_len = trial_attempts;
    }
    const EnsembleTable* cands[] = {ensemble_table_512(), ensemble_table_1024()};
    DevBuf tg;
    int rc = GNLSE_OK;
    double best_ms[kEnsMaxGroups] = {};
    double total = 0.0;
    for (int s = 0; s < e->ngroups && rc == GNLSE_OK; ++s) {
      pick[s] = nullptr;
      bool sel[kEnsMaxGroups] = {};
      sel[s] = true;
      for (const EnsembleTable* t : cands) {
        const EnsembleTable* tabs[kEnsMaxGroups];
        for (int q = 0; q < kEnsMaxGroups; ++q) tabs[q] = t;
        EnsLaunch L[2];
        int nl = 0;
        rc = ens_plan(e, trial, tabs, sel, tg, L, nl);
        double ms = 0.0;
        for (int rep = 0; rep < 2 && rc == GNLSE_OK; ++rep) {  // the first run loads the code object
          rc = ensemble_time(e, L, nl, (const EnsembleGroup*)tg.p, false);
          ms = e->last_ms;
        }
        if (rc != GNLSE_OK) break;
        if (!pick[s] || ms < best_ms[s]) {
          pick[s] = t;
          best_ms[s] = ms;
        }
      }
      total += best_ms[s];
    }
    release(tg);
    if (rc != GNLSE_OK) return rc;
    e->last_ms = total;
  }
  bool same = true;
  for (int s = 0; s < e->ngroups; ++s) same = same && pick[s] == e->tabs[s];
  if (!same) {
    drop_graph(e);
    for (int s = 0; s < e->ngroups; ++s) e->tabs[s] = pick[s];
    RET_IF(ens_replan(e));
  }
  if (chosen) *chosen = e->tabs[0]->t;
  return GNLSE_OK;
}

int gnlse_ensemble_threads(gnlse_ensemble* e, int* threads) {
  if (!e || !threads) return fail(GNLSE_ERR_ARG, "null argument");
  *threads = e->tabs[0]->t;
  return GNLSE_OK;
}

int gnlse_ensemble_set_group_threads(gnlse_ensemble* e, int group, int threads) {
  if (!e || group < 0 || group >= e->ngroups) return fail(GNLSE_ERR_ARG, "bad ensemble or group");
  const EnsembleTable* t = ensemble_for_threads(threads);
  if (!t) return fail(GNLSE_ERR_ARG, "threads per realisation must be 512 or 1024");
  HIPCHK(hipSetDevice(e->device));
  const int s = e->slot[group];
  if (e->tabs[s] == t) return GNLSE_OK;
  drop_graph(e);
  e->tabs[s] = t;
  return ens_replan(e);
}

int gnlse_ensemble_group_threads(gnlse_ensemble* e, int group, int* threads) {
  if (!e || !threads || group < 0 || group >= e->ngroups) return fail(GNLSE_ERR_ARG, "bad ensemble, group or output");
  *threads = e->tabs[e->slot[group]]->t;
  return GNLSE_OK;
}

int gnlse_ensemble_device_ptrs(gnlse_ensemble* e, int group, double** d_in, double** d_out, double** d_g12) {
  if (!e || group < 0 || group >= e->ngroups) return fail(GNLSE_ERR_ARG, "bad ensemble or group");
  const int s = e->slot[group];
  if (d_in) *d_in = (double*)((double2*)e->in.p + e->field_off[s]);
  if (d_out) *d_out = (double*)((double2*)e->out.p + e->field_off[s]);
  if (d_g12) *d_g12 = (double*)e->g12.p + e->g12_off[s];
  return GNLSE_OK;
}

int gnlse_ensemble_set_input(gnlse_ensemble* e, int group, const double* A_in) {
  if (!e || !A_in || group < 0 || group >= e->ngroups) return fail(GNLSE_ERR_ARG, "bad ensemble, group or input");
  HIPCHK(hipSetDevice(e->device));
  const int s = e->slot[group];
  HIPCHK(hipMemcpyAsync((double2*)e->in.p + e->field_off[s], A_in, (size_t)e->batch[group] * e->n[group] * sizeof(double2),
                        hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return GNLSE_OK;
}

int gnlse_ensemble_run(gnlse_ensemble* e, int use_graph) {
  if (!e) return fail(GNLSE_ERR_ARG, "null ensemble");
  HIPCHK(hipSetDevice(e->device));
  if (use_graph && !e->exec) {
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    hipError_t r = ensemble_enqueue(e, e->launch, e->nlaunch, (const EnsembleGroup*)e->dgroups.p);
    hipGraph_t g = nullptr;
    hipError_t r2 = hipStreamEndCapture(e->stream, &g);
    if (r != hipSuccess || r2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(GNLSE_ERR_HIP, std::string("ensemble capture: ") + hipGetErrorString(r != hipSuccess ? r : r2));
    }
    e->graph = g;
    if (hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0) != hipSuccess) {
      drop_graph(e);
      return fail(GNLSE_ERR_HIP, "ensemble graph instantiation failed");
    }
  }
  return ensemble_time(e, e->launch, e->nlaunch, (const EnsembleGroup*)e->dgroups.p, use_graph != 0);
}

int gnlse_ensemble_output(gnlse_ensemble* e, int group, double* A_out, double* g12, gnlse_instance_result* results) {
  if (!e || group < 0 || group >= e->ngroups) return fail(GNLSE_ERR_ARG, "bad ensemble or group");
  HIPCHK(hipSetDevice(e->device));
  const int s = e->slot[group];
  if (A_out)
    HIPCHK(hipMemcpyAsync(A_out, (double2*)e->out.p + e->field_off[s],
                          (size_t)e->batch[group] * e->n[group] * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
  if (g12)
    HIPCHK(hipMemcpyAsync(g12, (double*)e->g12.p + e->g12_off[s], (size_t)e->n[group] * sizeof(double),
                          hipMemcpyDeviceToHost, e->stream));
  if (results)
    HIPCHK(hipMemcpyAsync(results, (gnlse_instance_result*)e->res.p + e->res_off[s],
                          (size_t)e->batch[group] * sizeof(gnlse_instance_result), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return GNLSE_OK;
}

int gnlse_ensemble_last_ms(gnlse_ensemble* e, double* ms) {
  if (!e || !ms) return fail(GNLSE_ERR_ARG, "null argument");
  *ms = e->last_ms;
  return GNLSE_OK;
}

}  // extern "C"

