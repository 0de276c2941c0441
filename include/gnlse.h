/*
 * gnlse.h -- C ABI of the MI355X (gfx950) GNLSE propagation engine.
 *
 * The engine replaces pyofss's fibre stepping path (KarinaSaraeva/
 * pyofssChanged): Fibre.__call__ -> Stepper -> Solver.rk4ip/ss_symmetric ->
 * Linearity.exp_lin + Nonlinearity.non -> field.fft/ifft. pyofss is pure
 * Python and has no FFI of its own; the boundary is cut at
 * Stepper.__call__ (pyofss/modules/stepper.py:147-156) so that a whole
 * propagation runs on the device and only traces come back. Each entry below
 * names the reference function whose behaviour it reproduces.
 *
 * Conventions
 *   - complex arrays are interleaved (re, im) float64, C-contiguous, exactly
 *     numpy complex128; spectral arrays are in FFT order.
 *   - every call returns GNLSE_OK (0) or a negative error code and sets a
 *     thread-local message readable with gnlse_last_error().
 *   - "host" entry points take host pointers and copy through device buffers
 *     owned by the context; "_device" entry points take device pointers.
 *   - one context per thread at a time; calls on one context must be
 *     serialised by the caller.
 */
#ifndef GNLSE_H
#define GNLSE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gnlse_ctx gnlse_ctx;

enum {
  GNLSE_OK = 0,
  GNLSE_ERR_ARG = -1,
  GNLSE_ERR_HIP = -2,
  GNLSE_ERR_UNSUPPORTED = -3,
  GNLSE_ERR_STATE = -4
};

/* integrators, pyofss/modules/solver.py:339-354, :266-275 and :330-337
 * (ss_sym_rk4, also opencl_fibre.py:929-970 cl_ss_sym_rk4) */
enum { GNLSE_METHOD_RK4IP = 0, GNLSE_METHOD_SS_SYMMETRIC = 1, GNLSE_METHOD_SS_SYM_RK4 = 2 };

/* nonlinear operator variants, pyofss/modules/nonlinearity.py:131-152 */
enum {
  GNLSE_NL_KERR = 0,  /* default_f / default_exp_f                 :266-270 */
  GNLSE_NL_ALL = 1,   /* default_f_all / default_exp_f_all         :201-221 */
  GNLSE_NL_SS = 2,    /* default_f_with_ss / default_exp_f_with_ss :223-234 */
  GNLSE_NL_RS = 3,    /* default_f_with_rs / default_exp_f_with_rs :236-248 */
  GNLSE_NL_SS_RS = 4  /* default_f_with_ss_and_rs / exp variant    :250-264 */
};

/* down-sampling rules of gnlse_trace_powers */
enum {
  GNLSE_DOWNSAMPLE_PYOFSS = 0, /* field.get_downsampled, field.py:300-303 */
  GNLSE_DOWNSAMPLE_OPENCL = 1  /* cl_interpolate, opencl_fibre.py:207-216 */
};

/* per-instance termination status; map to pyofss/modules/stepper.py:33-46 */
enum {
  GNLSE_STATUS_OK = 0,
  GNLSE_STATUS_SMALL_STEP = 1,      /* SmallStepSizeError          stepper.py:350-351 */
  GNLSE_STATUS_SUITABLE_STEP = 2,   /* SuitableStepSizeError       stepper.py:353-354 */
  GNLSE_STATUS_MAX_STEPS = 3,       /* MaximumStepsAllocatedError  stepper.py:376 */
  GNLSE_STATUS_TRACE_OVERFLOW = 4,  /* trace buffer too small (engine-only) */
  GNLSE_STATUS_REPLAY_EXHAUSTED = 5, /* replay list shorter than the run (engine-only) */
  GNLSE_STATUS_PAIR_TIMEOUT = 6     /* two-workgroup schedule: partner never arrived (engine-only) */
};

/* Adaptive step-doubling controller constants, stepper.py:133-139, 256-274 */
typedef struct gnlse_adaptive_params {
  double length;          /* fibre length (km) */
  double h0;              /* initial step, stepper.py:256-264 */
  double local_error;     /* target relative local error */
  double safety;          /* 0.9 */
  double max_factor;      /* 10.0 */
  double min_factor;      /* 0.2 */
  double step_size_min;   /* 1e-37 */
  int64_t traces;         /* Stepper.traces (>= 1) */
  int64_t total_attempts; /* 1000 */
  int64_t steps_max;      /* 100000 */
  int32_t eta;            /* local error order of the method (5 rk4ip, 3 ss_symmetric) */
  int32_t reserved;
} gnlse_adaptive_params;

/* Per-instance results of an adaptive run. */
typedef struct gnlse_instance_result {
  double z;               /* final z */
  double h;               /* next step size proposed by the controller */
  int64_t n_attempts;     /* attempted steps (3 integrator calls each) */
  int64_t n_accepted;     /* accepted steps */
  int64_t n_traces;       /* traces written (including the initial one) */
  int32_t status;         /* GNLSE_STATUS_* */
  int32_t reserved;
} gnlse_instance_result;

const char* gnlse_last_error(void);
int gnlse_version(void);
/* 1 if grid size n has a compiled kernel plan */
int gnlse_grid_supported(int64_t n);
int gnlse_device_count(int* count);

int gnlse_ctx_create(int device, gnlse_ctx** out);
int gnlse_ctx_destroy(gnlse_ctx* ctx);

/* Grid size (samples). Replaces Domain.total_samples, pyofss/domain.py:202 */
int gnlse_set_grid(gnlse_ctx* ctx, int64_t n);

/* Linear factor i*sum_n beta_n Domega^n/n! - alpha/2 (complex128[n], FFT
 * order). Replaces Linearity.default_linearity, linearity.py:162-186 */
int gnlse_set_linear(gnlse_ctx* ctx, const double* factor);

/* Nonlinear operator. Replaces Nonlinearity.__call__/default_nonlinearity,
 * nonlinearity.py:160-192. omega: float64[n] (omega - omega_c, FFT order);
 * nl_mult: complex128[n] = i*gamma*(1 + omega*ss) (NL_ALL only); h_R:
 * complex128[n] = fft(h_R)*window_t (NL_ALL only). NULL where unused. */
int gnlse_set_nonlinear(gnlse_ctx* ctx, int kind, double gamma, double f_R, double ss, double rs,
                        const double* omega, const double* nl_mult, const double* h_R);

/* Two-channel WDM fields (pyofss sim_type="wdm"). Linear factors of the
 * two channels (complex128[n] each, FFT order): Linearity.wdm_linearity,
 * linearity.py:188-211. Nonlinear coefficients gamma_0, gamma_1 with the
 * cross-phase coupling N_c = i gamma_c (|A_c|^2 + 2 |A_o|^2) A_c:
 * Nonlinearity.wdm_nonlinearity / wdm_f / wdm_exp_f, nonlinearity.py:
 * 194-200, 287-304. After gnlse_set_nonlinear_wdm every field argument of the
 * operator, step and propagation entry points holds two channels per
 * instance, [batch][2][n] (traces [batch][max_traces][2][n]); the relative
 * local error of the adaptive stepper is taken over both channels (the norm
 * of the 2 x n array, stepper.py:230-245). gnlse_set_nonlinear returns to
 * single-channel fields. No gain model with WDM. */
int gnlse_set_linear_wdm(gnlse_ctx* ctx, const double* factor0, const double* factor1);
int gnlse_set_nonlinear_wdm(gnlse_ctx* ctx, double gamma0, double gamma1);

/* Saturable gain coupled into every linear sub-step. Replaces
 * Amplifier.factor (amplifier.py:149-158) as called from linearity.py:248-272.
 * gain_log = ln(10^(0.1*G_dB)); profile: float64[n] FFT order; simpson_w:
 * float64[n] weights with energy = sum_i w_i |A_i|^2 (field.py:161-172). */
int gnlse_set_gain(gnlse_ctx* ctx, int enabled, double gain_log, double length, double E_sat,
                   const double* profile, const double* simpson_w);

/* Two-level Yb rate-equation gain coupled into every linear sub-step.
 * Replaces Amplifier2LevelModel.factor (amplifier.py:368-402) as called from
 * linearity.py:248-272: for the array A passed to a linear call, with
 * Ps = |fft(A)|^2 * power_factor (power_factor = len(A) dt / Tr),
 *   N2 = N (ratio_p Pp/Psat_p + sum ratio_s Ps/Psat_s) / (1 + Pp/Psat_p + sum Ps/Psat_s),
 * the spectrum is multiplied by exp((alpha_s N2 - eta_s) h 1e3 / 2) and the
 * pump power becomes Pp exp((alpha_p N2 - eta_p) h 1e3) (amplifier.py:386-388).
 * Arrays float64[n], FFT order: ratio_s/Psat_s, 1/Psat_s, alpha_s, eta_s,
 * finite (0 where the cross sections vanish: the terms the reference drops as
 * NaN). gnlse_set_gain(ctx, 0, ...) switches any gain off. */
int gnlse_set_yb_gain(gnlse_ctx* ctx, const double* ratio_over_psat, const double* inv_psat, const double* alpha_s,
                      const double* eta_s, double power_factor, double ratio_p, double psat_p, double n_total,
                      double alpha_p, double eta_p);

/* Pump power Pp per instance (Amplifier2LevelModel._Pp, amplifier.py:201),
 * read by every launch with the Yb gain (instance b uses pump[b]; count must
 * cover the batch) and written back at its end. */
int gnlse_set_pump(gnlse_ctx* ctx, int64_t count, const double* pump);
int gnlse_get_pump(gnlse_ctx* ctx, int64_t count, double* pump);

/* sign=+1: pyofss field.fft (scipy ifft, 1/N); sign=-1: field.ifft
 * (scipy fft). Replaces field.py:109-136. */
int gnlse_fft(gnlse_ctx* ctx, int sign, int64_t batch, const double* in, double* out);

/* Linearity.exp_lin(A, h) = ifft(exp(factor*h)[*exp(gain)] * fft(A)),
 * linearity.py:241-256 */
int gnlse_apply_linear(gnlse_ctx* ctx, double h, int64_t batch, const double* in, double* out);

/* Linearity.lin(A, z) = ifft(factor * fft(A)), linearity.py:238-239 */
int gnlse_apply_linear_factor(gnlse_ctx* ctx, int64_t batch, const double* in, double* out);

/* Nonlinearity.non(A, z), nonlinearity.py:201-207, 223-228, 236-241, 250-256, 266-267 */
int gnlse_apply_nonlinear(gnlse_ctx* ctx, int64_t batch, const double* in, double* out);

/* Nonlinearity.exp_non(A, h, B), nonlinearity.py:209-221, 230-234, 243-248, 258-264, 269-270 */
int gnlse_apply_exp_nonlinear(gnlse_ctx* ctx, double h, int64_t batch, const double* A, const double* B,
                              double* out);

/* One integrator call Solver.<method>(A, z, h) (solver.py:266-275, 339-354)
 * with linear sub-steps exp(factor*h_exp) (linearity.py:241-272): h_exp <= 0
 * means h/2 (recomputed, as default_exp_f); a Linearity that caches passes
 * the step of its first call (default_exp_f_cached). */
int gnlse_step(gnlse_ctx* ctx, int method, double h, double h_exp, int64_t batch, const double* in, double* out);

/* Fixed-step propagation, Stepper.standard_stepper (stepper.py:164-222) with
 * the cached linear factor exp(factor*h_exp) of linearity.py:220-224/258-272
 * (h_exp <= 0 means h/2, the step of the first linear call). Traces: after
 * step i when trace_every > 0 and i % trace_every == 0, as stepper.py:206-209;
 * traces is complex128[batch][max_traces][n]. n_traces: int64[batch]. */
int gnlse_propagate_fixed(gnlse_ctx* ctx, int method, int64_t batch, double h, double h_exp, int64_t nsteps,
                          int64_t trace_every, int64_t max_traces, const double* A_in, double* A_out,
                          double* traces, int64_t* n_traces);

/* Adaptive propagation, Stepper.adaptive_stepper (stepper.py:247-376) with
 * per-instance state. Optional replay (replay_h != NULL): attempt k uses step
 * replay_h[k] and acceptance replay_accept[k] instead of the controller's.
 * Traces (max_traces > 0): trace_A complex128[batch][max_traces][n], trace_z,
 * step_z, step_h float64[batch][max_traces] (step_* valid where a trace was
 * stored after a step). Attempt log (log_cap > 0): log_h, log_delta
 * float64[batch][log_cap], log_accept uint8[batch][log_cap]. */
int gnlse_propagate_adaptive(gnlse_ctx* ctx, int method, int64_t batch, const gnlse_adaptive_params* p,
                             const double* replay_h, const uint8_t* replay_accept, int64_t replay_len,
                             const double* A_in, double* A_out, int64_t max_traces, double* trace_A,
                             double* trace_z, double* step_z, double* step_h, int64_t log_cap, double* log_h,
                             double* log_delta, uint8_t* log_accept, gnlse_instance_result* results);

/* Device-memory helpers and device-resident adaptive propagation (inputs
 * already in HBM; no traces, no log). results is a host array; d_A_in and
 * d_A_out must not overlap (GNLSE_ERR_ARG otherwise). */
int gnlse_malloc(gnlse_ctx* ctx, int64_t bytes, void** dptr);
int gnlse_free(gnlse_ctx* ctx, void* dptr);
int gnlse_memcpy_htod(gnlse_ctx* ctx, void* dst, const void* src, int64_t bytes);
int gnlse_memcpy_dtoh(gnlse_ctx* ctx, void* dst, const void* src, int64_t bytes);
int gnlse_memcpy_dtod(gnlse_ctx* ctx, void* dst, const void* src, int64_t bytes);
int gnlse_propagate_adaptive_device(gnlse_ctx* ctx, int method, int64_t batch, const gnlse_adaptive_params* p,
                                    const double* d_A_in, double* d_A_out, gnlse_instance_result* results);
/* Engine-only (no pyofss counterpart): gather/scatter of `count` device rows
 * of row_bytes (a multiple of 16) in one launch: row src_rows[r] of d_src to
 * row dst_rows[r] of d_dst (host index arrays); the surrogate's judge moves
 * flagged fields to and from the solver with it. */
int gnlse_copy_rows(gnlse_ctx* ctx, int64_t count, int64_t row_bytes, const int64_t* src_rows, const int64_t* dst_rows,
                    const void* d_src, void* d_dst);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the propagation of gnlse_propagate_adaptive_device with
 * the per-instance workspace in a library-owned device array of
 * GNLSE_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead
 * of context buffers, and the per-instance results written to device memory
 * d_results (batch entries). The call allocates nothing, enqueues one launch
 * on the context's stream and returns without waiting (gnlse_memcpy_dtoh on
 * the same context waits for it). Rejected with GNLSE_ERR_ARG when batch x
 * 176 n bytes (192 n on the 32768- and 65536-sample grids; twice that for
 * WDM fields) exceeds the array (45 instances at n = 8192, 10 at n = 32768)
 * and when d_A_in and d_A_out overlap. Every call on one GPU uses that GPU's one array: calls that
 * overlap on a GPU (different contexts or threads) race. The operator
 * set-up calls (gnlse_set_grid, _linear, _nonlinear, _gain, _pump) still
 * allocate and must come first. */
#define GNLSE_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)
int gnlse_propagate_adaptive_device_noalloc(gnlse_ctx* ctx, int method, int64_t batch,
                                            const gnlse_adaptive_params* p, const double* d_A_in, double* d_A_out,
                                            gnlse_instance_result* d_results);

/* Trace characteristics, Storage.update_characts (pyofss/modules/storage.py:
 * 194-209) for `count` fields A (host, complex128, time domain) in one
 * launch. out[4k..4k+3] = energy (sum of simpson_w * |A|^2, simpson_w the
 * old-Simpson weights of t * 1e-3, field.py:161-172), maximum power, width
 * of the highest temporal peak at half prominence in samples
 * (get_duration, field.py:207-229, 283-287) and spectral width in samples
 * (get_duration_spec, field.py:232-280; NaN where no spectral peak has
 * prominence >= max/100). n_peaks[k] and peaks[k * max_peaks ...] = the
 * temporal peaks with prominence >= max/10 in increasing order (get_peaks,
 * field.py:272-274; at most max_peaks stored, n_peaks counts all). */
int gnlse_trace_characteristics(gnlse_ctx* ctx, int64_t count, const double* A, const double* simpson_w, double* out,
                                int32_t* n_peaks, int32_t* peaks, int max_peaks);

/* Down-sampled temporal and spectral powers of `count` fields A (host,
 * complex128, time domain), n_out points each, into temporal[count][n_out]
 * and spectral[count][n_out] (spectral power in natural order, as
 * field.spectral_power). mode GNLSE_DOWNSAMPLE_PYOFSS: get_downsampled of
 * temporal_power / spectral_power as Storage.get_plot_data does
 * (storage.py:304-350); GNLSE_DOWNSAMPLE_OPENCL: cl_power, cl_fftshift and
 * cl_interpolate as OpenclFibre stores its traces (opencl_fibre.py:650-662). */
int gnlse_trace_powers(gnlse_ctx* ctx, int64_t count, const double* A, int64_t n_out, int mode, double* temporal,
                       double* spectral);

/* Duration in ms of the last propagation kernel, from HIP events recorded on
 * the context's stream around the launch. */
int gnlse_last_kernel_ms(gnlse_ctx* ctx, double* ms);

/* Schedules and their autotuner (engine-only; pyofss has one NumPy path).
 * The adaptive RK4IP launches with the full Raman + self-steepening operator
 * (no gain, or the saturable gain; Stepper.adaptive_stepper, stepper.py:
 * 247-376, over Solver.rk4ip, solver.py:339-354) are compiled in several
 * schedules per grid: threads per instance (waves per instance x 64), the
 * pass plan of the N-point transform and of the N/2-point transform of the
 * Raman convolution, and the block of the pointwise stage passes. Schedule 0
 * is the grid's default. All schedules compute the same arithmetic up to the
 * order of the transforms' roundings. */
/* Number of schedules of an n-sample grid (1 where no alternative is
 * compiled); no GPU needed. */
int gnlse_schedule_count(int64_t n, int* count);
/* threads per instance, CFG code and a description ("T512 N:32L.16.16
 * R:8.8.8.8 tile4"; radices of the passes, L = lane-pair pass) of schedule
 * `index` of an n-sample grid; no GPU needed. */
int gnlse_schedule_info(int64_t n, int index, int* threads, int* cfg, char* desc, int desc_len);
/* Select the schedule of the context's launches (0 = default) / read it. */
int gnlse_set_schedule(gnlse_ctx* ctx, int index);
int gnlse_get_schedule(gnlse_ctx* ctx, int* index);
/* Two workgroups per instance (engine-only). With the default schedule and
 * pairing on (the default), an adaptive propagation whose batch leaves room
 * on the device (every workgroup resident at once, e.g. a single pyofss
 * Fibre run, or 64 instances on 256 CUs) runs each instance on two
 * workgroups: the coarse step beside the half and fine steps of every
 * attempt, one meeting per attempt; results are equal bit for bit to one
 * workgroup per instance. Grids of 4096-16384 samples; rk4ip with the full
 * Raman + self-steepening operator (no gain or the saturable gain), rk4ip
 * and ss_symmetric with the Kerr operator; single-channel fields.
 * enable = 0 keeps one workgroup per instance. */
int gnlse_set_pairing(gnlse_ctx* ctx, int enable);
/* Autotuner: times every schedule of the context's grid on trial_attempts
 * replayed steps of p->h0 (all accepted) of `batch` instances of the host
 * inputs A_in, `repeats` runs each (the first one untimed), selects the
 * fastest for the context and returns its index in *chosen and, when
 * ms is not NULL, each schedule's time (ms[count]). GNLSE_ERR_UNSUPPORTED
 * when the context's operators are not a scheduled configuration. */
int gnlse_tune_adaptive(gnlse_ctx* ctx, int64_t batch, const gnlse_adaptive_params* p, const double* A_in,
                        int64_t trial_attempts, int repeats, int* chosen, double* ms);

/* Noise ensembles on mixed grids (engine-only: pyofss runs one System per
 * realisation; this replaces that loop for the stepping path). Groups are up
 * to four sets of realisations, each on the grid (4096, 8192 or 16384) and
 * operators of one context; every realisation runs Stepper.adaptive_stepper
 * (stepper.py:247-376) with its own step sizes, all in one launch per thread
 * count (groups at a second thread count run in a second launch beside the
 * first, in one graph when replayed). After the
 * propagation the spectral coherence of each group of M >= 2 realisations,
 * g12 = | |sum a_i|^2 - sum |a_i|^2 | / ((M-1) sum |a_i|^2) with a_i the
 * pyofss spectrum fft(A_i) (field.py:109-121), is formed on the device.
 * The contexts must outlive the ensemble. */
typedef struct gnlse_ensemble gnlse_ensemble;
int gnlse_ensemble_create(int method, const gnlse_adaptive_params* p, int ngroups, gnlse_ctx* const* ctxs,
                          const int64_t* batch, gnlse_ensemble** out);
/* threads = 512 or 1024 selects the launch shape of every group; 0 times
 * each group alone at each candidate on trial_attempts (1..64) replayed steps
 * of h0 per realisation and keeps the fastest for that group. chosen (and
 * gnlse_ensemble_threads) report the thread count of the largest grid's
 * group; gnlse_ensemble_group_threads that of any group (caller order). */
int gnlse_ensemble_tune(gnlse_ensemble* e, int threads, int trial_attempts, int* chosen);
int gnlse_ensemble_threads(gnlse_ensemble* e, int* threads);
int gnlse_ensemble_group_threads(gnlse_ensemble* e, int group, int* threads);
/* sets one group's thread count (512 or 1024) by hand */
int gnlse_ensemble_set_group_threads(gnlse_ensemble* e, int group, int threads);
/* device pointers of a group's inputs (batch x N complex128), outputs and g12 */
int gnlse_ensemble_device_ptrs(gnlse_ensemble* e, int group, double** d_in, double** d_out, double** d_g12);
int gnlse_ensemble_set_input(gnlse_ensemble* e, int group, const double* A_in);
/* One run (inputs are not modified, so runs repeat). use_graph = 1 replays a
 * hipGraph of the launch sequence, captured at the first such call. */
int gnlse_ensemble_run(gnlse_ensemble* e, int use_graph);
/* host copies of a group's outputs; any pointer may be NULL */
int gnlse_ensemble_output(gnlse_ensemble* e, int group, double* A_out, double* g12, gnlse_instance_result* results);
int gnlse_ensemble_last_ms(gnlse_ensemble* e, double* ms);
int gnlse_ensemble_destroy(gnlse_ensemble* e);

/* Spectral neural operator surrogate of a fibre segment (engine-only; no
 * pyofss counterpart). A 1-D Fourier neural operator (lift to `width` = 64
 * channels, `layers` Fourier layers keeping `modes` modes, GELU, a projection
 * through `hidden` channels to 2 outputs added to the first two inputs)
 * trained and evaluated with bf16 matrix-core kernels and FP32 accumulation.
 * Inputs x are B x n x in_channels FP32, outputs y B x n x 2 FP32, both in
 * device memory. Parameters are one FP32 vector (gnlse_sno_param_count). */
typedef struct gnlse_sno gnlse_sno;
int gnlse_sno_create(int device, int64_t n, int width, int modes, int layers, int in_channels, int hidden,
                     int64_t max_batch, gnlse_sno** out);
int gnlse_sno_param_count(gnlse_sno* s, int64_t* count);
int gnlse_sno_set_params(gnlse_sno* s, const float* params);
int gnlse_sno_get_params(gnlse_sno* s, float* params);
int gnlse_sno_forward(gnlse_sno* s, int64_t batch, const float* d_x, float* d_y);
/* Forward, relative L2 loss mean_b |y_b - t_b|^2 / |t_b|^2 against targets
 * d_t, and the gradient of every parameter (kept on the device; copied to
 * `grad` when not NULL). rel (host, batch entries, may be NULL) receives
 * |y_b - t_b| / |t_b|. batch must be a multiple of 32. */
int gnlse_sno_loss_grad(gnlse_sno* s, int64_t batch, const float* d_x, const float* d_t, double* loss, double* rel,
                        float* grad);
/* One Adam step on the last gradient. */
int gnlse_sno_adam(gnlse_sno* s, float lr, float beta1, float beta2, float eps);
int gnlse_sno_last_ms(gnlse_sno* s, double* ms);
int gnlse_sno_destroy(gnlse_sno* s);
/* Field encoding on the device with the context c of the same grid.
 * Spectral mode (time_domain = 0): d_out (B x n x channels FP32) = the pyofss
 * spectrum fft(A) (field.py:109-121) in natural frequency order, times
 * d_phase (complex128, natural order, may be NULL), divided by scale.
 * Time mode (time_domain = 1): d_out = ifft(d_phase * fft(A)) / scale on the
 * time grid (d_phase in FFT order; NULL means A itself). Channel 2
 * (channels = 3) is the grid coordinate in [-1, 1). decode inverts either
 * mode for predictions d_y (B x n x 2) back to fields d_A (complex128, time
 * domain). */
int gnlse_sno_encode(gnlse_sno* s, gnlse_ctx* c, int64_t batch, const double* d_A, double scale, const double* d_phase,
                     int channels, int time_domain, float* d_out);
int gnlse_sno_decode(gnlse_sno* s, gnlse_ctx* c, int64_t batch, const float* d_y, double scale, const double* d_phase,
                     int time_domain, double* d_A);
/* Per-field relative change of energy between inputs d_x and predictions d_y
 * (the judge's invariant gate for a lossless segment), to host `defect`. */
int gnlse_sno_energy_defect(gnlse_sno* s, int64_t batch, const float* d_x, const float* d_y, double* defect);
/* Two networks' predictions d_y1, d_y2 (B x n x 2 FP32) of the same fields:
 * d_ym = their mean and dist[b] (host) = |y1_b - y2_b| / |ym_b|, the
 * disagreement the judge gates on. */
int gnlse_sno_combine(gnlse_sno* s, int64_t batch, const float* d_y1, const float* d_y2, float* d_ym, double* dist);
/* C (FP32) = op(A) op(B) on bf16 operands (device pointers), batched with
 * element strides sA, sB, sC and split over K; at/bt select stored-transposed
 * operands (not both). K / ksplit must be a multiple of 32. */
int gnlse_gemm_bf16(int device, int at, int bt, int64_t M, int64_t N, int64_t K, const void* d_A, int64_t lda,
                    const void* d_B, int64_t ldb, float* d_C, int64_t ldc, int64_t batch, int64_t sA, int64_t sB,
                    int64_t sC, int ksplit);

/* Self-test of the engine's workgroup reductions (no counterpart in the
 * reference): for `blocks` workgroups of `threads` threads (64, 128, 256, 512
 * or 1024), each thread's m values (2, 3, 4, 9 or 18) from
 * in[(b m + j) threads + t] are summed over the workgroup once as m sums at
 * a time (wg_sum_n, the routine of the gain stages and the error norm) into
 * out_n and once one by one (wg_sum) into out_1, same layout; the two must
 * agree bit for bit in every thread. Host arrays of blocks m threads doubles. */
int gnlse_check_reductions(int device, int threads, int m, int64_t blocks, const double* in, double* out_n,
                           double* out_1);

#ifdef __cplusplus
}
#endif
#endif /* GNLSE_H */
