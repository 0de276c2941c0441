#!/usr/bin/env python3
"""Benchmark of the GNLSE engine on its headline workload.

Metric: GNLSE propagations/s at a 2^13 grid (FP64) on one GPU. A "step" is
one batched launch that propagates B independent realisations of the
supercontinuum workload (8192 samples over 12.5 ps at 835 nm; 50 fs FWHM
sech at 10 kW with one-photon-per-mode input noise seeded per realisation;
15 cm of photonic crystal fibre, gamma = 110 /W/km, beta2..beta10,
self-steepening and Blow-Wood Raman; adaptive RK4IP with step doubling,
local error 1e-6) through the whole fibre. Inputs are resident in HBM when
the timed region starts; every instance runs the full adaptive propagation.

Printed: one JSON line with the metric, the roofline of the propagation
kernel (the reference's fft_counter work, 72 FFTs x 5 N log2 N flops per
attempted step, over the kernel's HIP-event duration, against the 78.6
TFLOP/s FP64 peak; beside it the measured HBM traffic of the same launch
against 8 TB/s and the fft_counter work as bytes), a CPU baseline (the NumPy
oracle on one host core running the whole propagation of realisation 0,
unscaled; beside it, "all_cores": a bounded sample in one single-threaded
process per core of the GPU's 16-core share, scaled) and "error", the
relative L2 field error of realisation 0 against that oracle run. Before
the timed launches the autotuner picks the headline kernel's schedule
(threads per instance, transform pass plans, stage block) from trial runs on
the same inputs; "autotuner" records the choice and every candidate's trial
time. Further objects report the other workloads of the engine:
"ensemble" (512 noise-seeded soliton realisations on 4096/8192/16384-sample
grids in one graph-replayed launch with |g12| on the device), "amplifier"
(64 chirped pulses through the 5 m saturable-gain amplifier, 16384 samples),
"yb_attractor" (the two-level Yb amplifier run whose timing pyofss publishes
in examples/notebooks/nonlinear_attractor.ipynb, once as the notebook runs it
and batched over pump powers) and "surrogate" (the spectral neural operator
trained here on solver-generated data, its throughput against the FP64
solver on the same fields, and its spectral error after the judge re-solves
gated and audited fields).

    python bench.py [--steps K] [--warmup W] [--batch B] [--no-cpu]
                    [--no-ensemble] [--no-amplifier] [--no-yb] [--no-surrogate]
                    [--surrogate-steps S] [--no-tune | --schedule I]
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "GNLSE propagations/s at 2^13 grid (FP64) on 1 GPU; surrogate speedup at 1% err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64
FFTS_PER_ATTEMPT = 72  # 3 RK4IP calls x (4 linear x 2 + 4 Raman/SS nonlinear x 4) FFTs
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "supercontinuum_8192_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=0, help="realisations per launch (0: one per CU)")
    ap.add_argument("--length", type=float, default=1.5e-4, help="fibre length in km (default: the full 15 cm)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-workers", type=int, default=16,
                    help="processes of the all-core CPU baseline (at most the host's cores; 1: single core only)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the noise-ensemble measurement")
    ap.add_argument("--no-amplifier", action="store_true", help="skip the amplifier-batch measurement")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate training and measurement")
    ap.add_argument("--no-yb", action="store_true", help="skip the Yb attractor measurement")
    ap.add_argument("--no-tune", action="store_true", help="keep the default schedules (no autotuner)")
    ap.add_argument("--schedule", type=int, default=None, help="headline schedule index (skips its tuning)")
    ap.add_argument("--surrogate-steps", type=int, default=60000, help="training steps of the surrogate")
    ap.add_argument("--surrogate-train", type=int, default=20000, help="solver runs in the training set")
    ap.add_argument("--surrogate-length", type=float, default=1e-3, help="surrogate fibre segment (km)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy kernel of known size: FETCH_SIZE/WRITE_SIZE of this
    dispatch calibrate the counters read on the propagation kernel."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}))


def cu_count():
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.get_device_properties(0).multi_processor_count
    except Exception:
        pass
    return 256


def _oracle_attempts(A0, length, n, want_field=False):
    """Time the oracle's first n attempted steps of the supercontinuum
    propagation on one core (n = None: the whole propagation): (seconds,
    attempts done) and, with want_field, the field reached."""
    from threadpoolctl import threadpool_limits

    import oracle as O
    from pyofsschanged_amd.workloads import DUDLEY_BETA, C_NM_PS

    with threadpool_limits(1):
        dom = O.Domain(1, 8192, 12.5, C_NM_PS / 835.0)
        f = O.Fibre(length=length, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True,
                    method="ark4ip", local_error=1e-6)
        f.linearity(dom)
        f.nonlinearity(dom)
        f.stepper.storage(dom)
        f.stepper.storage.reset_fft_counter()
        t = time.perf_counter()
        A = f.stepper.adaptive_stepper(A0.copy(), None, max_attempts=n)
        dt = time.perf_counter() - t
        if want_field:
            return dt, len(f.stepper.attempt_log), A
        return dt, len(f.stepper.attempt_log)


def _oracle_worker(args):
    return _oracle_attempts(*args)


def host_description():
    """CPU model (lscpu "Model name") and the host's logical CPU count."""
    model = None
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in txt.splitlines():
            if ln.strip().startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def cpu_baseline(A0, length, gpu_out0, gpu_attempts0, attempts_full, budget_s, pool_workers):
    """The oracle (NumPy restatement of pyofss) on one host core: the whole
    propagation of realisation 0, unscaled, whose output also gives that
    realisation's error (GPU vs oracle, relative L2 over the full fibre).
    With pool_workers > 1 also the all-core throughput: that many
    single-threaded worker processes each run the first attempted steps of
    the same propagation (a bounded sample of about budget_s seconds) side
    by side; their attempt rates add and are scaled by the attempts of a full
    propagation. Returns (cpu_baseline, error)."""
    elapsed, done, ref = _oracle_attempts(A0, length, None, want_field=True)
    err = float(np.linalg.norm(gpu_out0 - ref) / np.linalg.norm(ref))
    error = {"realisation": 0, "rel_l2_vs_oracle": err, "attempts": int(gpu_attempts0), "oracle_attempts": done,
             "bound": 1e-9,
             "note": "relative L2 field error of the GPU's realisation 0 against the oracle over the full 15 cm; "
                     "every realisation of the launch runs the same kernel"}
    out = {
        "value": 1.0 / elapsed,
        "unit": "propagations/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"NumPy oracle, the whole propagation of realisation 0 ({done} attempted steps) in "
                   f"{elapsed:.1f} s on one core, unscaled"),
        "host": host_description(),
    }
    if pool_workers > 1:
        t_probe, n_probe = _oracle_attempts(A0, length, 8)
        n = max(8, int(budget_s / (t_probe / max(n_probe, 1))))
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(pool_workers) as pool:
            res = pool.map(_oracle_worker, [(A0, length, n)] * pool_workers)
        rate = sum(d / t for t, d in res)  # attempted steps per second over all workers
        out["all_cores"] = {
            "value": rate / attempts_full,
            "unit": "propagations/s",
            "cores": pool_workers,
            "sample": (f"{pool_workers} single-threaded worker processes (the CPU share of one GPU on this host), "
                       f"each the first {n} attempted steps of realisation 0 "
                       f"({1e3 * pool_workers / rate:.2f} ms/attempt per worker on average), scaled to the "
                       f"{attempts_full:.0f} attempted steps of a full propagation"),
        }
    return out, error


def ensemble_measurement():
    """512 realisations (128 x 16384, 256 x 8192, 128 x 4096 samples) of the
    soliton fibre with one-photon-per-mode noise, adaptive RK4IP, one launch
    replayed from a hipGraph, |g12| per grid on the device."""
    from pyofsschanged_amd import workloads
    ens, setups = workloads.noise_ensemble()
    ens.tune(0, trial_attempts=32)
    threads = {str(s["N"]): ens.group_threads(g) for g, s in enumerate(setups)}
    ens.run(graph=True)  # capture + first replay
    ms = ens.run(graph=True)
    att = {s["N"]: float(np.mean([r["n_attempts"] for r in ens.output(g)["results"]])) for g, s in enumerate(setups)}
    total = sum(len(s["seeds"]) for s in setups)
    print(f"ensemble: {total} realisations in {ms:.0f} ms (threads per realisation by grid: {threads})",
          file=sys.stderr, flush=True)
    return {"workload": "soliton fibre (pi/40 km, beta2 -20 ps^2/km, gamma 1.3 /W/km), sech T0 1 ps N=2 + "
                        "one-photon-per-mode noise, adaptive RK4IP local error 1e-6, 100 ps window",
            "realisations": {str(s["N"]): len(s["seeds"]) for s in setups},
            "mean_attempts": {str(k): v for k, v in att.items()},
            "launch_ms": ms, "propagations_per_s": total / (1e-3 * ms), "threads_per_realisation": threads,
            "threads_trial": "each grid's realisations alone at 512 and 1024 threads, 32 replayed steps of h0; "
                             "grids at different thread counts run in two launches side by side",
            "graph": True}


def tune(ctx, fields, params, trial_attempts=100):
    """The autotuner on a context: every compiled schedule of its grid runs
    trial_attempts replayed steps of h0 on these fields (three runs, the first
    untimed); the fastest stays selected. Returns the record for the JSON line."""
    from pyofsschanged_amd import _lib
    idx, trial = ctx.tune_adaptive(fields, params, trial_attempts=trial_attempts, repeats=3)
    sc = _lib.schedules(ctx.n)
    print(f"autotuner ({ctx.n} samples): {sc[idx]['desc']}", file=sys.stderr, flush=True)
    return {"chosen": sc[idx]["desc"], "index": idx,
            "trial_ms": {s["desc"]: round(m, 3) for s, m in zip(sc, trial)},
            "trial": f"{trial_attempts} replayed attempts of h0 per instance, {fields.shape[0]} instances, mean of "
                     f"two timed runs per schedule"}


def amplifier_measurement(do_tune=True):
    """64 chirped Gaussian pulses through the 5 m saturable-gain amplifier
    (16384 samples, Raman, Lorentzian gain) in one adaptive launch."""
    from pyofsschanged_amd import workloads
    from pyofsschanged_amd._lib import METHOD_RK4IP
    domain, fields, fibre = workloads.amplifier()
    ctx = fibre.prepare(domain)
    params = fibre.stepper.adaptive_params()
    B, N = fields.shape
    d_in = ctx.malloc(fields.nbytes)
    d_out = ctx.malloc(fields.nbytes)
    ctx.htod(d_in, fields)
    sched = tune(ctx, fields, params, trial_attempts=40) if do_tune else None
    ctx.propagate_adaptive_device(METHOD_RK4IP, B, params, d_in, d_out)  # warm-up
    t0 = time.perf_counter()
    res = ctx.propagate_adaptive_device(METHOD_RK4IP, B, params, d_in, d_out)
    wall = time.perf_counter() - t0
    ms = ctx.last_kernel_ms()
    ctx.free(d_in)
    ctx.free(d_out)
    if any(r["status"] != 0 for r in res):
        raise RuntimeError("amplifier propagation failed")
    att = float(np.mean([r["n_attempts"] for r in res]))
    # 72 FFTs x 32 B x N per attempted step plus 12 Simpson energy reads of 16 N bytes
    achieved = B * att * (72 * 32 * N + 12 * 16 * N) / (1e-3 * ms) / 1e9
    print(f"amplifier: {B} chirps in {ms:.0f} ms", file=sys.stderr, flush=True)
    return {"workload": "Yb-band amplifier, 16384 samples / 40 ps at 1030 nm, 200 fs 1 nJ Gaussian, chirp in "
                        "[-2, 2], 5 m, beta2 0.023 ps^2/m, gamma 5e-3 /W/m, Raman, 15 dB gain, E_sat 5 nJ, "
                        "Lorentzian 40 nm gain, adaptive RK4IP",
            "batch": B, "launch_ms": ms, "propagations_per_s": B / wall, "mean_attempts": att,
            "algorithmic_GBps": achieved, "hbm_frac": achieved / HBM_PEAK_GBS, "schedule": sched}


def yb_attractor_measurement(batch):
    """The Yb amplifier run of examples/notebooks/nonlinear_attractor.ipynb,
    whose timing is the reference's only published GPU figure: once through
    System/Fibre exactly as the notebook runs it (wall time of System.run,
    500 traces and their characteristics included), then `batch` instances
    with pump powers spread over 5-15 W in one fixed-step launch (output
    fields only)."""
    from pyofsschanged_amd import System, workloads
    from pyofsschanged_amd._lib import METHOD_RK4IP
    domain, pulse, fibre = workloads.yb_attractor()
    sys0 = System(domain)
    sys0.add(pulse)
    sys0.add(fibre)
    t0 = time.perf_counter()
    sys0.run()
    single = time.perf_counter() - t0
    peak = float(sys0.df_results["Peak Power [W]"].values[-1])
    # batched: the same fibre, one instance per pump power
    ctx = fibre.prepare(domain)
    A0 = pulse.generate(domain.t)
    fields = np.ascontiguousarray(np.broadcast_to(A0, (batch, domain.total_samples)))
    src, ctx.pump_source = ctx.pump_source, None
    pumps = np.linspace(5.0, 15.0, batch)
    ctx.set_pump(pumps)
    h = fibre.length / fibre.stepper.total_steps
    t0 = time.perf_counter()
    ctx.propagate_fixed(METHOD_RK4IP, fields, h, fibre.stepper.total_steps, h_exp=fibre.linear_step(0.5 * h))
    wall = time.perf_counter() - t0
    ms = ctx.last_kernel_ms()
    out_pump = ctx.get_pump(batch)
    ctx.pump_source = src
    print(f"yb attractor: one run {single:.1f} s, {batch} pump powers in {wall:.1f} s", file=sys.stderr, flush=True)
    return {"workload": "Yb two-level amplifier of nonlinear_attractor.ipynb: 16384 samples / 200 ps at 1028 nm, "
                        "600 W 2 ps Gaussian, 8 m, beta2 20 ps^2/km, gamma 5.1 /W/km, Hollenbeck Raman, Yb gain "
                        "(N 4.8, 80 MHz), 30000 rk4ip steps",
            "single_run_s": single, "single_run_traces": 500, "final_peak_W": peak,
            "published_cpu_s": 143.37, "published_opencl_rtx4090_s": 115.50,
            "speedup_vs_published_opencl": 115.50 / single,
            "batch": batch, "batch_pump_W": [float(pumps[0]), float(pumps[-1])], "batch_launch_ms": ms,
            "batch_wall_s": wall, "batch_propagations_per_s": batch / wall,
            "batch_final_pump_W": [float(out_pump.min()), float(out_pump.max())]}


def surrogate_measurement(train_count, steps, length):
    """Train two spectral neural operators (different initialisations and
    minibatch orders) on the same solver-generated propagations, calibrate
    the certified judge on 4096 held-out solver runs, then compare the
    surrogate with the FP64 adaptive solver on 4096 new fields: the certified
    pipeline (mean of the two networks; energy gate, disagreement gate and
    audit; flagged fields re-solved) and, beside it, one network with the
    energy gate and audit only."""
    from pyofsschanged_amd import surrogate as S
    t0 = time.perf_counter()
    prob = S.SegmentProblem(length=length)
    nets = []
    for k in range(2):
        net = S.SpectralOperator(n=prob.n, max_batch=1024, in_channels=4)
        net.init_params(seed=7 * k)
        nets.append(net)
    train = S.Dataset(prob, nets[0], train_count, seed=1)
    t1 = time.perf_counter()
    for k, net in enumerate(nets):
        print(f"surrogate: training network {k + 1} of 2", file=sys.stderr, flush=True)
        S.train(net, train, steps, batch=32, lr=1e-3, seed=5 * k, log_every=max(1, steps // 6), log_file=sys.stderr)
    t_train = time.perf_counter() - t1
    n_eval = 4096

    def solved(seed):
        fields, _ = prob.sample(n_eval, seed=seed)
        d_in = S.DeviceArray(prob.ctx, fields.shape, np.complex128)
        d_out = S.DeviceArray(prob.ctx, fields.shape, np.complex128)
        d_in.upload(fields)
        t = time.perf_counter()
        res = prob.solve(d_in.ptr, d_out.ptr, n_eval)
        return d_in, d_out, d_out.download(), time.perf_counter() - t, res

    c_in, c_out, c_ref, _, _ = solved(4)  # held-out calibration fields
    d_in, d_out, ref, t_solver, res = solved(3)
    out = {}
    for mode, partner in (("certified", nets[1]), ("single_network", None)):
        sur = S.FibreSurrogate(prob, nets[0], train.scale, gate=0.005, audit_every=64, tol=0.01, partner=partner)
        cal = sur.calibrate(c_in.ptr, c_out.ptr, n_eval, c_ref) if partner is not None else None
        sur.infer(d_in.ptr, d_out.ptr, n_eval)  # first call loads code objects
        t3 = time.perf_counter()
        sur.infer(d_in.ptr, d_out.ptr, n_eval)
        t_net = time.perf_counter() - t3
        raw = S.spectral_error(d_out.download(), ref)
        rep = sur.propagate(d_in.ptr, d_out.ptr, n_eval, reference=ref)
        t_ideal = t_net + t_solver * float(np.mean(raw > 0.01))  # a judge that re-solves exactly those above tol
        out[mode] = {"network_s": t_net, "raw_speedup": t_solver / t_net, "raw_err_median": float(np.median(raw)),
                     "raw_err_p90": float(np.quantile(raw, 0.9)), "raw_within_1pct": float(np.mean(raw <= 0.01)),
                     "pipeline_s": rep["t_total_s"], "speedup": t_solver / rep["t_total_s"],
                     "resolved": rep["resolved"], "within_1pct_after_judge": rep["within_tol"],
                     "err_median_after_judge": rep["err_median"], "err_max_after_judge": rep["err_max"],
                     "speedup_ideal_judge": t_solver / t_ideal, "calibration": cal}
        print(f"surrogate {mode}: solver {t_solver:.2f}s, network {t_net:.3f}s, pipeline {rep['t_total_s']:.2f}s, "
              f"within 1% after judge {rep['within_tol']:.4f}", file=sys.stderr, flush=True)
    c = out["certified"]
    return {"segment": (f"soliton fibre physics (beta2 -20 ps^2/km, gamma 1.3 /W/km) over {length * 1e3:g} m, "
                        "4096 samples / 25 ps at 1550 nm; sech 1-10 kW (log-uniform), 30-150 fs FWHM, chirp "
                        "[-1, 1], one-photon-per-mode noise"),
            "network": ("two spectral neural operators, 4 Fourier layers, width 64, 128 modes, bf16 MFMA / FP32 "
                        "accumulate; output their mean"),
            "train_fields": train_count, "train_steps_per_network": steps, "train_s": t_train,
            "data_solve_s": train.solve_s, "setup_s": t1 - t0, "fields": n_eval, "solver_s": t_solver,
            "solver_mean_attempts": float(np.mean([r["n_attempts"] for r in res])),
            "speedup": c["speedup"], "within_1pct_after_judge": c["within_1pct_after_judge"],
            "err_max_after_judge": c["err_max_after_judge"], "tol": 0.01,
            "judge": ("energy gate 0.005, disagreement gate calibrated on 4096 held-out solver runs (0.8 x the "
                      "smallest disagreement of a held-out field above 1 %), audit of every 64th field; flagged "
                      "fields re-solved by the FP64 solver"),
            "certified": c, "single_network": out["single_network"]}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (instances shard by realisation, "
              f"with no exchange until ensemble statistics)", file=sys.stderr)
        sys.exit(2)

    from pyofsschanged_amd import _lib, workloads
    from pyofsschanged_amd._lib import METHOD_RK4IP

    domain, pulse, fibre = workloads.supercontinuum(length=args.length)
    N = domain.total_samples
    ctx = fibre.prepare(domain)
    params = fibre.stepper.adaptive_params()
    B = args.batch or cu_count()

    A0 = pulse.generate(domain.t)
    inputs = np.stack([A0 + workloads.oppm_noise(domain, seed=k) for k in range(B)]).astype(np.complex128)
    nbytes = inputs.nbytes
    d_in = ctx.malloc(nbytes)
    d_out = ctx.malloc(nbytes)
    ctx.htod(d_in, inputs)
    sched = None
    if args.schedule is not None:
        ctx.set_schedule(args.schedule)
    elif not args.no_tune:
        sched = tune(ctx, inputs, params)

    for _ in range(args.warmup):
        ctx.propagate_adaptive_device(METHOD_RK4IP, B, params, d_in, d_out)
        print(f"warmup launch {ctx.last_kernel_ms():.1f} ms", file=sys.stderr, flush=True)

    kernel_ms = []
    attempts = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        res = ctx.propagate_adaptive_device(METHOD_RK4IP, B, params, d_in, d_out)
        kernel_ms.append(ctx.last_kernel_ms())
        attempts.append(sum(r["n_attempts"] for r in res))
        bad = [r["status"] for r in res if r["status"] != 0]
        if bad:
            raise RuntimeError(f"{len(bad)} instances failed with status {set(bad)}")
        print(f"step {k}: {kernel_ms[-1]:.1f} ms, {attempts[-1] / B:.1f} attempts/instance", file=sys.stderr,
              flush=True)
    wall = time.perf_counter() - t0
    out = np.empty_like(inputs)
    ctx.dtoh(out, d_out)
    ctx.free(d_in)
    ctx.free(d_out)
    if not np.all(np.isfinite(out)):
        raise RuntimeError("non-finite output field")

    ms_per_step = 1e3 * wall / args.steps
    value = B * args.steps / wall
    mean_kernel_s = 1e-3 * float(np.mean(kernel_ms))
    attempts_launch = float(np.mean(attempts))
    # the reference's accounting (field.fft_counter): 72 N-point FFTs per
    # attempted step, each 5 N log2 N flops and one HBM read + write of the field
    flops_per_launch = attempts_launch * FFTS_PER_ATTEMPT * 5.0 * N * math.log2(N)
    fft_bytes_per_launch = attempts_launch * FFTS_PER_ATTEMPT * 32 * N
    achieved_tflops = flops_per_launch / mean_kernel_s / 1e12
    # measured HBM bytes of the same kernel and batch over the full fibre
    # (rocprofv3 FETCH_SIZE / WRITE_SIZE, copy-calibrated), per attempted step,
    # times this launch's attempts (the inputs are seeded: the same attempts)
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
        if t.get("batch", B) == B and t.get("length_km") == args.length and "hbm_bytes_per_attempt" in t:
            traffic = t["hbm_bytes_per_attempt"] * attempts_launch

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "propagations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("supercontinuum: 8192 samples / 12.5 ps at 835 nm, 50 fs FWHM sech 10 kW + one-photon-"
                         "per-mode noise, 15 cm PCF gamma 110 /W/km beta2..beta10, self-steepening + Blow-Wood "
                         "Raman, adaptive RK4IP local error 1e-6"),
            "grid": N,
            "batch": B,
            "length_km": args.length,
            "mean_attempts_per_propagation": float(np.mean(attempts)) / B,
            "schedule": _lib.schedules(N)[ctx.get_schedule()]["desc"],
        },
        "autotuner": sched,
        "roofline": {
            "bound": "fp64",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "hbm_GBps": None if traffic is None else traffic / mean_kernel_s / 1e9,
            "hbm_frac": None if traffic is None else traffic / mean_kernel_s / 1e9 / HBM_PEAK_GBS,
            "algorithmic_GBps": fft_bytes_per_launch / mean_kernel_s / 1e9,
            "accounting": ("achieved: the reference's fft_counter work, 72 N-point FFTs of 5 N log2 N flops per "
                           "attempted step, over the kernel's HIP-event time, against the FP64 peak (the kernel "
                           "keeps every transform in registers and LDS, so FP64 issue and LDS-exchange latency "
                           "bind it, not HBM); traffic: measured HBM bytes of the launch (hbm_GBps, hbm_frac "
                           "against 8 TB/s); algorithmic_GBps: the same fft_counter work as bytes (32 N per "
                           "FFT), which exceeds the HBM peak because no transform goes through HBM"),
        },
        "cpu_baseline": None,
        "error": None,
    }
    if not args.no_ensemble:
        line["ensemble"] = ensemble_measurement()
    if not args.no_amplifier:
        line["amplifier"] = amplifier_measurement(not args.no_tune)
    if not args.no_yb:
        line["yb_attractor"] = yb_attractor_measurement(cu_count())
    if not args.no_surrogate:
        line["surrogate"] = surrogate_measurement(args.surrogate_train, args.surrogate_steps, args.surrogate_length)
    if not args.no_cpu:
        line["cpu_baseline"], line["error"] = cpu_baseline(
            inputs[0], args.length, out[0], res[0]["n_attempts"], attempts_launch / B, args.cpu_seconds,
            min(os.cpu_count() or 1, args.cpu_workers))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
