"""Throughput benchmark of the split-operator hot path on one MI355X.

Workload (headline, mixed precision): Wigner-Moyal split-operator propagation of a batch of
B independent 1024 x 1024 phase-space states (harmonic + quartic V, displaced Gaussian
initial states), Strang steps dt = 0.01, ending with the reference's normalisation through
the Wigner transform (propagate without Ehrenfest, wigner_moyal_cuda_1d.py:563-582).  The
same states are also propagated by the plain FP64 path for the speedup and the error.

A "step" is one Strang step of the whole batch.  value = B * 1024 * 1024 * steps / wall time
(grid-point-steps per second), inputs resident in HBM.

Prints ONE JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import contextlib
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-point-steps/s vs plain FP64 at stated error; p50/p99 solve latency"
HBM_PEAK_GBPS = 8000.0
HEADLINE_PRECISION = "mixed"


@contextlib.contextmanager
def no_gc():
    """Python's cyclic garbage collector off inside a timed region (as timeit does): a full
    collection of the interpreter's heap takes tens of ms and lands in whichever region it
    falls, which is host noise, not the work being measured."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def make_system(N, batch, precision, L=10.0, dt=0.01):
    from accuratewigner_amd import WignerMoyalCUDA1D
    sysm = WignerMoyalCUDA1D(X_gridDIM=N, X_amplitude=L, P_gridDIM=N, P_amplitude=L, dt=dt, t=0.,
                             K="0.5 * P * P", V="0.5 * X * X + 0.05 * X * X * X * X",
                             precision=precision, batch=batch)
    return sysm


def initial_states(sysm, batch, seed=7):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-2, 2, batch)
    p0 = rng.uniform(-1, 1, batch)
    X, P = sysm.X, sysm.P
    W = np.exp(-(X[None] - x0[:, None, None]) ** 2 - (P[None] - p0[:, None, None]) ** 2) / np.pi
    return W if batch > 1 else W[0]


def time_propagate(sysm, warmup, steps):
    """W warm-up steps, then the K timed steps right after them: the host's garbage
    collection runs before the warm-up, not between the two calls (a GPU left idle for its
    40-50 ms falls back towards its idle clocks, and a short timed call then runs through
    their transient again: 1.23 -> 1.275e11 gp-steps/s for the 20-step headline call)."""
    import torch
    with no_gc():
        sysm.propagate(warmup)
        torch.cuda.synchronize()
        sysm._pass_events = []
        t0 = time.perf_counter()
        sysm.propagate(steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    evs, sysm._pass_events = sysm._pass_events, None
    pass_ms = sum(a.elapsed_time(b) for a, b, _ in evs)
    npass = sum(n for _, _, n in evs)
    return t1 - t0, pass_ms / max(npass, 1)


def cpu_baseline(N, seconds, workers=None):
    """Oracle (NumPy complex128) WM step on one 1024^2 state with precomputed factors."""
    from oracle.awig_oracle import DFT, WignerMoyalOracle
    o = WignerMoyalOracle(X_gridDIM=N, X_amplitude=10., P_gridDIM=N, P_amplitude=10., dt=0.01,
                          V=lambda X, t: 0.5 * X * X + 0.05 * X ** 4, K=lambda P, t: 0.5 * P * P,
                          dft=DFT(workers))
    o.rho = np.exp(-(o.kXX[None, :] - 1.) ** 2 / 2 - (o.kXXp[:, None] - 1.) ** 2 / 2).astype(np.complex128)
    o.propagate_precomputed(1)
    n, t0 = 0, time.perf_counter()
    while True:
        o.propagate_precomputed(1)
        n += 1
        el = time.perf_counter() - t0
        if el > seconds and n >= 2:
            break
    return N * N * n / el, n, el


def copy_bandwidth(mib=1024, reps=10):
    """Achievable HBM bandwidth of a plain copy (libawig's k_convert, complex128 -> complex128)
    of ``mib`` MiB, timed with HIP events on the launch stream: the practical ceiling next to
    the 8 TB/s peak, and the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    from accuratewigner_amd import engine
    n = mib * 2 ** 20 // 16
    a = torch.ones(2 * n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    engine.convert(a.data_ptr(), 0, b.data_ptr(), 0, n)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        engine.convert(a.data_ptr(), 0, b.data_ptr(), 0, n)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2 * n * 16 / (ms * 1e-3) / 1e9, ms


def serving_run(n_jobs, seed=11, chunk=64):
    """Serving workload: a seeded Poisson stream (mean 200 req/s) of WM / WB jobs, grids from
    {128, 256, 512, 1024}^2, V = X^2/2 + c4 X^4 with c4 ~ U(0, 0.2), beta ~ U(0.5, 5),
    steps ~ U{200..5000}, FP32 / FP64 tiers, served in real time by continuous batching."""
    from accuratewigner_amd.serving import Job, Server, poisson_jobs
    srv = Server(hbm_budget_bytes=32 << 30, chunk=chunk, max_slots=64)
    # warm-up: one short job per pool key (hipRTC compiles, plans, pool allocations)
    warm = [Job(k, N, prec, steps=8, beta=0.04, job_id=-1) for k in ("wm", "wb")
            for N in srv.grids for prec in ("mixed", "fp64")]
    srv.run(warm, realtime=False)
    jobs = poisson_jobs(n_jobs, rate=200.0, seed=seed)
    srv.stats.clear()
    done = srv.run(jobs, realtime=True)
    lat = np.array([j.latency for j in done]) * 1e3
    work = float(sum(j.N * j.N * j.steps for j in done))
    wall = srv.stats["wall_s"]
    return {"workload": "Poisson stream, mean 200 req/s, %d jobs (seed %d), WM/WB, grids 128..1024, "
                        "FP32/FP64 tiers, steps 200..5000, beta 0.5..5, chunk %d" % (n_jobs, seed, chunk),
            "jobs": len(done), "value": work / wall, "unit": "gp-steps/s", "wall_s": wall,
            "arrival_span_s": float(jobs[-1].arrival), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max())}


def host_info():
    """CPU model, thread count and library versions of the host the CPU baseline runs on."""
    import platform
    import scipy
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "numpy": np.__version__, "scipy": scipy.__version__}


def leg_tier(N, B, K, W, precision):
    """One complex64-state tier on the WM batch: throughput and the roofline of its step
    launches (one read and one write of the complex64 states per pass)."""
    mixed = make_system(N, B, precision)
    W0 = initial_states(mixed, B)
    mixed.set_wignerfunction(W0)
    wall, pass_ms = time_propagate(mixed, W, K)
    # a pass over the batch is one launch per group of items (groups stay in the MALL
    # between passes; equal groups here): per launch, bytes and time are both 1/launches
    group = mixed.plan.step_group(min(K, 256), B)
    launch_ms = pass_ms * group / B
    bytes_per_launch = 2 * group * N * N * 8      # one read + one write of the group's c64 states
    return dict(sys=mixed, W0=W0, wall=wall, value=B * N * N * K / wall, pass_ms=pass_ms, group=group,
                launches=-(-B // group), launch_ms=launch_ms, bytes_per_launch=bytes_per_launch,
                bytes_per_pass=2 * B * N * N * 8,
                achieved=bytes_per_launch / (launch_ms * 1e-3) / 1e9)


def rel_err(sysm, ref, Bf):
    """Largest per-state relative L2 error of the first Bf states of sysm against ref."""
    a = sysm.wignerfunction.get()[:Bf].astype(np.complex128)
    return float(np.max(np.linalg.norm((a - ref).reshape(Bf, -1), axis=1)
                        / np.linalg.norm(ref.reshape(Bf, -1), axis=1)))


def leg_fp64(N, Bf, K, W, W0, tiers):
    """The plain FP64 path on the first Bf states: the speed-ups and errors of the tiers."""
    fp64 = make_system(N, Bf, "fp64")
    fp64.set_wignerfunction(W0[:Bf] if Bf > 1 else W0[0])
    wall64, pass_ms64 = time_propagate(fp64, W, K)
    out = {"value": Bf * N * N * K / wall64, "unit": "gp-steps/s", "batch": Bf, "ms_per_step": wall64 / K * 1e3,
           "avg_pass_ms": pass_ms64, "achieved_GBps": 2 * Bf * N * N * 16 / (pass_ms64 * 1e-3) / 1e9}
    b = fp64.wignerfunction.get()
    if Bf == 1:
        b = b[None]
    errs = {name: rel_err(t["sys"], b, Bf) for name, t in tiers.items()}
    return out, errs


def leg_latency():
    """Config 1 solve latency (WM FP64 256^2, x,p in [-10, 10), V = x^2/2, W0 = exp(-(x-2)^2
    - p^2)/pi, dt = 0.01, 1000 steps, norm/<x>/<p>/<H> recorded every step), 22 solves; and
    the same grid without observables in mixed precision."""
    import torch
    from accuratewigner_amd import WignerMoyalCUDA1D
    lat = []
    c1 = WignerMoyalCUDA1D(X_gridDIM=256, X_amplitude=10., P_gridDIM=256, P_amplitude=10., dt=0.01, t=0.,
                           K="0.5 * P * P", V="0.5 * X * X", diff_K="P", diff_V="X")
    for r in range(22):
        c1.t = 0.
        for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
            getattr(c1, name).clear()
        c1.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
        torch.cuda.synchronize()
        with no_gc():
            t0 = time.perf_counter()
            c1.propagate(1000)
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.array(lat[2:])
    lat_nr = []
    one = make_system(256, 1, "mixed")
    st = initial_states(one, 1)
    for r in range(12):
        one.set_wignerfunction(st)
        torch.cuda.synchronize()
        with no_gc():
            t0 = time.perf_counter()
            one.propagate(1000)
            torch.cuda.synchronize()
            lat_nr.append((time.perf_counter() - t0) * 1e3)
    lat_nr = np.array(lat_nr[2:])
    return {"workload": "config 1 solve: WM FP64 256x256, 1000 steps, norm/<x>/<p>/<H> every step "
                        "(20 timed solves)",
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "no_observables_mixed": {"workload": "WM mixed 256x256, 1000 steps, observables off (10 solves)",
                                     "p50_ms": float(np.percentile(lat_nr, 50)),
                                     "p99_ms": float(np.percentile(lat_nr, 99))}}


def leg_gibbs():
    """Config 2: Wigner-Bloch Gibbs state, 512^2, x, p in [-8, 8), V = x^2/2 + 0.1 x^4, beta
    0 -> 2 in d beta = 0.005 (400 steps), renormalised and <H>(beta) recorded every step; FP64
    and mixed, 5 timed solves each."""
    import torch
    from accuratewigner_amd import WignerBlochCUDA1D
    gibbs, h_rec = {}, {}
    for prec in ("fp64", "mixed"):
        wb = WignerBlochCUDA1D(X_gridDIM=512, X_amplitude=8., P_gridDIM=512, P_amplitude=8., kT=0.5, dbeta=0.005,
                               t_initial=0., t=0., V="0.5 * X * X + 0.1 * X * X * X * X", K="0.5 * P * P",
                               precision=prec)
        times = []
        for r in range(6):
            torch.cuda.synchronize()
            with no_gc():
                t0 = time.perf_counter()
                wb.get_gibbs_state(record=[wb.hamiltonian])
                torch.cuda.synchronize()
                times.append((time.perf_counter() - t0) * 1e3)
        h_rec[prec] = wb.gibbs_record[:, 0].copy()
        p50 = float(np.percentile(times[1:], 50))
        gibbs[prec] = {"p50_ms": p50, "p99_ms": float(np.percentile(times[1:], 99)),
                       "gp_steps_per_s": 512 * 512 * len(h_rec[prec]) / (p50 * 1e-3)}
        del wb
    gibbs["mixed"]["H_rel_err_vs_fp64"] = float(np.max(np.abs(h_rec["mixed"] - h_rec["fp64"]) / np.abs(h_rec["fp64"])))
    gibbs["workload"] = ("config 2: Wigner-Bloch Gibbs state 512x512, beta 0 -> 2 in 400 steps of 0.005, "
                         "<H>(beta) and purity every step (5 timed solves)")
    return gibbs


def leg_strong(max_steps=0):
    """Config 3: strong-field von Neumann rho(x, x'), 1024^2, x in [-200, 200), soft-core V,
    E(t) = 0.05 sin^2(pi t / T) cos(0.06 t) over 4 cycles (20944 steps of 0.02), cos^(1/8)
    absorbing edge, Ehrenfest lists (dipole acceleration -<V'>) every step; FP64 and mixed.
    max_steps > 0 cuts the pulse (profiling runs only)."""
    import torch
    from accuratewigner_amd import RhoVNeumannCUDA1D
    from accuratewigner_amd.strong_field import rv_ground_state_rho, rv_sys_params
    p3 = rv_sys_params()
    steps3 = int(round(p3["T_pulse"] / p3["dt"]))
    if max_steps > 0:
        steps3 = min(steps3, max_steps)
    rho0 = rv_ground_state_rho(p3, nsteps=5000)
    strong, accel = {}, {}
    for prec in ("fp64", "mixed"):
        g3 = RhoVNeumannCUDA1D(precision=prec, **p3)
        g3.set_rho(rho0)
        g3.propagate(16)                  # compiles and captures outside the timed run
        g3.set_rho(rho0)
        g3.t = 0.
        for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
            getattr(g3, name).clear()
        torch.cuda.synchronize()
        with no_gc():
            t0 = time.perf_counter()
            g3.propagate(steps3)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
        accel[prec] = np.array(g3.P_average_RHS)
        strong[prec] = {"s": el, "us_per_step": el / steps3 * 1e6, "gp_steps_per_s": 1024 ** 2 * steps3 / el}
        del g3
    strong["mixed"]["accel_err_of_max"] = float(np.max(np.abs(accel["mixed"] - accel["fp64"]))
                                                / np.max(np.abs(accel["fp64"])))
    strong["workload"] = ("config 3: von Neumann rho 1024x1024, 4-cycle pulse, %d steps of 0.02, "
                          "Ehrenfest lists every step (1 run each)" % steps3)
    return strong


def leg_adjoint():
    """Config 4: adjoint gradient d<x>(T)/dE_k, WM 512^2, x, p in [-6, 6), V = x^4/4 - x^2/2,
    64 field pieces ~ U(-0.1, 0.1), 2000 steps of 0.005, checkpoints every 50 steps."""
    import torch
    from accuratewigner_amd.adjoint import WignerMoyalDipoleControl
    c4 = WignerMoyalDipoleControl(X_gridDIM=512, X_amplitude=6., P_gridDIM=512, P_amplitude=6., dt=0.005,
                                  steps=2000, n_pieces=64, V0="0.25 * X * X * X * X - 0.5 * X * X",
                                  W0="exp(-2. * (X + 1.) * (X + 1.) - 0.5 * P * P) / M_PI", checkpoint=50)
    E4 = np.random.default_rng(4).uniform(-0.1, 0.1, 64)
    c4.gradient(E4)
    times = []
    for r in range(3):
        torch.cuda.synchronize()
        with no_gc():
            t0 = time.perf_counter()
            c4.forward(E4)
            c4.gradient()
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
    return {"workload": "config 4: forward + adjoint gradient, WM FP64 512x512, 2000 steps, 64 pieces, "
                        "checkpoint 50 (3 timed runs)",
            "p50_ms": float(np.percentile(times, 50)), "max_ms": float(np.max(times))}


def leg_cpu(N, seconds):
    """The oracle on the host: one core (numpy.fft) and all cores (scipy.fft workers)."""
    info = host_info()
    v1, n1, e1 = cpu_baseline(N, seconds)
    # the host's share of cores (OMP_NUM_THREADS on the GPU box), at most the CPUs it has
    ncpu = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "0")) or 16)
    va, na, ea = cpu_baseline(N, seconds, workers=ncpu)
    return {"value": v1, "unit": "gp-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle (NumPy complex128, numpy.fft) WM %dx%d single state, %d steps in %.1f s, "
                      "time-independent factors precomputed" % (N, N, n1, e1),
            "all_cores": {"value": va, "unit": "gp-steps/s", "cores": ncpu,
                          "sample": "the same loop with scipy.fft(workers=%d), %d steps in %.1f s" % (ncpu, na, ea)},
            **info}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=96)
    ap.add_argument("--fp64-batch", type=int, default=48)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--strong-steps", type=int, default=0,
                    help="cut the von Neumann leg to this many steps (profiling runs; 0: the whole pulse)")
    ap.add_argument("--legs", default="all",
                    help="comma list of the legs to run: headline, c64, fp64, latency, gibbs, strong, adjoint, "
                         "serving, cpu (default: all); a profiling run names the legs it measures")
    ap.add_argument("--copy-only", type=int, default=0, metavar="MIB",
                    help="only run the copy-bandwidth kernel on MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the propagators would shard over independent states, one batch slice per GPU)", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    N, B, K, W = args.grid, args.batch, args.steps, args.warmup
    if args.copy_only:
        gbps, ms = copy_bandwidth(args.copy_only, reps=3)
        return {"copy_MiB": args.copy_only, "copy_GBps": gbps, "ms": ms}
    legs = set("headline,c64,fp64,latency,gibbs,strong,adjoint,serving,cpu".split(",") if args.legs == "all"
               else args.legs.split(","))
    if args.no_cpu:
        legs.discard("cpu")
    if args.serve_jobs <= 0:
        legs.discard("serving")
    copy_gbps, _ = copy_bandwidth()
    done = {}

    def run(name, fn, *a):
        print("leg: %s" % name, file=sys.stderr, flush=True)
        done[name] = fn(*a) if name in legs else None
        return done[name]

    # the complex64-state tiers: "mixed" (FP32 arithmetic) and "c64" (FP64 arithmetic)
    tiers = {}
    for prec, leg in (("mixed", "headline"), ("c64", "c64")):
        if leg in legs:
            print("leg: %s (%s)" % (leg, prec), file=sys.stderr, flush=True)
            tiers[prec] = leg_tier(N, B, K, W, prec)
    Bf = min(args.fp64_batch, B)
    fp64_line, errs = None, {}
    if "fp64" in legs:
        print("leg: fp64", file=sys.stderr, flush=True)
        W0 = next(iter(tiers.values()))["W0"] if tiers else initial_states(make_system(N, 1, "fp64"), B)
        fp64_line, errs = leg_fp64(N, Bf, K, W, W0, tiers)
    for t in tiers.values():
        t.pop("sys")
        t.pop("W0")
    h = tiers.get(HEADLINE_PRECISION, {})
    err = errs.get(HEADLINE_PRECISION)
    run("latency", leg_latency)
    run("gibbs", leg_gibbs)
    run("strong", leg_strong, args.strong_steps)
    run("adjoint", leg_adjoint)
    run("serving", serving_run, args.serve_jobs)

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_wm1024_mixed_b%d.json" % B)
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
    cpu = run("cpu", leg_cpu, N, args.cpu_seconds)

    value = h.get("value")
    line = {
        "metric": METRIC, "value": value, "unit": "gp-steps/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": h["wall"] / K * 1e3 if h else None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "Wigner-Moyal split-operator propagate (no Ehrenfest, normalised at end), "
                               "%dx%d grid, batch of %d states, mixed precision (complex64 state, FP64 phases)"
                               % (N, N, B),
                   "grid": [N, N], "batch": B, "precision": "mixed", "dt": 0.01,
                   "V": "0.5*X*X + 0.05*X^4", "K": "0.5*P*P"},
        "fp64": fp64_line,
        "speedup_vs_fp64": value / fp64_line["value"] if (value and fp64_line) else None,
        "rel_l2_err_vs_fp64": err, "err_steps": W + K,
        "error_bound": 1e-6, "error_bound_met": (err <= 1e-6) if err is not None else None,
        "tiers": {name: {"value": t["value"], "ms_per_step": t["wall"] / K * 1e3, "avg_launch_ms": t["launch_ms"],
                         "roofline_frac": t["achieved"] / HBM_PEAK_GBPS,
                         "speedup_vs_fp64": t["value"] / fp64_line["value"] if fp64_line else None,
                         "rel_l2_err_vs_fp64": errs.get(name),
                         "error_bound_met": (errs[name] <= 1e-6) if errs.get(name) is not None else None,
                         "arithmetic": "FP32 butterflies and factors" if name == "mixed" else
                                       "FP64 butterflies and factors; line fixed point (int32, one power-of-two "
                                       "scale per line) between the passes of a step sequence"}
                  for name, t in tiers.items()},
        "roofline": None if not h else {
            "bound": "hbm", "achieved": h["achieved"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": h["achieved"] / HBM_PEAK_GBPS, "traffic": traffic,
            "kernel": "k_line_pass (row and column passes)", "avg_launch_ms": h["launch_ms"],
            "bytes_per_launch": h["bytes_per_launch"], "group_items": h["group"],
            "launches_per_pass": h["launches"], "avg_pass_ms": h["pass_ms"],
            "bytes_per_pass": h["bytes_per_pass"], "copy_GBps": copy_gbps,
            "frac_of_copy": h["achieved"] / copy_gbps,
            "step_level_frac": value * 32 / 1e9 / HBM_PEAK_GBPS},
        "latency": done["latency"],
        "gibbs": done["gibbs"],
        "strong_field": done["strong"],
        "adjoint": done["adjoint"],
        "serving": done["serving"],
        "cpu_baseline": cpu,
    }
    return line


if __name__ == "__main__":
    # the propagator constructors print a memory report (as the reference's do): send it and
    # everything else to stderr, so that stdout carries the one JSON line
    with contextlib.redirect_stdout(sys.stderr):
        result = main()
    print(json.dumps(result), flush=True)
