"""Continuous-batching serving of Wigner-Moyal / Wigner-Bloch jobs on one MI355X.

Jobs of different grid sizes, potentials and temperatures arrive over time.  For every
(grid, precision, kind) the server keeps a resident slot pool: one [capacity, N, N] state
tensor sized from the HBM budget, with per-slot factor tables.  Each scheduling round
advances every non-empty pool by one chunk of merged Strang steps in ONE batched launch
sequence over its active slots (aw_step_sequence with per-slot tables and constant-step
mode).  Pools run on their own HIP streams, so the work of different grid sizes overlaps
on the device.  Arrived jobs are admitted into free slots between rounds; finished jobs are
transformed back to W, copied out, and their slot is refilled by compaction (the last
active slot moves into the hole).

Job family (config 5 of the serving workload): K = P^2/2, V = X^2/2 + c4 X^4 on
x, p in [-L, L).  "wm" jobs propagate the Gaussian W0 = exp(-(x-x0)^2 - (p-p0)^2)/pi for
`steps` real-time steps of dt (wigner_moyal_cuda_1d.py:543-582, normalised once at the end
as propagate does without Ehrenfest).  "wb" jobs run the Bloch propagation of
wigner_bloch_cuda_1d.py:119-155 from the infinite-temperature state to beta = 1/kT in
steps of dbeta, normalised at the end: per-step renormalisation only rescales, so the
result is the same up to rounding.  The serving tier does not apply the per-step
purity < 1 guard of get_gibbs_state (use WignerBlochCUDA1D for that).
Since V >= 0 and K >= 0 vanish at X = P = 0, which are grid points, V_min = K_min = 0.
"""
import collections
import ctypes
import time

import numpy as np

from . import rtc
from .device import DeviceArray, current_stream, empty, require_gpu, to_device
from .engine import Plan
from .wm_ops import WMGrid, rho2wigner, wigner2rho

_SERVE_SRC = r"""
extern "C" __global__ void aw_serve_vk(const double *x, int n, double c4, int which, double *out)
{
    // which 0: V(X) = X^2/2 + c4 X^4;  which 1: K(P) = P^2/2
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const double X = x[k];
        out[k] = which == 0 ? 0.5 * X * X + c4 * X * X * X * X : 0.5 * X * X;
    }
}
extern "C" __global__ void aw_serve_init(%(st)s *W, int nx, int np_, double dX, double dP, int kind,
                                         double x0, double p0, double fill)
{
    // kind 0: Gaussian exp(-(X-x0)^2 - (P-p0)^2)/pi on the kernel grid (:791-801); 1: constant
    const long long total = (long long)nx * np_;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long i = idx / nx, j = idx %% nx;
        const double X = dX * (j - 0.5 * nx), P = dP * (i - 0.5 * np_);
        const double v = kind == 0 ? exp(-(X - x0) * (X - x0) - (P - p0) * (P - p0)) / M_PI : fill;
        aw_state_store<%(st)s>::put(W + idx, cuda_complex(v, 0.));
    }
}
"""

_MODULES = {}


def _serve_module(state_t):
    mod = _MODULES.get(state_t)
    if mod is None:
        mod = rtc.Module(rtc.PRELUDE + _SERVE_SRC % dict(st=state_t), "serving.hip")
        _MODULES[state_t] = mod
    return mod


class Job:
    """One serving request."""
    _next = 0

    def __init__(self, kind, N, precision="mixed", c4=0.1, steps=1000, beta=1.0, dt=0.01,
                 dbeta=0.005, x0=1.0, p0=0.0, arrival=0.0, keep_result=True, L=8.0, job_id=None):
        assert kind in ("wm", "wb")
        self.kind, self.N, self.precision = kind, int(N), precision
        self.c4, self.dt, self.dbeta, self.x0, self.p0, self.L = float(c4), float(dt), float(dbeta), x0, p0, float(L)
        if kind == "wm":
            self.steps = int(steps)
        else:
            # dbeta adjusted so that beta is reached in a whole number of steps
            # (get_dbeta_num_beta_steps, wigner_bloch_cuda_1d.py:94-117)
            n = 1. / ((1. / beta) * dbeta)
            if np.floor(n + 0.5) != n:
                n = np.floor(n + 0.5)
                self.dbeta = 1. / ((1. / beta) * n)
            self.steps = int(n)
        self.beta = beta
        self.arrival = float(arrival)
        self.keep_result = keep_result
        if job_id is None:
            job_id = Job._next
            Job._next += 1
        self.id = job_id
        self.result = None
        self.summary = None
        self.t_admit = self.t_done = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


class _Pool:
    """Resident slot pool for one (N, precision, kind) with per-slot factor tables."""

    def __init__(self, N, precision, kind, capacity, L):
        import torch
        self.N, self.precision, self.kind, self.capacity = N, precision, kind, capacity
        self.stream = torch.cuda.Stream()
        with torch.cuda.stream(self.stream):
            self.plan = Plan(N, N, capacity, precision)
            self.plan1 = Plan(N, N, 1, precision)
            self.dtype = self.plan.state_dtype
            self.state_t = "double2" if self.dtype == np.complex128 else "float2"
            self.mod = _serve_module(self.state_t)
            self.g = WMGrid(N, L, N, L)
            self.state = DeviceArray((capacity, N, N), self.dtype)
            self.W1 = DeviceArray((N, N), self.dtype)
            self.colL = empty((capacity, 3, N), self.dtype)
            self.colE = empty((capacity, 3, N), self.dtype)
            self.rowL = empty((capacity, 1, N), self.dtype)
            self.rowE = empty((capacity, 1, N), self.dtype)
            self.vk = empty((4, N), np.float64)
            self.dscale = empty((1, 2), np.float64)
            self.coords = [to_device(self.g.kXX), to_device(self.g.kXXp), to_device(self.g.kPP),
                           to_device(self.g.kPPp)]
        self.jobs = []
        self.remaining = []
        self.pending = []        # (job, pinned host W, event) of results still being copied out
        self._free_host = []     # pinned host buffers to reuse

    @property
    def slot_bytes(self):
        return self.N * self.N * np.dtype(self.dtype).itemsize

    def _factors(self, k, job):
        """Per-slot tables of the merged Strang step (constant-step layout 0 open, 1 mid, 2 close)."""
        N, p = self.N, self.plan
        esz = np.dtype(self.dtype).itemsize
        for q, (coord, which) in enumerate(zip(self.coords, (0, 0, 1, 1))):
            self.mod.launch("aw_serve_vk", (max(1, N // 256),), (256,),
                            [coord.ptr, ctypes.c_int(N), ctypes.c_double(job.c4), ctypes.c_int(which),
                             self.vk.ptr + q * N * 8], current_stream())
        Vj, Vi, Kj, Ki = (self.vk.ptr + q * N * 8 for q in range(4))
        cl = self.colL.ptr + k * 3 * N * esz
        ce = self.colE.ptr + k * 3 * N * esz
        rl = self.rowL.ptr + k * N * esz
        re = self.rowE.ptr + k * N * esz
        if job.kind == "wm":
            h = job.dt
            # expV (:763-778) split over columns j and rows i; expK (:742-754)
            p.factors(cl, N, 1, A=Vj, ca=-0.5 * h, alt_sign=True)
            p.factors(cl, N, 1, A=Vj, ca=-h, out_offset=1)
            p.factors(cl, N, 1, A=Vj, ca=-0.5 * h, alt_sign=True, out_offset=2)
            p.factors(ce, N, 1, A=Vi, ca=0.5 * h, alt_sign=True)
            p.factors(ce, N, 1, A=Vi, ca=h, scale=1.0 / N, out_offset=1)
            p.factors(ce, N, 1, A=Vi, ca=0.5 * h, alt_sign=True, scale=1.0 / N, out_offset=2)
            p.factors(rl, N, 1, A=Ki, ca=h, scale=1.0 / N)
            p.factors(re, N, 1, A=Kj, ca=-h)
        else:
            q, hb = -0.25 * job.dbeta, -0.5 * job.dbeta
            # bloch_expV / bloch_expK (wigner_bloch_cuda_1d.py:276-316) with V_min = K_min = 0
            p.factors(cl, N, 1, A=Vj, ca=q, alt_sign=True, mode=1)
            p.factors(cl, N, 1, A=Vj, ca=2 * q, mode=1, out_offset=1)
            p.factors(cl, N, 1, A=Vj, ca=q, alt_sign=True, mode=1, out_offset=2)
            p.factors(ce, N, 1, A=Vi, ca=q, alt_sign=True, mode=1)
            p.factors(ce, N, 1, A=Vi, ca=2 * q, mode=1, scale=1.0 / N, out_offset=1)
            p.factors(ce, N, 1, A=Vi, ca=q, alt_sign=True, mode=1, scale=1.0 / N, out_offset=2)
            p.factors(rl, N, 1, A=Ki, ca=hb, mode=1, scale=1.0 / N)
            p.factors(re, N, 1, A=Kj, ca=hb, mode=1)

    def _slot(self, k):
        return DeviceArray(None, self.dtype, tensor=self.state.tensor[k])

    def admit(self, job, now):
        import torch
        k = len(self.jobs)
        with torch.cuda.stream(self.stream):
            g = self.g
            kind, fill = (0, 0.0) if job.kind == "wm" else (1, 1.0)
            self.mod.launch("aw_serve_init", (max(1, self.N * self.N // 256),), (256,),
                            [self.W1.ptr, ctypes.c_int(self.N), ctypes.c_int(self.N), ctypes.c_double(g.dX),
                             ctypes.c_double(g.dP), ctypes.c_int(kind), ctypes.c_double(job.x0),
                             ctypes.c_double(job.p0), ctypes.c_double(fill)], current_stream())
            # W /= (complex) sum W dXdP (set_wignerfunction, :415), then wigner2rho (:418)
            st = self.plan1.stats(self.W1)
            self.plan1.scale_from_stats(st.ptr, 8, 0, 1, g.dXdP, self.dscale)
            self.plan1.scale(self.W1, self.dscale)
            wigner2rho(self.plan1, g, self.W1, self._slot(k))
            self._factors(k, job)
        self.jobs.append(job)
        self.remaining.append(job.steps)
        job.t_admit = now

    def tick(self, chunk):
        """Advance all active slots by n = min(chunk, min remaining) merged steps."""
        import torch
        if not self.jobs:
            return 0
        n = int(min(chunk, min(self.remaining)))
        N = self.N
        with torch.cuda.stream(self.stream):
            self.plan.steps(self.state, n, self.colL, self.colE, self.rowL, self.rowE,
                            col_line_bstride=3 * N, col_elem_bstride=3 * N, row_line_bstride=N,
                            row_elem_bstride=N, const_steps=True, batch=len(self.jobs))
        self.remaining = [r - n for r in self.remaining]
        return n

    def finish(self, clock):
        """Finalise slots whose steps are done: enqueue W = rho2wigner, W /= Re sum W dXdP and
        the copy of W to pinned host memory on the pool's stream, then compact the slots.  The
        host reads the result once the copy's event has completed (``collect``), so it keeps
        launching the other pools' work meanwhile."""
        import torch
        k = len(self.jobs) - 1
        while k >= 0:
            if self.remaining[k] == 0:
                job = self.jobs[k]
                with torch.cuda.stream(self.stream):
                    g = self.g
                    rho2wigner(self.plan1, g, self._slot(k), self.W1)
                    st = self.plan1.stats(self.W1)
                    self.plan1.scale_from_stats(st.ptr, 8, 0, 0, g.dXdP, self.dscale)
                    self.plan1.scale(self.W1, self.dscale)
                    host = (self._free_host.pop() if self._free_host
                            else torch.empty(self.W1.tensor.shape, dtype=self.W1.tensor.dtype, pin_memory=True))
                    host.copy_(self.W1.tensor, non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record(self.stream)
                self.pending.append((job, host, ev))
                last = len(self.jobs) - 1
                if k != last:
                    # stream order keeps these copies after the transform read slot k
                    with torch.cuda.stream(self.stream):
                        self.state.tensor[k].copy_(self.state.tensor[last])
                        for t in (self.colL, self.colE, self.rowL, self.rowE):
                            t.tensor[k].copy_(t.tensor[last])
                    self.jobs[k] = self.jobs[last]
                    self.remaining[k] = self.remaining[last]
                self.jobs.pop()
                self.remaining.pop()
            k -= 1

    def collect(self, clock, block=False):
        """Jobs whose result has reached host memory (all pending ones with ``block``)."""
        done, keep = [], []
        for job, host, ev in self.pending:
            if block:
                ev.synchronize()
            if not (block or ev.query()):
                keep.append((job, host, ev))
                continue
            W = host.numpy()
            g = self.g
            job.summary = dict(norm=float(np.sum(W.real) * g.dXdP),
                               x_mean=float(np.sum(W.real * g.kX[None, :]) * g.dXdP),
                               p_mean=float(np.sum(W.real * g.kP[:, None]) * g.dXdP))
            if job.keep_result:
                job.result = W.copy()
            job.t_done = clock()
            self._free_host.append(host)
            done.append(job)
        self.pending = keep
        return done


class Server:
    """Admit jobs by HBM budget into per-(grid, precision, kind) slot pools and advance them in
    batched, stream-overlapped rounds (continuous batching)."""

    def __init__(self, hbm_budget_bytes=32 << 30, chunk=32, max_slots=64, L=8.0, grids=(128, 256, 512, 1024)):
        require_gpu()
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk)
        self.max_slots = int(max_slots)
        self.L = float(L)
        self.grids = tuple(grids)
        self.pools = {}
        self.stats = collections.Counter()

    def _capacity(self, N, precision):
        """Slots per pool: an equal share of the HBM budget per grid size and precision."""
        slot = N * N * (16 if precision == "fp64" else 8)
        share = self.budget // (2 * len(self.grids) * 2)
        return int(max(1, min(self.max_slots, share // slot)))

    def _pool(self, job):
        key = (job.N, job.precision, job.kind)
        pool = self.pools.get(key)
        if pool is None:
            pool = _Pool(job.N, job.precision, job.kind, self._capacity(job.N, job.precision), job.L)
            self.pools[key] = pool
        return pool

    def run(self, jobs, realtime=True):
        """Serve ``jobs`` (arrival times in seconds from the start); returns them finished."""
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0)
        incoming = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        waiting = collections.deque()
        finished = []
        while incoming or waiting or any(p.jobs or p.pending for p in self.pools.values()):
            now = clock()
            while incoming and (incoming[0].arrival <= now or not realtime):
                waiting.append(incoming.popleft())
            for _ in range(len(waiting)):
                job = waiting.popleft()
                pool = self._pool(job)
                if len(pool.jobs) < pool.capacity:
                    pool.admit(job, now)
                    self.stats["admitted"] += 1
                else:
                    waiting.append(job)
            active = [p for p in self.pools.values() if p.jobs]
            for p in active:
                self.stats["gp_steps"] += p.tick(self.chunk) * len(p.jobs) * p.N * p.N
                self.stats["rounds"] += 1
            for p in active:
                p.finish(clock)
            # results that reached the host; wait for them only when nothing else can run
            idle = not any(p.jobs for p in self.pools.values()) and not waiting
            for p in self.pools.values():
                finished.extend(p.collect(clock, block=idle and not incoming))
            if idle and incoming and not any(p.pending for p in self.pools.values()):
                time.sleep(max(0.0, incoming[0].arrival - clock()))
        self.stats["wall_s"] = clock()
        return finished


def poisson_jobs(n, rate=200.0, seed=0, grids=(128, 256, 512, 1024), precisions=("mixed", "fp64"),
                 steps_range=(200, 5000), beta_range=(0.5, 5.0), keep_result=False):
    """The serving workload: Poisson arrivals (mean ``rate`` per second) of WM / WB jobs with
    grid in ``grids``, V = X^2/2 + c4 X^4 with c4 ~ U(0, 0.2), beta ~ U(0.5, 5), steps ~
    U{200..5000}, FP32/FP64 tiers; seeded."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        kind = "wm" if rng.random() < 0.5 else "wb"
        jobs.append(Job(kind, int(rng.choice(grids)), precision=str(rng.choice(precisions)),
                        c4=float(rng.uniform(0.0, 0.2)), steps=int(rng.integers(steps_range[0], steps_range[1] + 1)),
                        beta=float(rng.uniform(*beta_range)), x0=float(rng.uniform(-1.5, 1.5)),
                        p0=float(rng.uniform(-1.0, 1.0)), arrival=t, keep_result=keep_result, job_id=i))
    return jobs
