"""Continuous-batching serving of Wigner-Moyal / Wigner-Bloch jobs on one MI355X.

Jobs of different grid sizes, potentials and temperatures arrive over time.  For every
(grid, precision, kind) the server keeps a resident slot pool: one [capacity, N, N] state
tensor with per-slot factor tables.  Each scheduling round advances every non-empty pool by
one chunk of merged Strang steps in ONE batched launch sequence over its active slots
(aw_step_sequence with per-slot tables and constant-step mode).  Pools run on their own HIP
streams, so the work of different grid sizes overlaps on the device.  Arrived jobs are
admitted into free slots between rounds; finished jobs are transformed back to W, copied
out, and their slot is refilled by compaction (the last active slot moves into the hole).

Admission is by one shared HBM budget: a pool is created with one slot and grows (doubling,
up to ``max_slots``) from the free part of the budget when jobs wait for it; pools with no
work are released when another pool needs their memory.  So a grid size that is alone in
the mix can take the whole budget, and a 1024^2 job is admitted while idle pools of other
sizes hold memory.

Job family (config 5 of the serving workload): K = P^2/2, V = X^2/2 + c4 X^4 on
x, p in [-L, L).  "wm" jobs propagate the Gaussian W0 = exp(-(x-x0)^2 - (p-p0)^2)/pi for
`steps` real-time steps of dt (wigner_moyal_cuda_1d.py:543-582, normalised once at the end
as propagate does without Ehrenfest).  "wb" jobs run the Bloch propagation of
wigner_bloch_cuda_1d.py:119-155 from the infinite-temperature state to beta = 1/kT in
steps of dbeta, normalised at the end: per-step renormalisation only rescales, so the
result is the same up to rounding.  The purity guard of get_gibbs_state (stop with a
warning at the first step whose purity is not below 1, :146-153) is checked after every
chunk of steps from Re <L^H 1, rho> and sum |rho|^2 (one weighted-sum launch over the pool):
the purity of the normalised Bloch state Z(2 beta) / Z(beta)^2 does not decrease with beta,
so a step inside a chunk exceeds the limit only if the chunk's last step does.  A job that
exceeds it is taken out of its pool and run alone through WignerBlochCUDA1D, which stops at
the first violating step as the reference does (``job.gibbs_stopped_at``).
Since V >= 0 and K >= 0 vanish at X = P = 0, which are grid points, V_min = K_min = 0.
"""
import collections
import ctypes
import time

import numpy as np

from . import rtc
from .device import DeviceArray, current_stream, empty, require_gpu, to_device
from .engine import Plan
from .wm_ops import WMGrid, rho2wigner, rho2wigner_adjoint, wigner2rho

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
                 dbeta=0.005, x0=1.0, p0=0.0, arrival=0.0, keep_result=True, L=8.0, job_id=None,
                 purity_limit=1.0):
        assert kind in ("wm", "wb")
        self.kind, self.N, self.precision = kind, int(N), precision
        self.c4, self.dt, self.dbeta, self.x0, self.p0, self.L = float(c4), float(dt), float(dbeta), x0, p0, float(L)
        self.dbeta0 = self.dbeta
        self.purity_limit = float(purity_limit)     # the Gibbs loop's guard (wb jobs)
        self.gibbs_stopped_at = None
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
        self.guard_tripped = False

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


def _c4_str(c4):
    return "%.17g" % c4


class _Pool:
    """Resident slot pool for one (N, precision, kind) with per-slot factor tables.  It starts
    with no slots; ``resize`` grows it (the server decides by the shared budget)."""

    def __init__(self, N, precision, kind, L):
        import torch
        self.N, self.precision, self.kind, self.L = N, precision, kind, L
        self.capacity = 0
        self.stream = torch.cuda.Stream()
        with torch.cuda.stream(self.stream):
            self.plan = None
            self.plan1 = Plan(N, N, 1, precision)
            self.dtype = self.plan1.state_dtype
            self.fdtype = self.plan1.factor_dtype    # factor tables follow the arithmetic
            self.state_t = "double2" if self.dtype == np.complex128 else "float2"
            self.mod = _serve_module(self.state_t)
            self.g = WMGrid(N, L, N, L)
            self.state = None
            self.colL = self.colE = self.rowL = self.rowE = None
            self.W1 = DeviceArray((N, N), self.dtype)
            self.vk = empty((4, N), np.float64)
            self.dscale = empty((1, 2), np.float64)
            self.coords = [to_device(self.g.kXX), to_device(self.g.kXXp), to_device(self.g.kPP),
                           to_device(self.g.kPPp)]
            self.g1 = None
            if kind == "wb":
                # L^H 1 (FP64): Re <L^H 1, rho> = Re sum W, the norm of the purity guard
                g1 = DeviceArray((1, N, N), np.complex128)
                g1.fill(1.0)
                rho2wigner_adjoint(Plan(N, N, 1, "fp64"), self.g, g1)
                self.g1 = g1
        self.jobs = []
        self.remaining = []
        self.pending = []        # (job, pinned host W, event) of results still being copied out
        self._free_host = []     # pinned host buffers to reuse
        self.checks = []         # (event, pinned host sums, jobs by slot) of the purity guard
        self.rerun = []          # wb jobs that tripped the guard, to be run alone
        self.peak_active = 0

    @staticmethod
    def slot_cost(N, precision):
        """HBM bytes per slot: the state and its 8 factor vectors."""
        esz = 16 if precision == "fp64" else 8
        fsz = 8 if precision == "mixed" else 16
        return N * N * esz + 8 * N * fsz

    @staticmethod
    def fixed_cost(N, precision, kind):
        """HBM bytes a pool holds besides its slots: the W staging array (and L^H 1)."""
        esz = 16 if precision == "fp64" else 8
        return N * N * esz + (N * N * 16 if kind == "wb" else 0)

    @property
    def bytes(self):
        return self.fixed_cost(self.N, self.precision, self.kind) + self.capacity * self.slot_cost(self.N, self.precision)

    @property
    def idle(self):
        return not (self.jobs or self.pending or self.checks or self.rerun)

    def resize(self, capacity):
        """Grow to ``capacity`` slots; the active slots and their tables move along (on the
        pool's stream, so after every queued use of the old arrays)."""
        import torch
        N, n = self.N, len(self.jobs)
        with torch.cuda.stream(self.stream):
            state = DeviceArray((capacity, N, N), self.dtype)
            tabs = [empty((capacity, 3, N), self.fdtype), empty((capacity, 3, N), self.fdtype),
                    empty((capacity, 1, N), self.fdtype), empty((capacity, 1, N), self.fdtype)]
            if n:
                state.tensor[:n].copy_(self.state.tensor[:n])
                for new, old in zip(tabs, (self.colL, self.colE, self.rowL, self.rowE)):
                    new.tensor[:n].copy_(old.tensor[:n])
            self.plan = Plan(N, N, capacity, self.precision)
        self.state = state
        self.colL, self.colE, self.rowL, self.rowE = tabs
        self.capacity = capacity

    def _factors(self, k, job):
        """Per-slot tables of the merged Strang step (constant-step layout 0 open, 1 mid, 2 close)."""
        N, p = self.N, self.plan
        esz = p.fsize
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
        assert k < self.capacity
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
        self.peak_active = max(self.peak_active, len(self.jobs))
        job.t_admit = now

    def tick(self, chunk):
        """Advance all active slots by n = min(chunk, min remaining) merged steps; for Bloch
        jobs also queue the purity guard's sums of the new states."""
        import torch
        if not self.jobs:
            return 0
        n = int(min(chunk, min(self.remaining)))
        N, B = self.N, len(self.jobs)
        with torch.cuda.stream(self.stream):
            self.plan.steps(self.state, n, self.colL, self.colE, self.rowL, self.rowE,
                            col_line_bstride=3 * N, col_elem_bstride=3 * N, row_line_bstride=N,
                            row_elem_bstride=N, const_steps=True, batch=B)
            if self.kind == "wb":
                # [Re <L^H 1, rho>, sum |rho|^2] of every slot (inactive ones are ignored)
                sums = empty((self.capacity, 2), np.float64)
                self.plan.weighted_sums(self.state, self.g1, 1, sums, norm2=True)
                host = torch.empty((self.capacity, 2), dtype=torch.float64, pin_memory=True)
                host.copy_(sums.tensor, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(self.stream)
                self.checks.append((ev, host, list(self.jobs), sums))
        self.remaining = [r - n for r in self.remaining]
        return n

    def check_purity(self, block):
        """Read the guard's sums that have arrived (all of them with ``block``): a job whose
        purity 2 pi Np sum |rho|^2 dXdP / (Re sum W dXdP)^2 is not below its limit
        (wigner_bloch_cuda_1d.py:146-153 on the normalised W) is marked to run alone."""
        g, N = self.g, float(self.N)
        while self.checks:
            ev, host, jobs, _ = self.checks[0]
            if block:
                ev.synchronize()
            elif not ev.query():
                break
            self.checks.pop(0)
            s = host.numpy()
            for k, job in enumerate(jobs):
                norm = s[k, 0] * g.dXdP
                purity = 2. * np.pi * N * s[k, 1] * g.dXdP / (norm * norm)
                if not purity < job.purity_limit:
                    job.guard_tripped = True

    def _remove(self, k):
        """Compaction: the last active slot moves into slot k (stream order keeps the copies
        after every queued use of slot k)."""
        import torch
        last = len(self.jobs) - 1
        if k != last:
            with torch.cuda.stream(self.stream):
                self.state.tensor[k].copy_(self.state.tensor[last])
                for t in (self.colL, self.colE, self.rowL, self.rowE):
                    t.tensor[k].copy_(t.tensor[last])
            self.jobs[k] = self.jobs[last]
            self.remaining[k] = self.remaining[last]
        self.jobs.pop()
        self.remaining.pop()

    def finish(self, clock):
        """Finalise slots whose steps are done: enqueue W = rho2wigner, W /= Re sum W dXdP and
        the copy of W to pinned host memory on the pool's stream, then compact the slots.  The
        host reads the result once the copy's event has completed (``collect``), so it keeps
        launching the other pools' work meanwhile.  Bloch jobs are finalised only after the
        guard has seen their last chunk; those that tripped it leave for ``rerun``."""
        import torch
        if self.kind == "wb":
            self.check_purity(block=any(r == 0 for r in self.remaining))
        k = len(self.jobs) - 1
        while k >= 0:
            job = self.jobs[k]
            if job.guard_tripped:
                self.rerun.append(job)
                self._remove(k)
            elif self.remaining[k] == 0:
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
                self._remove(k)
            k -= 1

    def _summarise(self, job, W, clock):
        g = self.g
        job.summary = dict(norm=float(np.sum(W.real) * g.dXdP),
                           x_mean=float(np.sum(W.real * g.kX[None, :]) * g.dXdP),
                           p_mean=float(np.sum(W.real * g.kP[:, None]) * g.dXdP))
        if job.keep_result:
            job.result = np.array(W, copy=True)
        job.t_done = clock()

    def collect(self, clock, block=False):
        """Jobs whose result has reached host memory (all pending ones with ``block``)."""
        done, keep = [], []
        for job, host, ev in self.pending:
            if block:
                ev.synchronize()
            if not (block or ev.query()):
                keep.append((job, host, ev))
                continue
            self._summarise(job, host.numpy(), clock)
            self._free_host.append(host)
            done.append(job)
        self.pending = keep
        return done

    def run_alone(self, clock):
        """Bloch jobs that tripped the purity guard, each run alone through
        WignerBlochCUDA1D.get_gibbs_state (same grid, V, kT, dbeta and limit), which stops at
        the first violating step with the reference's warning."""
        from .wigner_bloch_cuda_1d import WignerBlochCUDA1D
        done = []
        while self.rerun:
            job = self.rerun.pop(0)
            wb = WignerBlochCUDA1D(X_gridDIM=job.N, X_amplitude=job.L, P_gridDIM=job.N, P_amplitude=job.L,
                                   kT=1. / job.beta, dbeta=job.dbeta0, t_initial=0., t=0.,
                                   V="0.5 * X * X + %s * X * X * X * X" % _c4_str(job.c4), K="0.5 * P * P",
                                   precision=job.precision)
            wb.purity_limit = job.purity_limit
            W = wb.get_gibbs_state().get()
            job.gibbs_stopped_at = wb.gibbs_stopped_at
            self._summarise(job, W, clock)
            done.append(job)
        return done


class Server:
    """Admit jobs into per-(grid, precision, kind) slot pools against one shared HBM budget and
    advance the pools in batched, stream-overlapped rounds (continuous batching)."""

    def __init__(self, hbm_budget_bytes=32 << 30, chunk=32, max_slots=64, L=8.0, grids=(128, 256, 512, 1024)):
        require_gpu()
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk)
        self.max_slots = int(max_slots)
        self.L = float(L)
        self.grids = tuple(grids)
        self.pools = {}
        self.stats = collections.Counter()

    def used_bytes(self):
        return sum(p.bytes for p in self.pools.values())

    def _release_idle(self, keep):
        """Drop pools that hold no work (their HBM returns to the budget)."""
        for key in [k for k, p in self.pools.items() if k != keep and p.idle]:
            del self.pools[key]
            self.stats["released"] += 1

    def _try_admit(self, job, now):
        """Admit ``job`` into its pool, growing the pool from the free budget (doubling, up to
        max_slots) and releasing idle pools when the budget is short; False: it waits."""
        key = (job.N, job.precision, job.kind)
        pool = self.pools.get(key)
        per = _Pool.slot_cost(job.N, job.precision)
        fixed = 0 if pool is not None else _Pool.fixed_cost(job.N, job.precision, job.kind)
        if pool is not None and len(pool.jobs) < pool.capacity:
            pool.admit(job, now)
            return True
        cap = pool.capacity if pool is not None else 0
        if cap >= self.max_slots:
            return False
        want = min(self.max_slots, max(1, 2 * cap)) - cap
        if self.budget - self.used_bytes() < fixed + want * per:
            self._release_idle(key)
        extra = min(want, (self.budget - self.used_bytes() - fixed) // per)
        if extra < 1:
            return False
        if pool is None:
            pool = self.pools[key] = _Pool(job.N, job.precision, job.kind, job.L)
        pool.resize(cap + int(extra))
        self.stats["grown"] += 1
        pool.admit(job, now)
        return True

    def run(self, jobs, realtime=True):
        """Serve ``jobs`` (arrival times in seconds from the start); returns them finished."""
        for j in jobs:
            need = _Pool.fixed_cost(j.N, j.precision, j.kind) + _Pool.slot_cost(j.N, j.precision)
            if need > self.budget:
                raise ValueError("job %s needs %d bytes of HBM, above the budget of %d" % (j.id, need, self.budget))
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0)
        incoming = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        waiting = collections.deque()
        finished = []
        while incoming or waiting or any(not p.idle for p in self.pools.values()):
            now = clock()
            while incoming and (incoming[0].arrival <= now or not realtime):
                waiting.append(incoming.popleft())
            for _ in range(len(waiting)):
                job = waiting.popleft()
                if self._try_admit(job, now):
                    self.stats["admitted"] += 1
                else:
                    waiting.append(job)
            active = [p for p in self.pools.values() if p.jobs]
            for p in active:
                self.stats["gp_steps"] += p.tick(self.chunk) * len(p.jobs) * p.N * p.N
                self.stats["rounds"] += 1
            for p in active:
                p.finish(clock)
            for p in list(self.pools.values()):
                if p.rerun:
                    self.stats["guard_reruns"] += len(p.rerun)
                    finished.extend(p.run_alone(clock))
                if not p.jobs and p.checks:
                    p.check_purity(block=False)
            # results that reached the host; wait for them only when nothing else can run
            idle = not any(p.jobs for p in self.pools.values()) and not waiting
            for p in self.pools.values():
                finished.extend(p.collect(clock, block=idle and not incoming))
                if idle and not incoming and p.checks:
                    p.check_purity(block=True)
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
