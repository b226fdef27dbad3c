"""Serving tier: continuous batching of Wigner-Moyal / Wigner-Bloch jobs.

Each job's result is checked against the oracle (FP64 tier, relative max error 1e-10) and
against the same job served alone (FP64: 1e-12 relative; mixed: 2e-5 relative).  A job is
not bit-identical alone and in a mix because the chunk boundaries (where a merged
exp(-i dt V) kick is split into two half kicks) follow the other jobs' step counts.
"""
import numpy as np
import pytest

from oracle.awig_oracle import WignerBlochOracle, WignerMoyalOracle

pytestmark = pytest.mark.gpu


def _oracle(job):
    V = (lambda c4: (lambda X, t: 0.5 * X * X + c4 * X ** 4))(job.c4)
    K = lambda P, t: 0.5 * P * P
    grid = dict(X_gridDIM=job.N, X_amplitude=job.L, P_gridDIM=job.N, P_amplitude=job.L, V=V, K=K)
    if job.kind == "wm":
        o = WignerMoyalOracle(dt=job.dt, **grid)
        o.set_wignerfunction(lambda X, P: np.exp(-(X - job.x0) ** 2 - (P - job.p0) ** 2) / np.pi)
        o.propagate(job.steps)
    else:
        o = WignerBlochOracle(kT=1. / job.beta, dbeta=0.005, **grid)
        o.get_gibbs_state()
        assert not o.warnings
    return o.wignerfunction


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _jobs(precision, mix=True):
    from accuratewigner_amd.serving import Job
    target = Job("wm", 128, precision, c4=0.12, steps=150, x0=0.8, p0=-0.4, arrival=0.0, job_id=0)
    if not mix:
        return target, [target]
    others = [
        Job("wb", 128, precision, c4=0.05, beta=0.9, arrival=0.0, job_id=1),
        Job("wm", 128, precision, c4=0.02, steps=37, x0=-1.0, arrival=0.0, job_id=2),
        Job("wm", 256, precision, c4=0.18, steps=61, x0=0.3, p0=0.5, arrival=0.0, job_id=3),
        Job("wm", 128, precision, c4=0.07, steps=90, x0=0.1, arrival=0.001, job_id=4),
        Job("wb", 256, precision, c4=0.15, beta=0.6, arrival=0.002, job_id=5),
        Job("wm", 128, precision, c4=0.19, steps=20, x0=1.2, p0=0.2, arrival=0.003, job_id=6),
    ]
    return target, [target] + others


def test_serving_fp64_matches_oracle():
    from accuratewigner_amd.serving import Server
    _, jobs = _jobs("fp64")
    srv = Server(hbm_budget_bytes=1 << 30, chunk=16, max_slots=8)
    done = srv.run(jobs, realtime=False)
    assert sorted(j.id for j in done) == list(range(len(jobs)))
    for j in done:
        print("job", j.id, j.kind, j.N, j.steps, flush=True)
        ref = _oracle(j)
        e = _rel(j.result, ref)
        assert e < 1e-10, (j.id, e)
        assert abs(j.summary["norm"] - 1.0) < 1e-12
        assert j.latency is not None and j.latency >= 0.0


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-5), ("c64", 2e-6)])
def test_serving_alone_equals_in_mix(precision, tol):
    from accuratewigner_amd.serving import Server
    target, jobs = _jobs(precision)
    Server(hbm_budget_bytes=1 << 30, chunk=16, max_slots=8).run(jobs, realtime=False)
    in_mix = target.result
    alone, one = _jobs(precision, mix=False)
    Server(hbm_budget_bytes=1 << 30, chunk=16, max_slots=8).run(one, realtime=False)
    assert _rel(in_mix, alone.result) < tol
    from accuratewigner_amd.serving import Job
    for j in jobs[1:]:
        j2 = Job(j.kind, j.N, precision, c4=j.c4, steps=j.steps, beta=j.beta, x0=j.x0, p0=j.p0, job_id=100 + j.id)
        Server(hbm_budget_bytes=1 << 30, chunk=16, max_slots=8).run([j2], realtime=False)
        assert _rel(j.result, j2.result) < tol, (j.id, _rel(j.result, j2.result))


def test_serving_c64_within_the_error_bound_of_fp64():
    """c64 pools (complex64 slots, FP64 arithmetic, line fixed point inside each chunk's step
    sequence): every job of the mix within 1e-6 of the same job served alone in FP64."""
    from accuratewigner_amd.serving import Job, Server
    _, jobs = _jobs("c64")
    Server(hbm_budget_bytes=1 << 30, chunk=16, max_slots=8).run(jobs, realtime=False)
    for j in jobs:
        ref = Job(j.kind, j.N, "fp64", c4=j.c4, steps=j.steps, beta=j.beta, x0=j.x0, p0=j.p0, job_id=200 + j.id)
        Server(hbm_budget_bytes=1 << 30, chunk=16, max_slots=8).run([ref], realtime=False)
        assert _rel(j.result, ref.result) < 1e-6, (j.id, _rel(j.result, ref.result))


def test_serving_admission_waits_for_free_slots():
    """With two slots per pool, later jobs of one shape queue until a slot frees up; each
    result still equals the job served alone."""
    from accuratewigner_amd.serving import Job, Server
    jobs = [Job("wm", 128, "fp64", c4=0.03 * k, steps=30 + 7 * k, x0=0.2 * k, arrival=0.0, job_id=k)
            for k in range(5)]
    srv = Server(hbm_budget_bytes=1 << 30, chunk=8, max_slots=2)
    done = srv.run(jobs, realtime=False)
    assert len(done) == 5
    assert max(len(p.jobs) for p in srv.pools.values()) == 0
    for j in jobs:
        j2 = Job("wm", 128, "fp64", c4=j.c4, steps=j.steps, x0=j.x0, job_id=50 + j.id)
        Server(hbm_budget_bytes=1 << 30, chunk=8, max_slots=2).run([j2], realtime=False)
        assert _rel(j.result, j2.result) < 1e-12


def _alone(j, precision=None):
    from accuratewigner_amd.serving import Job, Server
    j2 = Job(j.kind, j.N, precision or j.precision, c4=j.c4, steps=j.steps, beta=j.beta, x0=j.x0, p0=j.p0,
             job_id=1000 + j.id, purity_limit=j.purity_limit)
    Server(hbm_budget_bytes=4 << 30, chunk=16, max_slots=8).run([j2], realtime=False)
    return j2


def test_serving_shared_budget_grows_pool_and_releases_idle():
    """One shared budget: four 1024^2 jobs run side by side in one pool (grown 1 -> 2 -> 4
    slots) after the idle 128^2 pool of an earlier batch is released for them.  A static
    split of this budget over the 16 (grid, precision, kind) pools would give the 1024^2
    pool one slot."""
    from accuratewigner_amd.serving import Job, Server, _Pool
    MiB = 1 << 20
    srv = Server(hbm_budget_bytes=41 * MiB, chunk=16, max_slots=8)
    small = [Job("wb", 128, "fp64", c4=0.01 * k, beta=0.1 + 0.02 * k, arrival=0.0, job_id=k) for k in range(16)]
    done = srv.run(small, realtime=False)
    assert len(done) == 16
    p128 = srv.pools[(128, "fp64", "wb")]
    assert p128.idle and p128.capacity == 8          # 1 -> 2 -> 4 -> 8 (max_slots)
    big = [Job("wm", 1024, "mixed", c4=0.05 * k, steps=12 + 3 * k, x0=0.3 * k, arrival=0.0, job_id=20 + k)
           for k in range(4)]
    done = srv.run(big, realtime=False)
    assert len(done) == 4
    pool = srv.pools[(1024, "mixed", "wm")]
    assert pool.peak_active == 4 and pool.capacity == 4
    assert srv.stats["released"] >= 1 and (128, "fp64", "wb") not in srv.pools
    assert srv.used_bytes() <= srv.budget
    assert 4 * _Pool.slot_cost(1024, "mixed") > srv.budget // 16
    for j in big:
        assert _rel(j.result, _alone(j).result) < 2e-5, j.id


def test_serving_realtime_large_grids_match_alone_and_oracle():
    """Real-time arrivals of 512^2 and 1024^2 jobs (both kinds, both tiers) next to 256^2
    ones: each result equals the job served alone (FP64 1e-12, mixed 2e-5); the FP64 WM jobs
    also match the oracle (1e-10)."""
    from accuratewigner_amd.serving import Job, Server
    spec = [("wm", 512, "fp64", dict(steps=60, c4=0.1, x0=0.5)),
            ("wm", 1024, "mixed", dict(steps=40, c4=0.15, x0=-0.7, p0=0.3)),
            ("wb", 512, "mixed", dict(beta=0.3, c4=0.08)),
            ("wb", 1024, "fp64", dict(beta=0.2, c4=0.12)),
            ("wm", 256, "fp64", dict(steps=90, c4=0.05, x0=1.1)),
            ("wm", 1024, "fp64", dict(steps=30, c4=0.02, x0=0.2, p0=-0.5)),
            ("wb", 256, "mixed", dict(beta=0.5, c4=0.19))]
    jobs = [Job(k, N, prec, arrival=0.008 * i, job_id=i, **kw) for i, (k, N, prec, kw) in enumerate(spec)]
    srv = Server(hbm_budget_bytes=2 << 30, chunk=16, max_slots=8)
    done = srv.run(jobs, realtime=True)
    assert sorted(j.id for j in done) == list(range(len(jobs)))
    for j in done:
        print("job", j.id, j.kind, j.N, j.precision, j.steps, "latency %.1f ms" % (j.latency * 1e3), flush=True)
        assert j.t_admit >= j.arrival and j.latency > 0
        assert abs(j.summary["norm"] - 1.0) < (1e-12 if j.precision == "fp64" else 1e-6)
        tol = 1e-12 if j.precision == "fp64" else 2e-5
        assert _rel(j.result, _alone(j).result) < tol, j.id
        if j.precision == "fp64" and j.kind == "wm":
            assert _rel(j.result, _oracle(j)) < 1e-10, j.id


def test_serving_purity_guard_runs_violating_job_alone():
    """A Bloch job whose purity passes its limit leaves the mix and stops at the first
    violating step, as WignerBlochCUDA1D.get_gibbs_state does (wigner_bloch_cuda_1d.py:146-153);
    the other jobs are unaffected."""
    from accuratewigner_amd import WignerBlochCUDA1D
    from accuratewigner_amd.serving import Job, Server
    guarded = Job("wb", 128, "fp64", c4=0.0, beta=2.0, purity_limit=0.5, job_id=0)
    others = [Job("wb", 128, "fp64", c4=0.1, beta=1.0, job_id=1),
              Job("wm", 128, "fp64", c4=0.1, steps=40, job_id=2)]
    srv = Server(hbm_budget_bytes=1 << 30, chunk=16, max_slots=8)
    done = srv.run([guarded] + others, realtime=False)
    assert len(done) == 3 and srv.stats["guard_reruns"] == 1
    assert guarded.guard_tripped and not any(j.guard_tripped for j in others)
    wb = WignerBlochCUDA1D(X_gridDIM=128, X_amplitude=8., P_gridDIM=128, P_amplitude=8., kT=0.5, dbeta=0.005,
                           t_initial=0., t=0., V="0.5 * X * X + 0 * X * X * X * X", K="0.5 * P * P")
    wb.purity_limit = 0.5
    W = wb.get_gibbs_state().get()
    assert wb.gibbs_stopped_at is not None and guarded.gibbs_stopped_at == wb.gibbs_stopped_at
    # harmonic purity tanh(beta / 2) first reaches 1/2 at beta = 2 artanh(1/2) = 1.0986
    assert abs((guarded.gibbs_stopped_at + 1) * 0.005 - 2 * np.arctanh(0.5)) < 0.01
    assert _rel(guarded.result, W) < 1e-12
    for j in others:
        assert _rel(j.result, _alone(j).result) < 1e-12
