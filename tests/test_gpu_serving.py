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


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-5)])
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
