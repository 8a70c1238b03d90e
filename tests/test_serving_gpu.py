"""Continuous-batching server (r2splat_amd.serving): jobs admitted as they
arrive, stepped together, each job's result bitwise the same alone and in
the mix (Gaussian and Beta jobs alike); the served steps agree with the
eager explicit-API loop (torch Adam); admission respects the HBM budget."""

import pytest
import torch

from oracle import splat as O
from r2splat_amd.serving import ExplicitFit, FitRequest, FitServer, build_job, job_config
from r2splat_amd.train import AdamConfig, GaussianFit

pytestmark = pytest.mark.gpu

ADAM = AdamConfig(lr=1e-3, min_scale=1e-6)


def _requests():
    return [FitRequest(0, 0.0, 800, 3, 48, 101, False, 5), FitRequest(1, 0.0, 1500, 2, 64, 103, False, 4),
            FitRequest(2, 2.0, 600, 4, 40, 105, True, 3), FitRequest(3, 3.0, 1000, 3, 56, 107, False, 6)]


def _alone(req, device):
    """The request's job fitted alone, with the kernel configuration the server gives it."""
    job = build_job(req, device)
    cfg = job_config(req)
    fit = GaussianFit(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"], job["Ks"],
                      job["target"], adam=ADAM, beta=job["beta"], tile_size=cfg.tile_size, fwd_waves=cfg.fwd_waves,
                      bwd_waves=cfg.bwd_waves)
    fit.capture()
    fit.step(req.steps)
    fit.sync()
    return fit.params.clone()


@pytest.mark.parametrize("policy", [{}, {"max_active": 2}, {"active_ms": 0.2}])
def test_each_result_is_the_same_alone_and_in_the_mix(device, policy):
    """Every job equals the job fitted alone whatever steps with it; with an
    earliest-arrival active set (2 jobs, or ~0.2 ms of estimated work: the
    test jobs estimate 0.15-0.16 ms each, so 2 jobs) the early jobs finish first."""
    assert all(0.15 <= FitServer.estimate_ms(r) < 0.2 for r in _requests())
    server = FitServer(device, n_streams=2, adam=ADAM, **policy)
    stats = server.run(_requests(), clock="ticks", keep_results=True)
    assert sorted(f[0] for f in stats.finished) == [0, 1, 2, 3]
    assert max(f[3] - f[1] for f in stats.finished) > 0
    if policy:  # the earliest arrivals step first: they finish before the last arrival
        order = [f[0] for f in sorted(stats.finished, key=lambda f: f[3])]
        assert order.index(0) < order.index(3) and order.index(1) < order.index(3)
    for req in _requests():
        assert torch.equal(stats.results[req.rid], _alone(req, device)), req.rid


def test_jobs_run_with_the_configuration_of_their_shape(device):
    seen = {}
    server = FitServer(device, n_streams=2, adam=ADAM)
    admit = server._admit

    def record(req, now):
        ok = admit(req, now)
        if ok:
            fit = server.live[-1].fit
            seen[req.rid] = (fit.tile_size, fit.fwd_waves, fit.bwd_waves)
        return ok

    server._admit = record
    server.run(_requests(), clock="ticks")
    assert seen == {r.rid: (job_config(r).tile_size, job_config(r).fwd_waves, job_config(r).bwd_waves)
                    for r in _requests()}


@pytest.mark.parametrize("beta", [False, True])
def test_served_steps_agree_with_eager_explicit_loop(device, beta):
    req = FitRequest(0, 0.0, 700, 3, 48, 109, beta, 4)
    server = FitServer(device, n_streams=1, adam=ADAM)
    got = server.run([req], clock="ticks", keep_results=True).results[0]
    ref = ExplicitFit(build_job(req, device), ADAM)
    for _ in range(req.steps):
        ref.step()
    exp = torch.cat([p.detach().reshape(-1) for p in ref.params])
    assert O.relative_error(got, exp) <= 2e-6


def test_admission_waits_for_memory(device):
    reqs = _requests()[:2]
    budget = max(FitServer.estimate_bytes(r) for r in reqs) + 1  # one job at a time
    server = FitServer(device, n_streams=2, adam=ADAM, hbm_budget=budget)
    stats = server.run(reqs, clock="ticks", keep_results=True)
    (r0, _, adm0, fin0, _, _), (r1, _, adm1, fin1, _, _) = sorted(stats.finished)
    assert adm1 >= fin0 - 1e-9  # the second job was admitted only after the first finished
    assert torch.equal(stats.results[1], _alone(reqs[1], device))


def test_job_stalls_when_its_capacity_cannot_grow_and_resumes(device, monkeypatch):
    """A job whose capacity growth runs out of device memory stalls (its
    overflowed steps stay skipped), is retried at the next synchronisation and
    ends bitwise equal to the job fitted alone."""
    req = FitRequest(0, 0.0, 800, 3, 48, 113, False, 6)
    expected = _alone(req, device)
    monkeypatch.setattr(GaussianFit, "_measure_capacity", lambda self: 1024)  # the first step overflows
    calls = []
    regrow = GaussianFit.regrow

    def flaky(self):
        calls.append(len(calls))
        if len(calls) == 1:
            raise torch.OutOfMemoryError("simulated: no device memory to grow the capacity")
        return regrow(self)

    monkeypatch.setattr(GaussianFit, "regrow", flaky)
    server = FitServer(device, n_streams=1, adam=ADAM)
    stats = server.run([req], clock="ticks", keep_results=True)
    assert len(calls) >= 2  # one failed growth, then a successful retry
    assert [f[0] for f in stats.finished] == [0]
    assert torch.equal(stats.results[0], expected)


def test_job_survives_a_failed_recapture_after_growth(device, monkeypatch):
    """The capacity grows, then re-capturing the step runs out of device memory
    once: the skipped steps stay counted (not lost), the next synchronisation
    re-captures without growing again, replays them, and the job ends bitwise
    equal to the job fitted alone."""
    req = FitRequest(0, 0.0, 800, 3, 48, 117, True, 6)
    expected = _alone(req, device)
    monkeypatch.setattr(GaussianFit, "_measure_capacity", lambda self: 1024)  # the first step overflows
    events = []
    capture, regrow = GaussianFit.capture, GaussianFit.regrow

    def flaky_capture(self):
        if events == ["capture", "grow"]:  # the re-capture right after the growth
            events.append("capture failed")
            raise torch.OutOfMemoryError("simulated: no device memory to re-capture the step")
        events.append("capture")
        return capture(self)

    def counted_regrow(self):
        events.append("grow")
        return regrow(self)

    monkeypatch.setattr(GaussianFit, "capture", flaky_capture)
    monkeypatch.setattr(GaussianFit, "regrow", counted_regrow)
    server = FitServer(device, n_streams=1, adam=ADAM)
    stats = server.run([req], clock="ticks", keep_results=True)
    # one growth, the failed capture, then a re-capture without a second growth
    assert events[:4] == ["capture", "grow", "capture failed", "capture"], events
    assert [f[0] for f in stats.finished] == [0]
    assert torch.equal(stats.results[0], expected)
