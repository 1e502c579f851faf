"""Scheduler of the serving loop on the host: admission under an HBM budget,
resident phi reuse and eviction, bucketing, batch limits, continuous batching
across streams, cancellation.  A host-side backend stands in for the GPU: it
"lenses" a request by a known function of (x, phi, direction) and completes a
batch after a fixed number of polls, so every property is checked exactly."""
import importlib

import pytest
import torch

FWD, INV = 1, -1


@pytest.fixture(scope="module")
def srv(pkg):
    return importlib.import_module("cmblensing_amd.serving")


class HostBackend:
    """Backend with the GpuBackend interface; result = x * (phi_id + 1) * dir."""

    def __init__(self, nstreams=2, sizes=(16, 32), max_batch=None, latency_polls=3,
                 fixed=1000, phi_nbytes=None, free=None):
        self.nstreams, self.sizes = nstreams, tuple(sizes)
        self.mb = max_batch or {16: 4, 32: 2}
        self.latency_polls = latency_polls
        self.fixed = fixed
        self.phi_nbytes = phi_nbytes or {16: 5000, 32: 20000}
        self.free = free
        self.prepared, self.released, self.launches = [], [], []
        self.clock = 0.0
        self.concurrent_max = 0
        self.running = 0

    def free_bytes(self):
        return self.free

    def fixed_bytes(self):
        return self.fixed

    def request_bytes(self, req):
        return 2 * req.x.numel() * 4

    def phi_bytes(self, N):
        return self.phi_nbytes[N]

    def max_batch(self, key):
        return self.mb[key[0]]

    def prepare_phi(self, N, phi):
        h = ("phi", float(phi), N)
        self.prepared.append(h)
        return h

    def release_phi(self, h):
        self.released.append(h)

    def alloc_output(self, req):
        return torch.empty_like(req.x)

    def start_clock(self):
        import time
        return time.perf_counter()

    def launch(self, key, reqs, handles, si):
        assert len(reqs) <= self.mb[key[0]]
        assert all(r.key == key for r in reqs)
        for r, h in zip(reqs, handles):
            assert h[1] == r.phi_id and h[2] == r.N
            r.out.copy_(r.x * (r.phi_id + 1) * r.direction)
        self.running += 1
        self.concurrent_max = max(self.concurrent_max, self.running)
        tok = {"polls": self.latency_polls, "reqs": reqs, "stream": si}
        self.launches.append(tok)
        return tok

    def ready(self, tok):
        tok["polls"] -= 1
        if tok["polls"] <= 0 and not tok.get("fin"):
            tok["fin"] = True
            self.running -= 1
        return tok["polls"] <= 0

    def finish_time(self, tok):
        import time
        return time.perf_counter()

    # Wiener-filter lane: "solves" by scaling the data by 3
    wf_slots = 1

    def wf_bytes(self, req):
        return 4096

    def launch_wf(self, req, phi):
        self.running += 1
        self.concurrent_max = max(self.concurrent_max, self.running)
        tok = {"polls": self.latency_polls, "reqs": [req], "stream": -1, "wf": True}
        self.launches.append(tok)
        return tok

    def wf_result(self, tok, req):
        req.out = req.data * 3
        req.info = type("Info", (), {"iterations": [7]})()
        return None


def _req(srv, N, C, phi, d=FWD, seed=0):
    g = torch.Generator().manual_seed(seed)
    return srv.Request(x=torch.randn(C, N, N, generator=g), phi_id=phi, direction=d)


def _phi_source(phi_id, N):
    return float(phi_id)


def test_every_request_done_with_its_own_result(srv):
    be = HostBackend()
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9)
    reqs = [s.submit(_req(srv, [16, 32][i % 2], 1 + (i // 2) % 2, i % 3, [FWD, INV][i % 5 % 2], i))
            for i in range(40)]
    s.drain()
    assert all(r.state == "done" for r in reqs)
    for r in reqs:
        assert torch.equal(r.out, r.x * (r.phi_id + 1) * r.direction)
    st = s.stats()
    assert st["completed"] == 40
    # every request's bytes returned; the resident phi stay charged
    assert s.used == be.fixed + sum(e.nbytes for e in s.phis.values())
    assert all(e.refs == 0 for e in s.phis.values())
    # batches are homogeneous and within the bucket limit (checked in launch)
    assert sum(len(t["reqs"]) for t in be.launches) == 40


def test_phi_operators_resident_and_hit_rate(srv):
    be = HostBackend()
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9)
    for i in range(30):
        s.submit(_req(srv, 16, 1, i % 3, FWD, i))
    s.drain()
    assert len(be.prepared) == 3 and not be.released
    assert s.misses == 3 and s.hits == 27
    assert s.stats()["phi_hit_rate"] == pytest.approx(0.9)


def test_admission_respects_budget(srv):
    # each 16x16 T request holds 2*256*4 = 2048 B; phi 5000 B; fixed 1000 B
    be = HostBackend(latency_polls=2)
    budget = 1000 + 5000 + 3 * 2048
    s = srv.Server(be, _phi_source, hbm_budget=budget)
    reqs = [s.submit(_req(srv, 16, 1, 0, FWD, i)) for i in range(12)]
    peak_running = 0
    while s.busy():
        s.step()
        assert s.used <= budget
        peak_running = max(peak_running, sum(r.state in ("queued", "running") for r in reqs))
    assert all(r.state == "done" for r in reqs)
    assert peak_running == 3
    assert s.peak_used <= budget


def test_budget_evicts_unused_phi_lru(srv):
    be = HostBackend(latency_polls=1)
    budget = 1000 + 2 * 5000 + 2048
    s = srv.Server(be, _phi_source, hbm_budget=budget)
    for phi in (0, 1, 2, 0):
        r = s.submit(_req(srv, 16, 1, phi, FWD, phi))
        s.drain()
        assert r.state == "done"
    # phi 0 was least recently used when phi 2 arrived, then phi 1 for phi 0
    assert [h[1] for h in be.released] == [0.0, 1.0]
    assert s.evictions == 2 and s.misses == 4


def test_oversized_request_rejected(srv):
    be = HostBackend()
    s = srv.Server(be, _phi_source, hbm_budget=1000 + 5000 + 1000)
    r = s.submit(_req(srv, 16, 1, 0))
    s.drain()
    assert r.state == "rejected" and "budget" in r.error


def test_free_memory_caps_budget(srv):
    be = HostBackend(free=20000)
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9)
    assert s.budget == 20000


def test_continuous_batching_packs_late_arrivals(srv):
    """Requests that arrive while a batch runs are launched together in the
    next batch, and batches on both streams overlap."""
    be = HostBackend(nstreams=2, max_batch={16: 8, 32: 8}, latency_polls=4)
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9, max_inflight=1)
    first = [s.submit(_req(srv, 16, 1, 0, FWD, i)) for i in range(2)]
    s.step()                      # launches the first batch on stream 0
    assert len(be.launches) == 1 and len(be.launches[0]["reqs"]) == 2
    late = [s.submit(_req(srv, 16, 1, 1, FWD, 10 + i)) for i in range(5)]
    s.step()                      # stream 1 is free: the 5 late ones go together
    assert len(be.launches) == 2 and len(be.launches[1]["reqs"]) == 5
    assert be.launches[1]["stream"] != be.launches[0]["stream"]
    assert be.concurrent_max == 2
    s.drain()
    assert all(r.state == "done" for r in first + late)


def test_oldest_bucket_first_and_batch_cap(srv):
    be = HostBackend(nstreams=1, max_batch={16: 3, 32: 2}, latency_polls=1)
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9, max_inflight=1)
    a = s.submit(_req(srv, 32, 2, 0, INV, 0))
    bs = [s.submit(_req(srv, 16, 1, 0, FWD, i)) for i in range(7)]
    s.drain()
    sizes = [(t["reqs"][0].N, len(t["reqs"])) for t in be.launches]
    assert sizes == [(32, 1), (16, 3), (16, 3), (16, 1)]
    assert a.state == "done" and all(b.state == "done" for b in bs)


def test_cancel_waiting_queued_running(srv):
    be = HostBackend(nstreams=1, max_batch={16: 2, 32: 2}, latency_polls=3)
    budget = 1000 + 5000 + 3 * 2048
    s = srv.Server(be, _phi_source, hbm_budget=budget, max_inflight=1)
    reqs = [s.submit(_req(srv, 16, 1, 0, FWD, i)) for i in range(5)]
    s.step()  # admits 3 (budget), launches 2
    assert [r.state for r in reqs] == ["running", "running", "queued", "waiting", "waiting"]
    assert s.cancel(reqs[1].rid)       # running: result discarded
    assert s.cancel(reqs[2].rid)       # queued: dropped, memory back
    assert s.cancel(reqs[4].rid)       # waiting: dropped
    assert not s.cancel(12345)
    s.drain()
    assert [r.state for r in reqs] == ["done", "cancelled", "cancelled", "done", "cancelled"]
    assert reqs[1].out is None
    assert s.used == be.fixed + 5000
    assert len(s.completed) == 2


def test_serve_open_loop_latency_stats(srv):
    be = HostBackend(latency_polls=2)
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9)
    offs = srv.poisson_schedule(50, rate=20000.0, seed=2)
    desc = srv.mixed_workload(50, seed=3, sizes=(16, 32), weights=(0.5, 0.5), nphi=4)
    reqs = [_req(srv, N, C, phi, d, i) for i, (N, C, d, phi) in enumerate(desc)]
    st = s.serve(reqs, list(offs))
    assert st["completed"] == 50
    assert 0 <= st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    for r in reqs:
        assert r.done >= r.arrival
        assert torch.equal(r.out, r.x * (r.phi_id + 1) * r.direction)


def test_mixed_workload_distribution(srv):
    desc = srv.mixed_workload(20000, seed=2)
    import numpy as np
    Ns = np.array([d[0] for d in desc])
    for N, p in zip((64, 128, 256, 512), (0.2, 0.4, 0.3, 0.1)):
        assert abs((Ns == N).mean() - p) < 0.015
    assert abs(np.mean([d[1] == 2 for d in desc]) - 0.5) < 0.015
    assert abs(np.mean([d[2] == FWD for d in desc]) - 0.5) < 0.015
    assert {d[3] for d in desc} == set(range(32))
    offs = srv.poisson_schedule(20000, 2000.0, seed=2)
    assert abs(np.diff(offs).mean() - 1 / 2000) < 2e-5


def test_batch_class(srv):
    assert [srv._batch_class(c, 128) for c in (1, 2, 3, 5, 64, 65, 128)] == [1, 2, 4, 8, 64, 128, 128]
    assert srv._batch_class(20, 16) == 16


class _Field:
    """Minimal stand-in for a Fourier field (matrix names, scaling)."""

    def __init__(self, t):
        self.t = t

    def _matrix_names(self):
        return ["t"]

    def __mul__(self, k):
        return _Field(self.t * k)


def test_wiener_filter_requests_share_admission_and_overlap(srv):
    """WF requests pass the same budget, run one at a time in their own lane,
    and LenseFlow batches keep launching while one runs."""
    be = HostBackend(nstreams=1, latency_polls=4)
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9, max_inflight=1)
    F = _Field
    w1 = s.submit(srv.WFRequest(data=F(torch.ones(2, 16, 9)), phi_id=0, ops=(None,) * 4))
    w2 = s.submit(srv.WFRequest(data=F(torch.ones(2, 16, 9) * 2), phi_id=1, ops=(None,) * 4))
    lf = [s.submit(_req(srv, 16, 1, 0, FWD, i)) for i in range(3)]
    s.step()
    assert w1.state == "running" and w2.state == "queued"      # one WF at a time
    assert any(r.state == "running" for r in lf)               # LenseFlow batch alongside
    assert s.used == be.fixed + 2 * 4096 + 5000 + 3 * 2048
    s.drain()
    assert w1.state == w2.state == "done"
    assert torch.equal(w1.out.t, torch.ones(2, 16, 9) * 3)
    assert all(r.state == "done" for r in lf)
    st = s.stats()
    assert st["wf_completed"] == 2 and st["completed"] == 3
    assert s.used == be.fixed + 5000


def test_cancel_queued_wiener_filter(srv):
    be = HostBackend(nstreams=1, latency_polls=3)
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9)
    F = _Field
    a = s.submit(srv.WFRequest(data=F(torch.ones(1, 16, 9)), phi_id=0, ops=(None,) * 4))
    b = s.submit(srv.WFRequest(data=F(torch.ones(1, 16, 9)), phi_id=0, ops=(None,) * 4))
    s.step()
    assert s.cancel(b.rid) and b.state == "cancelled"
    s.drain()
    assert a.state == "done" and b.out is None
    assert s.used == be.fixed


def test_caustic_phi_rejected(srv):
    """A phi whose magnification determinant reaches 0 is refused: its
    requests are rejected (prepared once, not kept), others proceed."""
    class Handle(tuple):
        pass

    class CausticBackend(HostBackend):
        def prepare_phi(self, N, phi):
            h = Handle(("phi", float(phi), N))
            h.min_det = -0.3 if float(phi) == 1.0 else 0.8
            self.prepared.append(h)
            return h

    be = CausticBackend()
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9)
    bad = [s.submit(_req(srv, 16, 1, 1, FWD, i)) for i in range(3)]
    good = s.submit(_req(srv, 16, 1, 0, FWD, 9))
    s.drain()
    assert all(r.state == "rejected" and "caustic" in r.error for r in bad)
    assert good.state == "done"
    assert len([h for h in be.prepared if h[1] == 1.0]) == 1 and len(be.released) == 1
    assert s.used == be.fixed + 5000


def test_batches_formed_together_group_requests_by_phi(srv):
    """A bucket holding more than one batch while several stream slots are
    free: its batches are formed in the same step and ordered by phi before
    the split, so the requests of one phi fall into one batch (their p(t)
    rows are then read once per batch); every request still launches at once
    and gets its own result."""
    be = HostBackend(nstreams=2, max_batch={16: 8, 32: 8}, latency_polls=3)
    s = srv.Server(be, _phi_source, hbm_budget=10 ** 9, max_inflight=1)
    # 16 requests in arrival order with phi 0, 1, 2, 3, 0, 1, ... (4 each)
    reqs = [s.submit(_req(srv, 16, 1, i % 4, FWD, i)) for i in range(16)]
    s.step()
    assert len(be.launches) == 2                      # both streams, same step
    assert {t["stream"] for t in be.launches} == {0, 1}
    assert [len(t["reqs"]) for t in be.launches] == [8, 8]
    phis = [sorted({r.phi_id for r in t["reqs"]}) for t in be.launches]
    assert phis == [[0, 1], [2, 3]]                   # each phi in one batch
    for t in be.launches:                             # positions ordered by phi
        ids = [r.phi_id for r in t["reqs"]]
        assert ids == sorted(ids)
    s.drain()
    assert all(r.state == "done" for r in reqs)
    for r in reqs:
        assert torch.equal(r.out, r.x * (r.phi_id + 1) * r.direction)
