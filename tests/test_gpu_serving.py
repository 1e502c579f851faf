"""GPU serving: the graph executor against the plain path, graph replays, and
the alone-vs-mix equality of the continuous-batching server.

Tolerance of the alone-vs-mix and executor tests: bitwise equality (0).  A
request's maps go through the same kernels with the same per-map arithmetic
whatever batch it joins -- no operation mixes maps -- so a result computed
alone and one computed in a mixed batch must be identical.  Against the FP64
oracle the FP32 bar is 1e-5 relative L2."""
import importlib

import numpy as np
import pytest
import torch

from oracle import lenseflow as olf

pytestmark = pytest.mark.gpu

TH, NSTEPS = 2.0, 7


def _case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def srv(pkg):
    return importlib.import_module("cmblensing_amd.serving")


@pytest.fixture(scope="module")
def lib(pkg):
    return importlib.import_module("cmblensing_amd._lib")


def _inputs(sims, N, C, seeds):
    f = sims.t_map if C == 1 else sims.qu_map
    arr = [np.asarray(f(N, TH, s)).reshape(C, N, N) for s in seeds]
    return np.stack(arr)


@pytest.mark.parametrize("N,C", [(64, 1), (64, 2), (256, 2), (128, 1), (512, 1), (512, 2)])
@pytest.mark.parametrize("direction,adjoint", [(1, False), (-1, False), (1, True)])
def test_executor_matches_plain(pkg, lib, sims, dev, N, C, direction, adjoint):
    _case(f"executor N={N} C={C} dir={direction} adj={adjoint}")
    count, cls = 3, 4
    x = torch.as_tensor(_inputs(sims, N, C, [1, 2, 3]), dtype=torch.float32, device=dev)
    phis = [torch.as_tensor(sims.phi_map(N, TH, 50 + i), device=dev) for i in range(count)]
    plan = lib.Plan(N, TH, torch.float32, cls * C, dev)
    handles = [lib.PhiHandle(plan, p, NSTEPS) for p in phis]
    ex = lib.Executor(plan, cls, C, NSTEPS, direction, adjoint)
    out = torch.full_like(x, float("nan"))
    ex.run(handles, [x[b].data_ptr() for b in range(count)], [out[b].data_ptr() for b in range(count)],
           torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    for b in range(count):
        alone = pkg.lense_flow_matrices(x[b:b + 1], phis[b], TH, NSTEPS, direction, adjoint)
        assert torch.equal(out[b], alone[0]), f"request {b} differs from the plain path"
    # and the plain path itself is at the FP32 bar against the FP64 oracle
    if not adjoint:
        ref = olf.lense_flow(x[0].double().cpu().numpy().squeeze(), phis[0].cpu().numpy(),
                             olf.pix_width_from_theta(TH), NSTEPS, direction)
        got = out[0].double().cpu().numpy().squeeze()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
    ex.close()
    for h in handles:
        h.close()
    plan.close()


def test_graph_replays_stable(pkg, lib, sims, dev):
    """Every replay of one executor, with eager work and synchronisations in
    between and a different set of requests each time, equals the eager path."""
    _case("graph replays N=128 QU")
    N, C, cls = 128, 2, 8
    plan = lib.Plan(N, TH, torch.float32, cls * C, dev)
    phis = [torch.as_tensor(sims.phi_map(N, TH, 70 + i), device=dev) for i in range(3)]
    handles = [lib.PhiHandle(plan, p, NSTEPS) for p in phis]
    ex = lib.Executor(plan, cls, C, NSTEPS, 1, False)
    xs = torch.as_tensor(_inputs(sims, N, C, range(10, 18)), dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    for rep in range(6):
        count = [8, 5, 1, 8, 3, 7][rep]
        idx = [(rep + b) % 8 for b in range(count)]
        hs = [handles[(rep + b) % 3] for b in range(count)]
        out = torch.empty((count, C, N, N), device=dev)
        ex.run(hs, [xs[i].data_ptr() for i in idx], [out[b].data_ptr() for b in range(count)], s)
        junk = torch.randn(4096, device=dev).cumsum(0)  # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk).all()
        for b in range(count):
            alone = pkg.lense_flow_matrices(xs[idx[b]:idx[b] + 1], phis[(rep + b) % 3], TH,
                                            NSTEPS, 1, False)
            assert torch.equal(out[b], alone[0]), f"replay {rep} request {b}"
    ex.close()
    for h in handles:
        h.close()
    plan.close()


def _mixed_requests(srv, sims, dev, n, sizes, weights, nphi, seed):
    desc = srv.mixed_workload(n, seed=seed, sizes=sizes, weights=weights, nphi=nphi)
    pool = {}
    for N in sizes:
        for C in (1, 2):
            pool[(N, C)] = torch.as_tensor(_inputs(sims, N, C, range(4)), dtype=torch.float32,
                                           device=dev)
    reqs = [srv.Request(x=pool[(N, C)][i % 4], phi_id=phi, direction=d)
            for i, (N, C, d, phi) in enumerate(desc)]
    phi_maps = {(k, N): torch.as_tensor(sims.phi_map(N, TH, 1000 * N + k), device=dev)
                for N in sizes for k in range(nphi)}
    return reqs, phi_maps


def test_server_mix_equals_alone(pkg, srv, sims, dev):
    """Continuous batching of a mixed Poisson stream over every served size
    (N = 64 ... 512): every result equals the same request lensed alone by the
    plain path, bit for bit."""
    _case("server mix vs alone, 240 requests N in {64,128,256,512}")
    sizes = (64, 128, 256, 512)
    reqs, phi_maps = _mixed_requests(srv, sims, dev, 240, sizes, (0.25, 0.35, 0.25, 0.15), 4,
                                     seed=5)
    assert sum(r.N == 512 for r in reqs) >= 20
    be = srv.GpuBackend(dev, sizes=sizes, nstreams=2,
                        max_batch={64: 32, 128: 16, 256: 8, 512: 4})
    server = srv.Server(be, lambda k, N: phi_maps[(k, N)], hbm_budget=8 * srv.GiB)
    offs = srv.poisson_schedule(len(reqs), rate=20000.0, seed=6)
    st = server.serve(reqs, list(offs))
    torch.cuda.synchronize()
    print(f"[stats] {st}", flush=True)
    assert st["completed"] == len(reqs)
    assert st["mean_batch"] > 1.0  # requests were packed together
    assert st["phi_misses"] == len({(r.phi_id, r.N) for r in reqs})
    for i, r in enumerate(reqs):
        alone = pkg.lense_flow_matrices(r.x[None], phi_maps[(r.phi_id, r.N)], TH, NSTEPS,
                                        r.direction, False)
        assert torch.equal(r.out, alone[0]), f"request {i} {r.key} differs alone vs in the mix"
        assert r.done >= r.arrival
    # FP32 bar against the FP64 oracle on one T and one Q/U request of each size
    for N, C in ((N, C) for N in sizes for C in (1, 2)):
        r = next(q for q in reqs if q.N == N and q.C == C)
        ref = olf.lense_flow(r.x.double().cpu().numpy(), phi_maps[(r.phi_id, N)].cpu().numpy(),
                             olf.pix_width_from_theta(TH), NSTEPS, r.direction)
        got = r.out.double().cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"[oracle] N={N} C={C} dir={r.direction} rel L2 {err:.2e}", flush=True)
        assert err < 1e-5
    be.close()


def test_server_small_budget(pkg, srv, sims, dev):
    """Under a budget that holds a handful of requests, admission waits for
    memory, the budget is never exceeded and every result is still right."""
    _case("server small budget")
    sizes = (64, 128)
    reqs, phi_maps = _mixed_requests(srv, sims, dev, 60, sizes, (0.5, 0.5), 2, seed=8)
    be = srv.GpuBackend(dev, sizes=sizes, nstreams=2, max_batch={64: 8, 128: 8})
    budget = be.fixed_bytes() + 4 * be.phi_bytes(128) + 6 * 2 * 2 * 128 * 128 * 4
    server = srv.Server(be, lambda k, N: phi_maps[(k, N)], hbm_budget=budget)
    for r in reqs:
        server.submit(r)
    while server.busy():
        server.step()
        assert server.used <= server.budget
    torch.cuda.synchronize()
    assert all(r.state == "done" for r in reqs)
    for r in reqs:
        alone = pkg.lense_flow_matrices(r.x[None], phi_maps[(r.phi_id, r.N)], TH, NSTEPS,
                                        r.direction, False)
        assert torch.equal(r.out, alone[0])
    be.close()


def test_server_wiener_filter_requests(pkg, srv, sims, dev):
    """Wiener-filter requests admitted beside a LenseFlow stream: each solve
    runs in the Wiener-filter lane (worker thread, own stream), converges to
    the 1e-6 true residual and equals the same solve run directly; the
    LenseFlow requests still equal their plain results bit for bit."""
    from cmblensing_amd import matrix_operators as mo, wiener_filter as wfm
    _case("server with Wiener-filter requests N=64 QU")
    N, th = 64, TH
    sizes = (64, 128)
    reqs, phi_maps = _mixed_requests(srv, sims, dev, 80, sizes, (0.5, 0.5), 2, seed=11)
    wf_reqs, direct = [], []
    for k in range(2):
        prob = sims.wf_problem(N, th, pol=True, seed=30 + k, mask="fourier")
        fq = np.fft.irfft2(np.stack(olf.eb2qu(prob["f"][0], prob["f"][1], N, th)))
        lensed = olf.lense_flow(fq, phi_maps[(k, N)].cpu().numpy(), olf.pix_width_from_theta(th),
                                NSTEPS)
        prob = sims.wf_data(prob, lensed)
        T = lambda a, dt=torch.complex64: torch.as_tensor(a, device=dev).to(dt)
        op = lambda A: mo.DiagonalEB(matrix_EE=T(A[0], torch.float64),
                                     matrix_BB=T(A[1], torch.float64), theta_pix=th)
        data = pkg.FlatS2(polar_matrix_1=T(prob["d"][0]), polar_matrix_2=T(prob["d"][1]),
                          theta_pix=th)
        ops = (op(prob["Cf"]), op(prob["Cn"]), op(prob["M"]), op(prob["B"]))
        wf_reqs.append(srv.WFRequest(data=data, phi_id=k, ops=ops, n=NSTEPS, rtol=1e-6))
        phif = pkg.FlatS0(scalar_matrix=phi_maps[(k, N)], theta_pix=th)
        direct.append(wfm.wiener_filter(data * 0, phif, data, *ops, n=NSTEPS, tol=0.0, rtol=1e-6,
                                        recompute_every=20, maxiter=500))
    be = srv.GpuBackend(dev, sizes=sizes, nstreams=2, max_batch={64: 16, 128: 16})
    server = srv.Server(be, lambda k, N_: phi_maps[(k, N_)], hbm_budget=8 * srv.GiB)
    allreq = reqs[:40] + wf_reqs + reqs[40:]
    offs = srv.poisson_schedule(len(allreq), rate=5000.0, seed=12)
    st = server.serve(allreq, list(offs))
    torch.cuda.synchronize()
    print(f"[stats] {st}", flush=True)
    assert st["wf_completed"] == 2 and st["completed"] == len(reqs)
    for w, x in zip(wf_reqs, direct):
        assert w.state == "done" and all(w.info.converged)
        assert max(w.info.final_true_rel_residual) < 1e-6
        for nm in ("polar_matrix_1", "polar_matrix_2"):
            assert torch.equal(getattr(w.out, nm), getattr(x, nm))
    for r in reqs:
        alone = pkg.lense_flow_matrices(r.x[None], phi_maps[(r.phi_id, r.N)], TH, NSTEPS,
                                        r.direction, False)
        assert torch.equal(r.out, alone[0])
    be.close()
