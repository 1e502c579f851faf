"""LenseFlow throughput on one MI355X.

Workload: forward LenseFlow of Q/U map pairs, N = 256, 2 arcmin pixels, float32,
7 RK4 steps (the 256^2 QU configuration).  One "step" is one batched pass over
`--batch` requests; request r uses phi number r % 32 of 32 seeded lensing
potentials whose operators are prepared once and stay resident in HBM (the
serving form), and the requests of a step are ordered by phi before they
are split over the streams, as the serving backend forms the batches of a
bucket that become ready together.  The plain path is one request per call with its phi prepared
in the same call, as the reference's lense_flow does.

The timed step is the served form: the batch split over the serving
backend's two streams, each half one replay of a captured batched graph
(gather, 28 RK4 stages, scatter); the same batch through the direct batched
entry point on one stream is timed beside it.

Prints one JSON line (rank 0).  Extra keys: plain maps/s and the batched /
plain ratio, p50/p99 latency of plain single requests, relative L2 error of a
batched result against the FP64 oracle, the live roofline of the dominant
kernel, the CPU baseline (the FP64 oracle on this host's cores), and the
mixed serving stream (10k Poisson requests at 2k req/s of mixed N, T/QU,
lense/unlense: p50/p99 latency arrival -> done, maps/s, phi hit rate), and the
consumers of the path: gradients for 16 chains, one HMC proposal, the 512^2
Wiener filter with 8 right-hand sides, 64 simulated data sets and MAP-joint
(one BFGS run, 16 batched grid runs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def model_bytes_per_map(N, C, n, s):
    """Algorithmic bytes of one LenseFlow of one request under the canonical
    model: 4n [C (12H + 5P) + 5P] s."""
    P, H = N * N, N * (N // 2 + 1)
    return 4 * n * (C * (12 * H + 5 * P) + 5 * P) * s


def design_bytes_per_map(N, C, n, s):
    """Lower bound of the bytes one LenseFlow of one request moves in this
    two-pass design (every array read or written once per pass):
    4n [C (6.75 Q) + 2 Q] with Q = N^2 s.  Per component and stage: the
    column pass reads the stage spectra S and writes DX (2Q); the row pass
    reads S and DX and writes the next S (3Q) and moves 1.75Q of RK4 state on
    average (y read at every stage and written at stage 4, the accumulator
    written at stage 2 and read at stage 4); per request and stage the two
    p(t) maps (2Q), shared by the components."""
    Q = N * N * s
    return 4 * n * (C * 6.75 * Q + 2 * Q)


def row_bytes_per_launch(B, N, C, s):
    """Algorithmic bytes of one fused row-pass launch over B requests,
    averaged over the four RK4 stages: per request 2C packed half planes read
    (N^2 s each) and C written, the 2 maps of p(t), and per component 7 map
    transfers of the RK4 state per step (y read at every stage and written at
    stage 4, the accumulator written at stage 2 and read at stage 4), 1.75
    per stage."""
    return B * (4.75 * C + 2) * N * N * s


def adj_row_bytes_per_launch(B, N, C, s):
    """Algorithmic bytes of one adjoint-rate row-pass launch (low-storage
    bookkeeping) over B requests, averaged over the four RK4 stages: per
    component the x-derivative spectra U read and the next U written (2Q),
    the stage input z read and the next one written (2Q, the stage-1 read
    being y itself: -0.25Q), y read (1Q) and written at stage 4 (0.25Q), the
    accumulator written at stage 2 and read at stage 4 (0.5Q); per request
    the two p(t) maps (2Q).  (5.25 C + 2) Q with Q = N^2 s."""
    return B * (5.25 * C + 2) * N * N * s


def col_bytes_per_launch(B, N, C, s):
    """x-derivative column pass: per component one packed half plane (N^2 s
    bytes) read and one written."""
    return B * C * 2 * N * N * s


def cpu_baseline(f, phi, th, n, budget_s=10.0):
    from oracle import lenseflow as olf
    pw = olf.pix_width_from_theta(th)
    t0 = time.perf_counter()
    k = 0
    while True:
        olf.lense_flow(f, phi, pw, n)
        k += 1
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    return k / el, k, el


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads for the multi-core CPU baseline: the cores this process
    may run on, at most 16 (one GPU's share of the host)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline_threads(fs, phi, th, n, threads, budget_s=10.0):
    """The FP64 oracle over many requests at once: `threads` threads, each
    lensing one request at a time (numpy releases the interpreter lock in its
    FFTs and array arithmetic), until `budget_s` has passed.  Returns maps/s,
    the requests done and the elapsed time."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import lenseflow as olf
    pw = olf.pix_width_from_theta(th)
    t0 = time.perf_counter()
    done = [0] * threads

    def worker(w):
        i = w
        while time.perf_counter() - t0 < budget_s:
            olf.lense_flow(fs[i % len(fs)], phi, pw, n)
            done[w] += 1
            i += threads

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(worker, range(threads)))
    el = time.perf_counter() - t0
    return sum(done) / el, sum(done), el


LF_MODES = {"forward": (1, False), "inverse": (-1, False), "adjoint": (1, True),
            "inverse_adjoint": (-1, True)}


def config1_run(torch, _lib, plan, phis, x, B, C, n, steps, warmup, timed, sims, N, th, r_chk):
    """Q/U 256^2 FP32, 7 RK4 steps: batched maps/s of forward, inverse,
    adjoint and inverse adjoint LenseFlow (B requests, resident phi, direct
    batched entry point); the round trip ||L^-1 L f - f|| / ||f||; the adjoint
    dot-product test |<L y, x> - <y, L^T x>| / |<L y, x>| with x a second set
    of seeded Q/U fields (sums in float64).  Returns the leg and request
    r_chk's four outputs for the oracle check of the CPU leg."""
    out = torch.empty_like(x)
    leg, outs = {}, {}
    for name, (d, a) in LF_MODES.items():
        dt = timed(lambda: _lib.lenseflow(plan, phis, x, out, B, C, n, d, a), steps, warmup)
        leg[name] = {"batched_maps_s": round(B * steps / dt, 1),
                     "ms_per_batch": round(dt / steps * 1e3, 4)}
        outs[name] = out[r_chk].double().cpu().numpy()
    fwd, back = torch.empty_like(x), torch.empty_like(x)
    _lib.lenseflow(plan, phis, x, fwd, B, C, n, 1, False)
    _lib.lenseflow(plan, phis, fwd, back, B, C, n, -1, False)
    x64 = x.double().flatten(1)
    rt = ((back.double().flatten(1) - x64).norm(dim=1) / x64.norm(dim=1))
    xs = torch.as_tensor(np.stack([sims.qu_map(N, th, 300 + r % 8) for r in range(B)]),
                         dtype=x.dtype, device=x.device)
    ltx = torch.empty_like(x)
    _lib.lenseflow(plan, phis, xs, ltx, B, C, n, 1, True)
    lhs = (fwd.double().flatten(1) * xs.double().flatten(1)).sum(dim=1)
    rhs = (x64 * ltx.double().flatten(1)).sum(dim=1)
    dot_rel = ((lhs - rhs).abs() / lhs.abs())
    torch.cuda.synchronize()
    leg.update({
        "workload": "lenseflow_QU_256_2arcmin_n7_fp32 (forward, inverse, adjoint)",
        "batch": B,
        "round_trip_rel_l2_max": float(rt.max()),
        "round_trip_rel_l2_median": float(rt.median()),
        "adjoint_dot_test_rel_max": float(dot_rel.max()),
        "adjoint_dot_test_rel_median": float(dot_rel.median()),
    })
    return leg, outs


def serving_run(torch, sims, dev, n_requests, rate, seed=2, nphi=32, budget_gib=64.0,
                wf_every=0, pkg=None):
    """The mixed serving stream: Poisson arrivals at `rate`, N in {64, 128,
    256, 512} with weights {.2, .4, .3, .1}, T/QU and lense/unlense 50/50,
    phi ~ U{0..31} of 32 seeded potentials per N, float32, 2 arcmin pixels,
    7 RK4 steps, a 64 GiB HBM admission budget.  Inputs are resident in HBM
    (a pool of 8 seeded maps per (N, C)); phi operators start cold.  With
    wf_every = k, every k-th arrival is instead a lensed Wiener-filter
    request (Q/U at N = 64 or 128, Fourier-diagonal mask, FP32 operator with
    FP64 CG to a 1e-6 relative residual) on one of 4 seeded problems."""
    import importlib
    srv = importlib.import_module("cmblensing_amd.serving")
    th, n = 2.0, 7
    sizes = (64, 128, 256, 512)
    desc = srv.mixed_workload(n_requests, seed=seed, sizes=sizes, nphi=nphi)
    offs = srv.poisson_schedule(n_requests, rate, seed=seed)
    pool = {}
    for N in sizes:
        pool[(N, 1)] = torch.as_tensor(np.stack([sims.t_map(N, th, i) for i in range(8)]),
                                       dtype=torch.float32, device=dev)
        pool[(N, 2)] = torch.as_tensor(np.stack([sims.qu_map(N, th, i) for i in range(8)]),
                                       dtype=torch.float32, device=dev)
    phi_maps = {}
    for N in sizes:
        for k in range(nphi):
            phi_maps[(k, N)] = torch.as_tensor(sims.phi_map(N, th, 5000 + 100 * k + N % 97),
                                               device=dev)
    reqs = [srv.Request(x=pool[(N, C)][i % 8], phi_id=phi, direction=d)
            for i, (N, C, d, phi) in enumerate(desc)]
    if wf_every:
        mo = importlib.import_module("cmblensing_amd.matrix_operators")
        wf_pool = []
        for k, N in enumerate((64, 64, 128, 128)):
            pr = sims.wf_problem(N, th, pol=True, seed=60 + k, mask="fourier")
            pr["phi"] = phi_maps[(k, N)].cpu().numpy()      # the served phi lenses the truth
            pr = sims.wf_data(pr, _lensed_truth(torch, pkg, dev, pr, n))
            T = lambda a, dt=torch.complex64: torch.as_tensor(a, device=dev).to(dt)
            op = lambda A: mo.DiagonalEB(matrix_EE=T(A[0], torch.float64),
                                         matrix_BB=T(A[1], torch.float64), theta_pix=th)
            data = pkg.FlatS2(polar_matrix_1=T(pr["d"][0]), polar_matrix_2=T(pr["d"][1]),
                              theta_pix=th)
            wf_pool.append((k, data, (op(pr["Cf"]), op(pr["Cn"]), op(pr["M"]), op(pr["B"]))))
        for i in range(wf_every - 1, n_requests, wf_every):
            k, data, ops = wf_pool[(i // wf_every) % 4]
            reqs[i] = srv.WFRequest(data=data, phi_id=k, ops=ops, n=n, rtol=1e-6)
        # warm the Wiener-filter lane before traffic (plans, operator arrays,
        # phi tables and torch's kernels loaded on first use), as the graphs are
        wfm = importlib.import_module("cmblensing_amd.wiener_filter")
        fields = importlib.import_module("cmblensing_amd.fields")
        for k, data, ops in wf_pool:
            N = data.polar_matrix_1.shape[-2]
            wfm.wiener_filter(data * 0, fields.FlatS0(scalar_matrix=phi_maps[(k, N)], theta_pix=th),
                              data, *ops, n=n, tol=0.0, rtol=1e-6, recompute_every=20, maxiter=500)
        torch.cuda.synchronize()
    be = srv.GpuBackend(dev, theta_pix=th, nsteps=n, nstreams=2, sizes=sizes)
    t0 = time.perf_counter()
    be.warmup()
    warm_s = time.perf_counter() - t0
    server = srv.Server(be, lambda k, N: phi_maps[(k, N)], hbm_budget=int(budget_gib * srv.GiB))
    print(f"[bench] serving {n_requests} requests at {rate:.0f} req/s "
          f"(graphs captured in {warm_s:.1f} s)", file=sys.stderr, flush=True)
    st = server.serve(reqs, list(offs))
    torch.cuda.synchronize()
    # one served request of every (N, T/QU) against the same request lensed
    # alone by the plain path: must be bit-identical
    alone = {}
    if pkg is not None:
        for N in sizes:
            for C in (1, 2):
                r = next((q for q in reqs if isinstance(q, srv.Request) and q.N == N and q.C == C
                          and q.state == "done"), None)
                if r is not None:
                    a = pkg.lense_flow_matrices(r.x[None], phi_maps[(r.phi_id, N)], th, n,
                                                r.direction, False)[0]
                    alone[f"N{N}_{'QU' if C == 2 else 'T'}"] = bool(torch.equal(a, r.out))
    per_n = {}
    for N in sizes:
        lat = np.array([r.latency for r in reqs
                        if isinstance(r, srv.Request) and r.N == N and r.state == "done"])
        if len(lat):
            per_n[str(N)] = {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                             "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                             "requests": int(len(lat))}
    out = {"requests": n_requests, "offered_rate_req_s": rate,
           "completed": st["completed"], "maps_per_s": round(st.get("maps_per_s", 0.0), 1),
           "p50_ms": round(st.get("p50_ms", 0.0), 3), "p99_ms": round(st.get("p99_ms", 0.0), 3),
           "max_ms": round(st.get("max_ms", 0.0), 3),
           "batches": st["batches"], "mean_batch": round(st["mean_batch"], 2),
           "phi_hit_rate": round(st["phi_hit_rate"], 4), "phi_evictions": st["phi_evictions"],
           "hbm_budget_gib": round(st["hbm_budget_bytes"] / srv.GiB, 2),
           "hbm_peak_gib": round(st["hbm_peak_bytes"] / srv.GiB, 3),
           "graph_warmup_s": round(warm_s, 2), "per_nside": per_n,
           "served_equals_alone": alone}
    if wf_every:
        out.update({k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()
                     if k.startswith("wf_")})
    be.close()
    return out


def _gpu_problem(torch, pkg, sims, dev, prob, cdt):
    """GPU fields and operators of a numpy problem from sims.wf_problem /
    sims.wf_data (EB Fourier field, MAP phi, per-mode operators, pixel mask)."""
    import importlib
    mo = importlib.import_module("cmblensing_amd.matrix_operators")
    th = prob["theta_pix"]
    rdt = torch.float32 if cdt == torch.complex64 else torch.float64
    T = lambda a, dt=cdt: torch.as_tensor(a, device=dev).to(dt)
    op = lambda A: mo.DiagonalEB(matrix_EE=T(A[0], torch.float64), matrix_BB=T(A[1], torch.float64),
                                 theta_pix=th)
    f = pkg.FlatS2(polar_matrix_1=T(prob["f"][0]), polar_matrix_2=T(prob["f"][1]), theta_pix=th)
    data = pkg.FlatS2(polar_matrix_1=T(prob["d"][0], rdt), polar_matrix_2=T(prob["d"][1], rdt),
                      theta_pix=th)
    mask = mo.PixelMask(matrix=torch.as_tensor(prob["mask_pix"], device=dev))
    return f, data, op(prob["Cf"]), op(prob["Cn"]), mask, op(prob["B"])


def _lensed_truth(torch, pkg, dev, prob, n):
    """Q/U maps of L(phi) f for a problem's truth, lensed by the float64 GPU
    path (the data model d = M B L f + n)."""
    th = prob["theta_pix"]
    f = pkg.FlatS2(polar_matrix_1=torch.as_tensor(prob["f"][0], device=dev),
                   polar_matrix_2=torch.as_tensor(prob["f"][1], device=dev), theta_pix=th)
    q = pkg.map(pkg.eb2qu(f))
    x = torch.stack([q.polar_matrix_1, q.polar_matrix_2])[None]
    y = pkg.lense_flow_matrices(x, torch.as_tensor(prob["phi"], device=dev), th, n, 1, False)
    return y[0].cpu().numpy()


def _marker(torch, dev):
    """A launch that appears nowhere else in these workloads (row R2C of one
    16^2 map: k_row_r2c<float, 16>), to cut a leg out of a kernel trace."""
    import importlib
    fields = importlib.import_module("cmblensing_amd.fields")
    fields._rfft2(torch.zeros((16, 16), dtype=torch.float32, device=dev), 2.0)


def wf_run(torch, pkg, sims, dev, N=512, th=1.5, nrhs=8, n=7, rtol=1e-6, markers=False,
           fp64_solve=True):
    """Lensed Wiener filter, Q/U N=512 at 1.5 arcmin, 8 right-hand sides
    (signal, noise and phi seeded 100..107), 1 uK-arcmin white noise with a
    l_knee = 100 1/f term, 200 point-source holes of 5 arcmin, 1 arcmin beam.
    FP32 operator with FP64 CG vectors, dots, alpha/beta and FP64 residual
    recomputation, to a relative residual of 1e-6; compared with the same
    solve run entirely in FP64 to 1e-10."""
    import gc
    import importlib
    wfm = importlib.import_module("cmblensing_amd.wiener_filter")
    probs = []
    for i in range(nrhs):
        pr = sims.wf_problem(N, th, pol=True, seed=100 + i, mask="holes", noise_uk=1.0,
                             ell_knee=100.0, beam_fwhm=1.0)
        probs.append(sims.wf_data(pr, _lensed_truth(torch, pkg, dev, pr, n)))

    def batched(cdt):
        objs = [_gpu_problem(torch, pkg, sims, dev, p, cdt) for p in probs]
        f0, _, Cf, Cn, M, Bm = objs[0]
        st = lambda name, j: torch.stack([getattr(o[j], name) for o in objs])
        data = objs[0][1].replace(polar_matrix_1=st("polar_matrix_1", 1),
                                  polar_matrix_2=st("polar_matrix_2", 1))
        x0 = f0.replace(polar_matrix_1=torch.zeros_like(st("polar_matrix_1", 0)),
                        polar_matrix_2=torch.zeros_like(st("polar_matrix_2", 0)))
        phi = pkg.FlatS0(scalar_matrix=torch.stack([torch.as_tensor(p["phi"], device=dev)
                                                    for p in probs]), theta_pix=th)
        return x0, phi, data, Cf, Cn, M, Bm

    args32 = batched(torch.complex64)
    # the interpreter's garbage from earlier legs (the 10k-request stream's
    # objects) is collected before the timed solve, not during it
    gc.collect()
    # warm-up: one whole untimed solve (loads the kernels of every launch
    # shape the solve uses, including the shrinking active batches)
    wfm.wiener_filter(*args32, n=n, tol=0.0, rtol=rtol, recompute_every=20, maxiter=1000)
    torch.cuda.synchronize()
    if markers:
        _marker(torch, dev)
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    x32, info = wfm.wiener_filter(*args32, n=n, tol=0.0, rtol=rtol, recompute_every=20,
                                  maxiter=1000, return_info=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if markers:
        _marker(torch, dev)
        torch.cuda.synchronize()
    its = info.iterations
    leg = {"workload": "wf_QU_512_1.5arcmin_8rhs_holes_1overf", "nrhs": nrhs,
           "iterations": its, "converged": all(info.converged),
           "final_true_rel_residual_fp64": max(info.final_true_rel_residual),
           "time_s": round(dt, 4), "ms_per_iteration": round(dt / max(its) * 1e3, 3)}
    if not fp64_solve:
        return leg
    args64 = batched(torch.complex128)
    t1 = time.perf_counter()
    x64, info64 = wfm.wiener_filter(*args64, n=n, tol=0.0, rtol=1e-10, maxiter=3000,
                                    return_info=True)
    torch.cuda.synchronize()
    dt64 = time.perf_counter() - t1
    errs = []
    for i in range(nrhs):
        a = torch.stack([x32.polar_matrix_1[i], x32.polar_matrix_2[i]]).to(torch.complex128)
        b = torch.stack([x64.polar_matrix_1[i], x64.polar_matrix_2[i]])
        errs.append(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)))
    leg.update({"fp64_solve": {"iterations": info64.iterations, "time_s": round(dt64, 3),
                               "final_true_rel_residual": max(info64.final_true_rel_residual)},
                "rel_l2_vs_fp64_solve_max": max(errs)})
    return leg


def grad_run(torch, pkg, sims, dev, N=256, th=2.0, chains=16, n=7, reps=5):
    """d logpdf / d(f, phi) for 16 chains, Q/U N=256 at 2 arcmin, float32:
    1 arcmin beam, 5 uK-arcmin white noise, apodised square mask of 80% area;
    timed per evaluation of both gradients, compared with the float64 path."""
    import importlib
    grads = importlib.import_module("cmblensing_amd.gradients")
    stats = importlib.import_module("cmblensing_amd.statistics")
    mo = importlib.import_module("cmblensing_amd.matrix_operators")
    pr = sims.wf_problem(N, th, pol=True, seed=21, mask="square", noise_uk=5.0, beam_fwhm=1.0)
    pr = sims.wf_data(pr, _lensed_truth(torch, pkg, dev, pr, n))
    phi_m, Cphi = sims.phi_map_and_cov(N, th, 2)

    def objs(cdt):
        f, data, Cf, Cn, M, Bm = _gpu_problem(torch, pkg, sims, dev, pr, cdt)
        scale = torch.tensor([1 + 0.01 * k for k in range(chains)], device=dev).to(cdt)[:, None, None]
        f = f.replace(polar_matrix_1=f.polar_matrix_1[None] * scale,
                      polar_matrix_2=f.polar_matrix_2[None] * scale)
        phiF = pkg.FlatS0(scalar_matrix=torch.as_tensor(np.fft.rfft2(phi_m), device=dev).to(cdt),
                          theta_pix=th)
        Cp = mo.DiagonalScalar(scalar_matrix=torch.as_tensor(Cphi, device=dev), theta_pix=th)
        return f, phiF, data, Cf, Cn, M, Bm, Cp

    def both(o):
        # one forward LenseFlow and one gradient ODE give both gradients
        f, phiF, data, Cf, Cn, M, Bm, Cp = o
        return stats.grad_logpdf(f, phiF, data, Cn, Cp, Cf, M, Bm, n=n)

    def separate(o):
        f, phiF, data, Cf, Cn, M, Bm, Cp = o
        gf = grads.gradf_logpdf(f, phiF, data, Cf, Cn, M, Bm, n=n)
        gp = stats.grad_phi_logpdf(f, phiF, data, Cn, Cp, Cf, M, Bm, n=n)
        return gf, gp

    def timed(fn, o):
        fn(o)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            res = fn(o)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, res

    o32 = objs(torch.complex64)
    dt, (gf, gp) = timed(both, o32)
    dt_sep, _ = timed(separate, o32)
    # the FP64 reference through the separate calls (forward + adjoint
    # LenseFlow for d/d f): an independent route to the same gradients
    gf64, gp64 = separate(objs(torch.complex128))

    def rel(a, b):
        a = a.to(torch.complex128)
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    return {"workload": "grad_logpdf_QU_256_2arcmin_16chains", "chains": chains,
            "ms_per_evaluation": round(dt * 1e3, 3),
            "chain_gradients_per_s": round(chains / dt, 1),
            "ms_per_evaluation_separate_calls": round(dt_sep * 1e3, 3),
            "rel_l2_gradf_vs_fp64": max(rel(gf.polar_matrix_1, gf64.polar_matrix_1),
                                        rel(gf.polar_matrix_2, gf64.polar_matrix_2)),
            "rel_l2_grad_phi_vs_fp64": rel(gp.scalar_matrix, gp64.scalar_matrix)}


def hmc_run(torch, pkg, sims, dev, N=256, th=2.0, chains=16, n=7, num_steps=30,
            step_size=1e-3):
    """One HMC proposal (30 leapfrog steps) for 16 chains of the mixed-
    parametrisation phi posterior, T at N = 256, in FP64 like the reference
    (the accept test compares Hamiltonians of size |logpdf|): 31 mixed phi
    gradients and two Hamiltonians per chain.  Mass matrix G^-2 (Cphi^-1 +
    Nphi^-1) with Nphi = Cphi / 100 standing in for the quadratic-estimator
    noise; the step size is tuned to this posterior (0.01, the reference's
    default, diverges here)."""
    import importlib
    mixing = importlib.import_module("cmblensing_amd.mixing")
    sampling = importlib.import_module("cmblensing_amd.sampling")
    mo = importlib.import_module("cmblensing_amd.matrix_operators")
    pr = sims.wf_problem(N, th, pol=False, seed=41, mask="fourier")
    phim, Cphi = sims.phi_map_and_cov(N, th, 42)
    f0 = pkg.FlatS0(scalar_matrix=torch.as_tensor(pr["f"][0], device=dev), theta_pix=th)
    y = pkg.lense_flow_matrices(pkg.map(f0).scalar_matrix[None, None],
                                torch.as_tensor(phim, device=dev), th, n, 1, False)
    pr = sims.wf_data(pr, y[0].cpu().numpy())
    T = lambda a: torch.as_tensor(a, device=dev).to(torch.complex128)
    R = lambda a: mo.DiagonalScalar(scalar_matrix=torch.as_tensor(a, device=dev), theta_pix=th)
    D = np.sqrt(pr["Cf"] / (pr["Cf"] + pr["Cn"]))
    f = pkg.FlatS0(scalar_matrix=T(np.stack([pr["f"][0]] * chains)), theta_pix=th)
    d = pkg.FlatS0(scalar_matrix=T(pr["d"][0]), theta_pix=th)
    phis = np.stack([np.fft.rfft2(phim) * (1 + 0.01 * k) for k in range(chains)])
    phi = pkg.FlatS0(scalar_matrix=T(phis), theta_pix=th)
    ops = dict(Cn=R(pr["Cn"][0]), Cf=R(pr["Cf"][0]), M=R(pr["M"][0]), B=R(pr["B"][0]),
               D=R(D[0]), G=R(np.ones_like(Cphi)), Cphi=R(Cphi))
    mass = sampling.get_mass_matrix(ops["Cphi"], R(Cphi * 1e-2), ops["G"])
    fm, pm = mixing.mix(f, phi, ops["D"], ops["G"], n=n)
    args = (fm, d, ops["Cn"], ops["Cphi"], ops["Cf"], ops["M"], ops["B"], ops["D"], ops["G"])
    gen = torch.Generator(device=dev).manual_seed(5)
    sampling.hmc_step(pm, False, mass, *args, gen, num_steps=1, step_size=step_size, n=n)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, dh, acc = sampling.hmc_step(pm, False, mass, *args, gen, num_steps=num_steps,
                                   step_size=step_size, n=n)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ngrad = num_steps + 1
    return {"workload": "hmc_mixed_phi_T_256_2arcmin_16chains_30steps_fp64", "chains": chains,
            "step_size": step_size, "s_per_proposal": round(dt, 3),
            "mixed_gradients_per_s": round(chains * ngrad / dt, 1),
            "accept_rate": float(acc.double().mean()),
            "delta_h_median": float(dh.real.median())}


def sim_run(torch, sims, dev, N=256, th=2.0, pol="P", trials=64, n=10, reps=3):
    """Simulated data sets on the device (simulate.py's load_sim without CAMB,
    vmapped over trials like batch_simulated_trials): per trial one phi, the
    unlensed E/B fields, their lensing (10 RK4 steps, FP32) and the data
    M B L f + n.  The seed-independent set-up (covariances, D, the polar
    quadratic estimate) is timed once, separately."""
    import importlib
    simulate = importlib.import_module("cmblensing_amd.simulate")
    cls = simulate.extract_all_cls(*sims.camb_like_spectra(5000), 17000, 5000, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cfg = simulate.sim_config(N, th, pol, cls, device=dev)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    simulate.simulate(cfg, 0, trials, torch.float32, n)          # warm (plans, phi cache)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(reps):
        out = simulate.simulate(cfg, 1 + r, trials, torch.float32, n)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    # the draw kernel alone: E, B fields of every trial (one launch)
    covs = torch.stack([cfg.cf["EE"], cfg.cf["BB"]])
    simulate.grf_fourier(covs, trials, 9, 2 << 32, th, torch.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(20):
        simulate.grf_fourier(covs, trials, 9, 2 << 32, th, torch.float32)
    torch.cuda.synchronize()
    gdt = (time.perf_counter() - t0) / 20
    nmaps = trials * 2
    gbytes = nmaps * N * (N // 2 + 1) * (8 + 8) / 2 + nmaps * N * (N // 2 + 1) * 8
    return {"workload": f"simulated_datasets_{pol}_{N}_{th:g}arcmin_{trials}trials_n{n}_fp32",
            "trials": trials, "s_per_batch": round(dt, 4),
            "datasets_per_s": round(trials / dt, 1), "config_setup_s": round(setup_s, 3),
            "grf_draw_us": round(gdt * 1e6, 1),
            "grf_draw_maps_per_s": round(nmaps / gdt, 1),
            "finite": bool(torch.isfinite(torch.view_as_real(out[3])).all())}


def map_joint_run(torch, sims, dev, N=256, th=2.0, num_steps=30, trials=16, n=10):
    """MAP-joint (map_joint.py) on simulated temperature data sets, FP64 like
    the reference: one run with its BFGS line search (30 steps, one Wiener
    filter, a mix, a mixed phi gradient and the line search per step), and
    `trials` independent runs advanced together with the batched grid line
    search (every alpha of every trial in one LenseFlow batch)."""
    import importlib
    simulate = importlib.import_module("cmblensing_amd.simulate")
    mj = importlib.import_module("cmblensing_amd.map_joint")
    cls = simulate.extract_all_cls(*sims.camb_like_spectra(5000), 17000, 5000, device=dev)
    cfg = simulate.sim_config(N, th, "I", cls, uk_arcmin_t=1.0, device=dev)
    arrs = simulate.simulate(cfg, 3, trials, torch.float64, n)
    one = simulate.build_dataset(cfg, *[a[:1] for a in arrs])
    many = simulate.build_dataset(cfg, *arrs)
    mj.map_joint(one, num_steps=1, n=n)                          # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f, phi = mj.map_joint(one, num_steps=num_steps, n=n)
    torch.cuda.synchronize()
    dt1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    fb, phib = mj.map_joint(many, num_steps=num_steps, n=n, step="grid")
    torch.cuda.synchronize()
    dtb = time.perf_counter() - t0

    def corr(a, b):
        a, b = a.flatten(), b.flatten()
        return float(torch.real(torch.vdot(a, b)) / torch.sqrt(torch.vdot(a, a).real *
                                                             torch.vdot(b, b).real))
    # one Gibbs f-step (sampling_ar.py:14-46) for the 16 trials as chains:
    # device draws, one batched LenseFlow, a 16-RHS Wiener filter
    sampling = importlib.import_module("cmblensing_amd.sampling")
    gargs = dict(field_covariance=many.field_covariance, noise_covariance=many.noise_covariance,
                 mask=many.mask, beam=many.beam)
    sampling.gibbs_sample_f(fb, many.data, phib, gargs, seed=1, n=n)      # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sampling.gibbs_sample_f(fb, many.data, phib, gargs, seed=2, n=n)
    torch.cuda.synchronize()
    dtg = time.perf_counter() - t0
    return {"workload": f"map_joint_T_{N}_{th:g}arcmin_{num_steps}steps_fp64",
            "gibbs_f_step_16_chains_s": round(dtg, 3),
            "s_per_run_bfgs": round(dt1, 3),
            "batched_grid_trials": trials, "s_batched_grid": round(dtb, 3),
            "runs_per_s_batched_grid": round(trials / dtb, 2),
            "corr_phi_truth_bfgs": round(corr(phi.scalar_matrix, one.phi.scalar_matrix), 3),
            "corr_phi_truth_grid_mean": round(float(np.mean(
                [corr(phib.scalar_matrix[t], many.phi.scalar_matrix[t]) for t in range(trials)])), 3)}


def config0_run(torch, sims, dev, B=256, nphi=32, steps=20, plain=20):
    """The reference's CPU-runnable case: T, N = 128, 3 arcmin, float64, n = 7,
    forward LenseFlow.  Batched: B requests over `nphi` resident phi through
    cmbl_lenseflow; plain: one request per call with its phi prepared in the
    call.  Returns the leg and one batched result for the oracle check."""
    from cmblensing_amd import _lib
    f0, phi0, th, n = sims.lenseflow_inputs(0, seed=0)
    N = f0.shape[-1]
    phis = [torch.as_tensor(sims.phi_map(N, th, 200 + i), device=dev) for i in range(nphi)]
    phis[0] = torch.as_tensor(phi0, device=dev)
    x = torch.as_tensor(np.stack([f0] * B), device=dev)          # [B, 1, N, N] float64
    plan = _lib.get_plan(N, th, torch.float64, B, dev)
    hs = [_lib.get_phi(plan, p, n) for p in phis]
    req = [hs[r % nphi] for r in range(B)]
    out = torch.empty_like(x)
    for _ in range(3):
        _lib.lenseflow(plan, req, x, out, B, 1, n, 1, False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _lib.lenseflow(plan, req, x, out, B, 1, n, 1, False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    lat = []
    x1, o1 = x[:1].contiguous(), torch.empty_like(x[:1])
    for r in range(plain + 3):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        h = _lib.PhiHandle(plan, phis[r % nphi], n)
        _lib.lenseflow(plan, [h], x1, o1, 1, 1, n, 1, False)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
        h.close()
    lat = np.array(lat[3:]) * 1e3
    leg = {"workload": "lenseflow_fwd_T_128_3arcmin_n7_fp64", "batch": B,
           "batched_maps_s": round(B * steps / dt, 1), "plain_maps_s": round(1e3 / lat.mean(), 1),
           "plain_latency_ms_p50": round(float(np.percentile(lat, 50)), 4),
           "model_gbs": round(model_bytes_per_map(N, 1, n, 8) * B * steps / dt / 1e9, 1)}
    return leg, out[0].cpu().numpy(), f0, phi0, th, n


def copy_probe(torch):
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel give
    the scale factors of the HBM counters, and its time the achievable copy
    bandwidth."""
    n = 64 * 1024 * 1024
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 20
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                      "copy_us": round(ms * 1e3, 2),
                      "copy_gbs": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--plain-requests", type=int, default=40)
    ap.add_argument("--streams", type=int, default=2,
                    help="streams the served batch is split over (the serving backend's count)")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-wf-every", type=int, default=200,
                    help="second serving run (2000 requests): every k-th request a Wiener "
                         "filter (0: skip)")
    ap.add_argument("--no-wf", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-hmc", action="store_true")
    ap.add_argument("--no-sim", action="store_true")
    ap.add_argument("--no-config0", action="store_true")
    ap.add_argument("--no-map-joint", action="store_true")
    ap.add_argument("--wf-only", action="store_true",
                    help="only the 512^2 Q/U Wiener-filter leg, its timed solve between two "
                         "marker launches (k_row_r2c<float, 16>) for a kernel trace")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the LenseFlow path shards by independent "
              "requests with no communication, and only --gpus 1 is supported here.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        copy_probe(torch)
        return
    import __graft_entry__ as ge
    pkg = ge.load_package()
    from cmblensing_amd import _lib
    sims = ge._load_module("cmbl_sims_b", ROOT / "cmblensing.py_amd" / "sims.py")

    dev = torch.device("cuda:0")
    if args.wf_only:
        print(json.dumps({"wiener_filter": wf_run(torch, pkg, sims, dev, markers=True,
                                                  fp64_solve=False)}), flush=True)
        return
    N, th, n, C, B = 256, 2.0, 7, 2, args.batch
    NPHI = 32
    dtype = torch.float32
    print(f"[bench] workload QU N={N} theta={th} n={n} fp32 batch={B}", file=sys.stderr, flush=True)

    # synthetic inputs: 32 phi; request r (in arrival order) lenses Q/U draw
    # r % 8 of 8 by phi r % 32.  The serving backend forms the batches of a
    # step together and orders them by phi (requests that share
    # a phi then read the same p(t) rows together), so the batch positions
    # hold the requests in that order; every path below uses it.
    NS = args.streams
    if B % NS:
        sys.exit(f"--batch {B} must be a multiple of --streams {NS}")
    Bs = B // NS
    # the requests of a step ordered by phi across the whole step, then split
    # over the streams (the serving scheduler forms the batches of a bucket
    # that become ready together this way; stable, so arrival order among the
    # requests of one phi)
    order = sorted(range(B), key=lambda r: r % NPHI)
    req_f = [r % 8 for r in order]          # Q/U draw of batch position j
    req_p = [r % NPHI for r in order]       # phi of batch position j
    phis = [sims.phi_map(N, th, 1000 + i) for i in range(NPHI)]
    fq = [sims.qu_map(N, th, i) for i in range(8)]
    x = torch.as_tensor(np.stack([fq[req_f[j]] for j in range(B)]), dtype=dtype, device=dev)
    phi_t = [torch.as_tensor(p, device=dev) for p in phis]

    plan = _lib.get_plan(N, th, dtype, B * C, dev)
    handles = [_lib.get_phi(plan, p, n) for p in phi_t]
    req_phi = [handles[req_p[j]] for j in range(B)]
    out = torch.empty_like(x)

    # the served form, as the serving backend runs it: the B requests of a
    # step split over its streams (default 2, default priority), each half one
    # replay of a captured batched graph (gather, 57 stage kernels, scatter)
    # on a plan of its own, so the two batches overlap on the device
    lanes = []
    for k in range(NS):
        pk = _lib.Plan(N, th, dtype, Bs * C, dev)
        lanes.append((pk, _lib.Executor(pk, Bs, C, n, 1, False), torch.cuda.Stream(dev),
                      req_phi[k * Bs:(k + 1) * Bs],
                      [x[b].data_ptr() for b in range(k * Bs, (k + 1) * Bs)],
                      [out[b].data_ptr() for b in range(k * Bs, (k + 1) * Bs)]))

    def step():
        for _, ex, st, hs, ins, outs in lanes:
            ex.run(hs, ins, outs, st.cuda_stream)

    def direct_step():
        _lib.lenseflow(plan, req_phi, x, out, B, C, n, 1, False)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    direct_dt = timed(direct_step, args.steps, args.warmup)
    direct_maps_s = B * args.steps / direct_dt
    dt = timed(step, args.steps, args.warmup)
    ms_per_step = dt / args.steps * 1e3
    maps_s = B * args.steps / dt

    # one batched result, checked against the FP64 oracle in the cpu_baseline leg
    r_chk = 5
    got = out[r_chk].double().cpu().numpy()

    # live per-kernel timing (HIP events on the launch stream) over a few steps
    plan.set_timing(True)
    plan.get_timing(reset=True)
    ksteps = max(2, min(args.steps, 5))
    for _ in range(ksteps):
        direct_step()
    tim = plan.get_timing(reset=True)
    # the same for the adjoint mode (its own row kernel, the same column pass)
    for _ in range(ksteps):
        _lib.lenseflow(plan, req_phi, x, out, B, C, n, 1, True)
    tim_adj = plan.get_timing(reset=True)
    plan.set_timing(False)
    adj_row_avg_ms = tim_adj["row_ms"] / tim_adj["row_n"]
    adj_row_bytes = adj_row_bytes_per_launch(B, N, C, 4)
    adj_row_gbs = adj_row_bytes / (adj_row_avg_ms * 1e-3) / 1e9
    row_avg_ms = tim["row_ms"] / tim["row_n"]
    col_avg_ms = tim["col_ms"] / tim["col_n"]
    row_bytes = row_bytes_per_launch(B, N, C, 4)
    col_bytes = col_bytes_per_launch(B, N, C, 4)
    row_gbs = row_bytes / (row_avg_ms * 1e-3) / 1e9
    col_gbs = col_bytes / (col_avg_ms * 1e-3) / 1e9

    # plain single-request path: phi prepared per request, batch of one
    lat = []
    xs = x[:1].contiguous()
    o1 = torch.empty_like(xs)
    for r in range(args.plain_requests + 3):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        h = _lib.PhiHandle(plan, phi_t[r % NPHI], n)
        _lib.lenseflow(plan, [h], xs, o1, 1, C, n, 1, False)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
        h.close()
    lat = np.array(lat[3:]) * 1e3
    plain_maps_s = 1e3 / lat.mean()
    # plain with a resident phi (latency of the stage pipeline alone)
    lat_res = []
    for r in range(args.plain_requests):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _lib.lenseflow(plan, [handles[0]], xs, o1, 1, C, n, 1, False)
        torch.cuda.synchronize()
        lat_res.append(time.perf_counter() - t1)
    lat_res = np.array(lat_res) * 1e3

    # the 256^2 QU configuration in all four modes, its round trip and its
    # adjoint dot-product test (the oracle errors are added in the CPU leg)
    c1_leg, c1_chk = config1_run(torch, _lib, plan, req_phi, x, B, C, n, args.steps,
                                 args.warmup, timed, sims, N, th, r_chk)

    # HBM traffic per launch of the row pass, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated on a copy)
    traffic = trace_us = col_trace_us = adj_trace_us = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists() and B == 128:
        prof = json.loads(tfile.read_text())
        rowprof = prof.get("k_row_lf_std<float,256>", {})
        traffic = rowprof.get("bytes_per_launch")
        trace_us = rowprof.get("trace_avg_launch_us")
        col_trace_us = prof.get("k_col_lf_dx<float,256,16>", {}).get("trace_avg_launch_us")
        adj_trace_us = prof.get("k_row_lf_adj_ls<float,256>", {}).get("trace_avg_launch_us")
    model_b = model_bytes_per_map(N, C, n, 4)
    design_b = design_bytes_per_map(N, C, n, 4)
    result = {
        "metric": "LenseFlow maps/s @256^2 QU batched vs plain; p50/p99 latency; err vs FP64",
        "value": round(maps_s, 2),
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian random fields, 32 resident phi)",
        "config": {"workload": "lenseflow_fwd_QU_256_2arcmin_n7_fp32", "N": N, "theta_pix": th,
                   "nsteps": n, "components": "QU", "batch": B, "phi_resident": NPHI,
                   "streams": NS, "batch_order": "the step's requests ordered by phi, then split over the streams"},
        "direct_batched_maps_s": round(direct_maps_s, 2),
        "plain_maps_s": round(plain_maps_s, 2),
        "batched_over_plain": round(maps_s / plain_maps_s, 2),
        "plain_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                             "p99": round(float(np.percentile(lat, 99)), 4)},
        "plain_resident_phi_latency_ms": {"p50": round(float(np.percentile(lat_res, 50)), 4),
                                          "p99": round(float(np.percentile(lat_res, 99)), 4)},
        "byte_models": {
            "design_bytes_per_map": design_b,
            "design_gbs": round(design_b * maps_s / 1e9, 1),
            "design_frac_hbm": round(design_b * maps_s / 1e9 / HBM_PEAK_GBS, 4),
            "canonical_bytes_per_map": model_b,
            "canonical_rate_gbs": round(model_b * maps_s / 1e9, 1),
            "note": "design: 4n [C 6.75 + 2] N^2 s, the bytes this two-pass design moves per "
                    "map (its lower bound); canonical: 4n [C (12H + 5P) + 5P] s, which charges "
                    "transfers this design does not make, so its rate is no fraction of peak"},
        "roofline": {"bound": "hbm", "kernel": "k_row_lf_std<float,256>",
                     "achieved": round(row_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(row_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_us": round(row_avg_ms * 1e3, 3),
                     "algorithmic_bytes_per_launch": row_bytes,
                     "frac_events": round(row_gbs / HBM_PEAK_GBS, 4),
                     "frac_trace": (round(row_bytes / (trace_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
                                    if trace_us else None),
                     "trace_avg_launch_us": trace_us},
        "column_pass": {"kernel": "k_col_lf_dx<float,256,16>",
                        "achieved_gbs": round(col_gbs, 1),
                        "avg_launch_us": round(col_avg_ms * 1e3, 3),
                        "algorithmic_bytes_per_launch": col_bytes,
                        "frac_events": round(col_gbs / HBM_PEAK_GBS, 4),
                        "frac_trace": (round(col_bytes / (col_trace_us * 1e-6) / 1e9
                                             / HBM_PEAK_GBS, 4) if col_trace_us else None),
                        "trace_avg_launch_us": col_trace_us},
        "adjoint_row_pass": {"kernel": "k_row_lf_adj_ls<float,256>",
                             "achieved_gbs": round(adj_row_gbs, 1),
                             "avg_launch_us": round(adj_row_avg_ms * 1e3, 3),
                             "algorithmic_bytes_per_launch": adj_row_bytes,
                             "frac_events": round(adj_row_gbs / HBM_PEAK_GBS, 4),
                             "frac_trace": (round(adj_row_bytes / (adj_trace_us * 1e-6) / 1e9
                                                  / HBM_PEAK_GBS, 4) if adj_trace_us else None),
                             "trace_avg_launch_us": adj_trace_us},
        "config1": c1_leg,
    }
    for pk, ex, *_ in lanes:
        ex.close()
        pk.close()
    if not args.no_serve:
        result["serving"] = serving_run(torch, sims, dev, args.serve_requests, args.serve_rate,
                                        pkg=pkg)
        result["latency_ms"] = {"p50": result["serving"]["p50_ms"],
                                "p99": result["serving"]["p99_ms"]}
        if args.serve_wf_every:
            print("[bench] serving with Wiener-filter requests", file=sys.stderr, flush=True)
            result["serving_with_wf"] = serving_run(torch, sims, dev, 2000, args.serve_rate,
                                                    seed=3, wf_every=args.serve_wf_every,
                                                    pkg=pkg)
    if not args.no_grad:
        print("[bench] gradients, 16 chains", file=sys.stderr, flush=True)
        result["differentiation"] = grad_run(torch, pkg, sims, dev)
    if not args.no_hmc:
        print("[bench] HMC, 16 chains", file=sys.stderr, flush=True)
        result["hmc"] = hmc_run(torch, pkg, sims, dev)
    c0_leg = c0_chk = None
    if not args.no_config0:
        print("[bench] T 128^2 fp64 (CPU-runnable case)", file=sys.stderr, flush=True)
        c0_leg, *c0_chk = config0_run(torch, sims, dev)
        result["config0"] = c0_leg
    if not args.no_sim:
        print("[bench] simulated data sets, 64 trials 256^2 P", file=sys.stderr, flush=True)
        result["simulation"] = sim_run(torch, sims, dev)
    if not args.no_map_joint:
        print("[bench] MAP-joint 256^2 T, 30 steps", file=sys.stderr, flush=True)
        result["map_joint"] = map_joint_run(torch, sims, dev)
    if not args.no_wf:
        print("[bench] Wiener filter 512^2 QU, 8 RHS", file=sys.stderr, flush=True)
        result["wiener_filter"] = wf_run(torch, pkg, sims, dev)
    if not args.no_cpu:
        # the CPU leg: the FP64 oracle checks one batched result, then is timed
        from oracle import lenseflow as olf
        f_chk, p_chk = fq[req_f[r_chk]], phis[req_p[r_chk]]
        ref = olf.lense_flow(f_chk, p_chk, olf.pix_width_from_theta(th), n)
        result["rel_l2_vs_fp64_oracle"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        # the four modes of the 256^2 QU configuration against the oracle, and
        # the oracle's own round trip (the RK4 discretisation at n = 7)
        pw = olf.pix_width_from_theta(th)
        for name, (d, a) in LF_MODES.items():
            ref_m = olf.lense_flow(f_chk, p_chk, pw, n, d, a)
            c1_leg[name]["rel_l2_vs_fp64_oracle"] = float(
                np.linalg.norm(c1_chk[name] - ref_m) / np.linalg.norm(ref_m))
        back = olf.lense_flow(ref, p_chk, pw, n, -1, False)
        c1_leg["oracle_fp64_round_trip_rel_l2"] = float(
            np.linalg.norm(back - f_chk) / np.linalg.norm(f_chk))
        # the same workload on the host: Q/U 256^2 requests (8 seeded pairs,
        # one resident phi), forward, n = 7, FP64 oracle, threads over requests
        nth = cpu_threads()
        cps, k, el = cpu_baseline_threads(fq, phis[0], th, n, nth, args.cpu_budget)
        cps1, k1, el1 = cpu_baseline(fq[0], phis[0], th, n, min(args.cpu_budget, 5.0))
        result["cpu_baseline"] = {"value": round(cps, 4), "unit": "maps/s", "cores": nth,
                                  "kind": "port",
                                  "sample": f"{k} forward LenseFlows of 256^2 QU pairs (8 seeded "
                                            f"pairs, one phi), float64 numpy oracle, {nth} "
                                            f"threads over requests, {el:.1f} s",
                                  "one_core_maps_s": round(cps1, 4),
                                  "one_core_sample": f"{k1} LenseFlows on 1 thread, {el1:.1f} s",
                                  "host_cpu": cpu_model(),
                                  "os_cpu_count": os.cpu_count(),
                                  "affinity_cpus": len(os.sched_getaffinity(0))}
        if c0_leg is not None:
            # the T 128^2 FP64 case on the same core: its oracle check and CPU rate
            got0, f0, phi0, th0, n0 = c0_chk
            ref0 = olf.lense_flow(f0, phi0, olf.pix_width_from_theta(th0), n0)
            c0_leg["rel_l2_vs_fp64_oracle"] = float(np.linalg.norm(got0 - ref0) /
                                                    np.linalg.norm(ref0))
            cps0, k0, el0 = cpu_baseline(f0, phi0, th0, n0, min(args.cpu_budget, 5.0))
            c0_leg["cpu_maps_s"] = round(cps0, 3)
            c0_leg["cpu_sample"] = f"{k0} forward LenseFlows, float64 numpy oracle, 1 core, {el0:.1f} s"
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
