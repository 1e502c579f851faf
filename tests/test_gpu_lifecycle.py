"""Lifetimes of library objects under the package's caches, and the one-call
PCG step against the separate calls it replaces.

* A plan held by a caller keeps working after the cache replaces it with a
  larger one (a Wiener-filter process holds its plans across its yields).
* A phi handle held by a caller keeps working after the cache drops it.
* Plans are kept per stream: calls on two streams never share scratch.
* cmbl_cg_step equals cmbl_dot_f64 + the host-side alpha/beta +
  cmbl_cg_update_xrz + cmbl_cg_update_p bit for bit (the same float64
  operations), for all and for some active right-hand sides."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TH, NSTEPS = 2.0, 7


def _case(msg):
    print(f"[case] {msg}", flush=True)


def test_held_plan_survives_cache_growth(pkg, sims, dev):
    from cmblensing_amd import _lib
    N = 64
    _case("held plan after the cache grows")
    x = torch.as_tensor(np.stack([sims.qu_map(N, TH, i) for i in range(2)]), device=dev)
    phi = torch.as_tensor(sims.phi_map(N, TH, 3), device=dev)
    small = _lib.get_plan(N, TH, torch.float64, 4, dev)
    h = _lib.get_phi(small, phi, NSTEPS)
    ref = torch.empty_like(x)
    _lib.lenseflow(small, [h, h], x, ref, 2, 2, NSTEPS, 1, False)
    big = _lib.get_plan(N, TH, torch.float64, 4 * small.max_maps, dev)
    assert big is not small and small.handle          # replaced, not destroyed
    assert _lib.get_plan(N, TH, torch.float64, 1, dev) is big
    out = torch.empty_like(x)
    _lib.lenseflow(small, [h, h], x, out, 2, 2, NSTEPS, 1, False)
    out2 = torch.empty_like(x)
    _lib.lenseflow(big, [h, h], x, out2, 2, 2, NSTEPS, 1, False)
    torch.cuda.synchronize()
    assert torch.equal(out, ref) and torch.equal(out2, ref)


def test_held_phi_handle_survives_eviction(pkg, sims, dev):
    from cmblensing_amd import _lib
    N = 32
    _case("held phi handle after eviction from the cache")
    plan = _lib.get_plan(N, TH, torch.float64, 2, dev)
    phi = torch.as_tensor(sims.phi_map(N, TH, 5), device=dev)
    h = _lib.get_phi(plan, phi, NSTEPS)
    x = torch.as_tensor(sims.t_map(N, TH, 6)[None], device=dev)
    ref = torch.empty_like(x)
    _lib.lenseflow(plan, [h], x, ref, 1, 1, NSTEPS, 1, False)
    others = [torch.as_tensor(sims.phi_map(N, TH, 100 + i), device=dev)
              for i in range(_lib.PHI_CACHE_SIZE + 8)]
    for o in others:
        _lib.get_phi(plan, o, NSTEPS)
    assert all(v[2] is not h for v in _lib._phis.values())    # dropped from the cache
    assert h.handle                                            # but not released
    out = torch.empty_like(x)
    _lib.lenseflow(plan, [h], x, out, 1, 1, NSTEPS, 1, False)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


def test_plans_are_per_stream(pkg, sims, dev):
    from cmblensing_amd import _lib
    N = 32
    _case("plans per stream")
    p0 = _lib.get_plan(N, TH, torch.float32, 2, dev)
    s = torch.cuda.Stream(dev)
    with torch.cuda.stream(s):
        p1 = _lib.get_plan(N, TH, torch.float32, 2, dev)
    # a phi handle prepared on the first stream, used on the second
    phi = torch.as_tensor(sims.phi_map(N, TH, 7), device=dev)
    h = _lib.get_phi(p0, phi, NSTEPS)
    x = torch.as_tensor(sims.t_map(N, TH, 8)[None], dtype=torch.float32, device=dev)
    ref = torch.empty_like(x)
    _lib.lenseflow(p0, [h], x, ref, 1, 1, NSTEPS, 1, False)
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        h2 = _lib.get_phi(p1, phi, NSTEPS)       # the cached handle, ordered on s
        out = torch.empty_like(x)
        _lib.lenseflow(p1, [h2], x, out, 1, 1, NSTEPS, 1, False)
    torch.cuda.synchronize()
    assert p0 is not p1 and h2 is h
    assert torch.equal(out, ref)


@pytest.mark.parametrize("active", [None, [True, False, True]])
def test_cg_step_equals_separate_calls(pkg, dev, active):
    from cmblensing_amd import _lib
    N, C, B = 32, 2, 3
    _case(f"cg_step vs separate calls, active={active}")
    g = torch.Generator(device="cpu").manual_seed(3)
    shp = (B, C, N, N // 2 + 1)
    rnd = lambda: torch.complex(torch.randn(shp, generator=g, dtype=torch.float64),
                                torch.randn(shp, generator=g, dtype=torch.float64)).to(dev)
    x, r, p = rnd(), rnd(), rnd()
    Ap = rnd().to(torch.complex64)
    z = torch.zeros_like(r)
    pinv = torch.rand((C, N, N // 2 + 1), generator=g, dtype=torch.float64).to(dev)
    res = torch.rand(B, generator=g, dtype=torch.float64).to(dev) + 0.5
    bb = torch.rand(B, generator=g, dtype=torch.float64).to(dev) + 1.0
    act = None if active is None else torch.tensor(active, device=dev)
    plan = _lib.get_plan(N, TH, torch.float32, B * C, dev)
    V = torch.view_as_real
    # separate calls, scalars formed by torch (the previous host-side form)
    xa, ra, za, pa = x.clone(), r.clone(), z.clone(), p.clone()
    pAp = torch.empty(B, dtype=torch.float64, device=dev)
    _lib.dot_f64(plan, V(pa), V(Ap), B, C, pAp)
    act_d = torch.ones(B, dtype=torch.bool, device=dev) if act is None else act
    alpha = torch.where(act_d, res / torch.where(pAp != 0, pAp, torch.ones_like(pAp)),
                        torch.zeros_like(res))
    rz = torch.empty(B, dtype=torch.float64, device=dev)
    _lib.cg_update_xrz(plan, V(xa), V(ra), V(za), V(pa), V(Ap), pinv, alpha, B, C, rz)
    beta = torch.where(act_d, rz / torch.where(res != 0, res, torch.ones_like(res)),
                       torch.zeros_like(res))
    pop_a = torch.empty_like(Ap)
    _lib.cg_update_p(plan, V(pa), V(za), beta, V(pop_a), B, C)
    res_a = torch.where(act_d, rz, res)
    rr = torch.empty(B, dtype=torch.float64, device=dev)
    _lib.dot_f64(plan, V(ra), V(ra), B, C, rr)
    rel_a = torch.sqrt(rr / bb)
    # one call
    xb, rb, zb, pb, res_b = x.clone(), r.clone(), z.clone(), p.clone(), res.clone()
    pop_b = torch.empty_like(Ap)
    rel_b = torch.empty(B, dtype=torch.float64, device=dev)
    _lib.cg_step(plan, V(xb), V(rb), V(zb), V(pb), V(Ap), pinv, V(pop_b), res_b, act, bb, rel_b,
                 B, C)
    torch.cuda.synchronize()
    for a, b in ((xa, xb), (ra, rb), (za, zb), (pa, pb), (pop_a, pop_b), (res_a, res_b),
                 (rel_a, rel_b)):
        assert torch.equal(a, b)
