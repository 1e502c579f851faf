"""The allocation-free LenseFlow entry point (cmbl_lenseflow_noalloc) against
cmbl_lenseflow through the C ABI: bit-identical results (same kernels, same
arithmetic; the stage buffers only live elsewhere) on two calls in a row,
the second smaller than the first, in every mode; and a call larger than the
static scratch is refused before any work."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TH, NSTEPS = 2.0, 7


def _case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_matches_lenseflow(pkg, sims, dev, dtype):
    from cmblensing_amd import _lib
    N, C = 128, 2
    plan = _lib.Plan(N, TH, dtype, 16 * C, dev)
    phis_t = [torch.as_tensor(sims.phi_map(N, TH, 80 + i), device=dev) for i in range(4)]
    handles = [_lib.PhiHandle(plan, p, NSTEPS) for p in phis_t]
    x_all = torch.as_tensor(np.stack([sims.qu_map(N, TH, 90 + i) for i in range(16)]),
                            dtype=dtype, device=dev)
    for direction, adjoint in ((1, False), (-1, False), (1, True), (-1, True)):
        for batch in (16, 5):          # two calls in a row, the second smaller
            _case(f"noalloc N={N} QU {dtype} batch={batch} dir={direction} adj={adjoint}")
            x = x_all[:batch].contiguous()
            hs = [handles[b % 4] for b in range(batch)]
            ref = torch.full_like(x, float("nan"))
            got = torch.full_like(x, float("nan"))
            _lib.lenseflow(plan, hs, x, ref, batch, C, NSTEPS, direction, adjoint)
            _lib.lenseflow_noalloc(plan, hs, x, got, batch, C, NSTEPS, direction, adjoint)
            torch.cuda.synchronize()
            assert torch.isfinite(got).all()
            assert torch.equal(got, ref), "noalloc result differs from cmbl_lenseflow"
    # in place (in == out), T maps, and a batch above the plan's max_maps
    _case(f"noalloc in place T batch=40 {dtype}")
    xt = torch.as_tensor(np.stack([sims.t_map(N, TH, 100 + i)[0] for i in range(40)])[:, None],
                         dtype=dtype, device=dev)
    hs = [handles[b % 4] for b in range(40)]
    plan40 = _lib.Plan(N, TH, dtype, 40, dev)
    ref = torch.empty_like(xt)
    _lib.lenseflow(plan40, hs, xt, ref, 40, 1, NSTEPS, 1, False)
    got = xt.clone()
    _lib.lenseflow_noalloc(plan, hs, got, got, 40, 1, NSTEPS, 1, False)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    plan40.close()
    for h in handles:
        h.close()
    plan.close()


def test_noalloc_refuses_oversize(pkg, sims, dev):
    """A call whose stage buffers exceed the static scratch fails with a clear
    error before anything is launched (512^2 QU float64 needs 12.6 MB per
    request: 6 requests are more than 64 MiB)."""
    from cmblensing_amd import _lib
    N, C = 512, 2
    _case("noalloc oversize 512^2 QU fp64 batch 6")
    plan = _lib.Plan(N, TH, torch.float64, 2, dev)
    h = _lib.PhiHandle(plan, torch.as_tensor(sims.phi_map(N, TH, 3), device=dev), NSTEPS)
    x = torch.zeros((6, C, N, N), dtype=torch.float64, device=dev)
    out = torch.full_like(x, 7.0)
    with pytest.raises(_lib.CmblError, match="static scratch holds"):
        _lib.lenseflow_noalloc(plan, [h] * 6, x, out, 6, C, NSTEPS, 1, False)
    torch.cuda.synchronize()
    assert torch.equal(out, torch.full_like(x, 7.0))     # nothing was written
    # one request fits and matches the plan path
    ref = torch.empty_like(x[:1])
    x1 = torch.as_tensor(sims.qu_map(N, TH, 4)[None], device=dev)
    _lib.lenseflow(plan, [h], x1, ref, 1, C, NSTEPS, 1, False)
    got = torch.empty_like(x1)
    _lib.lenseflow_noalloc(plan, [h], x1, got, 1, C, NSTEPS, 1, False)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    h.close()
    plan.close()
