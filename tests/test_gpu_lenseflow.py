"""GPU parity of the HIP LenseFlow path against the FP64 oracle and the
reference's Julia fixtures.  Tolerances: FP64 1e-11 relative L2, FP32 1e-5
relative L2 against the FP64 oracle (the workload's stated bars)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import lenseflow as olf

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

MODES = [(1, False), (-1, False), (1, True), (-1, True)]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("N", [16, 32, 64, 128, 256, 512])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_rfft2_irfft2_match_numpy(pkg, dev, N, dtype):
    _case(f"rfft2/irfft2 N={N} {dtype}")
    rng = np.random.default_rng(N)
    x = rng.standard_normal((3, N, N))
    f = pkg.FlatS0(scalar_matrix=torch.as_tensor(x, dtype=dtype, device=dev))
    F = pkg.fourier(f).scalar_matrix.cpu().numpy()
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    assert rel(F, np.fft.rfft2(x)) < tol
    # irfft2 of a general (non-Hermitian-consistent) half plane: numpy semantics
    X = rng.standard_normal((3, N, N // 2 + 1)) + 1j * rng.standard_normal((3, N, N // 2 + 1))
    ct = torch.complex128 if dtype == torch.float64 else torch.complex64
    g = pkg.FlatS0(scalar_matrix=torch.as_tensor(X, dtype=ct, device=dev))
    xm = pkg.map(g).scalar_matrix.cpu().numpy()
    assert rel(xm, np.fft.irfft2(X)) < tol


@pytest.mark.parametrize("basis", ["map", "fourier"])
def test_phi_derivatives(pkg, dev, sims, basis):
    from cmblensing_amd import _lib
    N, th = 128, 3.0
    _case(f"phi derivatives N={N} basis={basis}")
    phi = sims.phi_map(N, th, 5)
    pw = olf.pix_width_from_theta(th)
    if basis == "map":
        pin, ref = phi, olf.primal_derivatives(phi, pw)
    else:
        rng = np.random.default_rng(1)
        P = np.fft.rfft2(phi)
        # anti-Hermitian content on the self-conjugate columns must be kept
        P[:, 0] += 1e-3 * np.abs(P).max() * rng.standard_normal(N) * 1j
        P[:, -1] += 1e-3 * np.abs(P).max() * rng.standard_normal(N) * 1j
        pin, ref = P, olf.primal_derivatives_from_fourier(P, pw)
    plan = _lib.get_plan(N, th, torch.float64, 1, dev)
    h = _lib.PhiHandle(plan, torch.as_tensor(pin, device=dev), 7)
    D = h.derivatives().cpu().numpy()
    for i in range(5):
        assert rel(D[i], ref[i]) < 1e-12, i
    if basis == "map":
        assert h.min_det == pytest.approx(sims.min_det_magnification(phi, th), rel=1e-9)


@pytest.mark.parametrize("direction,adjoint", MODES)
def test_config0_T_fp64(pkg, dev, sims, direction, adjoint):
    f, phi, th, n = sims.lenseflow_inputs(0)
    _case(f"config0 T N=128 fp64 dir={direction} adj={adjoint}")
    field = pkg.FlatS0(scalar_matrix=torch.as_tensor(f[0], device=dev), theta_pix=th)
    phif = pkg.FlatS0(scalar_matrix=torch.as_tensor(phi, device=dev), theta_pix=th)
    out = pkg.lense_flow(field, phif, n=n, direction=direction, adjoint=adjoint)
    ref = olf.lense_flow(f[0], phi, olf.pix_width_from_theta(th), n, direction, adjoint)
    assert rel(out.scalar_matrix.cpu().numpy(), ref) < 1e-11


@pytest.mark.parametrize("direction,adjoint", MODES)
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.float64, 1e-11)])
def test_config1_QU(pkg, dev, sims, direction, adjoint, dtype, tol):
    f, phi, th, n = sims.lenseflow_inputs(1)
    _case(f"config1 QU N=256 {dtype} dir={direction} adj={adjoint}")
    field = pkg.FlatS2(polar_matrix_1=torch.as_tensor(f[0], dtype=dtype, device=dev),
                       polar_matrix_2=torch.as_tensor(f[1], dtype=dtype, device=dev),
                       theta_pix=th)
    phif = pkg.FlatS0(scalar_matrix=torch.as_tensor(phi, device=dev), theta_pix=th)
    out = pkg.lense_flow(field, phif, n=n, direction=direction, adjoint=adjoint)
    got = np.stack([out.polar_matrix_1.cpu().numpy(), out.polar_matrix_2.cpu().numpy()])
    ref = olf.lense_flow(f, phi, olf.pix_width_from_theta(th), n, direction, adjoint)
    assert rel(got, ref) < tol


def test_full_size_linearity_round_trip_and_adjoint(pkg, dev, sims):
    """At the workload's full shape (256^2 QU FP32, a batch of 128 requests
    over 32 phis), properties that hold at any size: L is linear in the field
    (L(a x + b y) = a L x + b L y to FP32 rounding, the 1e-5 bar); the inverse
    undoes the forward to the RK4 error of n = 7 (below 3e-3 for every
    request, and request 0's residual field L^-1 L x - x equal to the FP64
    oracle's to 1e-5 of |x|: the residual is the method's, not rounding); and
    <L y, x> = <y, L^T x> per request to 1e-4."""
    N, th, n, B = 256, 2.0, 7, 128
    _case("full size 256^2 QU fp32 batch 128: linearity, round trip, dot test")
    phi_np = np.stack([sims.phi_map(N, th, 100 + k) for k in range(32)])
    P = torch.as_tensor(phi_np[np.arange(B) % 32], device=dev)
    x = torch.as_tensor(np.stack([sims.qu_map(N, th, 200 + s) for s in range(B)]),
                        dtype=torch.float32, device=dev)
    yv = torch.as_tensor(np.stack([sims.qu_map(N, th, 400 + s) for s in range(B)]),
                         dtype=torch.float32, device=dev)
    L = lambda f, d=1, adj=False: pkg.lense_flow_matrices(f, P, th, n, d, adj)
    a, b = 0.7, -1.3
    lin = L(a * x + b * yv).double()
    ref = a * L(x).double() + b * L(yv).double()
    assert float(torch.linalg.vector_norm(lin - ref) / torch.linalg.vector_norm(ref)) < 1e-5
    Lx = L(x)
    back = L(Lx, -1)
    res = (back.double() - x.double()).reshape(B, -1)
    rt = torch.linalg.vector_norm(res, dim=1) / torch.linalg.vector_norm(x.double().reshape(B, -1),
                                                                          dim=1)
    assert float(rt.max()) < 3e-3
    pw = olf.pix_width_from_theta(th)
    x0 = x[0].double().cpu().numpy()
    r0 = olf.lense_flow(olf.lense_flow(x0, phi_np[0], pw, n, 1, False), phi_np[0], pw, n, -1,
                        False) - x0
    assert np.linalg.norm(res[0].cpu().numpy().reshape(x0.shape) - r0) / np.linalg.norm(x0) < 1e-5
    Ly = L(yv).double().reshape(B, -1)
    LTx = L(x, 1, True).double().reshape(B, -1)
    lhs = (Ly * x.double().reshape(B, -1)).sum(1)
    rhs = (yv.double().reshape(B, -1) * LTx).sum(1)
    assert float(((lhs - rhs).abs() / lhs.abs()).max()) < 1e-4


def test_batched_mixed_phi_equals_single(pkg, dev, sims):
    """A batch with a different phi per element gives, element by element,
    exactly what the same element gives alone."""
    N, th, n = 64, 2.0, 7
    _case("batched mixed phi N=64 fp32 QU")
    fs = np.stack([sims.qu_map(N, th, s) for s in range(4)])
    phis = np.stack([sims.phi_map(N, th, 10 + (s % 2)) for s in range(4)])
    x = torch.as_tensor(fs, dtype=torch.float32, device=dev)
    P = torch.as_tensor(phis, device=dev)
    yb = pkg.lense_flow_matrices(x, P, th, n, 1, False)
    for s in range(4):
        ys = pkg.lense_flow_matrices(x[s:s + 1], P[s], th, n, 1, False)
        assert torch.equal(ys[0], yb[s])


def test_julia_fixture_T_fp64(pkg, dev):
    """FP64 GPU LenseFlow of the reference's Julia inputs reproduces the Julia
    lensed and adjoint-lensed maps (n = 7, N = 256, 2.5 arcmin)."""
    _case("julia fixture T N=256 fp64")
    J = GOLDEN / "julia"
    phi = np.fft.irfft2(np.load(J / "phi_I.npy"))
    t = np.fft.irfft2(np.load(J / "t_field_I.npy"))
    tl = np.fft.irfft2(np.load(J / "t_lensed_field_I.npy"))
    ta = np.fft.irfft2(np.load(J / "t_adjoint_lensed_field_I.npy"))
    field = pkg.FlatS0(scalar_matrix=torch.as_tensor(t, device=dev), theta_pix=2.5)
    phif = pkg.FlatS0(scalar_matrix=torch.as_tensor(phi, device=dev), theta_pix=2.5)
    out = pkg.lense_flow(field, phif, n=7).scalar_matrix.cpu().numpy()
    adj = pkg.lense_flow(field, phif, n=7, adjoint=True).scalar_matrix.cpu().numpy()
    assert rel(out, tl) < 1e-11
    assert rel(adj, ta) < 1e-11


@pytest.mark.parametrize("N", [16, 32, 64, 512])
@pytest.mark.parametrize("C", [1, 2])
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.float64, 1e-11)])
def test_sizes_components_steps(pkg, dev, sims, N, C, dtype, tol):
    """Every supported size beyond the two configurations (the smallest and
    the largest N), T and Q/U, forward with 7 steps and the adjoint with a
    single step, against the FP64 oracle."""
    th = 2.0
    _case(f"sweep N={N} C={C} {dtype}")
    f = (sims.t_map if C == 1 else sims.qu_map)(N, th, 3)
    phi = sims.phi_map(N, th, 4)
    x = torch.as_tensor(f, dtype=dtype, device=dev)[None]
    P = torch.as_tensor(phi, device=dev)
    for n, adjoint in ((7, False), (1, True)):
        y = pkg.lense_flow_matrices(x, P, th, n, 1, adjoint)[0].double().cpu().numpy()
        ref = olf.lense_flow(f, phi, olf.pix_width_from_theta(th), n, 1, adjoint)
        assert rel(y, ref) < tol, (n, adjoint)


def test_released_handle_memory_is_reused_correctly(pkg, dev, sims):
    """Handles released without waiting hand their memory to later
    preparations of the same size: lensing with phis prepared into reused
    memory equals lensing with the same phis prepared while nothing was
    released (bit for bit), the caustic guard read on demand equals the one
    read at preparation, and a handle used on a second stream still releases
    safely (after waiting for the device)."""
    from cmblensing_amd import _lib
    N, th, n, B = 64, 2.0, 7, 4
    _case("phi handle pool N=64 fp32 QU")
    plan = _lib.Plan(N, th, torch.float32, 2 * B, dev)
    x = torch.as_tensor(np.stack([sims.qu_map(N, th, s) for s in range(B)]),
                        dtype=torch.float32, device=dev)
    P = [torch.as_tensor(np.stack([sims.phi_map(N, th, 100 * r + s) for s in range(B)]),
                         device=dev) for r in range(3)]
    ref = []
    keep = []
    for r in range(3):                      # reference: nothing released in between
        hs = _lib.prepare_batch(plan, P[r], n)
        keep.append(hs)
        out = torch.empty_like(x)
        _lib.lenseflow(plan, hs, x, out, B, 2, n, 1, False)
        ref.append(out)
    guards = [[h.min_det for h in hs] for hs in keep]
    torch.cuda.synchronize()
    for hs in keep:
        for h in hs:
            h.close()                       # back to the pool, no wait
    for rep in range(2):
        for r in range(3):                  # prepared into reused memory
            hs = _lib.prepare_batch(plan, P[r], n)
            out = torch.empty_like(x)
            _lib.lenseflow(plan, hs, x, out, B, 2, n, 1, False)
            assert torch.equal(out, ref[r]), (rep, r)
            assert [h.min_det for h in hs] == guards[r]
            for h in hs:
                h.close()
    # eager guard == guard read on demand
    h1 = _lib.PhiHandle(plan, P[0][0], n, with_min_det=True)
    h2 = _lib.PhiHandle(plan, P[0][0], n)
    assert h1.min_det == h2.min_det == guards[0][0]
    # used on another stream: close() waits for the device instead
    side = torch.cuda.Stream(dev)
    with torch.cuda.stream(side):
        out = torch.empty_like(x[:1])
        _lib.lenseflow(plan, [h2], x[:1], out, 1, 2, n, 1, False)
    assert len(h2.streams) == 2
    h2.close()
    h1.close()
    torch.cuda.synchronize()
    assert torch.equal(out, ref[0][:1])
    plan.close()


def test_empty_batch_is_a_no_op(pkg, dev):
    from cmblensing_amd import _lib
    _case("empty batch")
    plan = _lib.get_plan(64, 2.0, torch.float32, 4, dev)
    x = torch.zeros((1, 2, 64, 64), device=dev)
    out = torch.full_like(x, 3.0)
    _lib.lenseflow(plan, [], x, out, 0, 2, 7, 1, False)
    torch.cuda.synchronize()
    assert torch.all(out == 3.0)


def test_negative_batch_and_zero_steps_rejected(pkg, dev):
    """The gradient, cotangent and VJP entry points refuse batch < 0 and
    nsteps < 1 with an error, before any launch (a negative batch would
    otherwise reach a launch as a huge grid); the CG p-update refuses
    nrhs < 1 and ncomp outside 1..3; the gradf last line refuses bad batch,
    ncomp and capacity."""
    from cmblensing_amd import _lib
    _case("argument validation")
    lib = _lib.load()
    plan = _lib.get_plan(64, 2.0, torch.float32, 4, dev)
    buf = torch.zeros(4 * 2 * 64 * 64 * 2, device=dev)
    ptr = ctypes.c_void_p(buf.data_ptr())
    phis = (ctypes.c_void_p * 1)(None)  # never read: the checks come first
    ops = _lib.WfOps()
    ops.mb = ops.bmcn = ops.cfinv = buf.data_ptr()
    for batch, nsteps, msg in [(-1, 7, "batch < 0"), (1, 0, "nsteps must be >= 1")]:
        for rc in [lib.cmbl_gradf_logpdf(plan.handle, phis, ptr, ptr, ctypes.byref(ops), ptr,
                                         batch, 2, nsteps, None),
                   lib.cmbl_lensed_cotangent(plan.handle, phis, ptr, ptr, ctypes.byref(ops), ptr,
                                             ptr, batch, 2, nsteps, None),
                   lib.cmbl_lenseflow_vjp(plan.handle, phis, ptr, ptr, ptr, ptr, batch, 2, nsteps,
                                          1, None),
                   lib.cmbl_lenseflow(plan.handle, phis, ptr, ptr, batch, 2, nsteps, 1, 0, None)]:
            assert rc != 0
            assert msg in lib.cmbl_last_error().decode()
    for nrhs, ncomp in [(0, 2), (-1, 2), (1, 0), (1, 4)]:
        assert lib.cmbl_cg_update_p(plan.handle, ptr, ptr, ptr, None, nrhs, ncomp, None) != 0
        assert "nrhs must be >= 1" in lib.cmbl_last_error().decode()
    # the last line of gradf from the gradient ODE's delta f: batch, ncomp
    # and the plan's capacity are checked before any launch
    for batch, ncomp, msg in [(-1, 2, "batch < 0"), (1, 3, "ncomp must be 1 (T) or 2 (EB)"),
                              (plan.max_maps // 2 + 1, 2, "exceeds the plan's max_maps")]:
        assert lib.cmbl_gradf_final(plan.handle, ptr, ptr, ptr, ptr, batch, ncomp, None) != 0
        assert msg in lib.cmbl_last_error().decode()
    torch.cuda.synchronize()


@pytest.mark.parametrize("adjoint", [False, True])
def test_small_and_large_launch_shapes_agree(pkg, dev, sims, adjoint):
    """A lone request runs with the small-grid launch shapes (64-thread row
    blocks, 4 columns per column block), the same request inside a batch of 64
    with the default shapes: the results are identical."""
    N, th, n = 256, 2.0, 7
    _case(f"launch shapes N=256 QU adjoint={adjoint}")
    fs = np.stack([sims.qu_map(N, th, s % 8) for s in range(64)])
    x = torch.as_tensor(fs, dtype=torch.float32, device=dev)
    P = torch.as_tensor(sims.phi_map(N, th, 9), device=dev)
    yb = pkg.lense_flow_matrices(x, P, th, n, 1, adjoint)
    for s in (0, 37):
        ys = pkg.lense_flow_matrices(x[s:s + 1].contiguous(), P, th, n, 1, adjoint)
        assert torch.equal(ys[0], yb[s])


@pytest.mark.parametrize("basis", ["map", "fourier"])
def test_phi_prepare_batch_equals_single(pkg, dev, basis):
    """cmbl_phi_prepare_batch (70 phis: two preparation chunks) gives the
    derivatives, min_det and LenseFlow results of one-at-a-time preparation,
    bit for bit; handles are released in any order."""
    print(f"[case] phi prepare batch {basis}", flush=True)
    from cmblensing_amd import _lib
    N, th, n, B = 32, 2.0, 5, 70
    rng = np.random.default_rng(11)
    maps = np.stack([sims_phi(N, th, rng) for _ in range(B)])
    phis = torch.as_tensor(maps if basis == "map" else np.fft.rfft2(maps), device=dev)
    plan = _lib.Plan(N, th, torch.float64, B, dev)
    batch = _lib.prepare_batch(plan, phis, n)
    single = [_lib.PhiHandle(plan, phis[i], n) for i in range(B)]
    x = torch.as_tensor(rng.standard_normal((B, 1, N, N)), device=dev)
    ya, yb = torch.empty_like(x), torch.empty_like(x)
    _lib.lenseflow(plan, batch, x, ya, B, 1, n, 1, False)
    _lib.lenseflow(plan, single, x, yb, B, 1, n, 1, False)
    assert torch.equal(ya, yb)
    for i in (0, 63, 64, 69):
        assert torch.equal(batch[i].derivatives(), single[i].derivatives())
        assert batch[i].min_det == single[i].min_det
    for i in list(range(1, B, 2)) + list(range(0, B, 2)):   # odd first, then even
        batch[i].close()
    for h in single:
        h.close()
    plan.close()


def sims_phi(N, th, rng):
    """A smooth random potential with a small deflection (no caustics)."""
    F = np.fft.rfft2(rng.standard_normal((N, N)))
    kx = np.fft.fftfreq(N)[:, None]
    ky = np.fft.rfftfreq(N)[None, :]
    F *= np.exp(-40 * (kx ** 2 + ky ** 2))
    m = np.fft.irfft2(F)
    return m * (1e-6 / np.abs(m).max())


_LENS_COMP = {"I": ["t"], "P": ["e", "b"], "IP": ["t", "e", "b"]}


def _julia_field(pkg, dev, pol, kind):
    """map(eb2qu(field)) of a CMBLensing.jl fixture set (kind: field,
    lensed_field or adjoint_lensed_field), as the reference's tests build it."""
    J = GOLDEN / "julia"
    L = lambda n: torch.as_tensor(np.load(J / f"{n}.npy", allow_pickle=False), device=dev)
    m = [L(f"{c}_{kind}_{pol}") for c in _LENS_COMP[pol]]
    kw = dict(theta_pix=2.5)
    if pol == "I":
        f = pkg.FlatS0(scalar_matrix=m[0], **kw)
    elif pol == "P":
        f = pkg.FlatS2(polar_matrix_1=m[0], polar_matrix_2=m[1], **kw)
    else:
        f = pkg.FlatS02(scalar_matrix=m[0], polar_matrix_1=m[1], polar_matrix_2=m[2], **kw)
    return pkg.map(pkg.eb2qu(f))


@pytest.mark.parametrize("pol", ["I", "P", "IP"])
@pytest.mark.parametrize("mode", ["forward", "adjoint", "inverse", "inverse_adjoint"])
def test_julia_lensing_matrix(pkg, dev, pol, mode):
    """test_lensing.py's twelve cases (FP64, n = 7 like the fixtures):
    forward and adjoint lensing of the unlensed field reproduce CMBLensing.jl's
    lensed / adjoint-lensed maps; inverse (adjoint) lensing of those maps
    returns the unlensed field up to the RK4 round trip.  The reference's bar
    is relative L2 <= 1 (MAX_NORM_DIFF); measured here 1e-13 to 5e-13 for
    forward / adjoint."""
    _case(f"julia lensing {pol} {mode} fp64")
    J = GOLDEN / "julia"
    phi = pkg.FlatS0(scalar_matrix=torch.as_tensor(np.load(J / f"phi_{pol}.npy"), device=dev),
                     theta_pix=2.5)
    phi = pkg.map(phi)
    unl = _julia_field(pkg, dev, pol, "field")
    st = lambda f: np.stack([getattr(f, n).cpu().numpy() for n in f._matrix_names()])
    if mode in ("forward", "adjoint"):
        out = pkg.lense_flow(unl, phi, n=7, adjoint=(mode == "adjoint"))
        ref = _julia_field(pkg, dev, pol, "lensed_field" if mode == "forward"
                           else "adjoint_lensed_field")
        tol = 1e-11
    else:
        src = _julia_field(pkg, dev, pol, "lensed_field" if mode == "inverse"
                           else "adjoint_lensed_field")
        out = pkg.lense_flow(src, phi, n=7, direction=pkg.INVERSE_LENSE,
                             adjoint=(mode == "inverse_adjoint"))
        ref = unl
        tol = 2e-4              # measured 3e-6 .. 6e-5: the 7-step RK4 round trip
    r = rel(st(out), st(ref))
    print(f"{pol} {mode}: rel L2 {r:.3g}", flush=True)
    assert r < tol


@pytest.mark.parametrize("N", [16, 32, 64])
@pytest.mark.parametrize("C", [1, 2])
def test_resident_small_maps(pkg, dev, sims, monkeypatch, N, C):
    """FP32 LenseFlows at N <= 64 run in one launch with each component's
    stage state in LDS (k_lf_resident); CMBL_RESIDENT=0, read when a plan is
    created, runs them through the two-pass stage sequence.  Both directions,
    n = 7 and n = 1, a batch of 3 with a phi per request: each request within
    1e-5 of the FP64 oracle both ways and the two ways within FP32 rounding of
    each other; the allocation-free entry point, and a request alone, equal
    the batched result bit for bit."""
    from cmblensing_amd import _lib
    th, B = 2.0, 3
    _case(f"resident N={N} C={C} batch {B}")
    pw = olf.pix_width_from_theta(th)
    fs = np.stack([(sims.t_map if C == 1 else sims.qu_map)(N, th, s).reshape(C, N, N)
                   for s in range(B)])
    phis = [sims.phi_map(N, th, 10 + s) for s in range(B)]
    x = torch.as_tensor(fs, dtype=torch.float32, device=dev)
    plans = {}
    for resident in (True, False):
        if resident:
            monkeypatch.delenv("CMBL_RESIDENT", raising=False)
        else:
            monkeypatch.setenv("CMBL_RESIDENT", "0")
        plans[resident] = _lib.Plan(N, th, torch.float32, B * C, dev)
    monkeypatch.delenv("CMBL_RESIDENT", raising=False)
    for n in (7, 1):
        hs = {k: [_lib.PhiHandle(p, torch.as_tensor(ph, device=dev), n) for ph in phis]
              for k, p in plans.items()}
        for direction in (1, -1):
            got = {}
            for k, plan in plans.items():
                out = torch.empty_like(x)
                _lib.lenseflow(plan, hs[k], x, out, B, C, n, direction, False)
                got[k] = out.double().cpu().numpy()
            for b in range(B):
                ref = olf.lense_flow(fs[b][0] if C == 1 else fs[b], phis[b], pw, n, direction,
                                     False)
                ref = np.asarray(ref).reshape(C, N, N)
                assert rel(got[True][b], ref) < 1e-5, (n, direction, b)
                assert rel(got[False][b], ref) < 1e-5, (n, direction, b)
            assert rel(got[True], got[False]) < 1e-5
            # the allocation-free entry point takes the same one-launch path
            na = torch.empty_like(x)
            _lib.lenseflow_noalloc(plans[True], hs[True], x, na, B, C, n, direction, False)
            assert torch.equal(na.cpu(), torch.as_tensor(got[True], dtype=torch.float32))
            # request 1 alone on the resident path equals its place in the batch
            one = torch.empty_like(x[1:2])
            _lib.lenseflow(plans[True], hs[True][1:2], x[1:2].contiguous(), one, 1, C, n,
                           direction, False)
            assert torch.equal(one[0].cpu(), torch.as_tensor(got[True][1], dtype=torch.float32))
        for k in hs:
            for h in hs[k]:
                h.close()
    for p in plans.values():
        p.close()


@pytest.mark.parametrize("N,B", [(256, 64), (512, 16)])
def test_streaming_accumulator_accesses(pkg, dev, sims, monkeypatch, N, B):
    """FP32 calls whose stage buffers (S0, S1, y, acc) take at least 128 MiB
    read and write the RK4 accumulator and the state as streaming
    (non-temporal) accesses, a cache policy only: forward and adjoint outputs
    are bit-identical to those of a plan with it off (CMBL_STREAM_ACC=0, read
    when a plan is created), the split into two halves included, and agree
    with the FP64 oracle (1e-5)."""
    from cmblensing_amd import _lib
    th, n, C = 2.0, 7, 2
    assert B * C * N * N * 4 * 4 >= 128 << 20  # the size that takes the streaming kernels
    _case(f"streaming accumulator N={N} QU batch {B} fp32")
    pw = olf.pix_width_from_theta(th)
    fs = np.stack([sims.qu_map(N, th, s % 8) for s in range(B)])
    phis = [sims.phi_map(N, th, 100 + k) for k in range(4)]
    x = torch.as_tensor(fs, dtype=torch.float32, device=dev)
    plans = {}
    for on in (True, False):
        if on:
            monkeypatch.delenv("CMBL_STREAM_ACC", raising=False)
        else:
            monkeypatch.setenv("CMBL_STREAM_ACC", "0")
        plans[on] = _lib.Plan(N, th, torch.float32, B * C, dev)
    monkeypatch.delenv("CMBL_STREAM_ACC", raising=False)
    hs = {k: [_lib.PhiHandle(p, torch.as_tensor(ph, device=dev), n) for ph in phis]
          for k, p in plans.items()}
    for direction, adjoint in ((1, False), (1, True)):
        got = {}
        for k, plan in plans.items():
            out = torch.empty_like(x)
            _lib.lenseflow(plan, [hs[k][b % 4] for b in range(B)], x, out, B, C, n, direction,
                           adjoint)
            got[k] = out
        torch.cuda.synchronize()
        assert torch.equal(got[True], got[False]), (direction, adjoint)
        for b in (0, B - 3):
            ref = olf.lense_flow(fs[b], phis[b % 4], pw, n, direction, adjoint)
            e = rel(got[True][b].double().cpu().numpy(), ref)
            print(f"[err] streaming dir={direction} adj={adjoint} request {b}: {e:.2e}",
                  flush=True)
            assert e < 1e-5
    for k in hs:
        for h in hs[k]:
            h.close()
    for p in plans.values():
        p.close()


@pytest.mark.parametrize("n", [1, 3, 10, 16])
@pytest.mark.parametrize("N,C,dtype,tol", [(64, 2, torch.float64, 1e-11),
                                           (32, 1, torch.float32, 1e-5),
                                           (128, 2, torch.float32, 1e-5)])
def test_other_step_counts(pkg, dev, sims, n, N, C, dtype, tol):
    """Step counts other than the workloads' n = 7 (the reference's default
    is n = 10, lense_flow.py:289): the phi tables' 2n+1 time points and each
    stage's time index, in all four modes, through the two-pass stage
    sequence (64^2 QU FP64, 128^2 QU FP32) and the resident kernel (32^2 T
    FP32: the standard rate), against the FP64 oracle."""
    th = 2.0
    _case(f"step count n={n} N={N} C={C} {dtype}")
    pw = olf.pix_width_from_theta(th)
    f = sims.qu_map(N, th, 11) if C == 2 else sims.t_map(N, th, 11)
    phi = sims.phi_map(N, th, 12)
    T = lambda a: torch.as_tensor(a, dtype=dtype, device=dev)
    if C == 2:
        field = pkg.FlatS2(polar_matrix_1=T(f[0]), polar_matrix_2=T(f[1]), theta_pix=th)
    else:
        field = pkg.FlatS0(scalar_matrix=T(f[0]), theta_pix=th)
    phif = pkg.FlatS0(scalar_matrix=torch.as_tensor(phi, device=dev), theta_pix=th)
    for direction, adjoint in MODES:
        out = pkg.lense_flow(field, phif, n=n, direction=direction, adjoint=adjoint)
        got = np.stack([getattr(out, nm).double().cpu().numpy() for nm in out._matrix_names()])
        ref = np.asarray(olf.lense_flow(f if C == 2 else f[0], phi, pw, n, direction, adjoint))
        e = rel(got.reshape(ref.shape), ref)
        print(f"[err] n={n} dir={direction} adj={adjoint}: {e:.2e}", flush=True)
        assert e < tol
