"""GPU parity of the hand-written LenseFlow backward (gradient ODE), the
autograd path, logpdf and grad_phi_logpdf."""
import numpy as np
import pytest
import torch

from oracle import lenseflow as olf
from oracle import wiener as owf

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("C", [1, 2])
@pytest.mark.parametrize("basis", ["map", "fourier"])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_vjp_matches_oracle(pkg, dev, sims, C, basis, dtype, tol):
    from cmblensing_amd.lense_flow import lense_flow_vjp_matrices
    N, th, n = 64, 2.0, 7
    _case(f"vjp N={N} C={C} phi={basis} {dtype}")
    pw = olf.pix_width_from_theta(th)
    f = sims.qu_map(N, th, 1) if C == 2 else sims.t_map(N, th, 1)
    phi = sims.phi_map(N, th, 2)
    phin = phi if basis == "map" else np.fft.rfft2(phi)
    y = olf.lense_flow(f, phi, pw, n)
    cot = np.random.default_rng(3).standard_normal(f.shape)
    _, dfs, dphi = olf.lensing_operator_gradients(phin, list(y), list(cot), -1, n, pw)
    T = lambda a: torch.as_tensor(a, device=dev)
    d_field, d_phi = lense_flow_vjp_matrices(T(y[None]).to(dtype), T(cot[None]).to(dtype), T(phin),
                                             th, n, 1)
    e_field = rel(d_field[0].double().cpu().numpy(), np.stack(dfs))
    got = d_phi[0].cpu().numpy()
    if basis == "map":
        got = np.fft.irfft2(got)
    e_phi = rel(got, dphi)
    print(f"[err] d_field {e_field:.2e} d_phi {e_phi:.2e}", flush=True)
    assert e_field < tol
    assert e_phi < tol


@pytest.mark.parametrize("n", [1, 10])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_vjp_other_step_counts(pkg, dev, sims, n, dtype, tol):
    """The gradient ODE at step counts other than n = 7 (the reference's
    default n = 10, lense_flow.py:10-51, and a single step): its stage times
    run backwards over the same 2n+1-point tables."""
    from cmblensing_amd.lense_flow import lense_flow_vjp_matrices
    N, th = 64, 2.0
    _case(f"vjp n={n} N={N} QU {dtype}")
    pw = olf.pix_width_from_theta(th)
    f = sims.qu_map(N, th, 21)
    phi = sims.phi_map(N, th, 22)
    y = olf.lense_flow(f, phi, pw, n)
    cot = np.random.default_rng(23).standard_normal(f.shape)
    _, dfs, dphi = olf.lensing_operator_gradients(phi, list(y), list(cot), -1, n, pw)
    T = lambda a: torch.as_tensor(a, device=dev)
    d_field, d_phi = lense_flow_vjp_matrices(T(np.asarray(y)[None]).to(dtype),
                                             T(cot[None]).to(dtype), T(phi), th, n, 1)
    e_field = rel(d_field[0].double().cpu().numpy(), np.stack(dfs))
    e_phi = rel(np.fft.irfft2(d_phi[0].cpu().numpy()), dphi)
    print(f"[err] n={n}: d_field {e_field:.2e} d_phi {e_phi:.2e}", flush=True)
    assert e_field < tol
    assert e_phi < tol


def test_autograd_matches_finite_differences(pkg, dev, sims):
    """torch.autograd through LenseFlowFunction: d/dx exact (linear), d/dphi
    within the continuous-adjoint gap of a central difference."""
    from cmblensing_amd.lense_flow import LenseFlowFunction
    N, th, n = 64, 2.0, 7
    _case("autograd fd N=64 QU fp64")
    T = lambda a: torch.as_tensor(a, device=dev)
    x = T(sims.qu_map(N, th, 4)[None]).requires_grad_(True)
    phi = T(sims.phi_map(N, th, 5)).requires_grad_(True)
    c = T(np.random.default_rng(6).standard_normal((1, 2, N, N)))
    L = (LenseFlowFunction.apply(x, phi, th, n, 1) * c).sum()
    gx, gphi = torch.autograd.grad(L, (x, phi))
    rng = np.random.default_rng(7)
    with torch.no_grad():
        for _ in range(3):
            v = T(rng.standard_normal(x.shape))
            e = 1.0
            fd = ((LenseFlowFunction.apply(x + e * v, phi, th, n, 1) * c).sum()
                  - (LenseFlowFunction.apply(x - e * v, phi, th, n, 1) * c).sum()) / (2 * e)
            assert abs(float((gx * v).sum()) - float(fd)) < 1e-9 * abs(float(fd))
            # a smooth, phi-like direction (white noise would be dominated by
            # the nonlinearity of M(t)^-1 in Hess phi at the pixel scale)
            w = T(sims.phi_map(N, th, int(rng.integers(1000))))
            e = 1e-3
            fd = ((LenseFlowFunction.apply(x, phi + e * w, th, n, 1) * c).sum()
                  - (LenseFlowFunction.apply(x, phi - e * w, th, n, 1) * c).sum()) / (2 * e)
            # continuous-adjoint gap at n = 7: ~1e-3, falling ~16x per doubling of n
            assert abs(float((gphi * w).sum()) - float(fd)) < 5e-3 * abs(float(fd))


def _julia_I(pkg, dev):
    from cmblensing_amd import matrix_operators as mo
    J = GOLDEN / "julia"
    L = lambda n: np.load(J / f"{n}.npy")
    T = lambda a: torch.as_tensor(a, device=dev)
    S = lambda a: pkg.FlatS0(scalar_matrix=T(a), theta_pix=2.5)
    D = lambda a: mo.DiagonalScalar(scalar_matrix=T(a), theta_pix=2.5)
    return L, S, D


def test_logpdf_julia_I(pkg, dev):
    from cmblensing_amd import statistics
    _case("logpdf julia I fp64")
    L, S, D = _julia_I(pkg, dev)
    lp = statistics.logpdf(S(L("t_field_I")), S(L("phi_I")), S(L("data_t_field_I")), D(L("cn_I")),
                           D(L("cphi_I")), D(L("cf_I")), D(L("m_I")), D(L("b_I")), n=7)
    ref = float((GOLDEN / "julia" / "logpdf_I.txt").read_text())
    assert abs(float(lp) - ref) < 1e-9 * abs(ref)


def test_grad_phi_julia_I(pkg, dev):
    from cmblensing_amd import statistics
    _case("grad_phi julia I fp64")
    L, S, D = _julia_I(pkg, dev)
    g = statistics.grad_phi_logpdf(S(L("t_field_I")), S(L("phi_I")), S(L("data_t_field_I")),
                                   D(L("cn_I")), D(L("cphi_I")), D(L("cf_I")), D(L("m_I")),
                                   D(L("b_I")), n=7).scalar_matrix.cpu().numpy()
    ref = owf.grad_phi_logpdf(L("t_field_I")[None], L("phi_I"), L("data_t_field_I")[None],
                              L("cn_I")[None], L("cphi_I"), L("cf_I")[None], L("m_I")[None],
                              L("b_I")[None], 2.5, n=7)
    assert rel(g, ref) < 1e-9
    assert rel(g, L("grad_phi_I")) < 2e-3


def test_logpdf_and_grad_phi_julia_P(pkg, dev):
    """logpdf and its phi gradient on CMBLensing.jl's polarisation problem
    (EB operators, FOURIER phi): logpdf to 1e-12 relative, the gradient
    against the FP64 oracle to 1e-9 and the Julia fixture to 2e-4."""
    from cmblensing_amd import statistics
    from test_gpu_wiener import _julia_P
    _case("logpdf / grad_phi julia P fp64")
    o, a, L = _julia_P(pkg, dev)
    lp = statistics.logpdf(o["f"], o["phi"], o["d"], o["Cn"], o["Cphi"], o["Cf"], o["M"], o["B"],
                           n=7)
    ref = float((GOLDEN / "julia" / "logpdf_P.txt").read_text())
    assert abs(float(lp) - ref) < 1e-12 * abs(ref)
    g = statistics.grad_phi_logpdf(o["f"], o["phi"], o["d"], o["Cn"], o["Cphi"], o["Cf"], o["M"],
                                   o["B"], n=7).scalar_matrix.cpu().numpy()
    gref = owf.grad_phi_logpdf(a["f"], a["phi"], a["d"], a["Cn"], a["Cphi"], a["Cf"], a["M"],
                               a["B"], 2.5, n=7)
    assert rel(g, gref) < 1e-9
    assert rel(g, L("grad_phi_t_P")) < 2e-4


def test_differentiation_workload_batch16(pkg, dev, sims):
    """QU N=256, 2 arcmin, float32 gradients of logpdf w.r.t. f and phi for 16
    chains (square pixel mask, 1 arcmin beam, 5 uK-arcmin noise), checked
    against float64 central differences on the GPU along random directions."""
    from cmblensing_amd import gradients, statistics, matrix_operators as mo
    N, th, n, B = 256, 2.0, 7, 16
    _case(f"differentiation workload N={N} QU batch={B}")
    prob = sims.wf_problem(N, th, pol=True, seed=21, mask="square", noise_uk=5.0, beam_fwhm=1.0)
    fq = owf.to_map(owf.eb2qu_field(prob["f"], N, th))
    prob = sims.wf_data(prob, olf.lense_flow(fq, prob["phi"], olf.pix_width_from_theta(th), n))
    phi_m, Cphi = sims.phi_map_and_cov(N, th, 2)
    rng = np.random.default_rng(22)

    def objs(dt, cdt):
        T = lambda a, d=cdt: torch.as_tensor(a, device=dev).to(d)
        # 16 chains: the truth plus small seeded offsets
        Fs = np.stack([prob["f"] * (1 + 0.01 * k) for k in range(B)])
        f = pkg.FlatS2(polar_matrix_1=T(Fs[:, 0]), polar_matrix_2=T(Fs[:, 1]), theta_pix=th)
        data = pkg.FlatS2(polar_matrix_1=T(prob["d"][0], dt), polar_matrix_2=T(prob["d"][1], dt),
                          theta_pix=th)
        op = lambda A: mo.DiagonalEB(matrix_EE=T(A[0], torch.float64),
                                     matrix_BB=T(A[1], torch.float64), theta_pix=th)
        phiF = pkg.FlatS0(scalar_matrix=T(np.fft.rfft2(phi_m)), theta_pix=th)
        Cp = mo.DiagonalScalar(scalar_matrix=T(Cphi, torch.float64), theta_pix=th)
        M = mo.PixelMask(matrix=torch.as_tensor(prob["mask_pix"], device=dev))
        return f, phiF, data, op(prob["Cf"]), op(prob["Cn"]), M, op(prob["B"]), Cp

    f32 = objs(torch.float32, torch.complex64)
    f, phiF, data, Cf, Cn, M, Bm, Cp = f32
    gf = gradients.gradf_logpdf(f, phiF, data, Cf, Cn, M, Bm, n=n)
    gp = statistics.grad_phi_logpdf(f, phiF, data, Cn, Cp, Cf, M, Bm, n=n)
    f64 = objs(torch.float64, torch.complex128)
    f6, phi6, data6, Cf6, Cn6, M6, B6, Cp6 = f64

    def lp(ff, pp):
        return statistics.logpdf(ff, pp, data6, Cn6, Cp6, Cf6, M6, B6, n=n)

    # 16 seeded directions, each scaled to the norm (under dot) of chain 0's
    # field or phi, central differences at eps = 1e-3 of that scale.  The f
    # directions are white; the phi directions are smooth phi-like fields
    # (white noise in phi is dominated by the nonlinearity of M(t)^-1 in
    # Hess phi at the pixel scale).
    f0 = f6.replace(polar_matrix_1=f6.polar_matrix_1[0], polar_matrix_2=f6.polar_matrix_2[0])
    fnorm = float(pkg.dot(f0, f0)) ** 0.5
    pnorm = float(pkg.dot(phi6, phi6)) ** 0.5
    e = 1e-3
    gp_norm = torch.sqrt(pkg.dot(gp, gp))                      # per chain
    rel_f, rel_p, cs_p = [], [], []
    for k in range(16):
        v = np.fft.rfft2(rng.standard_normal((2, N, N)))
        vt = [torch.as_tensor(v[c], device=dev) for c in range(2)]
        dirf = f6.replace(polar_matrix_1=vt[0], polar_matrix_2=vt[1])
        dirf = dirf * (fnorm / float(pkg.dot(dirf, dirf)) ** 0.5)
        dirf = dirf.replace(polar_matrix_1=dirf.polar_matrix_1.expand(B, -1, -1),
                            polar_matrix_2=dirf.polar_matrix_2.expand(B, -1, -1))
        fd = (lp(f6 + dirf * e, phi6) - lp(f6 - dirf * e, phi6)) / (2 * e)
        an = pkg.dot(gf, dirf)
        rel_f.append(torch.abs(an - fd) / torch.abs(fd))
        w = np.fft.rfft2(sims.phi_map(N, th, 500 + k))
        wt = phi6.replace(scalar_matrix=torch.as_tensor(w, device=dev))
        wt = wt * (pnorm / float(pkg.dot(wt, wt)) ** 0.5)
        fdp = (lp(f6, phi6 + wt * e) - lp(f6, phi6 - wt * e)) / (2 * e)
        anp = pkg.dot(gp, wt.replace(scalar_matrix=wt.scalar_matrix.expand(B, -1, -1)))
        rel_p.append(torch.abs(anp - fdp) / torch.abs(fdp))
        # the same gap relative to |g| |w| (Cauchy-Schwarz scale): directions
        # nearly orthogonal to the gradient make the plain ratio large
        cs_p.append(torch.abs(anp - fdp) / (gp_norm * pnorm))
    rel_f, rel_p, cs_p = (torch.stack(a).flatten() for a in (rel_f, rel_p, cs_p))
    print(f"[err] 16 directions x {B} chains: f max {float(rel_f.max()):.2e}; phi max "
          f"{float(rel_p.max()):.2e}, median {float(rel_p.median()):.2e}, max relative to "
          f"|g||w| {float(cs_p.max()):.2e}", flush=True)
    assert float(rel_f.max()) < 1e-5
    # the phi gradient is the discrete RK4 adjoint of the gradient ODE (as the
    # reference's): its gap to the derivative of the discrete logpdf is the
    # continuous-adjoint gap at n = 7, ~1e-3 of the directional derivative
    assert float(rel_p.median()) < 5e-3
    assert float(cs_p.max()) < 1e-5


def _autograd_vs_native(pkg, f, phi, d, Cn, Cphi, Cf, M, B, n=7):
    """torch.autograd.grad of logpdf with respect to every matrix of f and to
    phi, and the native gradf_logpdf / grad_phi_logpdf of the same problem."""
    from cmblensing_amd import gradients, statistics
    names = f._matrix_names()
    fm = {nm: getattr(f, nm).detach().clone().requires_grad_(True) for nm in names}
    pm = phi.scalar_matrix.detach().clone().requires_grad_(True)
    fa, pa = f.replace(**fm), phi.replace(scalar_matrix=pm)
    lp = statistics.logpdf(fa, pa, d, Cn, Cphi, Cf, M, B, n=n)
    grads = torch.autograd.grad(lp.sum(), [fm[nm] for nm in names] + [pm])
    gf_auto = torch.stack(grads[:-1]).cpu().numpy()
    gp_auto = grads[-1].cpu().numpy()
    gf = gradients.gradf_logpdf(f, phi, d, Cf, Cn, M, B, n=n)
    gf_nat = np.stack([getattr(gf, nm).cpu().numpy() for nm in names])
    gp_nat = statistics.grad_phi_logpdf(f, phi, d, Cn, Cphi, Cf, M, B, n=n).scalar_matrix
    return gf_auto, gp_auto, gf_nat, gp_nat.cpu().numpy(), float(lp.detach().sum())


def test_logpdf_autograd_julia_I(pkg, dev):
    """torch.autograd.grad(logpdf, (f, phi)) with a FOURIER phi on the
    reference's I problem equals gradf_logpdf and grad_phi_logpdf, and meets
    the CMBLensing.jl grad_phi_I fixture like grad_phi_logpdf does."""
    _case("logpdf autograd julia I fp64")
    L, S, D = _julia_I(pkg, dev)
    gfa, gpa, gfn, gpn, lp = _autograd_vs_native(
        pkg, S(L("t_field_I")), S(L("phi_I")), S(L("data_t_field_I")), D(L("cn_I")),
        D(L("cphi_I")), D(L("cf_I")), D(L("m_I")), D(L("b_I")))
    ef, ep, ej = rel(gfa, gfn), rel(gpa, gpn), rel(gpa, L("grad_phi_I"))
    print(f"[err] autograd vs gradf {ef:.2e}, vs grad_phi {ep:.2e}, vs julia {ej:.2e}", flush=True)
    assert abs(lp - float((GOLDEN / "julia" / "logpdf_I.txt").read_text())) < 1e-9 * abs(lp)
    assert ef < 1e-12
    assert ep < 1e-12
    assert ej < 2e-3


def test_logpdf_autograd_julia_P(pkg, dev):
    """The same on the reference's polarisation problem (EB field, FOURIER
    phi), against the grad_phi_t_P fixture."""
    from test_gpu_wiener import _julia_P
    _case("logpdf autograd julia P fp64")
    o, a, L = _julia_P(pkg, dev)
    gfa, gpa, gfn, gpn, _ = _autograd_vs_native(pkg, o["f"], o["phi"], o["d"], o["Cn"], o["Cphi"],
                                                o["Cf"], o["M"], o["B"])
    ef, ep, ej = rel(gfa, gfn), rel(gpa, gpn), rel(gpa, L("grad_phi_t_P"))
    print(f"[err] autograd vs gradf {ef:.2e}, vs grad_phi {ep:.2e}, vs julia {ej:.2e}", flush=True)
    assert ef < 1e-12
    assert ep < 1e-12
    assert ej < 2e-4


def test_logpdf_autograd_pixel_mask_batch(pkg, dev, sims):
    """Autograd through logpdf with a pixel mask (the residual formed on the
    Q/U maps) for a batch of 2 chains with a FOURIER phi each, against the
    native gradients and the FP64 oracle's grad_phi_logpdf."""
    from cmblensing_amd import matrix_operators as mo
    from test_gpu_wiener import make_problem
    N, th = 64, 2.0
    _case(f"logpdf autograd N={N} QU square mask, 2 chains fp64")
    probs = [make_problem(sims, N, th, "square", seed=40 + k) for k in range(2)]
    phi_m, Cphi = sims.phi_map_and_cov(N, th, 41)
    T = lambda a, dt=torch.complex128: torch.as_tensor(a, device=dev).to(dt)
    op = lambda A: mo.DiagonalEB(matrix_EE=T(A[0], torch.float64), matrix_BB=T(A[1], torch.float64),
                                 theta_pix=th)
    f = pkg.FlatS2(polar_matrix_1=T(np.stack([p["f"][0] for p in probs])),
                   polar_matrix_2=T(np.stack([p["f"][1] for p in probs])), theta_pix=th)
    d = pkg.FlatS2(polar_matrix_1=T(np.stack([p["d"][0] for p in probs]), torch.float64),
                   polar_matrix_2=T(np.stack([p["d"][1] for p in probs]), torch.float64),
                   theta_pix=th)
    phis = np.stack([np.fft.rfft2(phi_m), np.fft.rfft2(probs[1]["phi"])])
    phi = pkg.FlatS0(scalar_matrix=T(phis), theta_pix=th)
    Cp = mo.DiagonalScalar(scalar_matrix=T(Cphi, torch.float64), theta_pix=th)
    M = mo.PixelMask(matrix=torch.as_tensor(probs[0]["mask_pix"], device=dev))
    p0 = probs[0]
    gfa, gpa, gfn, gpn, _ = _autograd_vs_native(pkg, f, phi, d, op(p0["Cn"]), Cp, op(p0["Cf"]), M,
                                                op(p0["B"]))
    ef, ep = rel(gfa, gfn), rel(gpa, gpn)
    eo = max(rel(gpa[k], owf.grad_phi_logpdf(probs[k]["f"], phis[k], probs[k]["d"], p0["Cn"], Cphi,
                                             p0["Cf"], None, p0["B"], th, n=7,
                                             mask_pix=p0["mask_pix"])) for k in range(2))
    print(f"[err] autograd vs gradf {ef:.2e}, vs grad_phi {ep:.2e}, vs oracle {eo:.2e}", flush=True)
    assert ef < 1e-12
    assert ep < 1e-12
    assert eo < 1e-9


@pytest.mark.parametrize("pol", [True, False])
@pytest.mark.parametrize("mask", ["fourier", "square"])
@pytest.mark.parametrize("cdt,tol", [(torch.complex128, 1e-10), (torch.complex64, 1e-5)])
def test_grad_logpdf_equals_separate_gradients(pkg, dev, sims, pol, mask, cdt, tol):
    """statistics.grad_logpdf -- both gradients from one forward LenseFlow and
    one gradient ODE, d/d f from the ODE's delta f -- against gradf_logpdf
    (forward and adjoint LenseFlow) and grad_phi_logpdf, for 3 fields with a
    FOURIER phi each, with the prior and (Cf scaled by 1e30) without it, so
    that the data term is checked where the prior does not dominate: equal in
    FP64 (1e-10), the FP32 result within 1e-5 of the FP64 one, and the FP64
    result against the oracle."""
    from cmblensing_amd import gradients, statistics, matrix_operators as mo
    from test_gpu_wiener import make_problem
    N, th, nb = 64, 2.0, 3
    _case(f"grad_logpdf vs separate N={N} {'QU' if pol else 'T'} {mask} {cdt}")
    probs = [make_problem(sims, N, th, mask, pol=pol, seed=70 + k) for k in range(nb)]
    _, Cphi = sims.phi_map_and_cov(N, th, 71)
    rdt = torch.float32 if cdt == torch.complex64 else torch.float64
    T = lambda a, dt=cdt: torch.as_tensor(a, device=dev).to(dt)
    C = probs[0]["C"]
    names = ("polar_matrix_1", "polar_matrix_2") if C == 2 else ("scalar_matrix",)
    Fld = pkg.FlatS2 if C == 2 else pkg.FlatS0
    pixel = probs[0]["mask_pix"] is not None
    f = Fld(**{nm: T(np.stack([p["f"][c] for p in probs])) for c, nm in enumerate(names)},
            theta_pix=th)
    d = Fld(**{nm: T(np.stack([p["d"][c] for p in probs]), rdt if pixel else cdt)
               for c, nm in enumerate(names)}, theta_pix=th)
    phis = np.stack([np.fft.rfft2(p["phi"]) for p in probs])
    phi = pkg.FlatS0(scalar_matrix=T(phis), theta_pix=th)
    Cp = mo.DiagonalScalar(scalar_matrix=T(Cphi, torch.float64), theta_pix=th)

    def op(A):
        if C == 2:
            return mo.DiagonalEB(matrix_EE=T(A[0], torch.float64),
                                 matrix_BB=T(A[1], torch.float64), theta_pix=th)
        return mo.DiagonalScalar(scalar_matrix=T(A[0], torch.float64), theta_pix=th)

    p0 = probs[0]
    M = mo.PixelMask(matrix=torch.as_tensor(p0["mask_pix"], device=dev)) if pixel else op(p0["M"])
    stk = lambda fl: np.stack([getattr(fl, nm).cpu().numpy() for nm in names], axis=1)
    for scale in (1.0, 1e30):
        Cf = op(p0["Cf"] * scale)
        gf, gp = statistics.grad_logpdf(f, phi, d, op(p0["Cn"]), Cp, Cf, M, op(p0["B"]), n=7)
        gf_s = gradients.gradf_logpdf(f, phi, d, Cf, op(p0["Cn"]), M, op(p0["B"]), n=7)
        gp_s = statistics.grad_phi_logpdf(f, phi, d, op(p0["Cn"]), Cp, Cf, M, op(p0["B"]), n=7)
        ef = rel(stk(gf), stk(gf_s))
        ep = rel(gp.scalar_matrix.cpu().numpy(), gp_s.scalar_matrix.cpu().numpy())
        print(f"[err] Cf x{scale:g}: grad_logpdf vs gradf {ef:.2e}, vs grad_phi {ep:.2e}",
              flush=True)
        assert ep < tol
        if cdt == torch.complex128:
            assert ef < tol
        else:
            # two FP32 routes to d/d f (the adjoint LenseFlow, the gradient
            # ODE's delta f); the data term at the truth is a small residual
            # of large maps, so their FP32 rounding differs by up to ~2e-5 of
            # it for T (printed above): the combined route is held to the FP32
            # bar against FP64
            f64 = Fld(**{nm: getattr(f, nm).to(torch.complex128) for nm in names}, theta_pix=th)
            d64 = Fld(**{nm: getattr(d, nm).to(torch.float64 if pixel else torch.complex128)
                         for nm in names}, theta_pix=th)
            phi64 = pkg.FlatS0(scalar_matrix=T(phis, torch.complex128), theta_pix=th)
            op64 = lambda A: op(A)
            g64, p64 = statistics.grad_logpdf(f64, phi64, d64, op64(p0["Cn"]), Cp, Cf, M,
                                              op64(p0["B"]), n=7)
            e64 = rel(stk(gf), stk(g64))
            e64p = rel(gp.scalar_matrix.cpu().numpy(), p64.scalar_matrix.cpu().numpy())
            print(f"[err] Cf x{scale:g}: FP32 grad_logpdf vs FP64: f {e64:.2e}, phi {e64p:.2e}",
                  flush=True)
            assert e64 < tol and e64p < tol
        if cdt == torch.complex128:
            for k in (0, 2):
                pk = probs[k]
                mk = dict(mask_pix=pk["mask_pix"]) if pixel else {}
                rf = owf.gradf_logpdf(pk["f"], pk["phi"], pk["d"], p0["Cf"] * scale, p0["Cn"],
                                      None if pixel else p0["M"], p0["B"], th, n=7, **mk)
                assert rel(stk(gf)[k], rf) < 1e-9


def test_lense_flow_wrapper_shared_phi_batch(pkg, dev, sims):
    """A phi with a leading axis of 1 shared by a field batch of 2: the phi
    gradient is the sum of the two elements' gradients (MAP and FOURIER phi)."""
    from cmblensing_amd.lense_flow import LenseFlowFunction
    N, th, n = 64, 2.0, 7
    _case("wrapper shared phi [1,N,N] batch 2")
    T = lambda a: torch.as_tensor(a, device=dev)
    x = T(np.stack([sims.qu_map(N, th, 60 + k) for k in range(2)]))
    c = T(np.random.default_rng(61).standard_normal((2, 2, N, N)))
    for phi0 in (sims.phi_map(N, th, 62), np.fft.rfft2(sims.phi_map(N, th, 62))):
        phi = T(phi0[None]).requires_grad_(True)
        (g,) = torch.autograd.grad((LenseFlowFunction.apply(x, phi, th, n, 1) * c).sum(), (phi,))
        assert g.shape == phi.shape
        parts = []
        for k in range(2):
            pk = T(phi0).requires_grad_(True)
            (gk,) = torch.autograd.grad((LenseFlowFunction.apply(x[k:k + 1], pk, th, n, 1)
                                         * c[k:k + 1]).sum(), (pk,))
            parts.append(gk)
        assert rel(g[0].cpu().numpy(), (parts[0] + parts[1]).cpu().numpy()) < 1e-13


@pytest.mark.parametrize("dtype,tol,agree", [(torch.float64, 1e-11, 1e-12),
                                             (torch.float32, 1e-5, 1e-5)])
def test_adjoint_bookkeepings_agree(pkg, dev, sims, monkeypatch, dtype, tol, agree):
    """The adjoint rate's row pass with the low-storage RK4 bookkeeping (the
    default: the stage input stored in place of the V spectra) and with the
    classic accumulator (CMBL_ADJ_CLASSIC=1, read when a plan is created) run
    the same RK4: both adjoint directions and the gradient ODE (whose delta-f
    rate uses the same row pass) agree with each other and with the oracle."""
    from cmblensing_amd import _lib
    N, th, n, C, B = 64, 2.0, 7, 2, 3
    _case(f"adjoint bookkeepings N={N} QU batch {B} {dtype}")
    pw = olf.pix_width_from_theta(th)
    fs = np.stack([sims.qu_map(N, th, s) for s in range(B)])
    phi = sims.phi_map(N, th, 4)
    x = torch.as_tensor(fs, dtype=dtype, device=dev)
    P = torch.as_tensor(phi, device=dev)
    plans = {}
    for classic in (False, True):
        if classic:
            monkeypatch.setenv("CMBL_ADJ_CLASSIC", "1")
        else:
            monkeypatch.delenv("CMBL_ADJ_CLASSIC", raising=False)
        plans[classic] = _lib.Plan(N, th, dtype, B * C, dev)
    monkeypatch.delenv("CMBL_ADJ_CLASSIC", raising=False)
    handles = {k: _lib.PhiHandle(p, P, n) for k, p in plans.items()}
    for direction in (1, -1):
        got = {}
        for k, plan in plans.items():
            out = torch.empty_like(x)
            _lib.lenseflow(plan, [handles[k]] * B, x, out, B, C, n, direction, True)
            got[k] = out.double().cpu().numpy()
        ref = olf.lense_flow(fs[1], phi, pw, n, direction, True)
        e_ls, e_cl, e_ab = rel(got[False][1], ref), rel(got[True][1], ref), rel(got[False], got[True])
        print(f"[err] adjoint dir={direction}: low-storage {e_ls:.2e} classic {e_cl:.2e} "
              f"between {e_ab:.2e}", flush=True)
        assert e_ls < tol and e_cl < tol and e_ab < agree
    # the gradient ODE (lense_flow_wrapper backward) both ways
    y = np.stack([np.stack(olf.lense_flow(fs[b], phi, pw, n)) for b in range(B)])
    cot = np.random.default_rng(5).standard_normal(y.shape)
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    res = {}
    for k, plan in plans.items():
        yl = torch.as_tensor(y, dtype=dtype, device=dev)
        ct = torch.as_tensor(cot, dtype=dtype, device=dev)
        d_field = torch.empty_like(yl)
        d_phi = torch.empty((B, N, N // 2 + 1), dtype=cdt, device=dev)
        _lib.lenseflow_vjp(plan, [handles[k]] * B, yl, ct, d_field, torch.view_as_real(d_phi), B,
                           C, n, 1)
        res[k] = (d_field.double().cpu().numpy(), d_phi.cpu().numpy())
    _, dfs, dphi = olf.lensing_operator_gradients(phi, list(y[0]), list(cot[0]), -1, n, pw)
    for k in plans:
        e_f = rel(res[k][0][0], np.stack(dfs))
        e_p = rel(np.fft.irfft2(res[k][1][0]), dphi)
        print(f"[err] vjp classic={k}: d_field {e_f:.2e} d_phi {e_p:.2e}", flush=True)
        assert e_f < tol and e_p < tol
    assert rel(res[False][0], res[True][0]) < agree
    assert rel(res[False][1], res[True][1]) < agree
    for h in handles.values():
        h.close()
    for p in plans.values():
        p.close()
