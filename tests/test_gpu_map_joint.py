"""GPU parity of the MAP-joint driver (cmblensing_amd.map_joint) against the
oracle (oracle/map_joint.py) on data sets simulated on the device, and its
batched forms.  Tolerances (FP64): f and phi after two constant-step
iterations with fixed-length Wiener filters 1e-8 relative L2; line-search
values 1e-10 relative; batched trials vs single runs 1e-8."""
import numpy as np
import pytest
import torch

from oracle import map_joint as omj
from oracle import simulate as osim

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def psim(pkg):
    from cmblensing_amd import simulate
    return simulate


@pytest.fixture(scope="module")
def mj(pkg):
    from cmblensing_amd import map_joint
    return map_joint


@pytest.fixture(scope="module")
def cls(sims):
    return osim.extract_all_cls(*sims.camb_like_spectra(5000), 17000, 5000)


def _dataset(psim, cls, dev, N, th, pol, seed, ntrials=1):
    cfg = psim.sim_config(N, th, pol, cls, uk_arcmin_t=1.0, device=dev)
    return cfg, psim.build_dataset(cfg, *psim.simulate(cfg, seed, ntrials, torch.float64, n=10))


def _np(ds, pol):
    c = lambda t: t.detach().cpu().numpy()
    if pol == "I":
        arrs = lambda op: c(op.scalar_matrix)[None]
        d = c(ds.data.scalar_matrix)[None]
    else:
        arrs = lambda op: np.stack([c(op.matrix_EE), c(op.matrix_BB)])
        d = np.stack([c(ds.data.polar_matrix_1), c(ds.data.polar_matrix_2)])
    return dict(d=d, Cf=arrs(ds.field_covariance), Cn=arrs(ds.noise_covariance),
                M=arrs(ds.mask), B=arrs(ds.beam), D=arrs(ds.mixing_d),
                Cphi=c(ds.phi_covariance.scalar_matrix), QE=c(ds.quadratic_estimate.scalar_matrix))


def _fields(f, pol):
    c = lambda t: t.detach().cpu().numpy()
    if pol == "I":
        return c(f.scalar_matrix)[None]
    return np.stack([c(f.polar_matrix_1), c(f.polar_matrix_2)])


# alpha = 1 is stable for temperature only (the reference says as much): P
# takes the grid line search
@pytest.mark.parametrize("pol,step", [("I", "constant"), ("P", "grid")])
def test_map_joint_matches_oracle(psim, mj, dev, cls, pol, step):
    _case(f"map_joint {step} step vs oracle {pol}")
    N, th = 32, 2.0
    cfg, ds = _dataset(psim, cls, dev, N, th, pol, 4)
    # tol > 0: at phi = 0 the preconditioner is exact and PCG converges in one
    # iteration; iterating further on round-off (tol = 0) breaks down
    f, phi = mj.map_joint(ds, num_steps=2, step=step, n=10, maxiter=15, tol=1e-6)
    a = _np(ds, pol)
    fo, phio = omj.map_joint(a["d"], a["Cf"], a["Cn"], a["Cphi"], a["QE"], a["M"], a["B"], a["D"],
                             th, num_steps=2, step=step, n=10, maxiter=15, tol=1e-6)
    assert rel(_fields(f, pol), fo) < 1e-8
    assert rel(phi.scalar_matrix.cpu(), phio) < 1e-8


def test_line_search_values_match_oracle(psim, mj, pkg, dev, cls):
    _case("line search values vs oracle")
    from cmblensing_amd import mixing, wiener_filter as wfm
    from oracle import mixing as omx
    N, th = 32, 2.0
    cfg, ds = _dataset(psim, cls, dev, N, th, "I", 6)
    G = mj.get_scalar_identity(ds.mixing_d)
    H = mj.pinv(mj.phi_gradient_hessian(ds.phi_covariance, ds.quadratic_estimate))
    f = wfm.wiener_filter(0 * ds.data, mj.zero_scalar_field_like(ds.data), ds.data,
                          ds.field_covariance, ds.noise_covariance, ds.mask, ds.beam, n=10,
                          maxiter=10, tol=1e-6)
    phi = ds.phi.replace(scalar_matrix=0.3 * ds.phi.scalar_matrix)
    fm, phim = mixing.mix(f, phi, ds.mixing_d, G, 10)
    g = mixing.mixed_grad_phi_logpdf(fm, phim, ds.data, ds.noise_covariance, ds.phi_covariance,
                                     ds.field_covariance, ds.mask, ds.beam, ds.mixing_d, G, n=10)
    best, vals = mj.line_search(2.0, phim, H, g, fm, ds.mixing_d, G, ds.data, ds.noise_covariance,
                                ds.phi_covariance, ds.field_covariance, ds.mask, ds.beam, n=10,
                                return_values=True)
    a = _np(ds, "I")
    Go = np.ones((N, N // 2 + 1))
    Ho = H.scalar_matrix.cpu().numpy()
    fmo = fm.scalar_matrix.cpu().numpy()[None]
    go = omx.mixed_grad_phi_logpdf(fmo, phim.scalar_matrix.cpu().numpy(), a["d"], a["Cn"],
                                   a["Cphi"], a["Cf"], a["M"], a["B"], a["D"], Go, th, 10)
    assert rel(g.scalar_matrix.cpu(), go) < 1e-9
    bo, vo = omj.line_search(2.0, phim.scalar_matrix.cpu().numpy(), Ho, go, fmo, a["D"], Go, a["d"],
                             a["Cn"], a["Cphi"], a["Cf"], a["M"], a["B"], th, 10)
    assert np.allclose(vals.cpu().numpy(), vo, rtol=1e-10, atol=0)
    assert best == bo


def test_map_joint_batched_trials_match_single_runs(psim, mj, dev, cls):
    _case("map_joint batched trials vs single")
    N, th = 32, 2.0
    cfg = psim.sim_config(N, th, "I", cls, uk_arcmin_t=1.0, device=dev)
    arrs = psim.simulate(cfg, 8, 3, torch.float64, n=10)
    ds3 = psim.build_dataset(cfg, *arrs)
    f3, phi3 = mj.map_joint(ds3, num_steps=2, step="grid", n=10, maxiter=12, tol=1e-6)
    for t in (0, 2):
        ds1 = psim.build_dataset(cfg, *[x[t:t + 1] for x in arrs])
        f1, phi1 = mj.map_joint(ds1, num_steps=2, step="grid", n=10, maxiter=12, tol=1e-6)
        assert rel(phi3.scalar_matrix[t].cpu(), phi1.scalar_matrix.cpu()) < 1e-8
        assert rel(f3.scalar_matrix[t].cpu(), f1.scalar_matrix.cpu()) < 1e-8
    with pytest.raises(ValueError):
        mj.map_joint(ds3, num_steps=1)          # BFGS needs a single data set


def _corr(a, b):
    a, b = a.flatten().to(torch.complex128), b.flatten().to(torch.complex128)
    return float(torch.real(torch.vdot(a, b)) /
                 torch.sqrt(torch.vdot(a, a).real * torch.vdot(b, b).real))


# BFGS starts at alpha = 1 (the reference's choice); on the synthetic P
# spectra that first probe already folds the map and the search stalls there,
# so P takes the grid search, whose range includes alpha = 0
@pytest.mark.parametrize("pol,step", [("I", "bfgs"), ("P", "grid")])
def test_map_joint_on_simulated_data(psim, mj, dev, cls, pol, step):
    _case(f"map_joint {step} simulated {pol}")
    N, th = 64, 2.0
    cfg, ds = _dataset(psim, cls, dev, N, th, pol, 12)
    hist = []
    f, phi = mj.map_joint(ds, num_steps=4, n=10, step=step, history=hist)
    assert len(hist) == 4 and all(np.isfinite(float(h["neg_logpdf"])) for h in hist)
    assert all(0 <= float(h["alpha"]) < 10 for h in hist)
    c = _corr(phi.scalar_matrix, ds.phi.scalar_matrix)
    print(f"corr(phi_MAP, phi) = {c:.3f}, alphas {[round(float(h['alpha']), 3) for h in hist]}")
    assert c > 0.1          # the estimate points along the true phi


def _julia_dataset(pkg, psim, dev, pol):
    """The data set of the reference's test_map_joint.py (CMBLensing.jl
    fixtures, 256^2 at 2.5 arcmin, nphi as the quadratic estimate)."""
    from cmblensing_amd import matrix_operators as mo
    from conftest import GOLDEN
    L = lambda n: torch.as_tensor(np.load(GOLDEN / "julia" / f"{n}.npy", allow_pickle=False),
                                  device=dev)
    kw = dict(theta_pix=2.5)
    S = lambda n: mo.DiagonalScalar(scalar_matrix=L(n).real.contiguous(), **kw)
    if pol == "I":
        EB = S
        data = pkg.FlatS0(scalar_matrix=L("data_t_field_I"), **kw)
        truth = pkg.FlatS0(scalar_matrix=L("fJ_I"), **kw)
        ds = psim.DataSetT
        ops = dict(noise_covariance=S("cn_I"), field_covariance=S("cf_I"), mask=S("m_I"),
                   beam=S("b_I"), mixing_d=S("d_I"))
    elif pol == "P":
        EB = lambda o: mo.DiagonalEB(matrix_EE=L(f"{o}_ee_P"), matrix_BB=L(f"{o}_bb_P"), **kw)
        data = pkg.FlatS2(polar_matrix_1=L("data_e_field_P"), polar_matrix_2=L("data_b_field_P"),
                          **kw)
        truth = pkg.FlatS2(polar_matrix_1=L("fJ_e_P"), polar_matrix_2=L("fJ_b_P"), **kw)
        ds = psim.DataSetEB
        ops = {k: EB(o) for k, o in (("noise_covariance", "cn"), ("field_covariance", "cf"),
                                    ("mask", "m"), ("beam", "b"))}
        ops["mixing_d"] = mo.DiagonalEB(matrix_EE=L("d_ee_P"), matrix_BB=L("d_bb_P"), **kw)
    else:
        T = lambda o: mo.BlockTEB(matrix_TT=L(f"{o}_tt_IP"), matrix_TE=L(f"{o}_te_IP"),
                                  matrix_ET=L(f"{o}_te_IP"), matrix_EE=L(f"{o}_ee_IP"),
                                  matrix_BB=L(f"{o}_bb_IP"), **kw)
        data = pkg.FlatS02(scalar_matrix=L("data_t_field_IP"), polar_matrix_1=L("data_e_field_IP"),
                           polar_matrix_2=L("data_b_field_IP"), **kw)
        truth = pkg.FlatS02(scalar_matrix=L("fJ_t_IP"), polar_matrix_1=L("fJ_e_IP"),
                            polar_matrix_2=L("fJ_b_IP"), **kw)
        ds = psim.DataSetTEB
        ops = {k: T(o) for k, o in (("noise_covariance", "cn"), ("field_covariance", "cf"),
                                   ("mask", "m"), ("beam", "b"), ("mixing_d", "d"))}
    data_set = ds(data=data, phi_covariance=S(f"cphi_{pol}"),
                  quadratic_estimate=S(f"nphi_{pol}"), fid_r=0.2, fid_a_phi=1, nside=256,
                  theta_pix=2.5, **ops)
    return data_set, truth, L(f"phiJ_{pol}"), L(f"phi_{pol}")


@pytest.mark.parametrize("pol", ["I", "P", "IP"])
def test_map_joint_julia_fixtures(pkg, psim, mj, dev, pol):
    """test_map_joint.py of the reference: 30 BFGS steps on CMBLensing.jl's data
    set.  The reference's bar is relative L2 <= 1 per field component
    (MAX_NORM_DIFF); this build reaches 3e-3 (I) and 1e-3 (P, IP) for f and
    3e-2 / 4e-3 for phi, so the bars here are 2e-2 for f and 0.1 for phi."""
    _case(f"map_joint julia fixtures {pol}")
    data_set, fJ, phiJ, phi_true = _julia_dataset(pkg, psim, dev, pol)
    f, phi = mj.map_joint(data_set, num_steps=30)
    diffs = [rel(getattr(f, n).cpu(), getattr(fJ, n).cpu()) for n in f._matrix_names()]
    print(f"{pol}: rel L2 f vs Julia {['%.3g' % d for d in diffs]}, phi vs Julia "
          f"{rel(phi.scalar_matrix.cpu(), phiJ.cpu()):.3g}, corr(phi, phiJ) "
          f"{_corr(phi.scalar_matrix, phiJ):.3f}, corr(phi, phi_true) {_corr(phi.scalar_matrix, phi_true):.3f}, "
          f"corr(phiJ, phi_true) {_corr(phiJ, phi_true):.3f}")
    assert max(diffs) <= 2e-2
    assert rel(phi.scalar_matrix.cpu(), phiJ.cpu()) <= 0.1
    assert _corr(phi.scalar_matrix, phiJ) > 0.99
