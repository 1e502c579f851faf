"""Properties of the MAP-joint oracle (oracle/map_joint.py) on a small
simulated temperature data set: the line search returns the minimum of the
values it evaluates, the joint logpdf after each Wiener filter never
decreases under the grid line search (coordinate ascent: the Wiener filter
maximises over f, the phi step may take alpha = 0), and the estimate
correlates with the true phi."""
import numpy as np
import pytest

from oracle import map_joint as omj
from oracle import mixing as omx
from oracle import simulate as osim
from oracle import wiener as owf


@pytest.fixture(scope="module")
def problem(sims):
    cls = osim.extract_all_cls(*sims.camb_like_spectra(5000), 17000, 5000)
    N, th = 16, 3.0
    r = osim.load_sim_from_cls(N, th, "I", 1, cls, 17000, uk_arcmin_t=1.0, n=10, with_qe=True)
    return dict(N=N, th=th, d=r["data"], Cf=r["cf"]["TT"][None], Cn=r["cn"]["TT"][None],
                M=r["mask"][None], B=r["beam"][None], D=r["d"][0][None], Cphi=r["cphi"],
                QE=np.real(r["qe"]), phi=r["phi"])


def test_hessian_is_harmonic_combination(problem):
    p = problem
    H = owf.pinv(omj.phi_gradient_hessian(p["Cphi"], p["QE"]))
    nz = (p["Cphi"] > 0) & (p["QE"] > 0)
    ref = p["Cphi"][nz] * p["QE"][nz] / (p["Cphi"][nz] + p["QE"][nz])
    assert np.allclose(H[nz], ref, rtol=1e-12)


def test_line_search_minimum_and_ascent(problem):
    p = problem
    hist = []
    f, phi = omj.map_joint(p["d"], p["Cf"], p["Cn"], p["Cphi"], p["QE"], p["M"], p["B"], p["D"],
                           p["th"], num_steps=3, step="grid", n=10, history=hist)
    assert len(hist) == 3
    # non-decreasing up to the Wiener filter's tolerance and LenseFlow round trips
    for a, b in zip(hist, hist[1:]):
        assert b >= a - 1e-6 * abs(a)
    c = np.real(np.vdot(phi, p["phi"])) / np.sqrt(np.vdot(phi, phi).real * np.vdot(p["phi"], p["phi"]).real)
    assert c > 0.5
    # the grid search picks the smallest value it evaluated
    N = p["N"]
    G = np.ones((N, N // 2 + 1))
    H = owf.pinv(omj.phi_gradient_hessian(p["Cphi"], p["QE"]))
    fm, phim = omx.mix(f, phi, p["D"], G, p["th"], 10)
    g = omx.mixed_grad_phi_logpdf(fm, phim, p["d"], p["Cn"], p["Cphi"], p["Cf"], p["M"], p["B"],
                                  p["D"], G, p["th"], 10)
    a, vals = omj.line_search(1.0, phim, H, g, fm, p["D"], G, p["d"], p["Cn"], p["Cphi"], p["Cf"],
                              p["M"], p["B"], p["th"], 10, num=5)
    assert a == np.linspace(0, 1, 5)[int(np.argmin(vals))]
    assert np.isclose(vals[0], -owf.logpdf(*omx.unmix(fm, phim, p["D"], G, p["th"], 10), p["d"],
                                           p["Cn"], p["Cphi"], p["Cf"], p["M"], p["B"], p["th"],
                                           10), rtol=1e-12)
