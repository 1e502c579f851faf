"""GPU parity of the on-device simulation (cmbl_grf_fourier and
cmblensing_amd.simulate) against the oracle (oracle/simulate.py).
Tolerances: random draws FP64 1e-12 relative per mode (the device's log /
sincospi against numpy's), FP32 1e-6; set-up arrays FP64 1e-10 relative
(quadratic estimates through the library's FFTs); lensed fields and data FP64
1e-10 relative L2 (LenseFlow, n = 10)."""
import numpy as np
import pytest
import torch

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
def cls(sims):
    return osim.extract_all_cls(*sims.camb_like_spectra(5000), 17000, 5000)


@pytest.mark.parametrize("N,C", [(16, 1), (64, 2), (256, 3)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_grf_matches_oracle(psim, dev, N, C, dtype):
    _case(f"grf N={N} C={C} {dtype}")
    rng = np.random.default_rng(N + C)
    cov = rng.uniform(0.2, 3.0, (C, N, N // 2 + 1))
    seed, s0 = 2 ** 40 + 12345, (4 << 32) + 7
    out = psim.grf_fourier(torch.as_tensor(cov, device=dev), 3, seed, s0, 2.0, dtype)
    ref = osim.grf_fourier(cov, 3, seed, s0)
    got = out.cpu().numpy()
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    assert np.allclose(got, ref, rtol=tol, atol=tol * np.abs(ref).max())
    # self-conjugate columns hold Hermitian-consistent values
    assert np.allclose(got[..., 0, 0].imag, 0) and np.allclose(got[..., N // 2, 0].imag, 0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_grf_fused_data_model(psim, dev, dtype):
    _case(f"grf fused data model {dtype}")
    N, C, T = 64, 2, 4
    rng = np.random.default_rng(5)
    cov = rng.uniform(0.2, 3.0, (C, N, N // 2 + 1))
    mul = rng.uniform(0.0, 1.0, (C, N, N // 2 + 1))
    add = rng.standard_normal((T, C, N, N // 2 + 1)) + 1j * rng.standard_normal((T, C, N, N // 2 + 1))
    cdt = torch.complex128 if dtype == torch.float64 else torch.complex64
    add_t = torch.as_tensor(add, device=dev).to(cdt)
    out = psim.grf_fourier(torch.as_tensor(cov, device=dev), T, 9, 5 << 32, 2.0, dtype,
                           mul=torch.as_tensor(mul, device=dev), add=add_t)
    ref = mul * add_t.cpu().numpy().astype(np.complex128) + osim.grf_fourier(cov, T, 9, 5 << 32)
    tol = 1e-12 if dtype == torch.float64 else 1e-6
    assert np.allclose(out.cpu().numpy(), ref, rtol=tol, atol=tol * np.abs(ref).max())


def test_grf_power_statistics(psim, dev):
    _case("grf power statistics")
    N, T = 128, 64
    cov = torch.ones((1, N, N // 2 + 1), dtype=torch.float64, device=dev)
    G = psim.grf_fourier(cov, T, 3, 0, 2.0)
    p = (torch.abs(G[:, 0, :, 1:N // 2]) ** 2 / N ** 2).double()
    assert abs(float(p.mean()) - 1) < 5 / np.sqrt(p.numel())
    # distinct trials and distinct kinds are independent streams
    G2 = psim.grf_fourier(cov, T, 3, 1 << 32, 2.0)
    c = torch.sum(G[:, 0, :, 1:N // 2] * torch.conj(G2[:, 0, :, 1:N // 2])).abs() / (N ** 2 * p.numel())
    assert float(c) < 5 / np.sqrt(p.numel())


@pytest.mark.parametrize("pol", ["I", "P"])
def test_sim_config_matches_oracle(psim, dev, cls, pol):
    _case(f"sim config {pol}")
    N, th = 64, 2.0
    cfg = psim.sim_config(N, th, pol, cls, device=dev)
    ref = osim.load_sim_from_cls(N, th, pol, 0, cls, 17000, n=1, with_qe=True)
    for k in ("TT", "TE", "EE", "BB"):
        assert rel(cfg.cf[k].cpu(), ref["cf"][k]) < 1e-12 or np.all(ref["cf"][k] == 0)
        assert rel(cfg.cfl[k].cpu(), ref["cfl"][k]) < 1e-12
        if np.any(ref["cn"][k]):
            assert rel(cfg.cn[k].cpu(), ref["cn"][k]) < 1e-12
    assert rel(cfg.cphi.cpu(), ref["cphi"]) < 1e-12
    assert np.allclose(cfg.mask.cpu(), ref["mask"], atol=1e-15, rtol=0)
    assert np.array_equal(cfg.beam.cpu().numpy(), ref["beam"])
    for a, b in zip(cfg.d, ref["d"]):
        if np.any(b):
            assert rel(np.nan_to_num(a.cpu().numpy()), np.nan_to_num(b)) < 1e-12
    # compare 1 / qe (the normalisation sums): where a sum vanishes in exact
    # arithmetic (modes beyond twice the mask cut) both hold FFT round-off and
    # qe = 1 / round-off is implementation noise, in the reference too
    inv = lambda q: np.where(q != 0, 1 / np.where(q != 0, q, 1), 0.0)
    a, b = inv(cfg.qe.cpu().numpy()), inv(np.real(ref["qe"]))
    assert np.allclose(a, b, rtol=0, atol=1e-10 * np.abs(b).max())
    assert np.allclose(cfg.g.cpu().numpy(), np.real(ref["g"]), rtol=1e-12, atol=0)


@pytest.mark.parametrize("pol", ["I", "P", "IP"])
def test_simulate_matches_oracle(psim, dev, cls, pol):
    _case(f"simulate vs oracle {pol}")
    N, th, T = 32, 2.0, 2
    cfg = psim.sim_config(N, th, pol, cls, with_qe=False, device=dev)
    phi, field, lensed, data = psim.simulate(cfg, 77, T, torch.float64, n=10)
    for t in range(T):
        ref = osim.load_sim_from_cls(N, th, pol, 77, cls, 17000, n=10, trial=t, with_qe=False)
        assert rel(phi[t].cpu(), ref["phi"]) < 1e-12
        assert rel(field[t].cpu(), ref["field"]) < 1e-12
        assert rel(lensed[t].cpu(), ref["lensed"]) < 1e-10
        assert rel(data[t].cpu(), ref["data"]) < 1e-10
        assert rel(lensed[t].cpu(), ref["field"]) > 1e-4      # lensing did something


def test_simulate_trials_are_batch_invariant(psim, dev, cls):
    _case("simulate batch invariance")
    cfg = psim.sim_config(64, 2.0, "P", cls, with_qe=False, device=dev)
    a = psim.simulate(cfg, 5, 3, torch.float32, n=7)
    b = psim.simulate(cfg, 5, 5, torch.float32, n=7)
    assert torch.equal(a[0][2], b[0][2]) and torch.equal(a[1][2], b[1][2])
    for x, y in zip(a[2:], b[2:]):
        assert torch.allclose(x[2], y[2], rtol=0, atol=1e-6 * float(y[2].abs().max()))


def test_simulate_fp32_close_to_fp64(psim, dev, cls):
    _case("simulate fp32 vs fp64 256 P")
    cfg = psim.sim_config(256, 2.0, "P", cls, with_qe=False, device=dev)
    a = psim.simulate(cfg, 11, 4, torch.float64, n=10)
    b = psim.simulate(cfg, 11, 4, torch.float32, n=10)
    for x, y in zip(a[1:], b[1:]):
        assert rel(y.cpu().to(torch.complex128), x.cpu()) < 1e-5


def test_load_sim_dataset(psim, pkg, dev, cls):
    _case("load_sim_from_cls datasets")
    from cmblensing_amd import matrix_operators as mo
    ds = psim.load_sim_from_cls(64, 2.0, "P", 3, cls, device=dev)
    assert isinstance(ds, psim.DataSetEB) and isinstance(ds.data, pkg.FlatS2)
    assert isinstance(ds.mixing_d, mo.DiagonalEB) and isinstance(ds.mixing_g, mo.DiagonalScalar)
    assert ds.data.polar_matrix_1.shape == (64, 33) and ds.phi.scalar_matrix.shape == (64, 33)
    # data - mask beam lensed is the noise draw of kinds 5, 6
    mb = (ds.mask.matrix_EE * ds.beam.matrix_EE)
    noise = torch.stack([ds.data.polar_matrix_1 - mb * ds.lensed_field.polar_matrix_1,
                         ds.data.polar_matrix_2 - mb * ds.lensed_field.polar_matrix_2])
    cfg = psim.sim_config(64, 2.0, "P", cls, with_qe=False, device=dev)
    ref = psim.grf_fourier(torch.stack([cfg.cn["EE"], cfg.cn["BB"]]), 1, 3, 5 << 32, 2.0)[0]
    assert rel(noise.cpu(), ref.cpu()) < 1e-12
    ds_t = psim.load_sim_from_cls(32, 2.0, "IP", 3, cls, device=dev)
    assert isinstance(ds_t, psim.DataSetTEB) and isinstance(ds_t.mask, mo.BlockTEB)
    ds_i = psim.batch_simulated_trials(3, 32, 2.0, cls, pol="I", master_seed=3, device=dev)
    assert ds_i.data.scalar_matrix.shape == (3, 32, 17)
