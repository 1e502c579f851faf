"""Copy the subset of the reference's Julia (CMBLensing.jl) ground-truth data
used by these tests into tests/golden/julia/ as plain .npy files.

The reference stores them as .npy payloads with a .npz name
(tests/ground_truth_data/*.npz, written by tests/generate_julia_data/*.py with
seed 67, Nside 256, theta_pix 2.5 arcmin, Float64).  They are data, loaded with
numpy.load(allow_pickle=False).

    python tests/golden/copy_julia_fixtures.py [/path/to/reference]
"""
import shutil
import sys
from pathlib import Path

import numpy as np

NAMES = [
    # lensing, intensity and polarisation (test_lensing.py)
    "phi_I", "t_field_I", "t_lensed_field_I", "t_adjoint_lensed_field_I",
    "phi_P", "e_field_P", "b_field_P", "e_lensed_field_P", "b_lensed_field_P",
    "e_adjoint_lensed_field_P", "b_adjoint_lensed_field_P",
    # gradients / logpdf / Wiener filter, intensity (test_gradients.py, test_logpdf.py,
    # test_wiener_filter.py)
    "data_t_field_I", "cn_I", "cf_I", "m_I", "b_I", "cphi_I", "gradf_I", "grad_phi_I",
    "f_wf_ffpp_I", "f_wf_f0pp_I", "f_wf_f0p0_I",
    # gradients / logpdf / Wiener filter, polarisation (EB operators)
    "data_e_field_P", "data_b_field_P", "cn_ee_P", "cn_bb_P", "cf_ee_P", "cf_bb_P",
    "m_ee_P", "m_bb_P", "b_ee_P", "b_bb_P", "cphi_P", "gradf_e_P", "gradf_b_P",
    "grad_phi_t_P", "f_wf_ffpp_e_P", "f_wf_ffpp_b_P",
    # intensity and polarisation (T+EB, BlockTEB operators)
    "phi_IP", "t_field_IP", "e_field_IP", "b_field_IP", "t_lensed_field_IP",
    "e_lensed_field_IP", "b_lensed_field_IP", "data_t_field_IP", "data_e_field_IP",
    "data_b_field_IP", "cphi_IP",
] + [f"{op}_{blk}_IP" for op in ("cn", "cf", "m", "b") for blk in ("tt", "te", "ee", "bb")] + [
    "gradf_t_IP", "gradf_e_IP", "gradf_b_IP", "grad_phi_t_IP",
    "f_wf_ffpp_t_IP", "f_wf_ffpp_e_IP", "f_wf_ffpp_b_IP",
    # mixed parametrisation (mixing.py, gradients.py:40-85)
    "d_I", "g_I", "d_ee_P", "d_bb_P", "d_tt_IP", "d_te_IP", "d_ee_IP", "d_bb_IP",
    "mixed_grad_phi_I", "mixed_grad_phi_P", "mixed_grad_phi_IP",
    # quadratic-estimate noise and the binned spectra of simulated maps
    # (test_covariance_matrices.py, test_simulated_cls.py)
    "nphi_I", "nphi_P", "nphi_IP", "ell_I", "ell_P", "ell_IP", "cl_ll_avg_I", "cl_tt_avg_I",
] + [f"cl_{k}_{c}_avg_{p}" for k in ("ll", "tt") for c in ("e", "b") for p in ("P", "IP")] + [
    # adjoint lensing of the T+EB fields (test_lensing.py)
    "t_adjoint_lensed_field_IP", "e_adjoint_lensed_field_IP", "b_adjoint_lensed_field_IP",
    # the Wiener filter from every start (field or 0) and phi (phi or 0)
    # (test_wiener_filter.py): the rest of the matrix
    "f_wf_ffp0_I",
] + [f"f_wf_{s}_{c}_P" for s in ("ffp0", "f0pp", "f0p0") for c in ("e", "b")] + [
    f"f_wf_{s}_{c}_IP" for s in ("ffp0", "f0pp", "f0p0") for c in ("t", "e", "b")] + [
    # MAP-joint outputs after 30 steps (test_map_joint.py)
    "fJ_I", "phiJ_I", "fJ_e_P", "fJ_b_P", "phiJ_P", "fJ_t_IP", "fJ_e_IP", "fJ_b_IP", "phiJ_IP",
]
TEXT = ["logpdf_I.txt", "logpdf_P.txt", "logpdf_IP.txt", "n_side.txt", "theta_pix.txt"]


def main(ref_root="/root/reference"):
    src = Path(ref_root) / "tests" / "ground_truth_data"
    dst = Path(__file__).resolve().parent / "julia"
    dst.mkdir(exist_ok=True)
    for n in NAMES:
        a = np.load(src / f"{n}.npz", allow_pickle=False)
        np.save(dst / f"{n}.npy", a, allow_pickle=False)
    for t in TEXT:
        shutil.copyfile(src / t, dst / t)
    print(f"copied {len(NAMES) + len(TEXT)} files to {dst}")


if __name__ == "__main__":
    main(*sys.argv[1:])
