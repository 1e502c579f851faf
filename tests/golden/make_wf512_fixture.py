"""Write the FP64 oracle Wiener filter of the 512^2 Q/U point-source workload
(tests/golden/wf512/wf512_qu_holes.npz) for two of its eight right-hand sides.

Workload: Q/U, N = 512, 1.5 arcmin pixels, 7 RK4 steps, signal / noise / phi
seeded 100..107 (sims.wf_problem), 1 uK-arcmin white noise with an
l_knee = 100 1/f term, a 1 arcmin beam and the point-source mask of 200 holes
of 5 arcmin (seed 4).  The data d = m map(B L f) + n are made here with the
oracle's own LenseFlow; the solution is the oracle PCG
(wiener_filter.py:10-105, restated in oracle/wiener.py) run to a relative
residual of 1e-9 from a zero start.

Stored per right-hand side: the solution x (E, B half planes, rounded to
complex64: 6e-8 relative, far below the 1e-5 bar it is compared at), the
oracle's iteration count and final relative residual, and checksums of the
data the solution belongs to (sum of d^2, and d at a few pixels), so a test
that rebuilds the problem can show it solves the same one.

Runs in the build container only (numpy; about 3 minutes per right-hand side):

    python tests/golden/make_wf512_fixture.py
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402
from oracle import lenseflow as olf  # noqa: E402
from oracle import wiener as owf  # noqa: E402

N, TH, NSTEPS = 512, 1.5, 7
SEEDS = (100, 105)
RTOL = 1e-9
OUT = Path(__file__).resolve().parent / "wf512" / "wf512_qu_holes.npz"
PROBE_PIXELS = ((0, 0), (17, 300), (255, 256), (511, 3), (400, 400))


def problem(sims, seed):
    """The numpy problem of right-hand side `seed`, data lensed by the oracle."""
    pr = sims.wf_problem(N, TH, pol=True, seed=seed, mask="holes", noise_uk=1.0,
                         ell_knee=100.0, beam_fwhm=1.0)
    fq = owf.to_map(owf.eb2qu_field(pr["f"], N, TH))
    lensed = olf.lense_flow(fq, pr["phi"], olf.pix_width_from_theta(TH), NSTEPS)
    return sims.wf_data(pr, lensed)


def data_checksums(d):
    return (float(np.sum(d.astype(np.float64) ** 2)),
            np.array([[d[c, i, j] for (i, j) in PROBE_PIXELS] for c in range(d.shape[0])]))


def main():
    sims = __graft_entry__._load_module("cmbl_sims_fix", ROOT / "cmblensing.py_amd" / "sims.py")
    out = {"seeds": np.array(SEEDS), "rtol": np.array(RTOL), "nsteps": np.array(NSTEPS),
           "probe_pixels": np.array(PROBE_PIXELS)}
    for s in SEEDS:
        t0 = time.time()
        pr = problem(sims, s)
        hist = []
        x = owf.wiener_filter(0 * pr["f"], pr["phi"], pr["d"], pr["Cf"], pr["Cn"], None, pr["B"],
                              TH, n=NSTEPS, tol=0.0, rtol=RTOL, maxiter=2000, history=hist,
                              mask_pix=pr["mask_pix"])
        # the achieved relative residual, recomputed from x
        b = -owf.gradf_logpdf(0 * pr["f"], pr["phi"], pr["d"], pr["Cf"], pr["Cn"], None, pr["B"],
                              TH, n=NSTEPS, mask_pix=pr["mask_pix"])
        Ax = owf.gradf_logpdf(x, pr["phi"], np.zeros_like(pr["d"]), pr["Cf"], pr["Cn"], None,
                              pr["B"], TH, n=NSTEPS, mask_pix=pr["mask_pix"])
        r = b - Ax
        relres = float(np.sqrt(olf.field_dot(r, r, N) / olf.field_dot(b, b, N)))
        ss, probes = data_checksums(pr["d"])
        out[f"x_{s}"] = x.astype(np.complex64)
        out[f"iterations_{s}"] = np.array(len(hist))
        out[f"true_rel_residual_{s}"] = np.array(relres)
        out[f"d_sumsq_{s}"] = np.array(ss)
        out[f"d_probes_{s}"] = probes
        out[f"mask_frac_{s}"] = np.array(1.0 - float(pr["mask_pix"].mean()))
        print(f"seed {s}: {len(hist)} iterations, true relative residual {relres:.3e}, "
              f"masked {1 - pr['mask_pix'].mean():.4f}, {time.time() - t0:.0f} s", flush=True)
    OUT.parent.mkdir(parents=True, exist_ok=True)
    np.savez(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
