"""End-to-end parity on the reference's own CsPbI3 job
(filter_mpi/EXAMPLES/1x1x1_CsPbI3/filter: 15 atoms, SO + NL spinors, 44^3 grid,
16 starts x 16 targets, nCheby 2048).  The inputs are built by the oracle's
restatement of the reference input stage; everything after that runs through
the C ABI on the GPU: energy range, Newton filter, normalize_all, ortho,
Rayleigh-Ritz and sigma_E.

Pins:
* run.dat:160 prints the inflated range Emin = -0.756627, Emax = 25.2627
  (energy.c:366-372) from the fixed-seed start vectors;
* eval.dat holds 56 Ritz values at %.16lg; the 44 with sigma_E < 1e-6 are
  converged eigenvalues of H (error <= sigma^2/gap), so each must be found by
  the plain FP64 GPU path within 1e-8 Ha, matched by value (the reference's
  start states came from the clock, so only converged values are comparable);
* the mixed-precision path must reproduce the FP64 path's converged
  eigenvalues within 1e-5 Ha.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CSPBI3 = os.path.join(HERE, "golden", "cspbi3")


@pytest.fixture(scope="module")
def job():
    import torch
    from filterbse_amd import api
    from oracle import inputs as I

    j = I.reference_job(CSPBI3)
    ctx = api.Context(j["nx"], j["ny"], j["nz"], j["nspin"], j["is_complex"], j["dv"], j["ksqr"], j["pot_local"],
                      j["proj"])
    p0, p1 = api.energy_range_start_states(j["nx"], j["ny"], j["nz"], j["nspin"], j["is_complex"], j["dv"])
    rng = ctx.energy_range(p0, p1, max_iter=1500, nl_lambda_min=j["nl_lambda"][0],
                           vloc_min=float(j["pot_local"].min()))
    print("energy range", rng, flush=True)
    ref = np.loadtxt(os.path.join(CSPBI3, "eval.dat"))
    return dict(j=j, ctx=ctx, range=rng, ref=ref, dev=torch.device("cuda:0"))


def run_path(J, precision, refine=False):
    import torch
    from filterbse_amd import api

    j, ctx = J["j"], J["ctx"]
    Vmin, dE = J["range"][2], J["range"][3]
    M, ms, ns = 2048, 16, 16
    dt = (M / (2.5 * dE)) ** 2
    zn = api.leja_points(M)
    targets = api.ene_targets(10, 6, -0.50, -0.20, -0.167, -0.10)
    an = api.newton_coeff(zn, targets, Vmin, dE, dt)
    stlen = 2 * j["nspin"] * j["nx"] * j["ny"] * j["nz"]
    psi_rank, _ = api.init_filter_states(ns, ms, -1, j["nx"], j["ny"], j["nz"], j["nspin"], 1, j["dv"])
    d = torch.from_numpy(psi_rank).to(J["dev"])
    ptr = d.data_ptr()
    ctx.filter_dev(ptr, ns, ms, an, zn, Vmin, 4.0 / dE, precision)
    ctx.normalize_all_dev(ptr, ns * ms)
    rank, sv = ctx.ortho_dev(ptr, ns * ms)
    ev = ctx.rayleigh_ritz_dev(ptr, rank)
    sig, ev2 = ctx.sigma_dev(ptr, rank)
    if refine:
        sig_f = (1.0 / (2.0 * dt)) ** 0.5
        for _ in range(2):
            ev2, sig2, _ = ctx.refine_dev(ptr, rank, ev, sig, targets.min() - 4 * sig_f, targets.max() + 4 * sig_f,
                                          e_p=0.25, sigma_max=1e-2)
            assert (sig2 <= sig).all()
            ev, sig = ev2, sig2
    print("precision", precision, "refine", refine, "n", rank, "converged(1e-6)", int((sig < 1e-6).sum()),
          "converged(1e-4)", int((sig < 1e-4).sum()), flush=True)
    return ev, sig, rank


def test_energy_range_matches_reference_run(job):
    """Emax (power iteration, converged) agrees to the printed digits.  Emin
    comes from imaginary-time steps that stop when |dE/E| < 1e-6 or when the
    'positive step' guard (energy.c:253-259) has lowered tau to 0; where that
    guard fires depends on rounding, so the reference stopped at a slightly
    higher upper bound of the lowest eigenvalue than this run.  Both are upper
    bounds of the same minimum: require ours to be no higher and within 0.5%."""
    Vmin, dE = job["range"][2], job["range"][3]
    assert abs((Vmin + dE) - 25.2627) < 5e-6 * 25.2627 + 5e-5
    assert Vmin <= -0.756627 + 5e-7 and abs(Vmin - (-0.756627)) < 5e-3 * 0.756627


@pytest.fixture(scope="module")
def fp64_result(job):
    return run_path(job, 0)


def test_converged_eigenvalues_match_reference(job, fp64_result):
    ev, sig, rank = fp64_result
    ref = job["ref"]
    conv_ref = ref[ref[:, 2] < 1e-6, 1]
    assert len(conv_ref) == 44
    ours = ev[sig < 1e-4]
    assert len(ours) > 0
    worst = max(np.min(np.abs(ours - e)) for e in conv_ref)
    assert worst < 1e-8, worst
    # degeneracies (Kramers pairs and the 4-fold groups) are reproduced
    for e in np.unique(np.round(conv_ref, 9)):
        mult_ref = int(np.sum(np.abs(conv_ref - e) < 1e-8))
        mult = int(np.sum(np.abs(ours - e) < 1e-8))
        assert mult >= mult_ref, (e, mult, mult_ref)


def test_mixed_precision_matches_fp64(job, fp64_result):
    ev64, sig64, _ = fp64_result
    ev, sig, rank = run_path(job, 1, refine=True)
    a = ev[sig < 1e-4]
    b = ev64[sig64 < 1e-6]
    assert len(a) > 0.8 * len(b)
    worst = max(np.min(np.abs(a - e)) for e in b)
    print("mixed vs fp64 worst", worst, flush=True)
    assert worst < 1e-5, worst
