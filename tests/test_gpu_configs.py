"""The benchmark workloads on the GPU at their real grids, atoms and
projectors (filterbse_amd.dots.CONFIGS; the uneven dots of dots.UNEVEN):

* H on one ran2 state against the oracle (FP64, to FFT rounding);
* a short FP64 Newton filter of one start state against the oracle's
  filter_cycle, and the same filter in mixed precision (the FP32 FFT instances
  the benchmark runs: fft_fx_f32_n96 at 96^3, axis_fx_f32_n160 / n256 at
  160^3 / 256^3) against the oracle to FP32 rounding;
* the mixed filter of the workload's batch against the FP64 filter of the
  same batch (Newton length 40).

The oracle is the serial C restatement (oracle/fbse_oracle.c) up to 96^3;
at 160^3 and 256^3 it is the same restatement threaded (oracle/cpu_newton.py,
pinned against the serial one by tests/test_cpu_newton.py), whose scipy FFT
keeps an H application at a few seconds.  Spectrum bounds are a Weyl-type
bound (V_loc range + the largest k^2), which fixes the Newton map as the
filter tests need it; the energy range itself is tested elsewhere."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = ["c0_real_64", "c1_so_96", "c3_so_160", "c4_so_256"]
LARGE = {"c3_so_160", "c4_so_256"}


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module", params=CASES)
def workload(request):
    from filterbse_amd import api, dots
    name = request.param
    cfg = dict(dots.CONFIGS[name])
    print("case workload %s" % name, flush=True)
    job = dots.build_dot(cfg["radius_nm"], cfg["n_req"], cfg["dx"], so=cfg["so"], nl=cfg["nl"],
                         ke_max=cfg.get("ke_max", 20.0), builder=cfg.get("builder", "host"))
    ctx = api.Context(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
                      job["pot_local"], job["proj"])
    ngrid = job["nx"] * job["ny"] * job["nz"]
    Vmin = float(job["pot_local"].min()) - 1.0
    dE = float(job["ksqr"].max()) * ngrid + float(job["pot_local"].max() - job["pot_local"].min()) + 2.0
    yield name, cfg, job, ctx, Vmin, dE
    ctx.close()


def _oracle(name, job):
    if name in LARGE:
        from oracle.cpu_newton import CpuNewton
        return CpuNewton(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
                         job["pot_local"], job["proj"])
    return O.Ham(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
                 job["pot_local"], job["proj"])


def _oracle_h(name, orc, psi):
    if name in LARGE:
        z = psi.view(np.complex128).copy()
        out = np.empty_like(z)
        orc.hamiltonian(z, out)
        return out.view(np.float64)
    return orc.apply(psi)


def _oracle_filter(name, orc, psi_rank, ms, an, zn, Vmin, dE_1, stlen):
    if name in LARGE:
        M = len(zn)
        psi = psi_rank[:stlen].view(np.complex128).copy()
        block = np.zeros(ms * stlen)
        for t in range(ms):  # term 0 (filter.c:284-312)
            a = an[2 * M * t] + 1j * an[2 * M * t + 1]
            block[t * stlen:(t + 1) * stlen] = (a * psi).view(np.float64)
        orc.newton_steps(block, ms, an, zn, Vmin, dE_1, 1, M - 1, psi)
        return block
    return orc.filter_cycles(psi_rank, 1, ms, an, zn, Vmin, dE_1)


def _inputs(job, ms, M, Vmin, dE, n_starts=1):
    from filterbse_amd import api
    psi_rank, _ = api.init_filter_states(n_starts, ms, -1, job["nx"], job["ny"], job["nz"], job["nspin"],
                                         job["is_complex"], job["dv"])
    zn = api.leja_points(M)
    targets = np.linspace(-0.3, -0.1, ms)
    an = api.newton_coeff(zn, targets, Vmin, dE, (M / (2.5 * dE)) ** 2)
    return psi_rank, zn, an


def test_h_matches_oracle(workload):
    name, cfg, job, ctx, Vmin, dE = workload
    from filterbse_amd import api
    psi, _ = api.init_filter_states(1, 1, -1, job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"],
                                    job["dv"])
    out = ctx.apply_h(psi)
    ref = _oracle_h(name, _oracle(name, job), psi)
    err = rel(out, ref)
    print("%s H rel err %.2e" % (name, err))
    assert err < 1e-11, err


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 2e-5)])
def test_filter_matches_oracle(workload, precision, tol):
    """One start state, 3 targets, Newton length 6; mixed precision runs the
    FP32 FFT instance the benchmark uses for this grid."""
    name, cfg, job, ctx, Vmin, dE = workload
    ms, M = 3, 6
    psi_rank, zn, an = _inputs(job, ms, M, Vmin, dE)
    stlen = (job["is_complex"] + 1) * job["nspin"] * job["nx"] * job["ny"] * job["nz"]
    out = ctx.filter(psi_rank, 1, ms, an, zn, Vmin, 4.0 / dE, precision=precision)
    ref = _oracle_filter(name, _oracle(name, job), psi_rank, ms, an, zn, Vmin, 4.0 / dE, stlen)
    err = rel(out, ref)
    print("%s filter precision %d rel err %.2e (fft path %s)" % (name, precision, err, ctx.fft_choice(precision)))
    assert err < tol, err


def test_mixed_batch_matches_fp64(workload):
    """The workload's batch of start states (4 of 16 at 256^3, where two FP64
    blocks of all 256 filtered states would not fit together) through the
    mixed and the FP64 filter, Newton length 40."""
    import torch
    from filterbse_amd import api
    name, cfg, job, ctx, Vmin, dE = workload
    n_starts = 4 if name == "c4_so_256" else cfg["n_starts"]
    ms, M = cfg["ms"], 40
    psi_rank, zn, an = _inputs(job, ms, M, Vmin, dE, n_starts=n_starts)
    a = torch.from_numpy(psi_rank).cuda()
    b = a.clone()
    del psi_rank
    torch.cuda.synchronize()
    ctx.filter_dev(a.data_ptr(), n_starts, ms, an, zn, Vmin, 4.0 / dE, api.FBSE_PREC_FP64)
    ctx.filter_dev(b.data_ptr(), n_starts, ms, an, zn, Vmin, 4.0 / dE, api.FBSE_PREC_MIXED)
    ctx.synchronize()
    err = ((a - b).abs().max() / a.abs().max()).item()
    print("%s batch %d x %d mixed vs FP64 rel err %.2e, fft %s" % (name, n_starts, ms, err,
                                                                  ctx.fft_choice(api.FBSE_PREC_MIXED)))
    assert err < 3e-5, err
    del a, b
    ctx.trim()


def test_uneven_dots_filter_many():
    """Three of the uneven-dot workload's dots (the smallest, a middle and the
    largest radius of its seeded draw, each on its FFT-friendly grid) through
    fbse_filter_many: bit-identical to separate fbse_filter_dev calls, and the
    smallest dot's FP64 filter of one start against the oracle."""
    import torch
    from filterbse_amd import api, dots
    U = dots.UNEVEN
    radii = np.random.default_rng(U["seed"]).uniform(U["r_lo"], U["r_hi"], U["n_dots"])
    order = np.argsort(radii)
    pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    ms, M, ns = 4, 24, 2
    ctxs, psis, ans, zns, vmins, de1s, jobs = [], [], [], [], [], [], []
    for i in pick:
        n_req = dots.friendly_grid(float(radii[i]), U["dx"], seed=100 + i)
        job = dots.build_dot(float(radii[i]), n_req, U["dx"], so=True, nl=True, seed=100 + i, builder="gpu")
        print("case uneven dot %d: r %.2f nm grid %s" % (i, radii[i], n_req), flush=True)
        ctx = api.Context(job["nx"], job["ny"], job["nz"], 2, 1, job["dv"], job["ksqr"], job["pot_local"],
                          job["proj"])
        ngrid = job["nx"] * job["ny"] * job["nz"]
        Vmin = float(job["pot_local"].min()) - 1.0
        dE = float(job["ksqr"].max()) * ngrid + float(job["pot_local"].max() - job["pot_local"].min()) + 2.0
        psi_rank, zn, an = _inputs(job, ms, M, Vmin, dE, n_starts=ns)
        ctxs.append(ctx)
        psis.append(torch.from_numpy(psi_rank).cuda())
        ans.append(an)
        zns.append(zn)
        vmins.append(Vmin)
        de1s.append(4.0 / dE)
        jobs.append(job)
    sep = [p.clone() for p in psis]
    torch.cuda.synchronize()
    api.filter_many(ctxs, [p.data_ptr() for p in psis], [ns] * 3, ms, ans, zns, vmins, de1s, api.FBSE_PREC_MIXED)
    for k in range(3):
        ctxs[k].filter_dev(sep[k].data_ptr(), ns, ms, ans[k], zns[k], vmins[k], de1s[k], api.FBSE_PREC_MIXED)
    for c in ctxs:
        c.synchronize()
    for k in range(3):
        assert torch.equal(psis[k], sep[k]), k
    job = jobs[0]
    stlen = 2 * 2 * job["nx"] * job["ny"] * job["nz"]
    psi_rank = sep[0].cpu().numpy()[:1 * ms * stlen].copy()
    start, _ = api.init_filter_states(1, ms, -1, job["nx"], job["ny"], job["nz"], 2, 1, job["dv"])
    out = ctxs[0].filter(start, 1, ms, ans[0], zns[0], vmins[0], de1s[0], precision=0)
    ref = _oracle_filter("small", _oracle("small", job), start, ms, ans[0], zns[0], vmins[0], de1s[0], stlen)
    assert rel(out, ref) < 1e-10
    assert rel(psi_rank, ref) < 2e-5  # the mixed filter_many result of that start
    for c in ctxs:
        c.close()
