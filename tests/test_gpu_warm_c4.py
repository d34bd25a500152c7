"""The learned warm start cuts the Newton order at the 4.5 nm / 256^3
full-stack workload (c4_so_256: 13,740 atoms + 2,274 ligands, SO + NL
projectors, KEmax 5 Ha, 16 ran2 starts x 16 targets, the benchmark's
targets).  Measured when written (same ran2 starts, network trained on the
four small training dots):

| Newton length M                    | 768 | 896 | 1024 | 1152 | 1280 | 1408 | 1536 | 1792 | 2048 |
|------------------------------------|-----|-----|------|------|------|------|------|------|------|
| warm, mixed + one refinement sweep |  0  |  0  |  0   |  3   |  3   |  2   |  1   |  2   |  3   |
| plain ran2, mixed + refinement     |     |     |      |  0 (best sigma 9.7e-4) |  |  |  |  |   |
| plain ran2, FP64                   |     |     |      |  0 (9.8e-4) | 0 (8.1e-3) | | |  | 2 (the benchmark's FP64 pass) |

The FP64 pass at M = 2048 converges the conduction-band 1S level at
-0.160922746 Ha (sigma 9.4e-5) and one member of the 1P multiplet; the warm
mixed pass at M = 1152 converges the 1S Kramers pair and a 1P member.  The
test runs the warm and the plain mixed pass at M = 1152 and asserts that the
warm start converges band-edge pairs where the plain start converges none,
and that its 1S eigenvalues agree with the FP64 pass's within 1e-5 Ha (the
FP64 value is the benchmark's FP64 pass on this tree: no oracle run of a
256^3 filter fits a test)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E_1S_FP64 = -0.160922746   # conduction-band 1S level, plain FP64 filter at M = 2048
M_WARM = 1152


def test_warm_start_order_cut_at_c4():
    import torch
    from filterbse_amd import api, dots
    from filterbse_amd import warmstart as WS
    torch.cuda.empty_cache()
    cfg = dict(dots.CONFIGS["c4_so_256"])
    assert cfg["ncheby_warm"] == M_WARM
    mlp, info = WS.train_network()
    print("case warm start c4: training %s" % info, flush=True)
    job = dots.build_dot(cfg["radius_nm"], cfg["n_req"], cfg["dx"], so=True, nl=True, ke_max=cfg["ke_max"],
                         builder="gpu")
    ctx = api.Context(job["nx"], job["ny"], job["nz"], 2, 1, job["dv"], job["ksqr"], job["pot_local"], job["proj"])
    p0, p1 = api.energy_range_start_states(job["nx"], job["ny"], job["nz"], 2, 1, job["dv"])
    rr = ctx.energy_range(p0, p1, max_iter=1500)
    Vmin, dE = rr[2], rr[3]
    ns, ms = cfg["n_starts"], cfg["ms"]
    tg = api.ene_targets(*cfg["targets"])
    stlen = 2 * 2 * job["nx"] * job["ny"] * job["nz"]
    st, _ = api.init_filter_states(ns, 1, -1, job["nx"], job["ny"], job["nz"], 2, 1, job["dv"])
    starts = torch.from_numpy(st).cuda().view(ns, stlen)
    del st
    d_v = torch.from_numpy(job["pot_local"]).cuda()
    psi = torch.empty(ns * ms * stlen, dtype=torch.float64, device="cuda")
    M = M_WARM
    dt = (M / (2.5 * dE)) ** 2
    sf = math.sqrt(1.0 / (2.0 * dt))
    zn = api.leja_points(M)
    an = api.newton_coeff(zn, tg, Vmin, dE, dt)

    def mixed_pass(warm):
        blk = psi.view(ns, ms, stlen)
        blk[:, 0, :].copy_(starts)
        if warm:
            WS.warm_starts(ctx, mlp, job, d_v.data_ptr(), blk[:, 0, :], ms=ms)
        torch.cuda.synchronize()
        p = psi.data_ptr()
        ctx.filter_dev(p, ns, ms, an, zn, Vmin, 4.0 / dE, api.FBSE_PREC_MIXED)
        ctx.normalize_all_dev(p, ns * ms)
        rank, _ = ctx.ortho_dev(p, ns * ms)
        ev = ctx.rayleigh_ritz_dev(p, rank)
        sig, _ = ctx.sigma_dev(p, rank)
        ev2, sig2, _ = ctx.refine_dev(p, rank, ev, sig, tg.min() - 4 * sf, tg.max() + 4 * sf, e_p=0.25)
        assert (sig2 <= sig).all()
        ok = sig2 < 1e-4
        print("c4 mixed %s M=%d: converged %d %s, best sigma %.2e" % (
            "warm" if warm else "plain", M, ok.sum(), np.round(np.sort(ev2[ok]), 7).tolist(), sig2.min()), flush=True)
        return ev2[ok]

    ev_w = mixed_pass(True)
    ev_p = mixed_pass(False)
    ctx.close()
    assert len(ev_w) >= 2 and len(ev_p) == 0
    # the converged warm pairs lie at the conduction-band edge; its 1S Kramers
    # pair matches the FP64 filter's 1S level
    assert np.all(np.abs(ev_w - E_1S_FP64) < 0.005)
    one_s = ev_w[np.abs(ev_w - E_1S_FP64) < 1e-3]
    assert len(one_s) >= 1 and np.all(np.abs(one_s - E_1S_FP64) < 1e-5)
