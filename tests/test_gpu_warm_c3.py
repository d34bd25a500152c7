"""The learned warm start at the 3.5 nm / 160^3 workload (c3_so_160 at
KEmax 5 Ha, 8 starts x 16 targets spaced 0.0134 Ha >= 2 sigma_f around the
band edges).  Measured when written (same ran2 starts):

| start, precision                         | M = 2048 | M = 1536 | M = 1024 |
|------------------------------------------|----------|----------|----------|
| plain ran2, FP64                         | 0 (best sigma 2.1e-4) | 0 (1.1e-3) | 0 (4.2e-4) |
| warm, mixed + one refinement sweep       | 4 (3.7e-5)            | 0 (1.8e-4) | 0 (1.6e-4) |

So the warm start makes band-edge pairs converge at the configuration's
Newton length where the plain filter converges none, but it does not let the
order drop below 2048: at 1536 and 1024 the filter is wider than the target
spacing asks (sigma_f 0.0086 and 0.013 Ha) and no pair reaches 1e-4 Ha.
The test asserts exactly that outcome."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGE_VB, EDGE_CB = -0.24896, -0.15932   # the c3 dot's band edges at KEmax 5 (a warm mixed scan)


def test_warm_start_at_c3():
    import torch
    from filterbse_amd import api, dots
    from filterbse_amd import warmstart as WS
    cfg = dict(dots.CONFIGS["c3_so_160"])
    mlp, info = WS.train_network()
    print("case warm start c3: training %s" % info, flush=True)
    job = dots.build_dot(cfg["radius_nm"], cfg["n_req"], cfg["dx"], so=True, nl=True, ke_max=5.0, builder="gpu")
    ctx = api.Context(job["nx"], job["ny"], job["nz"], 2, 1, job["dv"], job["ksqr"], job["pot_local"], job["proj"])
    p0, p1 = api.energy_range_start_states(job["nx"], job["ny"], job["nz"], 2, 1, job["dv"])
    rr = ctx.energy_range(p0, p1, max_iter=1500)
    Vmin, dE = rr[2], rr[3]
    ns, ms = cfg["n_starts"], cfg["ms"]
    sp = 0.0134
    tg = api.ene_targets(8, 8, EDGE_VB + 0.042 - 8 * sp, EDGE_VB + 0.042, EDGE_CB - 0.029, EDGE_CB - 0.029 + 8 * sp)
    stlen = 2 * 2 * job["nx"] * job["ny"] * job["nz"]
    st, _ = api.init_filter_states(ns, 1, -1, job["nx"], job["ny"], job["nz"], 2, 1, job["dv"])
    starts = torch.from_numpy(st).cuda().view(ns, stlen)
    d_v = torch.from_numpy(job["pot_local"]).cuda()
    warm = starts.clone()
    WS.warm_starts(ctx, mlp, job, d_v.data_ptr(), warm)
    psi = torch.empty(ns * ms * stlen, dtype=torch.float64, device="cuda")

    def converged(s0, prec, M, refine):
        dt = (M / (2.5 * dE)) ** 2
        sf = math.sqrt(1.0 / (2.0 * dt))
        zn = api.leja_points(M)
        an = api.newton_coeff(zn, tg, Vmin, dE, dt)
        psi.view(ns, ms, stlen)[:, 0, :].copy_(s0)
        torch.cuda.synchronize()
        p = psi.data_ptr()
        ctx.filter_dev(p, ns, ms, an, zn, Vmin, 4.0 / dE, prec)
        ctx.normalize_all_dev(p, ns * ms)
        rank, _ = ctx.ortho_dev(p, ns * ms)
        ev = ctx.rayleigh_ritz_dev(p, rank)
        sig, _ = ctx.sigma_dev(p, rank)
        if refine:
            ev2, sig2, _ = ctx.refine_dev(p, rank, ev, sig, tg.min() - 4 * sf, tg.max() + 4 * sf, e_p=0.25)
            assert (sig2 <= sig).all()
            ev, sig = ev2, sig2
        n = int((sig < 1e-4).sum())
        print("c3 %s M=%d: converged %d, best sigma %.2e" % (
            "fp64" if prec == api.FBSE_PREC_FP64 else "mixed", M, n, sig.min()), flush=True)
        return n, ev[sig < 1e-4]

    plain, _ = converged(starts, api.FBSE_PREC_FP64, 2048, False)
    warm2048, ev_w = converged(warm, api.FBSE_PREC_MIXED, 2048, True)
    warm1024, _ = converged(warm, api.FBSE_PREC_MIXED, 1024, True)
    ctx.close()
    assert warm2048 >= 2 and warm2048 > plain
    assert warm1024 < warm2048
    # the converged warm pairs are the conduction-band edge (Kramers pairs)
    assert np.all(np.abs(ev_w - EDGE_CB) < 0.01)
