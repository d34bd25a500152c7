"""Learned warm start of the Newton filter (host side of include/fbse_mlp.h).

No counterpart in the reference, which starts every filter from ran2 noise
(init.c:1207-1305).  The pipeline:

1. band_edge_targets: on a training dot, the plain FP64 filter + ortho +
   Rayleigh-Ritz + sigma_E give converged Ritz pairs; those next to the
   largest spacing (the gap) inside the target window are the valence- and
   conduction-edge states.  Their densities (fbse_state_density_dev),
   normalised to mean 1 over the grid, are the regression targets
   log(rho + 1e-3) for channels 0 (valence edge), 1 (conduction edge) and 2
   (both).
2. dot_features: fbse_mlp_features_dev on the dot's grid (species A, B, P1, P2
   -> 0..3, V_loc, radius relative to the dot radius).
3. train: Adam on mini-batches of grid points of all training dots.
4. warm_starts: the ran2 start states are low-passed in k-space to envelopes
   on the scale of the dot, then multiplied by the square root of the
   network's channel-2 prediction on the new dot (fbse_warm_start_dev):
   psi_0 -> sqrt(rho) * lowpass(psi_0), renormalised.  The filter, ortho, RR
   and sigma_E then run unchanged and measure the result; the plain random
   start stays the fallback.
"""
import math

import numpy as np

from . import api

SPECIES = {"A": 0, "B": 1, "P1": 2, "P2": 3}
N_IN = 8 * len(SPECIES) + 2
N_OUT = 3


def dot_features(mlp, job, d_vloc):
    """BF16 feature matrix (torch int16 tensor, ngrid x k_in) of a dot."""
    import torch
    ng = job["nx"] * job["ny"] * job["nz"]
    d_x = torch.empty((ng, mlp.k_in), dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    sp = np.array([SPECIES[s] for s in job["syms"]], dtype=np.int32)
    center = np.asarray(job["R"][:job["n_core"]]).mean(axis=0)
    mlp.features_dev(job["nx"], job["ny"], job["nz"], job["origin"], job["dx"], job["R"], sp, len(SPECIES),
                     d_vloc, center, job["radius_bohr"], d_x.data_ptr())
    return d_x


def filter_pass(ctx, job, starts_dev, ms, M, targets, Vmin, dE, precision=api.FBSE_PREC_FP64):
    """Filter + normalize + ortho + RR + sigma from the given start states
    (device tensor n_starts x stlen).  Returns (psi tensor, eval, sigma, rank)."""
    import torch
    ns = starts_dev.shape[0]
    stlen = starts_dev.shape[1]
    dt = (M / (2.5 * dE)) ** 2
    zn = api.leja_points(M)
    an = api.newton_coeff(zn, targets, Vmin, dE, dt)
    psi = torch.zeros(ns * ms * stlen, dtype=torch.float64, device="cuda")
    psi.view(ns, ms, stlen)[:, 0, :].copy_(starts_dev)
    torch.cuda.synchronize()  # the library runs on its own stream
    p = psi.data_ptr()
    ctx.filter_dev(p, ns, ms, an, zn, Vmin, 4.0 / dE, precision)
    ctx.normalize_all_dev(p, ns * ms)
    rank, _ = ctx.ortho_dev(p, ns * ms)
    ev = ctx.rayleigh_ritz_dev(p, rank)
    sig, _ = ctx.sigma_dev(p, rank)
    return psi, ev, sig, rank


def edge_sets(ev, sig, lo, hi, resid=1e-4, n_edge=8):
    """Indices of the valence- and conduction-edge Ritz pairs: converged pairs
    in [lo, hi] split at the largest spacing between consecutive converged
    eigenvalues."""
    ok = np.nonzero((sig < resid) & (ev >= lo) & (ev <= hi))[0]
    if len(ok) < 2:
        return ok[:0], ok[:0]
    order = ok[np.argsort(ev[ok])]
    e = ev[order]
    cut = int(np.argmax(np.diff(e))) + 1
    return order[max(0, cut - n_edge):cut], order[cut:cut + n_edge]


def band_edge_targets(ctx, job, psi, vb, cb):
    """(ngrid x 3) float32 targets log(rho + 1e-3) with rho normalised to mean 1."""
    import torch
    ng = job["nx"] * job["ny"] * job["nz"]
    rho = torch.empty(ng, dtype=torch.float64, device="cuda")
    cols = []
    for idx in (vb, cb, np.concatenate([vb, cb])):
        torch.cuda.synchronize()
        ctx.state_density_dev(psi.data_ptr(), idx, rho.data_ptr())
        r = rho / rho.mean()
        cols.append(torch.log(r + 1e-3).float())
    return torch.stack(cols, dim=1).contiguous()


def train(mlp, xs, ts, steps=600, batch=65536, lr=2e-3, seed=0, log=None):
    """Adam on random mini-batches drawn from the concatenated points of all
    training dots (xs: BF16 feature tensors, ts: target tensors)."""
    import torch
    X = torch.cat(xs, dim=0)
    T = torch.cat(ts, dim=0)
    n = X.shape[0]
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    losses = []
    for it in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), device="cuda", generator=g)
        xb = X[idx].contiguous()
        tb = T[idx].contiguous()
        torch.cuda.synchronize()
        loss = mlp.loss_grad_dev(xb.data_ptr(), tb.data_ptr(), None, xb.shape[0])
        mlp.adam_step(lr * (0.1 ** (it / max(steps, 1))))
        losses.append(loss)
        if log and (it % 100 == 0 or it == steps - 1):
            log("  train step %d loss %.4f" % (it, loss))
    return losses


# Low-pass scale of the warm start (Ha): 1/(1 + 0.5|k|^2 / 0.01) keeps envelopes
# longer than ~2 pi / 0.14 bohr^-1 = 44 bohr, about a dot diameter.
SMOOTH_EP = 0.01


def warm_starts(ctx, mlp, job, d_vloc, starts_dev, channel=2, smooth_ep=SMOOTH_EP, ms=1):
    """Start states (device tensor n_starts x stlen, or the start-state view
    [:, 0, :] of a psi_rank block with ms targets per start), low-passed when
    smooth_ep > 0, times the predicted band-edge density's square root,
    renormalised (in place, returned)."""
    import torch
    x = dot_features(mlp, job, d_vloc)
    ng = x.shape[0]
    y = torch.empty((ng, N_OUT), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    mlp.forward_dev(x.data_ptr(), ng, y.data_ptr())
    ctx.warm_start_dev(y.data_ptr(), N_OUT, channel, starts_dev.data_ptr(), starts_dev.shape[0], ms, smooth_ep)
    return starts_dev


# Training dots: spin-orbit zincblende dots of 1.0-1.3 nm on a 48^3 grid at
# dx = 1.0 bohr (seeds 11-14), 16 ran2 starts x 16 targets straddling the gap,
# Newton length 2048, plain FP64 filter.
TRAIN_DOTS = ((1.0, 11), (1.1, 12), (1.2, 13), (1.3, 14))
TRAIN_TARGETS = (8, 8, -0.32, -0.22, -0.14, -0.04)


def train_network(width=128, steps=800, seed=1, log=None):
    """Trains the warm-start network on the band-edge states the plain FP64
    filter converges on the training dots, on the GPU (fbse_mlp_*).  Returns
    (mlp, info) with the per-dot converged counts and the loss history."""
    import torch
    from . import dots
    mlp = api.Mlp(N_IN, width, N_OUT, seed=seed)
    tg = api.ene_targets(*TRAIN_TARGETS)
    lo, hi = tg.min() - 0.05, tg.max() + 0.05
    xs, ts, info = [], [], []
    for r, sd in TRAIN_DOTS:
        job = dots.build_dot(r, (48, 48, 48), 1.0, so=True, nl=True, seed=sd)
        ctx = api.Context(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
                          job["pot_local"], job["proj"])
        p0, p1 = api.energy_range_start_states(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"],
                                               job["dv"])
        rng = ctx.energy_range(p0, p1, max_iter=1500)
        st, _ = api.init_filter_states(16, 1, -5, job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"],
                                       job["dv"])
        st = torch.from_numpy(st).cuda().view(16, -1)
        psi, ev, sig, rank = filter_pass(ctx, job, st, 16, 2048, tg, rng[2], rng[3])
        vb, cb = edge_sets(ev, sig, lo, hi)
        info.append(dict(radius_nm=r, converged=int((sig < 1e-4).sum()), vb=len(vb), cb=len(cb)))
        if log:
            log("warm-start training dot %.1f nm: %d converged, %d + %d band-edge states" % (
                r, info[-1]["converged"], len(vb), len(cb)))
        if len(vb) and len(cb):
            ts.append(band_edge_targets(ctx, job, psi, vb, cb))
            d_v = torch.from_numpy(job["pot_local"]).cuda()
            xs.append(dot_features(mlp, job, d_v.data_ptr()))
            del d_v
        del psi, st
        ctx.close()
    if not xs:
        raise RuntimeError("warm-start training: no training dot converged band-edge states")
    losses = train(mlp, xs, ts, steps=steps, log=log)
    return mlp, dict(dots=info, loss_first=float(losses[0]), loss_last=float(losses[-1]), steps=steps)

