"""Benchmark of the filter-diagonalization path on one MI355X.

Default workload: the full-stack configuration, a 4.5 nm spin-orbit
zincblende dot on a 256^3 grid, 16 ran2 start states x 16 energy targets.
Two passes of the path run with the start states in HBM:

* mixed + warm: the learned warm start (network trained beforehand, untimed,
  on small dots) shapes the start states, the Newton filter runs in FP32
  complex with FP64 coefficients and accumulators at the configuration's
  reduced Newton length (1152: the smallest at which the warm-started pass
  converges the band-edge pairs; plain starts converge none there), then
  normalize_all, FP64 orthogonalisation (rank cut 1e-10), Rayleigh-Ritz,
  sigma_E and one FP64 refinement sweep;
* plain FP64: ran2 start states, FP64 filter of Newton length 2048,
  normalize_all, ortho, RR, sigma_E.

value = eigenpairs with sigma_E < 1e-4 Ha per second of the mixed + warm pass;
the FP64 pass and the eigenvalue error of every FP64-converged state against
the mixed pass are reported beside it.  A step (--steps / --warmup) is one
Newton step of the mixed filter over the whole batch (H applied to every start
state, the accumulation into every target): the timed steps give ms_per_step
and H psi / s; the full passes give value.

    python bench.py [--steps K] [--warmup W] [--config c4_so_256]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "eigenpairs/s at resid<1e-4 Ha, mixed+warm vs plain FP64; Hψ/s, HBM roofline %"
HBM_PEAK_GBS = 8000.0
RESID = 1e-4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def filter_guard(targets, dE, M):
    """The reference's filter-width check (coeff.c:30-56): the Gaussian filter
    of width sigma_f = 1/sqrt(2 dt), dt = (M / (2.5 dE))^2 (mod_filter.c:114),
    must not exceed half the smallest spacing of consecutive targets."""
    dt = (M / (2.5 * dE)) ** 2
    sig_f = math.sqrt(1.0 / (2.0 * dt))
    d = np.abs(np.diff(targets))
    min_sp = float(d[d > 0].min()) if np.any(d > 0) else float("inf")
    return dict(sigma_f=round(sig_f, 6), min_target_spacing=round(min_sp, 6), dE=round(float(dE), 6), M=int(M),
                ok=bool(sig_f <= 0.5 * min_sp)), dt


def setup(cfg_name, dev, ncheby=0):
    import torch
    from filterbse_amd import api, dots

    cfg = dict(dots.CONFIGS[cfg_name])
    if ncheby:
        cfg["ncheby"] = ncheby
    t0 = time.perf_counter()
    job = dots.build_dot(cfg["radius_nm"], cfg["n_req"], cfg["dx"], so=cfg["so"], nl=cfg["nl"],
                         ke_max=cfg.get("ke_max", 20.0), builder=cfg.get("builder", "host"))
    log("dot: %d core atoms, %d ligands, grid %dx%dx%d, KEmax %.1f, %.1fs" % (
        job["n_core"], len(job["syms"]) - job["n_core"], job["nx"], job["ny"], job["nz"], cfg.get("ke_max", 20.0),
        time.perf_counter() - t0))
    ctx = api.Context(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
                      job["pot_local"], job["proj"])
    t0 = time.perf_counter()
    p0, p1 = api.energy_range_start_states(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"],
                                           job["dv"])
    rng_out = ctx.energy_range(p0, p1, max_iter=1500)
    t_range = time.perf_counter() - t0
    Emin_raw, Emax_raw, Vmin, dE = rng_out[:4]
    M, ms, n_starts = cfg["ncheby"], cfg["ms"], cfg["n_starts"]
    targets = api.ene_targets(*cfg["targets"])
    guard, dt = filter_guard(targets, dE, M)
    zn = api.leja_points(M)
    an = api.newton_coeff(zn, targets, Vmin, dE, dt)
    stlen = (job["is_complex"] + 1) * job["nspin"] * job["nx"] * job["ny"] * job["nz"]
    starts, _ = api.init_filter_states(n_starts, 1, -1, job["nx"], job["ny"], job["nz"], job["nspin"],
                                       job["is_complex"], job["dv"])
    starts_dev = torch.from_numpy(starts).to(dev).view(n_starts, stlen)
    del starts
    psi_rank = torch.empty(n_starts * ms * stlen, dtype=torch.float64, device=dev)
    log("energy range [%.6g, %.6g] (%d + %d iterations, %.2fs); Vmin %.6g dE %.6g; filter guard %s" % (
        Emin_raw, Emax_raw, rng_out[4], rng_out[5], t_range, Vmin, dE, guard))
    return dict(cfg=cfg, job=job, ctx=ctx, M=M, ms=ms, n_starts=n_starts, stlen=stlen, an=an, zn=zn, Vmin=Vmin,
                dE=dE, targets=targets, starts_dev=starts_dev, psi_rank=psi_rank, range=rng_out, t_range=t_range,
                sig_f=guard["sigma_f"], guard=guard)


def with_newton_length(S, M):
    """The pass state S with Newton length M (Leja points, coefficients and
    filter width for the same targets and spectrum bounds)."""
    from filterbse_amd import api
    Sm = dict(S)
    guard, dt = filter_guard(S["targets"], S["dE"], M)
    Sm.update(M=M, zn=api.leja_points(M), sig_f=guard["sigma_f"], guard=guard)
    Sm["an"] = api.newton_coeff(Sm["zn"], S["targets"], S["Vmin"], S["dE"], dt)
    return Sm


def warm_start_network(S):
    """Trains the warm-start network (untimed set-up: small training dots,
    plain FP64 filters, Adam on the GPU) and keeps the dot's V_loc on the device."""
    import torch
    from filterbse_amd import warmstart as WS
    t0 = time.perf_counter()
    mlp, info = WS.train_network(log=log)
    info["train_s"] = round(time.perf_counter() - t0, 2)
    S["mlp"], S["mlp_info"] = mlp, info
    S["d_vloc"] = torch.from_numpy(S["job"]["pot_local"]).to(S["psi_rank"].device)
    log("warm-start network trained: %s" % info)


def one_pass(S, precision, warm=False, refine=0, profile=False):
    """One pass of the path from the ran2 start states (warm: reshaped by the
    warm-start network inside the pass).  Returns (eval, sigma, rank, stage
    seconds [start states, filter, normalize+ortho, RR, sigma, refine])."""
    import torch
    from filterbse_amd import warmstart as WS
    ctx = S["ctx"]
    n = S["n_starts"] * S["ms"]
    blk = S["psi_rank"].view(S["n_starts"], S["ms"], S["stlen"])
    torch.cuda.synchronize()
    t = [time.perf_counter()]
    blk[:, 0, :].copy_(S["starts_dev"])
    if warm:
        WS.warm_starts(ctx, S["mlp"], S["job"], S["d_vloc"].data_ptr(), blk[:, 0, :], ms=S["ms"])
    torch.cuda.synchronize()
    t.append(time.perf_counter())
    ptr = S["psi_rank"].data_ptr()
    torch.cuda.empty_cache()  # memory torch's allocator caches is HBM the filter's ring cannot use
    free_gb = torch.cuda.mem_get_info()[0] / 1e9
    if profile:
        ctx.profile(True)
    ctx.filter_dev(ptr, S["n_starts"], S["ms"], S["an"], S["zn"], S["Vmin"], 4.0 / S["dE"], precision)
    if profile:
        S["prof"] = {k: ctx.profile_read(k) for k in range(5)}
        ctx.profile(False)
    S.setdefault("windows", []).append(dict(precision=int(precision), K=ctx.last_window(), runs=ctx.last_runs(),
                                            free_gb=round(free_gb, 1)))
    t.append(time.perf_counter())
    ctx.normalize_all_dev(ptr, n)
    rank, sv = ctx.ortho_dev(ptr, n)
    t.append(time.perf_counter())
    ev = ctx.rayleigh_ritz_dev(ptr, rank)
    t.append(time.perf_counter())
    sig, _ = ctx.sigma_dev(ptr, rank)
    t.append(time.perf_counter())
    S["refine_log"] = []
    for k in range(refine):
        # FP64 refinement over the Ritz pairs within 4 filter widths of the targets
        t_k = time.perf_counter()
        w = 4.0 * S["sig_f"]
        ev2, sig2, nrep = ctx.refine_dev(ptr, rank, ev, sig, S["targets"].min() - w, S["targets"].max() + w,
                                         e_p=S["cfg"].get("refine_ep", 0.25), sigma_max=1e-2)
        S["refine_log"].append(dict(replaced=int(nrep), converged_before=int((sig < RESID).sum()),
                                    converged_after=int((sig2 < RESID).sum()),
                                    best_before=float(sig.min()), best_after=float(sig2.min()),
                                    seconds=round(time.perf_counter() - t_k, 3)))
        ev, sig = ev2, sig2
    torch.cuda.synchronize()
    t.append(time.perf_counter())
    return ev, sig, rank, np.diff(t)


def converged(ev, sig):
    ok = sig < RESID
    return ev[ok], sig[ok]


def per_state_errors(ev64, sig64, ev_m, sig_m):
    """For every FP64-converged state: its eigenvalue, the nearest eigenvalue
    of the mixed pass, |difference| and both residuals."""
    out = []
    for i in np.argsort(ev64):
        if sig64[i] >= RESID:
            continue
        j = int(np.argmin(np.abs(ev_m - ev64[i])))
        out.append(dict(e_fp64=float(ev64[i]), e_mixed=float(ev_m[j]), abs_err=float(abs(ev_m[j] - ev64[i])),
                        sigma_fp64=float(sig64[i]), sigma_mixed=float(sig_m[j])))
    return out


def mixed_state_errors(ev_m, sig_m, ev64, sig64):
    """For every state the mixed pass converged: its eigenvalue, the nearest
    Ritz value of the FP64 pass (converged or not), |difference| and both
    residuals."""
    out = []
    for j in np.argsort(ev_m):
        if sig_m[j] >= RESID:
            continue
        i = int(np.argmin(np.abs(ev64 - ev_m[j])))
        out.append(dict(e_mixed=float(ev_m[j]), e_fp64=float(ev64[i]), abs_err=float(abs(ev_m[j] - ev64[i])),
                        sigma_mixed=float(sig_m[j]), sigma_fp64=float(sig64[i])))
    return out


def newton_steps(S, n, precision):
    """n Newton steps of the filter over the whole batch (a filter call of
    Newton length n + 1 on the ran2 start states); returns seconds."""
    import torch
    from filterbse_amd import api
    M = n + 1
    zn = api.leja_points(M)
    guard, dt = filter_guard(S["targets"], S["dE"], M)
    an = api.newton_coeff(zn, S["targets"], S["Vmin"], S["dE"], dt)
    S["psi_rank"].view(S["n_starts"], S["ms"], S["stlen"])[:, 0, :].copy_(S["starts_dev"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S["ctx"].filter_dev(S["psi_rank"].data_ptr(), S["n_starts"], S["ms"], an, zn, S["Vmin"], 4.0 / S["dE"],
                        precision)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def cpu_baseline(S, budget_s=20.0):
    """The oracle's threaded Newton step (oracle/cpu_newton.py: scipy FFT
    pair, OpenMP projectors / V_loc / accumulation) on the same workload, one
    start state, on the host cores this process may use."""
    from oracle import cpu_newton
    job = S["job"]
    start = S["starts_dev"][0].cpu().numpy()
    out = cpu_newton.baseline(job, S["ms"], S["an"], S["zn"], S["Vmin"], 4.0 / S["dE"], start, budget_s=budget_s)
    out["sample"] += " (%s)" % S["cfg_name"]
    return out


def measured_traffic(cfg_name, kname):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload, when it covers that
    kernel instance (else None)."""
    path = os.path.join(ROOT, "profiles", "%s_traffic.json" % cfg_name)
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("traffic_bytes_per_launch", {}).get(kname)


def traffic_sample(cfg_name, dev, M):
    """A short mixed filter of the workload (its FFT launches are the ones the
    roofline names) followed by a device-to-device copy of the bytes one FFT
    launch reads (and writes): the copy's FETCH_SIZE / WRITE_SIZE per byte
    calibrates the counters.  The spectrum bounds are a coarse Weyl-type bound
    (the traffic of a launch does not depend on them)."""
    import torch
    from filterbse_amd import api, dots

    cfg = dict(dots.CONFIGS[cfg_name])
    job = dots.build_dot(cfg["radius_nm"], cfg["n_req"], cfg["dx"], so=cfg["so"], nl=cfg["nl"],
                         ke_max=cfg.get("ke_max", 20.0), builder=cfg.get("builder", "host"))
    ctx = api.Context(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
                      job["pot_local"], job["proj"])
    ngrid = job["nx"] * job["ny"] * job["nz"]
    Vmin = float(job["pot_local"].min()) - 1.0
    dE = float(job["ksqr"].max()) * ngrid + float(job["pot_local"].max() - job["pot_local"].min()) + 2.0
    ms, n_starts = cfg["ms"], cfg["n_starts"]
    dt = (cfg["ncheby"] / (2.5 * dE)) ** 2
    zn = api.leja_points(M)
    an = api.newton_coeff(zn, api.ene_targets(*cfg["targets"]), Vmin, dE, dt)
    stlen = (job["is_complex"] + 1) * job["nspin"] * ngrid
    starts, _ = api.init_filter_states(n_starts, ms, -1, job["nx"], job["ny"], job["nz"], job["nspin"],
                                       job["is_complex"], job["dv"])
    psi = torch.from_numpy(starts).to(dev)
    del starts
    ctx.filter_dev(psi.data_ptr(), n_starts, ms, an, zn, Vmin, 4.0 / dE, api.FBSE_PREC_MIXED)
    torch.cuda.synchronize()
    n_items = n_starts if job["is_complex"] else (n_starts + 1) // 2
    nbytes = n_items * job["nspin"] * ngrid * 8
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    log("traffic sample: %d-step mixed filter; copy of %d bytes (one FFT launch reads and writes %d each)"
        % (M, nbytes, nbytes))


def io_sample(args, dev):
    """State-file streaming: one start block of the c1 workload (64 start
    states x 16 targets would be 29 GB; here the first n FP64 spinor states of
    96^3, 28.3 MB each) written from HBM in the reference's psi.dat layout
    (fbse_write_states: pinned double buffering) and read back as start states
    (fbse_read_start_states), against the same write through torch's
    device-to-host copy and numpy's tofile.  Files go to $TMPDIR (page cache:
    no fsync, as the reference's fwrite/fclose)."""
    import tempfile
    import torch
    from filterbse_amd import api
    cfg = dict(__import__("filterbse_amd.dots", fromlist=["CONFIGS"]).CONFIGS["c1_so_96"])
    n = args.io_sample
    nx, ny, nz = cfg["n_req"]
    nspin = 2
    ngrid = nx * ny * nz
    ksqr = np.zeros(ngrid)
    ctx = api.Context(nx, ny, nz, nspin, 1, cfg["dx"] ** 3, ksqr, np.zeros(nspin * ngrid), None)
    stlen = 2 * nspin * ngrid
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    d = torch.randn(n * stlen, dtype=torch.float64, device=dev, generator=g)
    torch.cuda.synchronize()
    nbytes = n * stlen * 8
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "psi.dat")
        for rep in range(2):  # the second repetition is reported (pinned buffers allocated)
            t0 = time.perf_counter()
            ctx.write_states(d.data_ptr(), n, path)
            out["write_gbs"] = nbytes / (time.perf_counter() - t0) / 1e9
        t0 = time.perf_counter()
        d.cpu().numpy().tofile(path + ".ref")
        out["write_torch_numpy_gbs"] = nbytes / (time.perf_counter() - t0) / 1e9
        same = open(path, "rb").read() == open(path + ".ref", "rb").read()
        r = torch.empty(n * stlen, dtype=torch.float64, device=dev)
        t0 = time.perf_counter()
        k = ctx.read_start_states(path, 0, n, 1, r.data_ptr(), n, 1)
        out["read_gbs"] = nbytes / (time.perf_counter() - t0) / 1e9
        ok = k == n and bool(torch.equal(r, d))
    line = dict(metric="state-file streaming GB/s (HBM <-> file, reference psi.dat layout)",
                value=round(out["write_gbs"], 3), unit="GB/s", n_gpus=1, higher_is_better=True,
                data="synthetic (seeded normal FP64 states)",
                config=dict(workload="%d FP64 spinor states of a 96^3 grid (%.2f GB), 64 MiB pinned chunks" % (
                    n, nbytes / 1e9)),
                write_gbs=round(out["write_gbs"], 3), write_torch_numpy_gbs=round(out["write_torch_numpy_gbs"], 3),
                read_gbs=round(out["read_gbs"], 3), bytes_identical_to_numpy=same, read_back_exact=ok)
    print(json.dumps(line), flush=True)


def mlp_sample(args, dev):
    """The warm-start network's own BF16 MFMA kernels (include/fbse_mlp.h) on
    n grid points of synthetic features: forward passes, then training steps
    (loss + gradient + Adam).  FLOPs are those of the MFMA tiles the kernels
    issue (input width padded to a multiple of 16): 2 n (k_in w + 2 w^2 + w
    n_out) per forward pass, three times that per training step (forward, data
    gradient, weight gradient); against the dense BF16 peak of 2.5 PFLOP/s."""
    import torch
    from filterbse_amd import api
    from filterbse_amd import warmstart as WS
    n = args.mlp_sample
    width = 128
    mlp = api.Mlp(WS.N_IN, width, WS.N_OUT, seed=1)
    k = mlp.k_in
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    x = (torch.randn((n, k), generator=g, device=dev) * (torch.arange(k, device=dev) < WS.N_IN)).to(
        torch.bfloat16).view(torch.int16).contiguous()
    t = torch.randn((n, WS.N_OUT), generator=g, device=dev)
    y = torch.empty((n, WS.N_OUT), dtype=torch.float32, device=dev)
    flops_fwd = 2.0 * n * (k * width + 2 * width * width + width * WS.N_OUT)
    torch.cuda.synchronize()
    mlp.forward_dev(x.data_ptr(), n, y.data_ptr())  # warm-up
    torch.cuda.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        mlp.forward_dev(x.data_ptr(), n, y.data_ptr())
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    mlp.loss_grad_dev(x.data_ptr(), t.data_ptr(), None, n)
    t0 = time.perf_counter()
    for _ in range(reps):
        mlp.loss_grad_dev(x.data_ptr(), t.data_ptr(), None, n)
        mlp.adam_step(1e-3)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / reps
    peak = 2500.0
    fwd = flops_fwd / t_fwd / 1e12
    trn = 3.0 * flops_fwd / t_step / 1e12
    line = dict(metric="warm-start network BF16 MFMA throughput", value=round(trn, 2), unit="TFLOP/s", n_gpus=1,
                higher_is_better=True, dtype="bf16 (fp32 accumulate)", data="synthetic features",
                config=dict(workload="%d grid points, %d -> %d -> %d -> %d -> %d MLP" % (
                    n, WS.N_IN, width, width, width, WS.N_OUT)),
                forward_ms=round(1e3 * t_fwd, 3), forward_tflops=round(fwd, 2),
                train_step_ms=round(1e3 * t_step, 3), train_tflops=round(trn, 2),
                mfma=dict(bound="mfma", achieved=round(trn, 2), peak=peak, unit="TFLOP/s", frac=round(trn / peak, 4)))
    print(json.dumps(line), flush=True)


def gemm_sample(args, dev):
    """The FP64 matrix-core subspace products (gemm_kernels.hip) at the
    workloads' shapes: a block of n random complex states of the workload's
    grid (n = its filtered states), the overlap dv A^H A (fbse_overlap_dev:
    the general kernel over all 64 x 64 tiles, split-K over row slabs) and the
    rotation A V in place (fbse_rotate_dev); FLOPs 8 m n^2 each (complex
    multiply-add = 8 real operations), against the dense FP64 peak of
    78.6 TFLOP/s."""
    import torch
    from filterbse_amd import api, dots
    shapes = {"c1_so_96": (96, 2, 1024), "c4_so_256": (256, 2, 256)}
    out = []
    for name in args.gemm_sample.split(","):
        g, nspin, n = shapes[name]
        ng = g ** 3
        ctx = api.Context(g, g, g, nspin, 1, 1.0, np.zeros(ng), np.zeros(nspin * ng), None)
        m = nspin * ng
        a = torch.randn(n * 2 * m, dtype=torch.float64, device=dev)
        a /= math.sqrt(2.0 * m)
        V = np.linalg.qr(np.random.default_rng(1).standard_normal((n, n)))[0].astype(np.complex128)
        torch.cuda.synchronize()
        ctx.overlap_dev(a.data_ptr(), n, a.data_ptr(), n)  # warm-up
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.overlap_dev(a.data_ptr(), n, a.data_ptr(), n)
        t_g = (time.perf_counter() - t0) / reps
        ctx.rotate_dev(a.data_ptr(), V)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.rotate_dev(a.data_ptr(), V)
        ctx.synchronize()
        t_r = (time.perf_counter() - t0) / reps
        fl = 8.0 * m * n * n
        out.append(dict(shape=name, rows=m, states=n, gram_ms=round(1e3 * t_g, 3),
                        gram_tflops=round(fl / t_g / 1e12, 2), rotate_ms=round(1e3 * t_r, 3),
                        rotate_tflops=round(fl / t_r / 1e12, 2)))
        log("gemm sample %s: %s" % (name, out[-1]))
        del a
        ctx.close()
        torch.cuda.empty_cache()
    peak = 78.6
    best = max(o["gram_tflops"] for o in out)
    line = dict(metric="FP64 matrix-core subspace products (Gram A^H A, rotation A V)", value=best,
                unit="TFLOP/s", n_gpus=1, higher_is_better=True, dtype="f64 (complex)", data="synthetic",
                config=dict(workload="random complex state blocks of the c1 / c4 shapes"), shapes=out,
                mfma=dict(bound="mfma", peak=peak, unit="TFLOP/s",
                          frac={o["shape"]: dict(gram=round(o["gram_tflops"] / peak, 4),
                                                 rotate=round(o["rotate_tflops"] / peak, 4)) for o in out}))
    print(json.dumps(line), flush=True)


def bench_uneven(args, dev):
    """One step = the filters of all dots (fbse_filter_many: every dot on its
    own stream, launches interleaved) plus normalize / ortho / RR / sigma_E of
    each dot, start states in HBM.  Dots are processed in as few groups as the
    HBM holds (psi_rank of a group + its FP32 rings)."""
    import torch
    from filterbse_amd import api, dots

    cfg = dots.UNEVEN
    rng = np.random.default_rng(cfg["seed"])
    radii = rng.uniform(cfg["r_lo"], cfg["r_hi"], cfg["n_dots"])
    D = []
    t_setup = time.perf_counter()
    t_range = 0.0
    targets = api.ene_targets(*cfg["targets"])
    M, ms, ns = cfg["ncheby"], cfg["ms"], cfg["n_starts"]
    zn = api.leja_points(M)
    for i, r in enumerate(radii):
        # the grid the reference's rule needs, rounded up to FFT-friendly axes
        n_req = dots.friendly_grid(float(r), cfg["dx"], seed=100 + i)
        job = dots.build_dot(float(r), n_req, cfg["dx"], so=True, nl=True, seed=100 + i,
                             ke_max=cfg.get("ke_max", 20.0), builder="gpu")
        ctx = api.Context(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
                          job["pot_local"], job["proj"])
        t0 = time.perf_counter()
        p0, p1 = api.energy_range_start_states(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"],
                                               job["dv"])
        rr = ctx.energy_range(p0, p1, max_iter=1500)
        t_range += time.perf_counter() - t0
        ctx.trim()
        Vmin, dE = rr[2], rr[3]
        an = api.newton_coeff(zn, targets, Vmin, dE, (M / (2.5 * dE)) ** 2)
        stlen = 2 * job["nspin"] * job["nx"] * job["ny"] * job["nz"]
        st, _ = api.init_filter_states(ns, 1, -(100 + i), job["nx"], job["ny"], job["nz"], job["nspin"], 1,
                                       job["dv"])
        D.append(dict(r=float(r), n=(job["nx"], job["ny"], job["nz"]), atoms=len(job["syms"]), ctx=ctx, an=an,
                      Vmin=Vmin, dE=dE, stlen=stlen, starts=torch.from_numpy(st).to(dev).view(ns, stlen)))
        log("dot %2d: r %.2f nm, grid %dx%dx%d, %d atoms" % (i, r, job["nx"], job["ny"], job["nz"], len(job["syms"])))
        del job, st
    t_setup = time.perf_counter() - t_setup
    # groups that fit: psi_rank (FP64) + about four FP32 ring slots per dot
    free, _ = torch.cuda.mem_get_info()
    budget = 0.55 * free
    # dots in grid-size order, so that dots of one size share a filter_many
    # call and with it the grouped FFT launches of their size
    groups, cur, used = [], [], 0.0
    for k in sorted(range(len(D)), key=lambda i: (D[i]["n"], i)):
        d = D[k]
        need = ns * ms * d["stlen"] * 8 + 4 * ns * d["stlen"] * 4
        if cur and (used + need > budget or len(cur) == cfg["group_max"]):
            groups.append(cur)
            cur, used = [], 0.0
        cur.append(k)
        used += need
    if cur:
        groups.append(cur)
    log("%d dots in %d group(s); setup %.1fs (energy ranges %.1fs)" % (len(D), len(groups), t_setup, t_range))
    mixed = api.FBSE_PREC_MIXED
    conv_total, t_filter = 0, 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for gi, grp in enumerate(groups):
        psis = []
        for k in grp:
            d = D[k]
            p = torch.zeros(ns * ms * d["stlen"], dtype=torch.float64, device=dev)
            p.view(ns, ms, d["stlen"])[:, 0, :].copy_(d["starts"])
            psis.append(p)
        torch.cuda.synchronize()
        tf = time.perf_counter()
        api.filter_many([D[k]["ctx"] for k in grp], [p.data_ptr() for p in psis], [ns] * len(grp), ms,
                        [D[k]["an"] for k in grp], [zn] * len(grp), [D[k]["Vmin"] for k in grp],
                        [4.0 / D[k]["dE"] for k in grp], mixed)
        t_filter += time.perf_counter() - tf
        for k in grp:
            D[k]["launch_group"] = D[k]["ctx"].last_group()
        for k, p in zip(grp, psis):
            c = D[k]["ctx"]
            c.normalize_all_dev(p.data_ptr(), ns * ms)
            rank, _ = c.ortho_dev(p.data_ptr(), ns * ms)
            ev = c.rayleigh_ritz_dev(p.data_ptr(), rank)
            sig, _ = c.sigma_dev(p.data_ptr(), rank)
            D[k]["conv"] = int((sig < RESID).sum())
            D[k]["rank"] = int(rank)
            conv_total += D[k]["conv"]
            c.trim()
        del psis
        log("group %d (%d dots) done: %.1fs so far" % (gi, len(grp), time.perf_counter() - t0))
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    hpsi = sum(ns * (M - 1) for _ in D) / t_filter
    pts = sum(ns * 2 * d["n"][0] * d["n"][1] * d["n"][2] for d in D) * (M - 1) / t_filter
    line = {
        "metric": METRIC,
        "value": round(conv_total / t_total, 4),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(1e3 * t_total, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-complex filter, f64 ortho/RR/sigma",
        "data": "synthetic (seeded zincblende dots, ran2 start states)",
        "config": {"workload": "c2_uneven: %d spin-orbit dots, radii %.2f-%.2f nm, grids %s, %d starts x %d targets "
                               "each, Newton length %d" % (len(D), min(radii), max(radii),
                                                           "%d^3..%d^3" % (min(d["n"][0] for d in D),
                                                                          max(d["n"][0] for d in D)), ns, ms, M),
                   "parallelism": "single GPU; %d filter_many call(s); dots of one grid size share grouped FFT "
                                  "launches (%d dots in %d launch groups), others interleaved over per-dot streams" % (
                                      len(groups), sum(1 for d in D if d["launch_group"] > 1),
                                      round(sum(1.0 / d["launch_group"] for d in D if d["launch_group"] > 1)))},
        "converged": conv_total,
        "hpsi_per_s": round(hpsi, 1),
        "grid_points_per_s": round(pts, 1),
        "filter_s": round(t_filter, 3),
        "energy_range_s": round(t_range, 2),
        "dots": [dict(r_nm=round(d["r"], 3), grid=d["n"], atoms=d["atoms"], rank=d["rank"], converged=d["conv"],
                      launch_group=d["launch_group"]) for d in D],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed Newton steps of the mixed filter")
    ap.add_argument("--warmup", type=int, default=5, help="untimed Newton steps before them")
    ap.add_argument("--window", type=int, default=0,
                    help="accumulation window depth K (0 = as deep as 75%% of the free HBM allows)")
    ap.add_argument("--config", default="c4_so_256")
    ap.add_argument("--no-warm", action="store_true", help="mixed pass from the plain ran2 start states")
    ap.add_argument("--refine", type=int, default=1, help="FP64 refinement sweeps in the mixed pass")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--traffic-sample", type=int, default=0, metavar="M",
                    help="only run an M-step mixed filter of the workload and one device copy of the bytes of "
                         "one FFT launch (for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs), then exit")
    ap.add_argument("--ncheby", type=int, default=0, metavar="M",
                    help="Newton length override of both passes (short filters for counter runs of the other stages)")
    ap.add_argument("--ncheby-warm", type=int, default=-1, metavar="M",
                    help="Newton length of the warm-started mixed pass (-1: the configuration's, 0: the FP64 pass's)")
    ap.add_argument("--mlp-sample", type=int, default=0, metavar="N",
                    help="only measure the warm-start network's MFMA kernels on N grid points, then exit")
    ap.add_argument("--gemm-sample", default="", metavar="SHAPES",
                    help="only time the FP64 matrix-core Gram and rotation at c1_so_96,c4_so_256 shapes, then exit")
    ap.add_argument("--io-sample", type=int, default=0, metavar="N",
                    help="only measure state-file streaming of N FP64 spinor states of 96^3, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the filter runs on a single MI355X; "
              "--gpus %d is not supported (start states would shard across GPUs with no exchange "
              "until one all-to-all before ortho)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from filterbse_amd import api
    dev = torch.device("cuda:0")
    if args.traffic_sample:
        traffic_sample(args.config, dev, args.traffic_sample)
        return
    if args.io_sample:
        io_sample(args, dev)
        return
    if args.gemm_sample:
        gemm_sample(args, dev)
        return
    if args.mlp_sample:
        mlp_sample(args, dev)
        return
    if args.config == "c2_uneven":
        bench_uneven(args, dev)
        return
    S = setup(args.config, dev, args.ncheby)
    S["cfg_name"] = args.config
    cfg, job, ctx = S["cfg"], S["job"], S["ctx"]
    if args.window:
        ctx.set_window(args.window)
    mixed, fp64 = api.FBSE_PREC_MIXED, api.FBSE_PREC_FP64
    warm = not args.no_warm
    if warm:
        warm_start_network(S)
    S["windows"] = []
    # the warm-started mixed pass runs at its own (reduced) Newton length
    M_w = args.ncheby_warm if args.ncheby_warm >= 0 else cfg.get("ncheby_warm", 0)
    Sw = with_newton_length(S, M_w) if warm and M_w > 0 and M_w != S["M"] and not args.ncheby else S

    # steps: Newton steps of the mixed filter over the whole batch
    if args.warmup:
        log("warmup: %d Newton steps (mixed)" % args.warmup)
        newton_steps(S, args.warmup, mixed)
    log("timed: %d Newton steps (mixed)" % args.steps)
    t_steps = newton_steps(S, args.steps, mixed)
    ms_per_step = 1e3 * t_steps / args.steps

    # the headline passes
    log("mixed%s pass, Newton length %d, filter guard %s" % ("+warm" if warm else "", Sw["M"], Sw["guard"]))
    ev, sig, rank, st = one_pass(Sw, mixed, warm=warm, refine=args.refine)
    t_mixed = float(st.sum())
    refine_log = Sw["refine_log"]
    ev_c, sig_c = converged(ev, sig)
    log("mixed: %.3fs (start states %.3f filter %.3f ortho %.3f RR %.3f sigma %.3f refine %.3f), rank %d, "
        "converged %d, refine %s" % (t_mixed, *st, rank, len(ev_c), refine_log))
    fp = None
    if not args.no_fp64:
        log("fp64 pass")
        ev64, sig64, rank64, st64 = one_pass(S, fp64)
        t64 = float(st64.sum())
        ev64_c, _ = converged(ev64, sig64)
        errs = per_state_errors(ev64, sig64, ev, sig)
        errs_m = mixed_state_errors(ev, sig, ev64, sig64)
        fp = dict(seconds_per_pass=round(t64, 3), eigenpairs_per_s=len(ev64_c) / t64, converged=int(len(ev64_c)),
                  rank=int(rank64), stage_s=[round(float(x), 4) for x in st64],
                  hpsi_per_s=round(S["n_starts"] * (S["M"] - 1) / st64[1], 1),
                  converged_eigenvalues=[round(float(x), 9) for x in np.sort(ev64_c)])
        fp["max_eig_err_mixed_vs_fp64"] = max([e["abs_err"] for e in errs]) if errs else None
        log("fp64: %.3fs/pass, rank %d, converged %d, per-state errors %s" % (t64, rank64, len(ev64_c), errs))
    else:
        errs, errs_m = [], []

    # per-kernel HIP events over a short profiled mixed filter (its launches
    # are those of the full filter; every launch on the library stream)
    Mp = max(3 * (max(ctx.last_window(), 1) + 1) + 2, 64)
    Sp = dict(S)
    Sp["M"] = Mp
    Sp["zn"] = api.leja_points(Mp)
    _, dtp = filter_guard(S["targets"], S["dE"], Mp)
    Sp["an"] = api.newton_coeff(Sp["zn"], S["targets"], S["Vmin"], S["dE"], dtp)
    blk = S["psi_rank"].view(S["n_starts"], S["ms"], S["stlen"])
    blk[:, 0, :].copy_(S["starts_dev"])
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.filter_dev(S["psi_rank"].data_ptr(), S["n_starts"], S["ms"], Sp["an"], Sp["zn"], S["Vmin"], 4.0 / S["dE"],
                   mixed)
    prof = {k: ctx.profile_read(k) for k in range(5)}
    ctx.profile(False)
    K_prof = ctx.last_window()
    runs_prof = max(ctx.last_runs(), 1)
    fft_path, fft_inst = ctx.fft_choice(mixed)
    steps_p = Mp - 1
    ngrid = job["nx"] * job["ny"] * job["nz"]
    s_ = job["nspin"]
    n_items = S["n_starts"] if job["is_complex"] else (S["n_starts"] + 1) // 2
    items_run = -(-n_items // runs_prof)  # batch items of one launch (the filter's sub-batch runs)
    # algorithmic bytes (FP32 complex psi, FP32 V_loc): B_Hpsi per state and
    # the windowed accumulation per state and step (FP64 accumulators)
    b_hpsi = 2 * s_ * ngrid * 8 + ngrid * 4
    b_acc = 2 * S["ms"] * s_ * ngrid * 16 / max(K_prof, 1) + s_ * ngrid * 8
    t_h = (prof[0][1] + prof[1][1] + prof[2][1]) / steps_p          # ms per Newton step, H psi kernels
    t_all = prof[4][1] / steps_p
    ach = b_hpsi * n_items / (t_h * 1e-3) / 1e9
    whole = (b_hpsi + b_acc) * n_items / (t_all * 1e-3) / 1e9
    # one FFT sweep on its own bytes: it reads and writes the batch
    l0, m0 = prof[0]
    own_bytes = 2.0 * items_run * s_ * ngrid * 8
    own = own_bytes / (m0 / max(l0, 1) * 1e-3) / 1e9
    # the x combine on its own bytes (per-axis passes, fused: it reads the work
    # buffer, psi_{j-1}, V_loc and the projector records, and writes psi_j and
    # the next step's forward x transform)
    l1, m1 = prof[1]
    n_rec = int(np.sum(job["proj"]["n_gpt"])) if job.get("proj") is not None else 0
    comb_bytes = 4.0 * items_run * s_ * ngrid * 8 + s_ * ngrid * 4 + n_rec * 40
    comb = comb_bytes / (m1 / max(l1, 1) * 1e-3) / 1e9
    if fft_path == 1:
        kname = ("fft_fx_f32_n%d<*>" % job["nx"]) if fft_inst == 4 else "fft_tile_f32_e*<*>"
    else:
        kname = "axis_fx_f32_n%d_l*" % job["nx"] if job["nx"] in (160, 256) else "fft_axis_kernel<float>"
    roofline = dict(
        bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4),
        traffic=measured_traffic(args.config, "step"),
        kernel="H psi of one Newton step: %s FFT passes + combine pass + projector passes" % kname,
        algorithmic="B_Hpsi = 2 s N b_psi + N b_V = %d bytes per state (b_psi 8, b_V 4) x %d states" % (
            b_hpsi, n_items),
        avg_ms_per_step=round(t_h, 4), profiled_steps=steps_p,
        kernel_own=dict(kernel=kname + " (one FFT sweep)", bytes_per_launch=own_bytes,
                        avg_launch_ms=round(m0 / max(l0, 1), 4), achieved=round(own, 1),
                        frac=round(own / HBM_PEAK_GBS, 4), traffic=measured_traffic(args.config, "fft_sweep")),
        combine_own=dict(kernel="x backward + V_loc + Newton shift + projector expansion + next forward x "
                                "(the longest launch of the step)",
                         bytes_per_launch=comb_bytes, records=n_rec, avg_launch_ms=round(m1 / max(l1, 1), 4),
                         achieved=round(comb, 1), frac=round(comb / HBM_PEAK_GBS, 4),
                         traffic=measured_traffic(args.config, "x_bwd_combine")),
        whole_filter=dict(bytes_per_step=round((b_hpsi + b_acc) * n_items), window_K=K_prof, runs=runs_prof,
                          avg_ms_per_step=round(t_all, 4), achieved=round(whole, 1),
                          frac=round(whole / HBM_PEAK_GBS, 4)),
        measured="HIP events on the library stream over a %d-step profiled mixed filter after the passes" % Mp)
    kernels = {name: dict(launches=prof[k][0], total_ms=round(prof[k][1], 3))
               for k, name in enumerate(["fft_passes", "x_backward_newton_combine", "projectors",
                                         "window_accumulate", "all"])}
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(S)
        log("cpu baseline: %s" % cpu)
    hpsi_pass = S["n_starts"] * (Sw["M"] - 1) / st[1]
    fp64_eps = fp["eigenpairs_per_s"] if fp else None
    line = {
        "metric": METRIC,
        "value": round(len(ev_c) / t_mixed, 5),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-complex filter; f64 coefficients, accumulators, ortho/RR/sigma/refine",
        "data": "synthetic (seeded zincblende dot, ran2 start states%s)" % (", learned warm start" if warm else ""),
        "config": {
            "workload": "%s: %.1f nm dot, %d atoms + %d ligands, %s, grid %dx%dx%d (dx %.3g bohr), KEmax %.1f Ha, "
                        "%d starts x %d targets %s, Newton length %s" % (
                            args.config, cfg["radius_nm"], job["n_core"], len(job["syms"]) - job["n_core"],
                            "SO+NL spinors" if cfg["so"] else "real states", job["nx"], job["ny"], job["nz"],
                            cfg["dx"], cfg.get("ke_max", 20.0), S["n_starts"], S["ms"],
                            "(setTargets %d %d, VB %.4f..%.4f, CB %.4f..%.4f)" % tuple(cfg["targets"]),
                            "%d (FP64 pass), %d (mixed + warm pass)" % (S["M"], Sw["M"])),
            "parallelism": "single GPU; start states batched in one filter",
            "step": "one Newton step of the mixed filter over the batch (H on every start state + accumulation)",
        },
        "converged": int(len(ev_c)),
        "rank": int(rank),
        "pass_s": round(t_mixed, 3),
        "stage_s": [round(float(x), 4) for x in st],
        "hpsi_per_s": round(hpsi_pass, 1),
        "hpsi_per_s_steps": round(S["n_starts"] / (ms_per_step * 1e-3), 1),
        "converged_eigenvalues": [round(float(x), 9) for x in np.sort(ev_c)],
        "refine": refine_log,
        "fp64": fp,
        "mixed_warm_over_fp64": (round((len(ev_c) / t_mixed) / fp64_eps, 3) if fp64_eps else None),
        "per_state_error_vs_fp64": errs,
        "mixed_converged_vs_fp64": errs_m,
        "filter_guard": S["guard"],
        "ncheby_mixed_warm": Sw["M"],
        "filter_guard_mixed_warm": Sw["guard"],
        "warm_start": S.get("mlp_info"),
        "energy_range_s": round(S["t_range"], 3),
        "roofline": roofline,
        "kernels_ms": kernels,
        "filter_windows": S.get("windows"),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
