"""Benchmark of the filter-diagonalization path on one MI355X.

One step = one full pass of the path over the workload with the start states
already in HBM: Newton filter of all start states at all energy targets,
normalize_all, orthogonalisation (rank cut 1e-10), Rayleigh-Ritz, and sigma_E
for every Ritz pair.  Metric: eigenpairs with sigma_E < 1e-4 Ha per second of
that pass (headline: mixed precision), next to the plain FP64 pass, Hpsi/s of
the filter and the HBM roofline of the dominant kernel.

    python bench.py [--steps K] [--warmup W] [--config c1_so_96]
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


def setup(cfg_name, dev, ncheby=0):
    import torch
    from filterbse_amd import api, dots

    cfg = dict(dots.CONFIGS[cfg_name])
    if ncheby:
        cfg["ncheby"] = ncheby
    t0 = time.perf_counter()
    job = dots.build_dot(cfg["radius_nm"], cfg["n_req"], cfg["dx"], so=cfg["so"], nl=cfg["nl"],
                         builder=cfg.get("builder", "host"))
    log("dot: %d core atoms, %d ligands, grid %dx%dx%d, %.1fs" % (
        job["n_core"], len(job["syms"]) - job["n_core"], job["nx"], job["ny"], job["nz"], time.perf_counter() - t0))
    ctx = api.Context(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
                      job["pot_local"], job["proj"])
    t0 = time.perf_counter()
    p0, p1 = api.energy_range_start_states(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"],
                                           job["dv"])
    rng_out = ctx.energy_range(p0, p1, max_iter=1500)
    t_range = time.perf_counter() - t0
    Emin_raw, Emax_raw, Vmin, dE = rng_out[:4]
    M, ms, n_starts = cfg["ncheby"], cfg["ms"], cfg["n_starts"]
    dt = (M / (2.5 * dE)) ** 2  # mod_filter.c:114
    sig_f = math.sqrt(1.0 / (2.0 * dt))
    targets = api.ene_targets(*cfg["targets"])
    zn = api.leja_points(M)
    an = api.newton_coeff(zn, targets, Vmin, dE, dt)
    stlen = (job["is_complex"] + 1) * job["nspin"] * job["nx"] * job["ny"] * job["nz"]
    starts, _ = api.init_filter_states(n_starts, 1, -1, job["nx"], job["ny"], job["nz"], job["nspin"],
                                       job["is_complex"], job["dv"])
    starts_dev = torch.from_numpy(starts).to(dev).view(n_starts, stlen)
    del starts
    psi_rank = torch.empty(n_starts * ms * stlen, dtype=torch.float64, device=dev)
    log("energy range [%.6g, %.6g] (%d + %d iterations, %.2fs); Vmin %.6g dE %.6g; sigma_filter %.4g" % (
        Emin_raw, Emax_raw, rng_out[4], rng_out[5], t_range, Vmin, dE, sig_f))
    return dict(cfg=cfg, job=job, ctx=ctx, M=M, ms=ms, n_starts=n_starts, stlen=stlen, an=an, zn=zn, Vmin=Vmin,
                dE=dE, targets=targets, starts_dev=starts_dev, psi_rank=psi_rank, range=rng_out, t_range=t_range,
                sig_f=sig_f)


def one_pass(S, precision, stage_times=None, refine=False, profile=False):
    """One pass of the path.  profile: per-kernel HIP events (on the library's
    stream) of this pass's filter are left in S["prof"]."""
    import torch
    ctx = S["ctx"]
    n = S["n_starts"] * S["ms"]
    S["psi_rank"].view(S["n_starts"], S["ms"], S["stlen"])[:, 0, :].copy_(S["starts_dev"])
    torch.cuda.synchronize()
    ptr = S["psi_rank"].data_ptr()
    free_gb = torch.cuda.mem_get_info()[0] / 1e9
    held_gb = [round(b / 1e9, 1) for b in ctx.held_bytes()]
    t = [time.perf_counter()]
    if profile:
        ctx.profile(True)
    ctx.filter_dev(ptr, S["n_starts"], S["ms"], S["an"], S["zn"], S["Vmin"], 4.0 / S["dE"], precision)
    # accumulation window depth the filter took, and the HBM free before it
    S.setdefault("windows", []).append(dict(precision=int(precision), profiled=bool(profile), K=ctx.last_window(),
                                            free_gb=round(free_gb, 1), held_gb=held_gb))
    if profile:
        S["prof"] = {k: ctx.profile_read(k) for k in range(5)}
        ctx.profile(False)
    t.append(time.perf_counter())
    mem = lambda: round(torch.cuda.mem_get_info()[0] / 1e9, 1)  # noqa: E731
    free_after = dict(filter=mem())
    ctx.normalize_all_dev(ptr, n)
    rank, sv = ctx.ortho_dev(ptr, n)
    t.append(time.perf_counter())
    free_after["ortho"] = mem()
    ev = ctx.rayleigh_ritz_dev(ptr, rank)
    t.append(time.perf_counter())
    free_after["rr"] = mem()
    sig, ev2 = ctx.sigma_dev(ptr, rank)
    free_after["sigma"] = mem()
    if refine:
        # one FP64 residual-refinement sweep over the Ritz pairs near the
        # targets that are within two decades of the threshold
        w = 4.0 * S["sig_f"]
        ev_r, sig_r = ctx.refine_dev(ptr, rank, ev, S["targets"].min() - w, S["targets"].max() + w, n,
                                     sigma=sig, sigma_max=100.0 * RESID)
        if len(ev_r):
            ev, sig, rank = ev_r, sig_r, len(ev_r)
        free_after["refine"] = mem()
    t.append(time.perf_counter())
    # HBM free after each stage (GB): what the next pass's filter can use
    S["windows"][-1]["free_after_gb"] = free_after
    if stage_times is not None:
        stage_times.append(np.diff(t))
    return ev, sig, rank


def converged(ev, sig):
    ok = sig < RESID
    return ev[ok], sig[ok]


def match_error(ev_a, ev_b):
    """Largest distance from each converged eigenvalue of a to the nearest of b."""
    if len(ev_a) == 0 or len(ev_b) == 0:
        return None
    return float(max(np.min(np.abs(ev_b - x)) for x in ev_a))


def cpu_baseline(S, budget_s=15.0):
    """The oracle's H-psi (a direct restatement of p_hamiltonian, one core) on
    the same Hamiltonian, over a bounded number of applications."""
    from oracle import oracle as O
    job = S["job"]
    h = O.Ham(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
              job["pot_local"], job["proj"])
    psi = S["starts_dev"][0].cpu().numpy()
    n = 0
    t0 = time.perf_counter()
    while True:
        h.apply(psi)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 50:
            break
    return dict(value=n / el, unit="Hψ/s", cores=1, kind="port",
                sample="%d oracle H-psi applications (T+V_loc+SO+NL, one state) on the %s Hamiltonian, %.1f s"
                       % (n, S["cfg_name"], el))


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
    job = dots.build_dot(cfg["radius_nm"], cfg["n_req"], cfg["dx"], so=cfg["so"], nl=cfg["nl"])
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


# Batched uneven dots: 32 spin-orbit dots with seeded radii uniform in
# 1.0-3.0 nm at dx = 1.0 bohr (the reference grid rule needs ~52^3 ... ~120^3;
# each axis is rounded up to a size with prime factors 2, 3, 5), 32 starts x 8
# targets each, Newton length 1024, mixed precision; at most 8 dots per
# fbse_filter_many call (progress is logged between calls).
C2 = dict(n_dots=32, r_lo=1.0, r_hi=3.0, dx=1.0, n_starts=32, ms=8, ncheby=1024,
          targets=(4, 4, -0.32, -0.22, -0.14, -0.04), seed=2, group_max=8)


def bench_uneven(args, dev):
    """One step = the filters of all dots (fbse_filter_many: every dot on its
    own stream, launches interleaved) plus normalize / ortho / RR / sigma_E of
    each dot, start states in HBM.  Dots are processed in as few groups as the
    HBM holds (psi_rank of a group + its FP32 rings)."""
    import torch
    from filterbse_amd import api, dots

    cfg = C2
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
        job = dots.build_dot(float(r), n_req, cfg["dx"], so=True, nl=True, seed=100 + i, builder="gpu")
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
    groups, cur, used = [], [], 0.0
    for k, d in enumerate(D):
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
                   "parallelism": "single GPU; dots interleaved over per-dot streams, %d group(s)" % len(groups)},
        "converged": conv_total,
        "hpsi_per_s": round(hpsi, 1),
        "grid_points_per_s": round(pts, 1),
        "filter_s": round(t_filter, 3),
        "energy_range_s": round(t_range, 2),
        "dots": [dict(r_nm=round(d["r"], 3), grid=d["n"], atoms=d["atoms"], rank=d["rank"], converged=d["conv"])
                 for d in D],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--window", type=int, default=0,
                    help="accumulation window depth K (0 = as deep as 60%% of the free HBM allows)")
    ap.add_argument("--config", default="c1_so_96")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--traffic-sample", type=int, default=0, metavar="M",
                    help="only run an M-step mixed filter of the workload and one device copy of the bytes of "
                         "one FFT launch (for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs), then exit")
    ap.add_argument("--ncheby", type=int, default=0, metavar="M",
                    help="Newton length override (short filters for counter runs of the other stages)")
    ap.add_argument("--mlp-sample", type=int, default=0, metavar="N",
                    help="only measure the warm-start network's MFMA kernels on N grid points, then exit")
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
    if args.mlp_sample:
        mlp_sample(args, dev)
        return
    if args.config == "c2_uneven":
        bench_uneven(args, dev)
        return
    S = setup(args.config, dev, args.ncheby)
    S["cfg_name"] = args.config
    if args.window:
        S["ctx"].set_window(args.window)
    mixed, fp64 = api.FBSE_PREC_MIXED, api.FBSE_PREC_FP64

    for _ in range(args.warmup):
        log("warmup pass (mixed)")
        one_pass(S, mixed, refine=True)
    log("timed mixed pass%s" % ("es" if args.steps > 1 else ""))
    stages = []
    ctx = S["ctx"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev, sig, rank = one_pass(S, mixed, stages, refine=True)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    # per-kernel HIP events of one more (untimed) pass: the library then runs
    # the projection pass on the main stream, so every launch is timed alone
    log("profiled mixed pass (untimed)")
    one_pass(S, mixed, None, refine=True, profile=True)
    prof = S["prof"]
    fft_path, fft_inst = ctx.fft_choice(mixed)
    ev_c, sig_c = converged(ev, sig)
    st = np.mean(stages, axis=0)
    log("mixed: %.3fs/pass (filter %.3f ortho %.3f RR %.3f refine+sigma %.3f), n %d, converged %d"
        % (t_mixed, *st, rank, len(ev_c)))

    fp = None
    if not args.no_fp64:
        fst = []
        log("fp64 pass")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev64, sig64, rank64 = one_pass(S, fp64, fst)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        ev64_c, _ = converged(ev64, sig64)
        fp = dict(seconds_per_pass=t64, eigenpairs_per_s=len(ev64_c) / t64, converged=int(len(ev64_c)),
                  rank=int(rank64), stage_s=[round(float(x), 4) for x in fst[0]],
                  hpsi_per_s=S["n_starts"] * (S["M"] - 1) / fst[0][0],
                  max_eig_err_mixed_vs_fp64=match_error(ev_c, ev64_c))
        log("fp64: %.3fs/pass, rank %d, converged %d, max |e_mixed - e_fp64| %s"
            % (t64, rank64, len(ev64_c), fp["max_eig_err_mixed_vs_fp64"]))

    job = S["job"]
    ngrid = job["nx"] * job["ny"] * job["nz"]
    nsg = job["nspin"] * ngrid
    n_items = S["n_starts"] if job["is_complex"] else (S["n_starts"] + 1) // 2
    elem = 8  # complex float
    grid_bytes = n_items * nsg * elem  # one pass over the batch of spinor grids
    pj = job["proj"]
    chained = pj is None or ((pj["n_so_proj"] if pj["so"] else 0) <= 1 and (pj["n_nl_proj"] if pj["nl"] else 0) <= 1)
    if fft_path == 1:
        e = {1: 12, 2: 16, 3: 32}.get(fft_inst, 12)
        # instance 4: the fixed-schedule persistent instance of a cubic grid
        kfmt = ("fft_fx_f32_n%d<%%d>" % job["nx"]) if fft_inst == 4 else ("fft_tile_f32_e%d<%%d>" % e)
        # dominant kernel: the combine pass.  Algorithmic bytes per launch:
        # the z-slab output and psi_{j-1} read, psi_j written, V_loc read
        # once, and (chained) the next step's forward xy transform written
        launches, ms_tot = prof[1]
        bytes_per_launch = (4.0 if chained else 3.0) * grid_bytes + job["nspin"] * ngrid * 4
        kname = kfmt % 2
        kdesc = ("xy-plane backward FFT, V_loc, Newton combine, projector expansion%s"
                 % (", next step's forward xy FFT" if chained else ""))
        zl, zms = prof[0]
        zslab = dict(kernel=kfmt % 3, desc="xz-slab forward z FFT, k^2, backward z FFT",
                     avg_launch_ms=round(zms / max(zl, 1), 4), bytes_per_launch=2.0 * grid_bytes,
                     achieved=round(2.0 * grid_bytes / (zms / max(zl, 1) * 1e-3) / 1e9, 1))
        zslab["frac"] = round(zslab["achieved"] / HBM_PEAK_GBS, 4)
        zslab["traffic"] = measured_traffic(args.config, zslab["kernel"])
    else:
        launches, ms_tot = prof[0]
        bytes_per_launch = 2.0 * grid_bytes
        kname = "per-axis FFT passes (axis_fx_f32_n%d_l* where the axis length has a fixed instance, else fft_axis_kernel<float>)" % job["nx"]
        kdesc = "x/y forward, z kinetic, y backward passes"
        zslab = None
    avg_ms = ms_tot / max(launches, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(args.config, kname),
                    kernel=kname, desc=kdesc, avg_launch_ms=round(avg_ms, 4), bytes_per_launch=bytes_per_launch,
                    launches=launches, measured="HIP events on the library stream, filter of one profiled pass after the timed ones",
                    zslab=zslab)
    kernels = {name: dict(launches=prof[k][0], total_ms=round(prof[k][1], 3))
               for k, name in enumerate(["fft_passes", "x_backward_newton_combine", "projectors",
                                         "window_accumulate", "all"])}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(S)
    cfg = S["cfg"]
    hpsi = S["n_starts"] * (S["M"] - 1) / st[0]
    line = {
        "metric": METRIC,
        "value": round(len(ev_c) / t_mixed, 4),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mixed, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-complex filter, f64 ortho/RR/sigma",
        "data": "synthetic (seeded zincblende dot, ran2 start states)",
        "config": {
            "workload": "%s: %.1f nm dot, %d atoms + %d ligands, %s, grid %dx%dx%d (dx %.3g bohr), %d starts x %d "
                        "targets, Newton length %d" % (
                            args.config, cfg["radius_nm"], job["n_core"], len(job["syms"]) - job["n_core"],
                            "SO+NL spinors" if cfg["so"] else "real states", job["nx"], job["ny"], job["nz"],
                            cfg["dx"], S["n_starts"], S["ms"], S["M"]),
            "parallelism": "single GPU; start states batched in one filter",
        },
        "converged": int(len(ev_c)),
        "rank": int(rank),
        "hpsi_per_s": round(hpsi, 1),
        "stage_s": [round(float(x), 4) for x in st],
        "fp64": fp,
        "energy_range_s": round(S["t_range"], 3),
        "roofline": roofline,
        "kernels_ms": kernels,
        "filter_windows": S.get("windows"),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
