"""Batched uneven dots (fbse_filter_many): several contexts of different grid
sizes filtered in one call, each on its own stream with interleaved launches
and graph replays.  With a common window depth the result of every dot is
bit-identical to a separate fbse_filter call on that dot, and the FP64 result
matches the oracle's Newton filter."""
import numpy as np
import pytest

from systems import gpu_context, oracle_ham, small_system

pytestmark = pytest.mark.gpu

SHAPES = [((16, 18, 20), dict(nspin=2, is_complex=1, so=1, nl=1)),
          ((20, 16, 24), dict(nspin=2, is_complex=1)),
          ((24, 20, 16), dict(nspin=2, is_complex=1, so=1, nl=1, seed=3)),
          ((18, 18, 18), dict(nspin=2, is_complex=1, nl=1, seed=5))]


def inputs(s, n_starts, ms, M, seed):
    from oracle import oracle as O
    psi_rank, _ = O.init_filter_states(n_starts, ms, -seed, s["nx"], s["ny"], s["nz"], s["nspin"], s["is_complex"],
                                       s["dv"])
    Vmin = s["pot_local"].min() - 3.0
    dE = s["pot_local"].max() + 25.0 - Vmin
    zn = O.leja_points(M)
    targets = np.linspace(s["pot_local"].min() + 0.5, s["pot_local"].min() + 2.0, ms)
    an = O.newton_coeff(zn, targets, Vmin, dE, (M / (2.5 * dE)) ** 2)
    return psi_rank, an, zn, Vmin, 4.0 / dE


@pytest.fixture(scope="module")
def dots():
    out = []
    for k, (n, kw) in enumerate(SHAPES):
        s = small_system(n=n, **kw)
        out.append((s, gpu_context(s)))
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_filter_many_matches_single(dots, precision):
    import torch
    from filterbse_amd import api
    ms, M = 3, 42
    n_starts = [2, 3, 1, 2]
    per = [inputs(s, n_starts[k], ms, M, 7 + k) for k, (s, g) in enumerate(dots)]
    ref = []
    for (s, g), (psi, an, zn, Vmin, dE_1), ns in zip(dots, per, n_starts):
        g.set_window(3)
        ref.append(g.filter(psi, ns, ms, an, zn, Vmin, dE_1, precision=precision))
    dev = torch.device("cuda:0")
    d = [torch.from_numpy(p[0]).to(dev) for p in per]
    api.filter_many([g for s, g in dots], [x.data_ptr() for x in d], n_starts, ms, [p[1] for p in per],
                    [p[2] for p in per], [p[3] for p in per], [p[4] for p in per], precision)
    torch.cuda.synchronize()
    for k, (s, g) in enumerate(dots):
        assert g.graph_replays() == 9, k
        out = d[k].cpu().numpy()
        assert np.array_equal(out, ref[k]), (k, np.abs(out - ref[k]).max())
    if precision == 0:
        s, g = dots[0]
        psi, an, zn, Vmin, dE_1 = per[0]
        o = oracle_ham(s).filter_cycles(psi, n_starts[0], ms, an, zn, Vmin, dE_1)
        assert np.abs(ref[0] - o).max() < 1e-10 * np.abs(o).max()
    for s, g in dots:
        g.set_window(0)


def test_filter_many_rejects_shared_context(dots):
    import torch
    from filterbse_amd import api
    s, g = dots[0]
    psi, an, zn, Vmin, dE_1 = inputs(s, 1, 2, 8, 3)
    d = torch.from_numpy(psi).to("cuda:0")
    with pytest.raises(api.FbseError) as ei:
        api.filter_many([g, g], [d.data_ptr(), d.data_ptr()], [1, 1], 2, [an, an], [zn, zn], [Vmin, Vmin],
                        [dE_1, dE_1])
    assert ei.value.code == api.FBSE_ERR_ARG


def test_filter_many_grouped_launches():
    """Mixed precision, dots of one size on the fixed FP32 plane instance
    (three 60^3 dots and two 64^3 dots, spinors with SO + NL projectors) share
    grouped launches (one launch per FFT pass for the whole group), a 20^3
    dot runs on its own: every dot's result is bit-identical to its separate
    filter on the same instance and window, graph replays included."""
    import torch
    from filterbse_amd import api
    spec = [(60, 1, 2), (60, 2, 1), (64, 3, 2), (60, 4, 3), (64, 5, 1), (20, 6, 2)]
    dots_ = []
    for n, seed, ns in spec:
        s = small_system(n=(n, n, n), dx=0.5, nspin=2, is_complex=1, so=1, nl=1, natoms=6, n_so_proj=1,
                         n_nl_proj=1, seed=seed)
        g = gpu_context(s)
        if n != 20:
            g.set_fft_kernel(1, 4)  # the fixed FP32 instance
        g.set_window(3)
        dots_.append((s, g, ns))
    ms, M = 3, 42
    per = [inputs(s, ns, ms, M, 11 + k) for k, (s, g, ns) in enumerate(dots_)]
    ref = [g.filter(p[0], ns, ms, p[1], p[2], p[3], p[4], precision=1) for (s, g, ns), p in zip(dots_, per)]
    d = [torch.from_numpy(p[0]).to("cuda:0") for p in per]
    api.filter_many([g for s, g, ns in dots_], [x.data_ptr() for x in d], [ns for s, g, ns in dots_], ms,
                    [p[1] for p in per], [p[2] for p in per], [p[3] for p in per], [p[4] for p in per], 1)
    torch.cuda.synchronize()
    groups = [g.last_group() for s, g, ns in dots_]
    print("group sizes", groups)
    assert groups == [3, 3, 2, 3, 2, 1]
    for k, (s, g, ns) in enumerate(dots_):
        assert g.graph_replays() == 9, k
        out = d[k].cpu().numpy()
        assert np.array_equal(out, ref[k]), (k, np.abs(out - ref[k]).max())
    for s, g, ns in dots_:
        g.close()
