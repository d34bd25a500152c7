"""GPU parity: every C-ABI entry point against the CPU oracle on the same
seeded inputs (small grids so the oracle runs in seconds).  Tolerances:
H, filter, normalisation, energies and sigma in FP64 are limited only by FFT
rounding (the oracle's FFT is an independent DFT code), so 1e-11 relative;
the mixed-precision filter is held to 3e-5 relative to the FP64 result."""
import numpy as np
import pytest

from oracle import oracle as O
from systems import VARIANTS, gpu_context, oracle_ham, random_states, small_system

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module", params=sorted(VARIANTS))
def system(request):
    s = small_system(**VARIANTS[request.param])
    print("case", request.param, flush=True)
    return request.param, s, oracle_ham(s), gpu_context(s)


@pytest.mark.parametrize("variant,lines", [(0, 0), (0, 4), (1, 0), (1, 1), (1, 2), (-1, 0)])
def test_apply_h(system, variant, lines):
    """Per-axis passes at several tile heights and every FP64 plane/slab
    kernel instance give the same H."""
    name, s, h, g = system
    psi = random_states(s, 3, seed=11)
    ref = h.apply(psi)
    g.set_fft_kernel(variant, lines)
    out = g.apply_h(psi)
    g.set_fft_kernel(-1, 0)
    assert rel(out, ref) < 1e-11, name


def spectrum_bounds(s):
    vmin = s["pot_local"].min() - 3.0
    vmax = s["pot_local"].max() + 20.0 + 5.0
    return vmin, vmax - vmin


def filter_inputs(s, n_starts, ms, M):
    stlen = (s["is_complex"] + 1) * s["nspin"] * s["nx"] * s["ny"] * s["nz"]
    psi_rank, _ = O.init_filter_states(n_starts, ms, -3, s["nx"], s["ny"], s["nz"], s["nspin"], s["is_complex"],
                                       s["dv"])
    Vmin, dE = spectrum_bounds(s)
    zn = O.leja_points(M)
    targets = np.linspace(s["pot_local"].min() + 0.5, s["pot_local"].min() + 2.0, ms)
    dt = (M / (2.5 * dE)) ** 2
    an = O.newton_coeff(zn, targets, Vmin, dE, dt)
    return psi_rank, stlen, an, zn, Vmin, 4.0 / dE


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 3e-5)])
def test_filter_more_than_16_targets(precision, tol):
    """ms = 21 energy targets per start (the accumulation takes them in
    launches of at most 16): FP64 and mixed filters against the oracle."""
    s = small_system(**VARIANTS["spinor_so_nl_1p"])
    h, g = oracle_ham(s), gpu_context(s)
    n_starts, ms, M = 2, 21, 30
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, n_starts, ms, M)
    ref = h.filter_cycles(psi_rank, n_starts, ms, an, zn, Vmin, dE_1)
    g.set_window(4)
    out = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    g.close()
    assert rel(out, ref) < tol


@pytest.mark.parametrize("window", [3, 0])
def test_filter_fp64(system, window):
    name, s, h, g = system
    n_starts, ms, M = 3, 4, 40
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, n_starts, ms, M)
    ref = h.filter_cycles(psi_rank, n_starts, ms, an, zn, Vmin, dE_1)
    g.set_window(window)
    out = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1)
    g.set_window(0)
    assert rel(out, ref) < 1e-10, name


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 3e-5)])
def test_filter_complex_coefficients(system, precision, tol):
    """Newton coefficients with nonzero imaginary parts (the general path of
    filter.c:331-357: complex states accumulate a*psi, real states only
    Re(a)*psi): the accumulation's complex-coefficient kernel against the
    oracle, FP64 to rounding and mixed to FP32 rounding; window 3 so that
    several accumulations run."""
    name, s, h, g = system
    n_starts, ms, M = 2, 3, 30
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, n_starts, ms, M)
    an = an.copy()
    an[1::2] = 0.4 * an[0::2] * np.cos(np.arange(ms * M))
    ref = h.filter_cycles(psi_rank, n_starts, ms, an, zn, Vmin, dE_1)
    g.set_window(3)
    out = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    g.set_window(0)
    assert rel(out, ref) < tol, name


@pytest.mark.parametrize("variant,lines", [(-1, 0), (0, 0), (1, 1), (1, 2), (1, 3)])
def test_filter_mixed(system, variant, lines):
    """The FP32 recurrence with FP64 accumulation stays within 3e-5 of the FP64
    filter, through the per-axis passes and every FP32 plane/slab instance."""
    name, s, h, g = system
    n_starts, ms, M = 2, 3, 40
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, n_starts, ms, M)
    ref = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1)
    g.set_fft_kernel(variant, lines)
    out = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=1)
    g.set_fft_kernel(-1, 0)
    assert rel(out, ref) < 3e-5, name


@pytest.mark.parametrize("precision", [0, 1])
def test_filter_sub_batches(system, precision):
    """A filter whose start states run as consecutive sub-batches (fbse_set_run_starts:
    2 starts per run over 5 starts, real states in packed pairs) gives the same
    bits as one batch at the same window depth, against the oracle as well."""
    name, s, h, g = system
    n_starts, ms, M = 5, 3, 30
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, n_starts, ms, M)
    g.set_window(4)
    one = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    assert g.last_runs() == 1
    g.set_run_starts(2)
    sub = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    assert g.last_runs() == 3
    g.set_run_starts(0)
    g.set_window(0)
    assert np.array_equal(sub, one), (name, rel(sub, one))
    if precision == 0:
        ref = h.filter_cycles(psi_rank, n_starts, ms, an, zn, Vmin, dE_1)
        assert rel(sub, ref) < 1e-10, name


@pytest.mark.parametrize("precision", [0, 1])
def test_filter_graph_replay(system, precision):
    """Full accumulation windows replayed from one captured hipGraph give the
    same bits as launching every kernel from the host; the graph path is run
    twice with eager work and synchronisations in between (window 3 -> ring of
    4 slots, M = 42 -> 9 replays per filter)."""
    name, s, h, g = system
    n_starts, ms, M = 2, 3, 42
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, n_starts, ms, M)
    g.set_window(3)
    g.set_graph(0)
    ref = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    assert g.graph_replays() == 0
    g.set_graph(1)
    out1 = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    assert g.graph_replays() == 9
    eager = g.apply_h(random_states(s, 2, seed=3))  # unrelated eager work between replays
    out2 = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    # profiling keeps every launch on the main stream (no side-stream projections)
    g.profile(True)
    out3 = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    g.profile(False)
    g.set_window(0)
    assert np.array_equal(out3, ref), (name, rel(out3, ref))
    assert np.isfinite(eager).all()
    assert np.array_equal(out1, ref), (name, rel(out1, ref))
    assert np.array_equal(out2, ref), (name, rel(out2, ref))
    if precision == 0:
        assert rel(ref, h.filter_cycles(psi_rank, n_starts, ms, an, zn, Vmin, dE_1)) < 1e-10


def test_normalize_energy_sigma(system):
    name, s, h, g = system
    psi = random_states(s, 3, seed=5)
    ref, rc = h.normalize_all(psi)
    assert rc == 0
    out = g.normalize_all(psi)
    assert rel(out, ref) < 1e-13
    e_ref = h.energies(ref)
    e = g.energies(ref)
    assert np.abs(e - e_ref).max() < 1e-11 * np.abs(e_ref).max()
    s_ref, ev_ref = h.sigma(ref)
    sg, ev = g.sigma(ref, 3)
    assert np.abs(ev - ev_ref).max() < 1e-11 * np.abs(ev_ref).max()
    assert np.abs(sg - s_ref).max() < 1e-9 * np.abs(s_ref).max()


def test_normalize_zero_norm_is_an_error(system):
    name, s, h, g = system
    from filterbse_amd import api
    psi = random_states(s, 2, seed=5)
    psi[psi.size // 2:] = 0.0
    with pytest.raises(api.FbseError) as ei:
        g.normalize_all(psi)
    assert ei.value.code == -3


def test_ortho_and_rayleigh_ritz(system):
    name, s, h, g = system
    stlen = (s["is_complex"] + 1) * s["nspin"] * s["nx"] * s["ny"] * s["nz"]
    rng = np.random.default_rng(2)
    # 12 states spanning an 8-dimensional space: one direction at ~1e-8 of the
    # largest singular value (kept by the 1e-10 cut) and one at ~1e-13 (cut)
    base = random_states(s, 8, seed=9).reshape(8, stlen)
    mix = rng.standard_normal((12, 8))
    mix[:, 6] *= 1e-8
    mix[:, 7] *= 1e-13
    psi = (mix @ base).ravel()
    U_ref, sv_ref, k_ref = O.ortho(psi, 12, stlen, s["is_complex"], s["dv"])
    out, k, sv = g.ortho(psi, 12)
    assert k == k_ref == 7, (k, k_ref, sv, sv_ref)
    assert np.allclose(sv[:6], sv_ref[:6], rtol=1e-9, atol=0)
    assert np.allclose(sv[6], sv_ref[6], rtol=1e-5, atol=0)
    assert sv[7] / sv[0] < 1e-10 and sv_ref[7] / sv_ref[0] < 1e-10
    U = O.as_matrix(out, k, stlen, s["is_complex"])
    # orthonormal under the dv weight and spanning the same space
    gram = s["dv"] * (U.conj().T @ U)
    assert np.abs(gram - np.eye(k)).max() < 1e-10
    P_ref = U_ref @ U_ref.conj().T * s["dv"]
    P = U @ U.conj().T * s["dv"]
    assert np.abs(P - P_ref).max() < 1e-6
    # Rayleigh-Ritz in that basis
    HU = O.as_matrix(h.apply(out[: k * stlen]), k, stlen, s["is_complex"])
    w_ref, _ = O.rayleigh_ritz(U, HU, s["dv"])
    rot, w = g.rayleigh_ritz(out[: k * stlen], k)
    assert np.abs(w - w_ref).max() < 1e-10 * np.abs(w_ref).max()
    # the rotated vectors are Ritz vectors: <psi_i|H|psi_j> dv = diag(w)
    V = O.as_matrix(rot, k, stlen, s["is_complex"])
    HV = O.as_matrix(h.apply(rot), k, stlen, s["is_complex"])
    Hs = s["dv"] * (V.conj().T @ HV)
    assert np.abs(Hs - np.diag(w)).max() < 1e-9 * np.abs(w).max()


def test_energy_range(system):
    name, s, h, g = system
    from filterbse_amd import api
    ref = h.energy_range(max_iter=60)
    p0, p1 = api.energy_range_start_states(s["nx"], s["ny"], s["nz"], s["nspin"], s["is_complex"], s["dv"])
    out = g.energy_range(p0, p1, max_iter=60)
    assert out[4] == ref[4] and out[5] == ref[5]
    assert np.abs(out[:4] - ref[:4]).max() < 1e-9 * np.abs(ref[:4]).max()


def test_refine_guarded():
    """fbse_refine after a mixed filter: every Ritz pair stays in its slot,
    none gets a larger residual, the best residuals drop (the FP32 noise of
    the recurrence is what the sweep removes), the returned residuals are the
    vectors' own (recomputed by fbse_sigma), and a second sweep keeps going."""
    import torch
    s = small_system(**VARIANTS["spinor_so_nl_1p"])
    g = gpu_context(s)
    n_starts, ms, M = 4, 6, 600
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, n_starts, ms, M)
    n = n_starts * ms
    d = torch.from_numpy(psi_rank.copy()).cuda()
    p = d.data_ptr()
    g.filter_dev(p, n_starts, ms, an, zn, Vmin, dE_1, 1)
    g.normalize_all_dev(p, n)
    rank, _ = g.ortho_dev(p, n)
    ev = g.rayleigh_ritz_dev(p, rank)
    sig, _ = g.sigma_dev(p, rank)
    lo, hi = ev.min() - 1.0, np.sort(ev)[rank // 2]
    sel = (ev >= lo) & (ev <= hi) & (sig < 1.0)
    ev1, sig1, nrep = g.refine_dev(p, rank, ev, sig, lo, hi, e_p=0.25, sigma_max=1.0)
    assert len(ev1) == rank and nrep > 0
    assert np.isfinite(ev1).all() and np.isfinite(sig1).all()
    assert (sig1 <= sig).all()
    assert np.array_equal(sig1[~sel], sig[~sel]) and np.array_equal(ev1[~sel], ev[~sel])
    assert sig1.min() < 0.5 * sig.min()
    chk, eck = g.sigma_dev(p, rank)
    assert np.abs(chk - sig1).max() < 1e-9 + 1e-6 * sig1.max()
    assert np.abs(eck - ev1).max() < 1e-9
    ev2, sig2, _ = g.refine_dev(p, rank, ev1, sig1, lo, hi, e_p=0.25, sigma_max=1.0)
    assert (sig2 <= sig1).all()
    print("best sigma: %.2e -> %.2e -> %.2e, replaced %d" % (sig.min(), sig1.min(), sig2.min(), nrep))
    g.close()
