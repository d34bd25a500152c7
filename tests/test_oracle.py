"""Pins the CPU oracle against the reference's own golden data and against what
the mathematics fixes (closed forms, invariants, brute force), so that a
dropped term, wrong sign or index in the restatement fails here."""
import math
import os

import numpy as np
import pytest

from oracle import inputs as I
from oracle import oracle as O
from systems import VARIANTS, oracle_ham, random_states, small_system

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CSPBI3 = os.path.join(GOLD, "cspbi3")


# ---------------------------------------------------------------- ran2 ----
def ran2_direct(seed, n):
    """ran2 written with exact modular products (no Schrage decomposition),
    i.e. the published L'Ecuyer generators s1 <- 40014 s1 mod 2147483563,
    s2 <- 40692 s2 mod 2147483399 with a 32-entry Bays-Durham shuffle."""
    m1, m2 = 2147483563, 2147483399
    s = max(1, -seed)
    s2 = s
    tab = [0] * 32
    for j in range(39, -1, -1):
        s = (40014 * s) % m1
        if j < 32:
            tab[j] = s
    y = tab[0]
    out = []
    for _ in range(n):
        s = (40014 * s) % m1
        s2 = (40692 * s2) % m2
        j = y // (1 + (m1 - 1) // 32)
        y = tab[j] - s2
        tab[j] = s
        if y < 1:
            y += m1 - 1
        v = (1.0 / m1) * y
        out.append(min(v, 1.0 - 1.2e-7))
    return np.array(out)


@pytest.mark.parametrize("seed", [-1, -874917403, -12345, 0])
def test_ran2_matches_exact_modular_restatement(seed):
    got, _ = O.ran_sequence(seed, 2000)
    assert np.array_equal(got, ran2_direct(seed, 2000))


def test_ran2_statistics():
    got, _ = O.ran_sequence(-7, 200000)
    assert got.min() > 0.0 and got.max() < 1.0
    assert abs(got.mean() - 0.5) < 3e-3
    assert abs(got.var() - 1.0 / 12.0) < 2e-3


def test_start_states_normalised_per_spin():
    psi, _ = O.init_filter_states(2, 3, -5, 6, 7, 8, 2, 1, 0.3)
    ng = 6 * 7 * 8
    stlen = 2 * 2 * ng
    for jns in range(2):
        blk = psi[jns * 3 * stlen: jns * 3 * stlen + stlen]
        for sp in range(2):
            comp = blk[sp * 2 * ng:(sp + 1) * 2 * ng]
            assert abs((comp ** 2).sum() * 0.3 - 1.0) < 1e-12
        # the other ms-1 slots stay untouched
        assert not psi[jns * 3 * stlen + stlen: (jns + 1) * 3 * stlen].any()


# ---------------------------------------------------------------- grid ----
def test_cspbi3_grid_matches_reference_run():
    """run.dat of the reference CsPbI3 job: 44^3, xmin = -11, dx = 0.5,
    n_NL_gridpts = 2325, Vmin/Vmax of the local potential (run.dat:95-107,131)."""
    job = I.reference_job(CSPBI3)
    assert (job["nx"], job["ny"], job["nz"]) == (44, 44, 44)
    assert np.allclose(job["mins"], -11.0, atol=0, rtol=0)
    assert job["proj"]["n_gridpts"] == 2325
    assert f"{job['pot_local'].min():.6g}" == "-1.28863"
    assert f"{job['pot_local'].max():.6g}" == "0.750581"
    conf = np.loadtxt(os.path.join(CSPBI3, "conf.dat"), skiprows=1, usecols=(1, 2, 3))
    assert np.abs(conf - job["R"]).max() < 1e-14


def test_grid_auto_growth_rule():
    R = np.array([[0.0, 0.0, 0.0], [10.3, -4.1, 7.7]])
    n, mins, maxs, dk = O.init_grid_params(R, (8, 8, 8), (0.5, 0.5, 0.5))
    for ax, ext in enumerate((10.3, 4.1, 7.7)):
        half = round(0.5 * ext + 5.0)
        need = int(((half - 0.5) + half) / 0.5)
        expect = need + (need % 2) if need > 8 else 8
        assert n[ax] == expect and n[ax] % 2 == 0
        assert mins[ax] == -n[ax] * 0.5 / 2 and abs(dk[ax] - 2 * math.pi / (n[ax] * 0.5)) < 1e-15


@pytest.mark.parametrize("n", [8, 9, 44])
def test_kvec_closed_form(n):
    dk = 0.37
    k = O.init_kvec(n, dk)
    expect = np.array([j if j <= n // 2 else j - n for j in range(n)], dtype=float) * dk
    if n % 2 == 0:
        expect[n // 2] = -(n // 2) * dk  # k[n-j] is written after k[j] (init.c:1388-1393)
    assert np.allclose(k, expect, rtol=0, atol=1e-15)


def test_ksqr_closed_form_and_cap():
    n = np.array([10, 12, 14])
    d = np.array([0.7, 0.6, 0.5])
    box = n * d
    dk = 2 * np.pi / box
    ksqr, gx, gy, gz = O.build_grid_ksqr(n, d, -box / 2, dk, ke_max=3.0)
    N = n.prod()
    kk = [np.fft.fftfreq(int(m), 1.0 / int(m)) * dki for m, dki in zip(n, dk)]
    KZ, KY, KX = np.meshgrid(kk[2], kk[1], kk[0], indexing="ij")
    expect = np.minimum(0.5 * (KX ** 2 + KY ** 2 + KZ ** 2), 3.0) / N
    assert np.allclose(ksqr.reshape(n[2], n[1], n[0]), expect, rtol=1e-13, atol=0)
    assert abs(gx[0] + box[0] / 2) < 1e-15 and abs(gx[1] - gx[0] - d[0]) < 1e-15


# -------------------------------------------------------- Newton / Leja ----
def test_leja_points_match_reference_zn_dat():
    """zn.dat written by the reference's gen_newton_coeff (coeff.c:66-71) for
    nCheby = 2048, printed with %lg: every node agrees to the printed digits,
    which pins the greedy argmax sequence of samp_points_ashkenazy."""
    ref = np.loadtxt(os.path.join(GOLD, "zn_M2048_2.2nm_CdSe.dat"))
    z = O.leja_points(2048)
    assert np.array_equal(ref[:, 0], np.arange(2048))
    assert np.all(np.abs(z - ref[:, 1]) <= 5e-6 * np.abs(ref[:, 1]) + 1e-12)


def test_leja_properties():
    z = O.leja_points(256)
    assert z[0] == -2.0 and z[1] == 2.0 and z[2] == 0.0
    assert len(np.unique(z)) == 256 and np.all(np.abs(z) <= 2.0)


def newton_eval(a, z, x):
    """Newton form p(x) = a0 + sum_j a_j prod_{i<j}(x - z_i)."""
    p = np.full_like(x, a[0])
    w = np.ones_like(x)
    for j in range(1, len(a)):
        w = w * (x - z[j - 1])
        p = p + a[j] * w
    return p


def test_newton_coefficients_interpolate_at_nodes():
    M = 48
    z = O.leja_points(M)
    Vmin, dE = -1.0, 6.0
    targets = np.array([0.2, 1.1])
    dt = (M / (2.5 * dE)) ** 2
    an = O.newton_coeff(z, targets, Vmin, dE, dt).reshape(2, M, 2)
    assert not an[..., 1].any()
    for t in range(2):
        x = (z + 2.0) * dE / 4.0 + Vmin
        f = math.sqrt(dt / math.pi) * np.exp(-dt * (x - targets[t]) ** 2)
        p = newton_eval(an[t, :, 0], z, z)
        assert np.abs(p - f).max() < 1e-9 * np.abs(f).max()


def test_newton_filter_reproduces_gaussian():
    """check_function (coeff.c:297-325): the series approximates the Gaussian."""
    M = 512
    z = O.leja_points(M)
    Vmin, dE = -1.0, 26.0
    t = np.array([0.3])
    dt = (M / (2.5 * dE)) ** 2
    a = O.newton_coeff(z, t, Vmin, dE, dt)[0::2]
    xs = np.linspace(-2, 2, 801)
    f = math.sqrt(dt / math.pi) * np.exp(-dt * ((xs + 2) * dE / 4 + Vmin - 0.3) ** 2)
    assert np.abs(newton_eval(a, z, xs) - f).max() < 1e-6 * f.max()


def test_targets_match_reference_run():
    t = O.ene_targets(10, 6, -0.5, -0.2, -0.167, -0.1)
    listed = [-0.47, -0.44, -0.41, -0.38, -0.35, -0.32, -0.29, -0.26, -0.23, -0.2, -0.167, -0.155833, -0.144667,
              -0.1335, -0.122333, -0.111167]  # run.dat:113-128
    assert [float(f"{v:g}") for v in t] == listed


# ------------------------------------------------------------------ FFT ----
@pytest.mark.parametrize("shape", [(6, 10, 12), (7, 5, 11), (4, 4, 44)])
def test_fft_matches_numpy(shape):
    rng = np.random.default_rng(0)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert np.abs(O.fft3d(a, -1) - np.fft.fftn(a)).max() < 1e-11 * np.abs(a).sum() ** 0.5
    assert np.abs(O.fft3d(a, +1) - np.fft.ifftn(a) * a.size).max() < 1e-11 * np.abs(a).sum() ** 0.5


# ----------------------------------------------------------- Hamiltonian ----
def test_kinetic_on_plane_waves():
    s = small_system(n=(12, 10, 8), dx=0.6, nspin=1, is_complex=1)
    s["pot_local"] = np.zeros_like(s["pot_local"])
    h = oracle_ham(s)
    nx, ny, nz = 12, 10, 8
    x = np.arange(nx) * 0.6
    y = np.arange(ny) * 0.6
    zz = np.arange(nz) * 0.6
    for (mx, my, mz) in [(1, 0, 0), (2, -3, 1), (0, 0, 4)]:
        k = np.array([mx * 2 * np.pi / (nx * 0.6), my * 2 * np.pi / (ny * 0.6), mz * 2 * np.pi / (nz * 0.6)])
        Z, Y, X = np.meshgrid(zz, y, x, indexing="ij")
        psi = np.exp(1j * (k[0] * X + k[1] * Y + k[2] * Z)).ravel()
        out = h.apply(psi.view(np.float64)).view(np.complex128)
        assert np.abs(out - 0.5 * (k @ k) * psi).max() < 1e-12


def test_constant_potential_shift_and_hermiticity():
    s = small_system(**VARIANTS["spinor_so_nl"])
    h = oracle_ham(s)
    psi = random_states(s, 2, seed=3)
    a, b = psi[: psi.size // 2], psi[psi.size // 2:]
    ha, hb = h.apply(a), h.apply(b)
    ca, cb = a.view(np.complex128), b.view(np.complex128)
    lhs = np.vdot(ca, hb.view(np.complex128))
    rhs = np.conj(np.vdot(cb, ha.view(np.complex128)))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    s2 = dict(s)
    s2["pot_local"] = s["pot_local"] + 0.75
    d = oracle_ham(s2).apply(a) - ha
    assert np.abs(d - 0.75 * a).max() < 1e-12


def test_ls_spectrum():
    ls = O.def_LS().reshape(36, 2)
    L = (ls[:, 0] + 1j * ls[:, 1]).reshape(6, 6)  # [jp, j]
    assert np.allclose(L, L.conj().T)
    w = np.linalg.eigvalsh(L)
    assert np.allclose(np.sort(w), [-1, -1, 0.5, 0.5, 0.5, 0.5], atol=1e-14)


@pytest.mark.parametrize("variant", ["real_nl", "spinor_so_nl"])
def test_projector_operator_matches_dense_definition(variant):
    """V_SO + V_NL equals sum over atoms of |beta> M <beta| written densely."""
    s = small_system(n=(8, 8, 8), dx=0.8, natoms=2, **VARIANTS[variant])
    s0 = dict(s)
    s0["proj"] = None
    h, h0 = oracle_ham(s), oracle_ham(s0)
    psi = random_states(s, 1, seed=4)
    vproj = h.apply(psi) - h0.apply(psi)
    p = s["proj"]
    ng = 512
    cplx = s["is_complex"]
    x = psi.view(np.complex128) if cplx else psi.astype(np.complex128)
    out = np.zeros(s["nspin"] * ng, dtype=np.complex128)
    ls = O.def_LS().reshape(36, 2)
    LS = (ls[:, 0] + 1j * ls[:, 1]).reshape(6, 6)
    for a in range(p["n_atoms"]):
        idx = a * p["n_gridpts"] + np.arange(p["n_gpt"][a])
        j = p["jxyz"][idx]
        Y = p["y1"].reshape(-1, 3, 2)[idx]
        Y = Y[..., 0] + 1j * Y[..., 1]
        for q in range(5):
            if p["so"] and q < p["n_so_proj"]:
                beta = p["proj"].reshape(-1, 5)[idx, q][:, None] * Y  # (g, m)
                c = np.array([[s["dv"] * np.vdot(beta[:, m], x[sp * ng + j]) for m in range(3)] for sp in range(2)])
                cflat = c.reshape(6)
                d = LS.T @ cflat  # d_j = sum_jp LS[jp, j] c_jp
                for jj in range(6):
                    sp, m = divmod(jj, 3)
                    np.add.at(out, sp * ng + j, beta[:, m] * d[jj])
            if p["nl"] and q < p["n_nl_proj"]:
                beta = p["nl_proj"].reshape(-1, 5)[idx, q][:, None] * Y
                sg = p["nl_sign"][a * 5 + q]
                for sp in range(s["nspin"]):
                    for m in range(3):
                        c = sg * s["dv"] * np.vdot(beta[:, m], x[sp * ng + j])
                        np.add.at(out, sp * ng + j, beta[:, m] * c)
    got = vproj.view(np.complex128) if cplx else vproj
    want = out if cplx else out.real
    assert np.abs(got - want).max() < 1e-12 * max(np.abs(want).max(), 1e-300)


def dense_h(s):
    h = oracle_ham(s)
    stlen = (s["is_complex"] + 1) * s["nspin"] * s["nx"] * s["ny"] * s["nz"]
    n = stlen // (s["is_complex"] + 1)
    Hd = np.zeros((n, n), dtype=np.complex128 if s["is_complex"] else np.float64)
    for i in range(n):
        e = np.zeros(n, dtype=Hd.dtype)
        e[i] = 1
        v = e.view(np.float64) if s["is_complex"] else e
        r = h.apply(v)
        Hd[:, i] = r.view(np.complex128) if s["is_complex"] else r
    return h, Hd


def test_filter_cycle_is_the_newton_polynomial_of_h():
    """On an eigenvector psi of H (dense brute force on a 6^3 grid), the filter
    returns p_t(H_s(lambda)) psi with p_t the Newton polynomial (filter.c:187-371
    with p_hnorm filter.c:375-410)."""
    s = small_system(n=(6, 6, 6), dx=0.9, natoms=2, nspin=1, is_complex=0, nl=1)
    h, Hd = dense_h(s)
    assert np.abs(Hd - Hd.T).max() < 1e-10 * np.abs(Hd).max()
    w, V = np.linalg.eigh(0.5 * (Hd + Hd.T))
    Vmin, dE = w[0] - 1.0, (w[-1] - w[0]) + 2.0
    M, ms = 32, 2
    z = O.leja_points(M)
    targets = np.array([w[3], w[10]])
    dt = (M / (2.5 * dE)) ** 2
    an = O.newton_coeff(z, targets, Vmin, dE, dt)
    for k in (0, 3, 50):
        psi_rank = np.zeros(ms * 216)
        psi_rank[:216] = V[:, k]
        out = h.filter_cycles(psi_rank, 1, ms, an, z, Vmin, 4.0 / dE)
        xs = 4.0 * (w[k] - Vmin) / dE - 2.0
        for t in range(ms):
            expect = newton_eval(an[2 * M * t: 2 * M * (t + 1): 2], z, np.array([xs]))[0]
            assert np.abs(out[t * 216:(t + 1) * 216] - expect * V[:, k]).max() < 1e-9 * max(abs(expect), 1e-3)


def test_sigma_and_energy_on_eigenstates():
    s = small_system(n=(6, 6, 6), dx=0.9, natoms=2, nspin=1, is_complex=0)
    h, Hd = dense_h(s)
    w, V = np.linalg.eigh(0.5 * (Hd + Hd.T))
    psi = (V[:, :3].T / math.sqrt(s["dv"])).ravel()
    sig, ev = h.sigma(psi)
    assert np.abs(ev - w[:3]).max() < 1e-10 and sig.max() < 1e-6
    assert np.abs(h.energies(psi) - w[:3]).max() < 1e-10
    # a mixed state: sigma = ||H psi - E psi||
    mix = (V[:, 0] + V[:, 1]) / math.sqrt(2 * s["dv"])
    sg, e = h.sigma(mix)
    assert abs(e[0] - 0.5 * (w[0] + w[1])) < 1e-10
    assert abs(sg[0] - 0.5 * abs(w[1] - w[0])) < 1e-6


def test_normalize_all_error_codes():
    psi = np.zeros(2 * 64)
    psi[:64] = 1.0
    assert O.lib().or_normalize_all(O._dp(psi), 2, 64, 0, 1.0) == -1
    psi[64:] = np.nan
    psi[:64] = 1.0
    assert O.lib().or_normalize_all(O._dp(psi), 2, 64, 0, 1.0) == -2


def test_energy_range_brackets_spectrum():
    s = small_system(n=(6, 6, 6), dx=0.9, natoms=2, nspin=1, is_complex=0)
    h, Hd = dense_h(s)
    w = np.linalg.eigvalsh(0.5 * (Hd + Hd.T))
    out = h.energy_range(max_iter=1500)
    assert out[0] == pytest.approx(w[0], rel=1e-4) and out[1] == pytest.approx(w[-1], rel=1e-3)
    assert out[2] < w[0] and out[2] + out[3] > w[-1]


def test_svd_rank_cut_and_rayleigh_ritz():
    rng = np.random.default_rng(1)
    base = rng.standard_normal((5, 300))
    mix = rng.standard_normal((9, 5))
    psi = (mix @ base).ravel()
    U, sv, k = O.ortho(psi, 9, 300, 0, 0.5)
    assert k == 5 and sv[5] / sv[0] < 1e-10
    assert np.allclose(0.5 * U.T @ U, np.eye(5), atol=1e-12)
