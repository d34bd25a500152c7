"""Small Hamiltonians for parity tests, built with the oracle's restatement of
the reference's input code (grid: init.c:259-349, V_loc: init.c:481-703,
projector tables: init.c:848-1068).  Sizes are chosen so the oracle finishes
in seconds."""
import math

import numpy as np

from oracle import oracle as O

PROJ_LEN = 1024


def _pot_tables(ntypes, rng):
    """Per-type tabulated V(r) on a uniform grid (the pot*.par layout, read.c:1801-1833)."""
    nlen = 600
    dr = 0.02
    r = np.arange(nlen) * dr
    tabs = []
    for t in range(ntypes):
        a = -1.2 - 0.6 * t + 0.1 * rng.standard_normal()
        v = a * np.exp(-r ** 2 / (1.4 + 0.3 * t)) + 0.35 * np.exp(-(r - 1.8) ** 2 / 0.5)
        tabs.append(v)
    pot_r = np.tile(r, ntypes)
    pot_v = np.concatenate(tabs)
    return pot_r, pot_v, np.full(ntypes, nlen, dtype=np.int64), np.full(ntypes, dr), nlen


def small_system(n=(16, 18, 20), dx=0.55, nspin=1, is_complex=0, so=0, nl=0, natoms=4, seed=0, n_so_proj=5,
                 n_nl_proj=2):
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    d = np.array([dx, dx, dx])
    n_arr = np.array(n)
    box = np.where(n_arr % 2 == 0, n_arr * dx, (n_arr - 1) * dx)
    mins = -box / 2.0
    dk = 2.0 * math.pi / box
    ksqr, gx, gy, gz = O.build_grid_ksqr(n_arr, d, mins, dk, ke_max=20.0)
    R = rng.uniform(-1.5, 1.5, size=(natoms, 3))
    atype = (np.arange(natoms) % 2).astype(np.int32)
    pot_r, pot_v, lens, pdr, mlen = _pot_tables(2, rng)
    pot = np.zeros(nspin * nx * ny * nz)
    O.lib().or_build_local_pot(O._dp(pot), nspin, O._dp(gx), O._dp(gy), O._dp(gz), nx, ny, nz,
                               O._dp(np.ascontiguousarray(R)), O._ip(atype), natoms, O._dp(pot_r), O._dp(pot_v),
                               O._lp(lens), O._dp(pdr), mlen)
    proj = None
    if so or nl:
        R2 = 1.5 + 6.0 * math.log(10.0) + 3.0 * dx  # read.c:908
        stride = O.lib().or_count_nl_gridpts(O._dp(np.ascontiguousarray(R)), natoms, O._dp(gx), O._dp(gy),
                                            O._dp(gz), nx, ny, nz, R2)
        vr = np.arange(PROJ_LEN) * (math.sqrt(R2) / PROJ_LEN)
        so_tab = np.zeros(5 * PROJ_LEN)
        for p in range(5):
            so_tab[p * PROJ_LEN:(p + 1) * PROJ_LEN] = (vr ** (1 + (p % 2))) * np.exp(-((vr - 0.4 * p) / 0.9) ** 2)
        nl_tab = np.zeros(natoms * 5 * PROJ_LEN)
        for a in range(natoms):
            for p in range(5):
                w = 0.8 + 0.2 * p + 0.05 * a
                nl_tab[(a * 5 + p) * PROJ_LEN:(a * 5 + p + 1) * PROJ_LEN] = vr * np.exp(-(vr / w) ** 2)
        sgn = np.array([[1, -1, 1, -1, 1]] * natoms, dtype=np.int32).ravel()
        so_par = 0.2 + 0.1 * np.arange(natoms)
        n_gpt = np.zeros(natoms, dtype=np.int64)
        jxyz = np.zeros(natoms * stride, dtype=np.int64)
        y1 = np.zeros(natoms * stride * 6)
        prj = np.zeros(natoms * stride * 5)
        nlp = np.zeros(natoms * stride * 5)
        sign_out = np.zeros(natoms * 5, dtype=np.int32)
        O.lib().or_init_nl_projectors(O._lp(n_gpt), O._lp(jxyz), O._dp(y1), O._dp(prj), O._dp(nlp),
                                      O._ip(sign_out), O._dp(np.ascontiguousarray(R)), natoms, O._dp(gx),
                                      O._dp(gy), O._dp(gz), nx, ny, nz, R2, stride, 5, O._dp(so_tab),
                                      O._dp(nl_tab), O._ip(sgn), O._dp(so_par))
        proj = dict(so=so, nl=nl, n_atoms=natoms, n_gridpts=stride, n_gpt=n_gpt, jxyz=jxyz, y1=y1, proj=prj,
                    nl_proj=nlp, nl_sign=sign_out, n_so_proj=n_so_proj if so else 0,
                    n_nl_proj=n_nl_proj if nl else 0, LS=O.def_LS())
    dv = dx ** 3
    return dict(nx=nx, ny=ny, nz=nz, nspin=nspin, is_complex=is_complex, dv=dv, ksqr=ksqr, pot_local=pot,
                proj=proj)


def oracle_ham(s):
    return O.Ham(s["nx"], s["ny"], s["nz"], s["nspin"], s["is_complex"], s["dv"], s["ksqr"], s["pot_local"],
                 s["proj"])


def gpu_context(s):
    from filterbse_amd import api
    return api.Context(s["nx"], s["ny"], s["nz"], s["nspin"], s["is_complex"], s["dv"], s["ksqr"], s["pot_local"],
                       s["proj"])


def random_states(s, n, seed=1):
    rng = np.random.default_rng(seed)
    stlen = (s["is_complex"] + 1) * s["nspin"] * s["nx"] * s["ny"] * s["nz"]
    return rng.standard_normal(n * stlen)


VARIANTS = {
    "real_vloc": dict(nspin=1, is_complex=0),
    "real_nl": dict(nspin=1, is_complex=0, nl=1),
    "spinor_so_nl": dict(nspin=2, is_complex=1, so=1, nl=1),
    "spinor_vloc": dict(nspin=2, is_complex=1),
    # one radial projector of each kind: the plane passes add the expansion
    "spinor_so_nl_1p": dict(nspin=2, is_complex=1, so=1, nl=1, n_so_proj=1, n_nl_proj=1),
}
