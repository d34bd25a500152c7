"""Input side of the reference job, restated for the oracle: conf.par, pot*.par
and SO_*.par reading, the local potential on the grid, and the SO/NL projector
tables.  TEST INFRASTRUCTURE ONLY (see oracle/oracle.py).

Citations (filter_mpi/): read_conf read.c:998-1171 (COM centring at
read.c:1146-1158), read_pot/read_pot_file read.c:1175-1884, SO_*.par fields
read.c:1759-1772, R_NLcut2 read.c:908, build_local_pot init.c:481-703,
gen_SO_projectors projectors.c:12-84, gen_nlc_projectors projectors.c:87-228,
init_NL_projectors init.c:848-1135.  dsyev is replaced by numpy.linalg.eigh
(both return ascending eigenvalues; eigenvector signs are arbitrary in both
and cancel because every projector enters H quadratically).
"""
import math
import os

import numpy as np

from . import oracle as O

PROJ_LEN = 1024
EPS = 1e-10


def read_conf(path, center=True):
    with open(path) as f:
        toks = f.read().split()
    n = int(toks[0])
    syms, R = [], []
    for i in range(n):
        s, x, y, z = toks[1 + 4 * i: 5 + 4 * i]
        syms.append(s)
        R.append([float(x), float(y), float(z)])
    R = np.array(R)
    if center:
        # xd accumulated in file order, then divided (read.c:1106-1108,1146-1158)
        com = np.zeros(3)
        for i in range(n):
            com = com + R[i]
        com = com / float(n)
        R = R - com
    return syms, R


def read_pot(path):
    a = np.loadtxt(path)
    return a[:, 0].copy(), a[:, 1].copy()


def read_so_par(path):
    vals = []
    with open(path) as f:
        for line in f:
            if line.strip():
                vals.append(float(line.split()[0]))
    return vals[0], vals[1], vals[2]  # SO_par, NL_par[0], NL_par[1]


def _bessel1(x):
    """calcBessel(x, 1/(x+EPS)) of projectors.c:8: sin(x)/x^2 - cos(x)/x with the EPS guard."""
    x1 = 1.0 / (x + EPS)
    return np.where(x < EPS, 0.0, np.sin(x) * (x1 * x1) - np.cos(x) * x1)


def _kernel_matrix(dx, rcut, weight):
    N = PROJ_LEN
    kmax = 1.5 * math.pi / dx
    dk = kmax / N
    dr = rcut / N
    vr = np.arange(N) * dr
    k = np.arange(N) * dk
    r = vr[1:]
    J = _bessel1(np.outer(k, r))                 # J[i, rpoint]
    wr = (r ** 2) * weight(r) * dr
    A = (J * wr) @ J.T                           # A[i, j] = sum_r r^2 j1(ki r) w(r) j1(kj r) dr
    return A, k, dk, vr


def _radial_from_vector(vec, k, dk, vr, W):
    pre = (2.0 / math.pi) * math.sqrt(W)
    J = _bessel1(np.outer(vr, k))                # [rpoint, kpoint]
    return (J * (k * k * vec * dk)).sum(axis=1) * pre


def gen_so_projectors(dx, rcut, nproj=5):
    A, k, dk, vr = _kernel_matrix(dx, rcut, lambda r: np.exp(-1.0 * (r / 0.7) ** 2))
    W, V = np.linalg.eigh(A)
    N = PROJ_LEN
    out = np.zeros(nproj * N)
    for p in range(nproj):
        out[p * N:(p + 1) * N] = _radial_from_vector(V[:, N - p - 1], k, dk, vr, W[N - p - 1])
    return out


def gen_nlc_projectors(dx, rcut, lam1, lam2, nproj=5):
    N = PROJ_LEN
    if lam1 == 0 and lam2 == 0:
        return np.zeros(nproj * N), np.zeros(nproj, dtype=np.int32)
    A, k, dk, vr = _kernel_matrix(
        dx, rcut, lambda r: lam1 * np.exp(-1.0 * (r / 1.0) ** 2) + lam2 * np.exp(-(((r - 1.5) / 1.0) ** 2)))
    W, V = np.linalg.eigh(A)
    # the nproj largest |W| in descending order (insertion at projectors.c:168-186)
    eigs = [0.0] * nproj
    idx = [0] * nproj
    for i in range(N):
        if abs(W[i]) > eigs[nproj - 1]:
            for j in range(nproj - 1, -1, -1):
                if j == 0 or eigs[j - 1] > abs(W[i]):
                    eigs[j] = abs(W[i])
                    idx[j] = i
                    break
                eigs[j] = eigs[j - 1]
                idx[j] = idx[j - 1]
    out = np.zeros(nproj * N)
    sgn = np.zeros(nproj, dtype=np.int32)
    for p in range(nproj):
        w = W[idx[p]]
        sgn[p] = -1 if w < 0.0 else 1
        out[p * N:(p + 1) * N] = _radial_from_vector(V[:, idx[p]], k, dk, vr, abs(w))
    return out, sgn


def reference_job(directory, n_req=(44, 44, 44), dx=0.5, so=1, nl=1, nproj=5, ke_max=20.0):
    """Builds the grid, V_loc and projector tables of the job in `directory`
    (conf.par, pot<Sym>.par, SO_<Sym>.par) as the reference does for
    spinOrbit/NonLocal runs without interpolation, strain or LR terms."""
    syms, R = read_conf(os.path.join(directory, "conf.par"))
    natoms = len(syms)
    n, mins, maxs, dk = O.init_grid_params(R, n_req, (dx, dx, dx))
    nx, ny, nz = (int(v) for v in n)
    ksqr, gx, gy, gz = O.build_grid_ksqr(n, (dx, dx, dx), mins, dk, ke_max)
    types = []
    for s in syms:
        if s not in types:
            types.append(s)
    atype = np.array([types.index(s) for s in syms], dtype=np.int32)
    tabs = [read_pot(os.path.join(directory, "pot%s.par" % t)) for t in types]
    mlen = max(len(r) for r, _ in tabs)
    pot_r = np.zeros(len(types) * mlen)
    pot_v = np.zeros(len(types) * mlen)
    lens = np.zeros(len(types), dtype=np.int64)
    pdr = np.zeros(len(types))
    for t, (r, v) in enumerate(tabs):
        pot_r[t * mlen: t * mlen + len(r)] = r
        pot_v[t * mlen: t * mlen + len(v)] = v
        lens[t] = len(r)
        pdr[t] = r[1] - r[0]
    nspin = 2 if so else 1
    pot = np.zeros(nspin * nx * ny * nz)
    Rc = np.ascontiguousarray(R)
    O.lib().or_build_local_pot(O._dp(pot), nspin, O._dp(gx), O._dp(gy), O._dp(gz), nx, ny, nz, O._dp(Rc),
                               O._ip(atype), natoms, O._dp(pot_r), O._dp(pot_v), O._lp(lens), O._dp(pdr), mlen)
    job = dict(nx=nx, ny=ny, nz=nz, nspin=nspin, is_complex=1 if so else 0, dv=dx ** 3, ksqr=ksqr, pot_local=pot,
               proj=None, R=R, syms=syms, mins=mins, dk=dk, gx=gx, gy=gy, gz=gz,
               pot_tables=dict(R=Rc, atype=atype, pot_r=pot_r, pot_v=pot_v, lens=lens, dr=pdr, n_types=len(types),
                               max_len=mlen))
    if so or nl:
        R2 = 1.5 + 6.0 * math.log(10.0) + 3.0 * dx
        rcut = math.sqrt(R2)
        stride = O.lib().or_count_nl_gridpts(O._dp(Rc), natoms, O._dp(gx), O._dp(gy), O._dp(gz), nx, ny, nz, R2)
        pars = {t: read_so_par(os.path.join(directory, "SO_%s.par" % t)) for t in types}
        so_tab = gen_so_projectors(dx, rcut, nproj)
        nl_tab = np.zeros(natoms * nproj * PROJ_LEN)
        sgn = np.zeros(natoms * 5, dtype=np.int32)
        nl_cache = {}
        for a, s in enumerate(syms):
            if s not in nl_cache:
                nl_cache[s] = gen_nlc_projectors(dx, rcut, pars[s][1], pars[s][2], nproj)
            tab, sg = nl_cache[s]
            nl_tab[a * nproj * PROJ_LEN:(a + 1) * nproj * PROJ_LEN] = tab
            sgn[a * 5: a * 5 + nproj] = sg
        so_par = np.array([pars[s][0] for s in syms])
        n_gpt = np.zeros(natoms, dtype=np.int64)
        jxyz = np.zeros(natoms * stride, dtype=np.int64)
        y1 = np.zeros(natoms * stride * 6)
        prj = np.zeros(natoms * stride * 5)
        nlp = np.zeros(natoms * stride * 5)
        sign_out = np.zeros(natoms * 5, dtype=np.int32)
        O.lib().or_init_nl_projectors(O._lp(n_gpt), O._lp(jxyz), O._dp(y1), O._dp(prj), O._dp(nlp),
                                      O._ip(sign_out), O._dp(Rc), natoms, O._dp(gx), O._dp(gy), O._dp(gz), nx, ny,
                                      nz, R2, stride, nproj, O._dp(so_tab), O._dp(nl_tab), O._ip(sgn),
                                      O._dp(so_par))
        job["nl_inputs"] = dict(R2=R2, stride=stride, nproj=nproj, so_tab=so_tab, nl_tab=nl_tab, sgn=sgn,
                                so_par=so_par)
        job["proj"] = dict(so=so, nl=nl, n_atoms=natoms, n_gridpts=stride, n_gpt=n_gpt, jxyz=jxyz, y1=y1,
                           proj=prj, nl_proj=nlp, nl_sign=sign_out, n_so_proj=5 if so else 0,
                           n_nl_proj=nproj if nl else 0, LS=O.def_LS())
        # NL spectral bound estimate (init.c:1087-1127)
        lam_min = lam_max = 0.0
        for a in range(natoms):
            for p in range(nproj):
                s_ = sign_out[a * 5 + p]
                base = a * stride
                cnt = n_gpt[a]
                b2 = nlp[(base + np.arange(cnt)) * 5 + p] ** 2
                for m in range(3):
                    yr = y1[(base + np.arange(cnt)) * 6 + 2 * m]
                    yi = y1[(base + np.arange(cnt)) * 6 + 2 * m + 1]
                    lam = float(s_) * (dx ** 3) * float((b2 * (yr * yr + yi * yi)).sum())
                    lam_min = min(lam_min, lam)
                    lam_max = max(lam_max, lam)
        job["nl_lambda"] = (lam_min, lam_max)
    return job
