"""ctypes front end of the CPU oracle (oracle/fbse_oracle.c) plus the dense
linear algebra steps restated with numpy/LAPACK.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the filterbse_amd package.  It is the
checker, not the thing measured or shipped.

Reference citations are relative to the FilterBSE tree (filter_mpi/).

Parity unpinned: the exact output of the reference's MKL ?gesvd/?heev calls
(eigenvector phases, the basis U of the SVD); numpy's LAPACK stands in for
them and only the composed eigenvalues against the reference's eval.dat pin
them.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force=False):
    """Compile liboracle.so with the Makefile in this directory."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


DP = C.POINTER(C.c_double)
LP = C.POINTER(C.c_long)
IP = C.POINTER(C.c_int)


class OrHam(C.Structure):
    """Mirror of or_ham in fbse_oracle.c."""
    _fields_ = [
        ("nx", C.c_long), ("ny", C.c_long), ("nz", C.c_long), ("ngrid", C.c_long),
        ("nspin", C.c_long), ("nspinngrid", C.c_long),
        ("is_complex", C.c_int), ("so", C.c_int), ("nl", C.c_int),
        ("dv", C.c_double),
        ("ksqr", DP), ("pot_local", DP),
        ("n_atoms", C.c_long), ("n_gridpts", C.c_long),
        ("n_gpt", LP), ("jxyz", LP), ("y1", DP), ("proj", DP), ("nl_proj", DP),
        ("nl_sign", IP), ("n_so_proj", C.c_int), ("n_nl_proj", C.c_int), ("LS", DP),
    ]


def _declare(L):
    L.or_ran_nrc.restype = C.c_double
    L.or_ran_nrc.argtypes = [LP]
    L.or_init_filter_states.argtypes = [DP, C.c_long, C.c_long, LP, C.c_long, C.c_long, C.c_long, C.c_long,
                                        C.c_int, C.c_double]
    L.or_init_psi.argtypes = [DP, LP, C.c_long, C.c_long, C.c_long, C.c_int, C.c_double]
    L.or_init_grid_params.argtypes = [DP, C.c_long, LP, DP, DP, DP, DP]
    L.or_init_kvec.argtypes = [DP, C.c_int, C.c_double, C.c_double]
    L.or_build_grid_ksqr.argtypes = [DP, DP, DP, DP, C.c_long, C.c_long, C.c_long, DP, DP, DP, C.c_double]
    L.or_build_local_pot.argtypes = [DP, C.c_long, DP, DP, DP, C.c_long, C.c_long, C.c_long, DP, IP, C.c_long,
                                     DP, DP, LP, DP, C.c_long]
    L.or_count_nl_gridpts.restype = C.c_long
    L.or_count_nl_gridpts.argtypes = [DP, C.c_long, DP, DP, DP, C.c_long, C.c_long, C.c_long, C.c_double]
    L.or_init_nl_projectors.argtypes = [LP, LP, DP, DP, DP, IP, DP, C.c_long, DP, DP, DP, C.c_long, C.c_long,
                                        C.c_long, C.c_double, C.c_long, C.c_long, DP, DP, IP, DP]
    L.or_samp_points_ashkenazy.restype = C.c_double
    L.or_samp_points_ashkenazy.argtypes = [DP, DP, C.c_double, C.c_double, C.c_long]
    L.or_gen_newton_coeff.argtypes = [DP, DP, DP, C.c_long, C.c_long, C.c_double, C.c_double, C.c_double]
    L.or_set_ene_targets.argtypes = [DP, C.c_long, C.c_long, C.c_double, C.c_double, C.c_double, C.c_double]
    L.or_fft3d.argtypes = [DP, C.c_long, C.c_long, C.c_long, C.c_int]
    L.or_def_LS.argtypes = [DP]
    L.or_apply_h.argtypes = [DP, DP, C.c_long, C.POINTER(OrHam)]
    L.or_filter_cycle.argtypes = [DP, C.c_long, C.c_long, C.c_long, DP, DP, C.c_double, C.c_double,
                                  C.POINTER(OrHam)]
    L.or_normalize_all.restype = C.c_int
    L.or_normalize_all.argtypes = [DP, C.c_long, C.c_long, C.c_int, C.c_double]
    L.or_energy_all.argtypes = [DP, C.c_long, DP, C.POINTER(OrHam)]
    L.or_calc_sigma_E.argtypes = [DP, C.c_long, DP, DP, C.POINTER(OrHam)]
    L.or_get_energy_range.argtypes = [C.POINTER(OrHam), C.c_long, C.c_double, C.c_double, DP]


def _dp(a):
    return a.ctypes.data_as(DP)


def _lp(a):
    return a.ctypes.data_as(LP)


def _ip(a):
    return a.ctypes.data_as(IP)


# --------------------------------------------------------------------------
# random numbers and start states (rand.c:22-62, init.c:1138-1305)
# --------------------------------------------------------------------------
def ran_sequence(seed, n):
    """n successive ran2 values from a (re-initialising, non-positive) seed."""
    s = C.c_long(seed)
    return np.array([lib().or_ran_nrc(C.byref(s)) for _ in range(n)]), s.value


def init_filter_states(n_starts, ms, seed, nx, ny, nz, nspin, is_complex, dv):
    stlen = (is_complex + 1) * nspin * nx * ny * nz
    psi_rank = np.zeros(n_starts * ms * stlen)
    s = C.c_long(seed)
    lib().or_init_filter_states(_dp(psi_rank), n_starts, ms, C.byref(s), nx, ny, nz, nspin, is_complex, dv)
    return psi_rank, s.value


def init_psi(seed, nx, ny, nz, is_complex, dv):
    psi = np.zeros(2 * nx * ny * nz)
    s = C.c_long(seed)
    lib().or_init_psi(_dp(psi), C.byref(s), nx, ny, nz, is_complex, dv)
    return psi.view(np.complex128), s.value


# --------------------------------------------------------------------------
# grid (init.c:5-397, 1383-1394)
# --------------------------------------------------------------------------
def init_grid_params(R, n_req, d):
    R = np.ascontiguousarray(R, dtype=np.float64)
    io = np.array(n_req, dtype=np.int64)
    d = np.array(d, dtype=np.float64)
    mn, mx, dk = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().or_init_grid_params(_dp(R), len(R), _lp(io), _dp(d), _dp(mn), _dp(mx), _dp(dk))
    return io, mn, mx, dk


def build_grid_ksqr(n, d, mins, dk, ke_max=20.0):
    nx, ny, nz = (int(v) for v in n)
    ksqr = np.zeros(nx * ny * nz)
    gx, gy, gz = np.zeros(nx), np.zeros(ny), np.zeros(nz)
    d = np.array(d, dtype=np.float64)
    lib().or_build_grid_ksqr(_dp(ksqr), _dp(gx), _dp(gy), _dp(gz), nx, ny, nz, _dp(d),
                             _dp(np.array(mins, dtype=np.float64)), _dp(np.array(dk, dtype=np.float64)), ke_max)
    return ksqr, gx, gy, gz


def init_kvec(n, dk):
    k = np.zeros(n)
    lib().or_init_kvec(_dp(k), n, dk, 1.0)
    return k


# --------------------------------------------------------------------------
# Newton interpolation (coeff.c:8-118, 227-293)
# --------------------------------------------------------------------------
def leja_points(M):
    re, im = np.zeros(M), np.zeros(M)
    lib().or_samp_points_ashkenazy(_dp(re), _dp(im), -2.0, 2.0, M)
    return re


def newton_coeff(zn, targets, Vmin, dE, dt):
    zn = np.ascontiguousarray(zn, dtype=np.float64)
    targets = np.ascontiguousarray(targets, dtype=np.float64)
    M, ms = len(zn), len(targets)
    an = np.zeros(2 * ms * M)
    lib().or_gen_newton_coeff(_dp(an), _dp(zn), _dp(targets), ms, M, Vmin, dE, dt)
    return an


def ene_targets(nVB, nCB, VBmin, VBmax, CBmin, CBmax):
    t = np.zeros(nVB + nCB)
    lib().or_set_ene_targets(_dp(t), nVB, nCB, VBmin, VBmax, CBmin, CBmax)
    return t


def fft3d(a, sign):
    """In-place-style unnormalised 3D DFT of a (nz, ny, nx) complex array."""
    b = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nz, ny, nx = b.shape
    lib().or_fft3d(b.ctypes.data_as(DP), nx, ny, nz, sign)
    return b


def def_LS():
    ls = np.zeros(72)
    lib().or_def_LS(_dp(ls))
    return ls


# --------------------------------------------------------------------------
# Hamiltonian
# --------------------------------------------------------------------------
class Ham:
    """Keeps the numpy arrays alive behind an OrHam."""

    def __init__(self, nx, ny, nz, nspin, is_complex, dv, ksqr, pot_local, proj=None):
        self.arrays = {}
        h = OrHam()
        h.nx, h.ny, h.nz = nx, ny, nz
        h.ngrid = nx * ny * nz
        h.nspin = nspin
        h.nspinngrid = nspin * h.ngrid
        h.is_complex = is_complex
        h.dv = dv
        self._keep("ksqr", np.ascontiguousarray(ksqr, dtype=np.float64))
        self._keep("pot", np.ascontiguousarray(pot_local, dtype=np.float64))
        assert self.arrays["pot"].size == h.nspinngrid
        h.ksqr = _dp(self.arrays["ksqr"])
        h.pot_local = _dp(self.arrays["pot"])
        if proj is None:
            h.so = h.nl = 0
            h.n_atoms = 0
        else:
            h.so, h.nl = int(proj["so"]), int(proj["nl"])
            h.n_atoms = int(proj["n_atoms"])
            h.n_gridpts = int(proj["n_gridpts"])
            for k, dt in (("n_gpt", np.int64), ("jxyz", np.int64), ("y1", np.float64), ("proj", np.float64),
                          ("nl_proj", np.float64), ("nl_sign", np.int32), ("LS", np.float64)):
                self._keep(k, np.ascontiguousarray(proj[k], dtype=dt))
            h.n_gpt = _lp(self.arrays["n_gpt"])
            h.jxyz = _lp(self.arrays["jxyz"])
            h.y1 = _dp(self.arrays["y1"])
            h.proj = _dp(self.arrays["proj"])
            h.nl_proj = _dp(self.arrays["nl_proj"])
            h.nl_sign = _ip(self.arrays["nl_sign"])
            h.LS = _dp(self.arrays["LS"])
            h.n_so_proj = int(proj["n_so_proj"])
            h.n_nl_proj = int(proj["n_nl_proj"])
        self.h = h
        self.stlen = (is_complex + 1) * h.nspinngrid

    def _keep(self, k, a):
        self.arrays[k] = a

    def apply(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        n = psi.size // self.stlen
        out = np.zeros_like(psi)
        lib().or_apply_h(_dp(psi), _dp(out), n, C.byref(self.h))
        return out

    def filter_cycles(self, psi_rank, n_starts, ms, an, zn, Vmin, dE_1):
        psi_rank = np.ascontiguousarray(psi_rank, dtype=np.float64).copy()
        M = len(zn)
        for jns in range(n_starts):
            lib().or_filter_cycle(_dp(psi_rank), jns, ms, M, _dp(np.ascontiguousarray(an)),
                                  _dp(np.ascontiguousarray(zn)), Vmin, dE_1, C.byref(self.h))
        return psi_rank

    def normalize_all(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64).copy()
        rc = lib().or_normalize_all(_dp(psi), psi.size // self.stlen, self.h.nspinngrid, self.h.is_complex,
                                    self.h.dv)
        return psi, rc

    def energies(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        n = psi.size // self.stlen
        e = np.zeros(n)
        lib().or_energy_all(_dp(psi), n, _dp(e), C.byref(self.h))
        return e

    def sigma(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        n = psi.size // self.stlen
        s, e = np.zeros(n), np.zeros(n)
        lib().or_calc_sigma_E(_dp(psi), n, _dp(s), _dp(e), C.byref(self.h))
        return s, e

    def energy_range(self, max_iter=1500, nl_lambda_min=0.0, vloc_min=0.0):
        out = np.zeros(6)
        lib().or_get_energy_range(C.byref(self.h), max_iter, nl_lambda_min, vloc_min, _dp(out))
        return out


# --------------------------------------------------------------------------
# Dense steps restated with LAPACK through numpy: ortho_real/ortho_cplx
# (ortho.c:5-185; ?gesvd JOBU='O' and the s_i/s_0 < SVDEPS cut, fd.h:535)
# and diag_H (Hmat.c:5-297; H_ij = <psi_i|H|psi_j> dv, ?syev/?heev).
# --------------------------------------------------------------------------
def as_matrix(psitot, n, stlen, is_complex):
    A = np.asarray(psitot[: n * stlen]).reshape(n, stlen)
    if is_complex:
        A = A.view(np.complex128)
    return A.T  # columns are states


def ortho(psitot, n, stlen, is_complex, dv, svd_eps=1.0e-10):
    A = as_matrix(psitot, n, stlen, is_complex)
    U, s, _ = np.linalg.svd(A, full_matrices=False)
    cut = n
    for i in range(n):
        if s[i] / s[0] < svd_eps:
            cut = i
            break
    U = U[:, :cut] / np.sqrt(dv)  # normalize_all (mod_ortho.c:84)
    return U, s, cut


def rayleigh_ritz(U, HU, dv):
    """U, HU: (rows x n) with columns orthonormal under the dv weight."""
    Hs = dv * (U.conj().T @ HU)
    Hs = 0.5 * (Hs + Hs.conj().T)
    w, Y = np.linalg.eigh(Hs)
    return w, U @ Y
