"""ctypes bindings of libfbse_gpu.so (include/fbse_gpu.h) and
libfbse_host.so (include/fbse_host.h).

The GPU path has no fallback: if the HIP library is missing or fails to load,
every call raises.  Arrays follow the reference layouts (filter_mpi/): a state
is stlen = nspin*ngrid*(is_complex+1) float64 values, complex interleaved.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GPU_LIB = os.path.join(HERE, "libfbse_gpu.so")
HOST_LIB = os.path.join(HERE, "libfbse_host.so")

FBSE_PREC_FP64 = 0
FBSE_PREC_MIXED = 1
FBSE_OK, FBSE_ERR_ARG, FBSE_ERR_HIP, FBSE_ERR_NORM_ZERO, FBSE_ERR_NORM_NAN, FBSE_ERR_SOLVER, FBSE_ERR_NOMEM, \
    FBSE_ERR_UNSUPPORTED, FBSE_ERR_IO = 0, -1, -2, -3, -4, -5, -6, -7, -8

DP = C.POINTER(C.c_double)
LP = C.POINTER(C.c_long)
IP = C.POINTER(C.c_int)

ERRORS = {0: "FBSE_OK", -1: "FBSE_ERR_ARG", -2: "FBSE_ERR_HIP", -3: "FBSE_ERR_NORM_ZERO",
          -4: "FBSE_ERR_NORM_NAN", -5: "FBSE_ERR_SOLVER", -6: "FBSE_ERR_NOMEM", -7: "FBSE_ERR_UNSUPPORTED",
          -8: "FBSE_ERR_IO"}


class FbseError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_long), ("ny", C.c_long), ("nz", C.c_long), ("nspin", C.c_long),
                ("is_complex", C.c_int), ("dv", C.c_double), ("ksqr", DP), ("pot_local", DP)]


class ProjDesc(C.Structure):
    _fields_ = [("so", C.c_int), ("nl", C.c_int), ("n_atoms", C.c_long), ("n_gridpts", C.c_long),
                ("n_gpt", LP), ("jxyz", LP), ("y1", DP), ("proj", DP), ("nl_proj", DP), ("nl_sign", IP),
                ("n_so_proj", C.c_int), ("n_nl_proj", C.c_int), ("LS", DP)]


_gpu = None
_host = None

# every entry point of include/fbse_gpu.h with its ctypes signature
GPU_SIGNATURES = {
    "fbse_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(ProjDesc), C.c_int, C.POINTER(C.c_void_p)]),
    "fbse_destroy": (None, [C.c_void_p]),
    "fbse_last_error": (C.c_char_p, [C.c_void_p]),
    "fbse_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fbse_synchronize": (C.c_int, [C.c_void_p]),
    "fbse_apply_h": (C.c_int, [C.c_void_p, DP, DP, C.c_long]),
    "fbse_apply_h_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]),
    "fbse_apply_h_dev_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]),
    "fbse_filter": (C.c_int, [C.c_void_p, DP, C.c_long, C.c_long, C.c_long, DP, DP, C.c_double, C.c_double,
                              C.c_int]),
    "fbse_filter_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_long, DP, DP, C.c_double,
                                  C.c_double, C.c_int]),
    "fbse_normalize_all": (C.c_int, [C.c_void_p, DP, C.c_long]),
    "fbse_normalize_all_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long]),
    "fbse_energy_all": (C.c_int, [C.c_void_p, DP, C.c_long, DP]),
    "fbse_energy_all_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, DP]),
    "fbse_ortho": (C.c_int, [C.c_void_p, DP, C.c_long, C.c_double, LP, DP]),
    "fbse_ortho_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_double, LP, DP]),
    "fbse_rayleigh_ritz": (C.c_int, [C.c_void_p, DP, C.c_long, DP]),
    "fbse_rayleigh_ritz_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, DP]),
    "fbse_sigma": (C.c_int, [C.c_void_p, DP, C.c_long, DP, DP]),
    "fbse_sigma_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, DP, DP]),
    "fbse_refine": (C.c_int, [C.c_void_p, DP, C.c_long, DP, DP, C.c_double, C.c_double, C.c_double, C.c_double,
                              DP, DP, LP]),
    "fbse_refine_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, DP, DP, C.c_double, C.c_double, C.c_double,
                                  C.c_double, DP, DP, LP]),
    "fbse_overlap_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p, C.c_long, DP]),
    "fbse_rotate_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, DP, C.c_long]),
    "fbse_energy_range": (C.c_int, [C.c_void_p, DP, DP, C.c_long, C.c_double, C.c_double, DP]),
    "fbse_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "fbse_profile_read": (C.c_int, [C.c_void_p, C.c_int, LP, DP]),
    "fbse_set_window": (C.c_int, [C.c_void_p, C.c_int]),
    "fbse_set_fft_kernel": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "fbse_fft_choice": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fbse_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "fbse_set_axis_fusion": (C.c_int, [C.c_void_p, C.c_int]),
    "fbse_set_combine_projection": (C.c_int, [C.c_void_p, C.c_int]),
    "fbse_last_group": (C.c_int, [C.c_void_p]),
    "fbse_build_local_pot": (C.c_int, [C.c_int, C.c_long, DP, DP, DP, C.c_long, C.c_long, C.c_long, DP,
                                       C.POINTER(C.c_int), C.c_long, DP, DP, LP, DP, C.c_long, C.c_long, DP]),
    "fbse_count_nl_gridpts": (C.c_long, [C.c_int, DP, C.c_long, DP, DP, DP, C.c_long, C.c_long, C.c_long,
                                         C.c_double]),
    "fbse_init_nl_projectors": (C.c_int, [C.c_int, LP, LP, DP, DP, DP, C.POINTER(C.c_int), DP, C.c_long, DP, DP,
                                          DP, C.c_long, C.c_long, C.c_long, C.c_double, C.c_long, C.c_long, DP,
                                          DP, C.POINTER(C.c_int), DP]),
    "fbse_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)]),
    "fbse_mlp_destroy": (None, [C.c_void_p]),
    "fbse_mlp_last_error": (C.c_char_p, [C.c_void_p]),
    "fbse_mlp_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fbse_mlp_input_width": (C.c_int, [C.c_void_p]),
    "fbse_mlp_num_params": (C.c_long, [C.c_void_p]),
    "fbse_mlp_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "fbse_mlp_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "fbse_mlp_forward_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p]),
    "fbse_mlp_loss_grad_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long,
                                         C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    "fbse_mlp_adam_step": (C.c_int, [C.c_void_p, C.c_float]),
    "fbse_mlp_features_dev": (C.c_int, [C.c_void_p, C.c_long, C.c_long, C.c_long, DP, C.c_double, DP,
                                        C.POINTER(C.c_int), C.c_long, C.c_int, C.c_void_p, DP, C.c_double,
                                        C.c_void_p]),
    "fbse_state_density_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_long), C.c_long, C.c_void_p]),
    "fbse_warm_start_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                      C.c_long, C.c_long]),
    "fbse_filter_many": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_long),
                                   C.c_long, C.c_long, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]),
    "fbse_graph_replays": (C.c_long, [C.c_void_p]),
    "fbse_last_window": (C.c_int, [C.c_void_p]),
    "fbse_last_runs": (C.c_long, [C.c_void_p]),
    "fbse_set_run_starts": (C.c_int, [C.c_void_p, C.c_long]),
    "fbse_held_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "fbse_trim": (C.c_int, [C.c_void_p]),
    "fbse_fwrite_states": (C.c_int, [C.c_void_p, C.c_void_p, LP, C.c_long, C.c_int, C.c_void_p]),
    "fbse_write_states": (C.c_int, [C.c_void_p, C.c_void_p, LP, C.c_long, C.c_int, C.c_char_p]),
    "fbse_read_start_states": (C.c_int, [C.c_void_p, C.c_char_p, C.c_long, C.c_long, C.c_int, C.c_void_p,
                                         C.c_long, C.c_long, LP]),
    "fbse_set_io_chunk": (C.c_int, [C.c_void_p, C.c_size_t]),
    "fbse_set_slab_split": (C.c_int, [C.c_void_p, C.c_int]),
    "fbse_dev_alloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "fbse_dev_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fbse_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "fbse_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
}

HOST_SIGNATURES = {
    "fbse_ran_nrc": (C.c_double, [LP]),
    "fbse_init_psi": (None, [DP, LP, C.c_long, C.c_long, C.c_long, C.c_int, C.c_double]),
    "fbse_init_filter_states": (None, [DP, C.c_long, C.c_long, LP, C.c_long, C.c_long, C.c_long, C.c_long,
                                       C.c_int, C.c_double]),
    "fbse_init_grid_params": (None, [DP, C.c_long, LP, DP, DP, DP, DP]),
    "fbse_build_grid_ksqr": (None, [DP, DP, DP, DP, LP, DP, DP, DP, C.c_double]),
    "fbse_samp_points_ashkenazy": (C.c_double, [DP, C.c_long, C.c_int]),
    "fbse_gen_newton_coeff": (None, [DP, DP, DP, C.c_long, C.c_long, C.c_double, C.c_double, C.c_double,
                                     C.c_int]),
    "fbse_set_ene_targets": (None, [DP, C.c_long, C.c_long, C.c_double, C.c_double, C.c_double, C.c_double]),
    "fbse_def_LS": (None, [DP]),
}


def _load(path, sigs):
    if not os.path.exists(path):
        raise RuntimeError("%s is not built: run `python -m filterbse_amd.build` (no CPU fallback exists)" % path)
    lib = C.CDLL(path)
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def gpu():
    global _gpu
    if _gpu is None:
        # One HIP runtime per process: PyTorch ships its own libamdhip64 with
        # the same soname.  Loading torch first lets this library bind to that
        # runtime (device buffers and streams are then shared); loading
        # /opt/rocm's first would leave torch with a second runtime that sees
        # no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _gpu = _load(GPU_LIB, GPU_SIGNATURES)
    return _gpu


def host():
    global _host
    if _host is None:
        gpu()  # the host library links against the GPU one
        _host = _load(HOST_LIB, HOST_SIGNATURES)
    return _host


def _dp(a):
    return a.ctypes.data_as(DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def nthreads():
    return max(1, min(16, os.cpu_count() or 1))


# --------------------------------------------------------------------------
# host-side inputs
# --------------------------------------------------------------------------
def init_filter_states(n_starts, ms, seed, nx, ny, nz, nspin, is_complex, dv):
    stlen = (is_complex + 1) * nspin * nx * ny * nz
    psi_rank = np.zeros(n_starts * ms * stlen)
    s = C.c_long(seed)
    host().fbse_init_filter_states(_dp(psi_rank), n_starts, ms, C.byref(s), nx, ny, nz, nspin, is_complex, dv)
    return psi_rank, s.value


def init_psi(seed, nx, ny, nz, is_complex, dv):
    psi = np.zeros(2 * nx * ny * nz)
    s = C.c_long(seed)
    host().fbse_init_psi(_dp(psi), C.byref(s), nx, ny, nz, is_complex, dv)
    return psi, s.value


def energy_range_start_states(nx, ny, nz, nspin, is_complex, dv):
    """The two start vectors of get_energy_range (energy.c:207,228-231,270-273):
    init_psi per spin component from the fixed seed -874917403."""
    seed = -874917403
    ng = nx * ny * nz
    out = []
    for _ in range(2):
        comps = []
        for _s in range(nspin):
            p, seed = init_psi(seed, nx, ny, nz, is_complex, dv)
            comps.append(p if is_complex else p[0::2].copy())
        out.append(np.concatenate(comps))
    return out


def grid_params(R, n_req, d):
    R = _f64(R)
    n = np.array(n_req, dtype=np.int64)
    d = _f64(d)
    mins, maxs, dk = np.zeros(3), np.zeros(3), np.zeros(3)
    host().fbse_init_grid_params(_dp(R), len(R), n.ctypes.data_as(LP), _dp(d), _dp(mins), _dp(maxs), _dp(dk))
    return n, mins, maxs, dk


def grid_ksqr(n, d, mins, dk, ke_max=20.0):
    n = np.array(n, dtype=np.int64)
    ksqr = np.zeros(int(np.prod(n)))
    gx, gy, gz = np.zeros(n[0]), np.zeros(n[1]), np.zeros(n[2])
    host().fbse_build_grid_ksqr(_dp(ksqr), _dp(gx), _dp(gy), _dp(gz), n.ctypes.data_as(LP), _dp(_f64(d)),
                                _dp(_f64(mins)), _dp(_f64(dk)), ke_max)
    return ksqr, gx, gy, gz


def leja_points(M):
    zn = np.zeros(M)
    host().fbse_samp_points_ashkenazy(_dp(zn), M, nthreads())
    return zn


def newton_coeff(zn, targets, Vmin, dE, dt):
    zn, targets = _f64(zn), _f64(targets)
    an = np.zeros(2 * len(targets) * len(zn))
    host().fbse_gen_newton_coeff(_dp(an), _dp(zn), _dp(targets), len(targets), len(zn), Vmin, dE, dt, nthreads())
    return an


def ene_targets(nVB, nCB, VBmin, VBmax, CBmin, CBmax):
    t = np.zeros(nVB + nCB)
    host().fbse_set_ene_targets(_dp(t), nVB, nCB, VBmin, VBmax, CBmin, CBmax)
    return t


def def_LS():
    ls = np.zeros(72)
    host().fbse_def_LS(_dp(ls))
    return ls


# --------------------------------------------------------------------------
# GPU context
# --------------------------------------------------------------------------
class Context:
    """One fbse_ctx: grid, local potential and projector tables on one GPU."""

    def __init__(self, nx, ny, nz, nspin, is_complex, dv, ksqr, pot_local, proj=None, device=-1):
        L = gpu()
        self._keep = {}
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nspin, self.is_complex, self.dv = nspin, is_complex, dv
        self.ngrid = nx * ny * nz
        self.stlen = (is_complex + 1) * nspin * self.ngrid
        ksqr = self._hold("ksqr", _f64(ksqr))
        pot = self._hold("pot", _f64(pot_local))
        if pot.size != nspin * self.ngrid:
            raise ValueError("pot_local must hold nspin*ngrid values (the reference duplicates it per spin)")
        g = GridDesc(nx, ny, nz, nspin, is_complex, dv, _dp(ksqr), _dp(pot))
        pd = None
        if proj is not None:
            arrs = {}
            for k, dt in (("n_gpt", np.int64), ("jxyz", np.int64), ("y1", np.float64), ("proj", np.float64),
                          ("nl_proj", np.float64), ("nl_sign", np.int32), ("LS", np.float64)):
                arrs[k] = self._hold(k, np.ascontiguousarray(proj[k], dtype=dt))
            pd = ProjDesc(int(proj["so"]), int(proj["nl"]), int(proj["n_atoms"]), int(proj["n_gridpts"]),
                          arrs["n_gpt"].ctypes.data_as(LP), arrs["jxyz"].ctypes.data_as(LP), _dp(arrs["y1"]),
                          _dp(arrs["proj"]), _dp(arrs["nl_proj"]), arrs["nl_sign"].ctypes.data_as(IP),
                          int(proj["n_so_proj"]), int(proj["n_nl_proj"]), _dp(arrs["LS"]))
        ctx = C.c_void_p()
        rc = L.fbse_create(C.byref(g), C.byref(pd) if pd is not None else None, device, C.byref(ctx))
        if rc != 0:
            raise FbseError(rc, "fbse_create failed")
        self.ctx = ctx
        self._keep.clear()  # the library copied everything

    def _hold(self, k, a):
        self._keep[k] = a
        return a

    def close(self):
        if getattr(self, "ctx", None):
            gpu().fbse_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise FbseError(rc, gpu().fbse_last_error(self.ctx).decode())

    # host-array entry points -------------------------------------------------
    def apply_h(self, psi):
        psi = _f64(psi)
        out = np.zeros_like(psi)
        self._check(gpu().fbse_apply_h(self.ctx, _dp(psi), _dp(out), psi.size // self.stlen))
        return out

    def filter(self, psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=FBSE_PREC_FP64):
        psi_rank = _f64(psi_rank).copy()
        an, zn = _f64(an), _f64(zn)
        self._check(gpu().fbse_filter(self.ctx, _dp(psi_rank), n_starts, ms, len(zn), _dp(an), _dp(zn), Vmin,
                                      dE_1, precision))
        return psi_rank

    def normalize_all(self, psi):
        psi = _f64(psi).copy()
        self._check(gpu().fbse_normalize_all(self.ctx, _dp(psi), psi.size // self.stlen))
        return psi

    def energies(self, psi):
        psi = _f64(psi)
        n = psi.size // self.stlen
        e = np.zeros(n)
        self._check(gpu().fbse_energy_all(self.ctx, _dp(psi), n, _dp(e)))
        return e

    def ortho(self, psi, n_cols, svd_eps=1.0e-10):
        psi = _f64(psi).copy()
        rank = C.c_long(0)
        sv = np.zeros(n_cols)
        self._check(gpu().fbse_ortho(self.ctx, _dp(psi), n_cols, svd_eps, C.byref(rank), _dp(sv)))
        return psi, rank.value, sv

    def rayleigh_ritz(self, psi, n):
        psi = _f64(psi).copy()
        ev = np.zeros(n)
        self._check(gpu().fbse_rayleigh_ritz(self.ctx, _dp(psi), n, _dp(ev)))
        return psi, ev

    def sigma(self, psi, n):
        psi = _f64(psi)
        s, e = np.zeros(n), np.zeros(n)
        self._check(gpu().fbse_sigma(self.ctx, _dp(psi), n, _dp(s), _dp(e)))
        return s, e

    def energy_range(self, psi0_min, psi0_max, max_iter=1500, nl_lambda_min=0.0, vloc_min=0.0):
        out = np.zeros(6)
        self._check(gpu().fbse_energy_range(self.ctx, _dp(_f64(psi0_min)), _dp(_f64(psi0_max)), max_iter,
                                            nl_lambda_min, vloc_min, _dp(out)))
        return out

    # device-pointer entry points (integers from torch .data_ptr()) -----------
    def set_stream(self, stream_ptr):
        self._check(gpu().fbse_set_stream(self.ctx, C.c_void_p(stream_ptr) if stream_ptr else None))

    def synchronize(self):
        self._check(gpu().fbse_synchronize(self.ctx))

    def apply_h_dev(self, d_in, d_out, nstates):
        self._check(gpu().fbse_apply_h_dev(self.ctx, C.c_void_p(d_in), C.c_void_p(d_out), nstates))

    def apply_h_dev_noalloc(self, d_in, d_out, nstates):
        self._check(gpu().fbse_apply_h_dev_noalloc(self.ctx, C.c_void_p(d_in), C.c_void_p(d_out), nstates))

    def filter_dev(self, d_psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=FBSE_PREC_FP64):
        an, zn = _f64(an), _f64(zn)
        self._check(gpu().fbse_filter_dev(self.ctx, C.c_void_p(d_psi_rank), n_starts, ms, len(zn), _dp(an),
                                          _dp(zn), Vmin, dE_1, precision))

    def normalize_all_dev(self, d_psi, n):
        self._check(gpu().fbse_normalize_all_dev(self.ctx, C.c_void_p(d_psi), n))

    def ortho_dev(self, d_psi, n_cols, svd_eps=1.0e-10):
        rank = C.c_long(0)
        sv = np.zeros(n_cols)
        self._check(gpu().fbse_ortho_dev(self.ctx, C.c_void_p(d_psi), n_cols, svd_eps, C.byref(rank), _dp(sv)))
        return rank.value, sv

    def rayleigh_ritz_dev(self, d_psi, n):
        ev = np.zeros(n)
        self._check(gpu().fbse_rayleigh_ritz_dev(self.ctx, C.c_void_p(d_psi), n, _dp(ev)))
        return ev

    def sigma_dev(self, d_psi, n):
        s, e = np.zeros(n), np.zeros(n)
        self._check(gpu().fbse_sigma_dev(self.ctx, C.c_void_p(d_psi), n, _dp(s), _dp(e)))
        return s, e

    def overlap_dev(self, d_a, n1, d_b, n2):
        """dv * A^H B (n1 x n2; complex for complex states) of two device state blocks."""
        cw = 2 if self.is_complex else 1
        s = np.zeros(n1 * n2 * cw)
        self._check(gpu().fbse_overlap_dev(self.ctx, C.c_void_p(d_a), n1, C.c_void_p(d_b), n2, _dp(s)))
        if self.is_complex:
            return s.view(np.complex128).reshape(n2, n1).T
        return s.reshape(n2, n1).T

    def rotate_dev(self, d_a, v):
        """In place: the first nout states of the device block become A V (V nin x nout)."""
        v = np.asarray(v)
        nin, nout = v.shape
        if self.is_complex:
            buf = np.ascontiguousarray(v.T, dtype=np.complex128).view(np.float64).ravel()
        else:
            buf = np.ascontiguousarray(v.T, dtype=np.float64).ravel()
        self._check(gpu().fbse_rotate_dev(self.ctx, C.c_void_p(d_a), nin, _dp(buf), nout))

    def refine(self, psi, n_ritz, eval_, sigma, e_lo, e_hi, e_p=1.0, sigma_max=1e-2):
        """One guarded FP64 refinement sweep (fbse_refine) of host Ritz vectors:
        returns (psi, eval, sigma, replaced), the pairs in their slots."""
        buf = _f64(psi)[: n_ritz * self.stlen].copy()
        ev, sg, n = np.zeros(n_ritz), np.zeros(n_ritz), C.c_long(0)
        self._check(gpu().fbse_refine(self.ctx, _dp(buf), n_ritz, _dp(_f64(eval_)), _dp(_f64(sigma)), sigma_max,
                                      e_lo, e_hi, e_p, _dp(ev), _dp(sg), C.byref(n)))
        return buf, ev, sg, n.value

    def refine_dev(self, d_psi, n_ritz, eval_, sigma, e_lo, e_hi, e_p=1.0, sigma_max=1e-2):
        """fbse_refine_dev: returns (eval, sigma, replaced) of the n_ritz pairs."""
        ev, sg, n = np.zeros(n_ritz), np.zeros(n_ritz), C.c_long(0)
        self._check(gpu().fbse_refine_dev(self.ctx, C.c_void_p(d_psi), n_ritz, _dp(_f64(eval_)), _dp(_f64(sigma)),
                                          sigma_max, e_lo, e_hi, e_p, _dp(ev), _dp(sg), C.byref(n)))
        return ev, sg, n.value

    def set_fft_kernel(self, variant, lines=0):
        self._check(gpu().fbse_set_fft_kernel(self.ctx, variant, lines))

    def set_graph(self, enable):
        """Newton-window hipGraph replay on (1, default) or off (0)."""
        self._check(gpu().fbse_set_graph(self.ctx, int(enable)))

    def set_axis_fusion(self, enable):
        """Fused per-axis combine (projector expansion and the next forward x
        transform inside the x combine pass) on (1, default) or off (0)."""
        self._check(gpu().fbse_set_axis_fusion(self.ctx, int(enable)))

    def set_combine_projection(self, enable):
        """Projections summed inside the spin-paired combine and reduced per
        atom by the next Newton step (1) or the projection pass every step (0,
        default)."""
        self._check(gpu().fbse_set_combine_projection(self.ctx, int(enable)))

    def trim(self):
        """Release the context's work buffers (allocated again on demand)."""
        self._check(gpu().fbse_trim(self.ctx))

    @staticmethod
    def _idx(idx):
        if idx is None:
            return None, None
        ix = np.ascontiguousarray(idx, dtype=np.int64)
        return ix, ix.ctypes.data_as(LP)

    def write_states(self, d_psi, n, path, idx=None, n_every=1):
        """States of the device block d_psi (idx order, or 0..n-1) to a new
        file at path in the reference's raw layout (fbse_write_states)."""
        ix, p = self._idx(idx)
        self._check(gpu().fbse_write_states(self.ctx, C.c_void_p(d_psi), p, len(ix) if ix is not None else n,
                                            n_every, os.fsencode(path)))

    def fwrite_states(self, d_psi, n, stream, idx=None, n_every=1):
        """fbse_fwrite_states to an open C stream (FILE *, as an integer)."""
        ix, p = self._idx(idx)
        self._check(gpu().fbse_fwrite_states(self.ctx, C.c_void_p(d_psi), p, len(ix) if ix is not None else n,
                                             n_every, C.c_void_p(stream)))

    def read_start_states(self, path, start, end, n_every, d_psi_rank, n_starts, ms):
        """Start states from a file into the filter's start blocks
        (fbse_read_start_states); returns the number placed."""
        n = C.c_long(0)
        self._check(gpu().fbse_read_start_states(self.ctx, os.fsencode(path), start, end, n_every,
                                                 C.c_void_p(d_psi_rank), n_starts, ms, C.byref(n)))
        return n.value

    def set_slab_split(self, split):
        """z-slab tiles per slab along x (0 = autotune)."""
        self._check(gpu().fbse_set_slab_split(self.ctx, split))

    def set_io_chunk(self, nbytes):
        self._check(gpu().fbse_set_io_chunk(self.ctx, nbytes))

    def held_bytes(self):
        """Bytes held between calls: (ring/work, second work buffer, partial Grams, rocBLAS workspace)."""
        out = (C.c_size_t * 4)()
        self._check(gpu().fbse_held_bytes(self.ctx, out))
        return tuple(int(v) for v in out)

    def last_window(self):
        """Accumulation window depth of the last filter call."""
        return gpu().fbse_last_window(self.ctx)

    def set_run_starts(self, n):
        """Start states per filter run (0: chosen by the filter)."""
        self._check(gpu().fbse_set_run_starts(self.ctx, n))

    def last_runs(self):
        """Sub-batch runs of the start states in the last filter call."""
        return gpu().fbse_last_runs(self.ctx)

    def last_group(self):
        """Dots that shared grouped FFT launches with this context in its last
        filter_many call (1: its own launches)."""
        return gpu().fbse_last_group(self.ctx)

    def graph_replays(self):
        """hipGraph replays of the last filter call."""
        return gpu().fbse_graph_replays(self.ctx)

    def state_density_dev(self, d_psi, idx, d_rho):
        """d_rho (ngrid doubles) = sum over the listed states and spins of |psi|^2."""
        ix = np.ascontiguousarray(idx, dtype=np.int64)
        self._check(gpu().fbse_state_density_dev(self.ctx, C.c_void_p(d_psi), ix.ctypes.data_as(C.POINTER(C.c_long)),
                                                 len(ix), C.c_void_p(d_rho)))

    def warm_start_dev(self, d_y, ld, channel, d_psi_rank, n_starts, ms, smooth_ep=0.0):
        """Start states jns of psi_rank: low-passed (smooth_ep > 0), times
        sqrt(exp(y[:, channel])), renormalised."""
        self._check(gpu().fbse_warm_start_dev(self.ctx, C.c_void_p(d_y), ld, channel, smooth_ep,
                                              C.c_void_p(d_psi_rank), n_starts, ms))

    def fft_choice(self, precision=FBSE_PREC_FP64):
        """(path, instance): path 1 = plane/slab passes, 0 = per-axis passes."""
        p, i = C.c_int(), C.c_int()
        self._check(gpu().fbse_fft_choice(self.ctx, precision, C.byref(p), C.byref(i)))
        return p.value, i.value

    def set_window(self, k):
        self._check(gpu().fbse_set_window(self.ctx, k))

    def profile(self, enable=True):
        self._check(gpu().fbse_profile(self.ctx, 1 if enable else 0))

    def profile_read(self, kind):
        n, ms = C.c_long(0), C.c_double(0)
        self._check(gpu().fbse_profile_read(self.ctx, kind, C.byref(n), C.byref(ms)))
        return n.value, ms.value


def filter_many(contexts, d_psi_ranks, n_starts, ms, ans, zns, Vmins, dE_1s, precision=FBSE_PREC_FP64):
    """fbse_filter_many: the Newton filters of several dots (one Context each,
    device pointers to each dot's psi_rank) in one call, their launches
    interleaved over the contexts' streams.  ms and the Newton length are
    common; an / zn / Vmin / dE_1 / n_starts are per dot."""
    n = len(contexts)
    if not (len(d_psi_ranks) == len(n_starts) == len(ans) == len(zns) == len(Vmins) == len(dE_1s) == n):
        raise ValueError("per-dot arguments must have one entry per context")
    ans = [_f64(a) for a in ans]
    zns = [_f64(z) for z in zns]
    M = len(zns[0])
    if any(len(z) != M for z in zns):
        raise ValueError("every dot needs the same Newton length")
    P = C.c_void_p * n
    ctxs = P(*[c.ctx for c in contexts])
    psis = P(*[C.c_void_p(p) for p in d_psi_ranks])
    a_p = P(*[a.ctypes.data for a in ans])
    z_p = P(*[z.ctypes.data for z in zns])
    ns = (C.c_long * n)(*n_starts)
    vm = (C.c_double * n)(*Vmins)
    de = (C.c_double * n)(*dE_1s)
    rc = gpu().fbse_filter_many(n, ctxs, psis, ns, ms, M, a_p, z_p, vm, de, precision)
    if rc != 0:
        raise FbseError(rc, gpu().fbse_last_error(contexts[0].ctx).decode())


class Mlp:
    """The warm-start network (include/fbse_mlp.h): n_in -> width x3 -> n_out,
    BF16 weights and activations, FP32 accumulation, FP32 master weights."""

    def __init__(self, n_in, width=128, n_out=2, seed=1, device=0):
        p = C.c_void_p()
        rc = gpu().fbse_mlp_create(device, n_in, width, n_out, seed, C.byref(p))
        if rc != 0:
            raise FbseError(rc, "fbse_mlp_create(%d, %d, %d)" % (n_in, width, n_out))
        self.m = p
        self.n_in, self.width, self.n_out = n_in, width, n_out
        self.k_in = gpu().fbse_mlp_input_width(self.m)
        self.n_params = gpu().fbse_mlp_num_params(self.m)

    def __del__(self):
        if getattr(self, "m", None) and _gpu is not None:
            try:
                _gpu.fbse_mlp_destroy(self.m)
            except Exception:  # interpreter shutdown
                pass
            self.m = None

    def _check(self, rc):
        if rc != 0:
            raise FbseError(rc, gpu().fbse_mlp_last_error(self.m).decode())

    def params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        self._check(gpu().fbse_mlp_get_params(self.m, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(gpu().fbse_mlp_set_params(self.m, p.ctypes.data_as(C.POINTER(C.c_float))))

    def layers(self, p=None):
        """[(W, b)] views of a packed parameter vector (W is fan_in x fan_out)."""
        p = self.params() if p is None else p
        dims = [(self.n_in, self.width), (self.width, self.width), (self.width, self.width), (self.width, self.n_out)]
        out, o = [], 0
        for k, n in dims:
            w = p[o:o + k * n].reshape(k, n)
            o += k * n
            b = p[o:o + n]
            o += n
            out.append((w, b))
        return out

    def forward_dev(self, d_x, n, d_y):
        self._check(gpu().fbse_mlp_forward_dev(self.m, C.c_void_p(d_x), n, C.c_void_p(d_y)))

    def loss_grad_dev(self, d_x, d_t, d_w, n, want_grads=False):
        loss = C.c_double(0)
        g = np.zeros(self.n_params, dtype=np.float32) if want_grads else None
        self._check(gpu().fbse_mlp_loss_grad_dev(self.m, C.c_void_p(d_x), C.c_void_p(d_t),
                                                 C.c_void_p(d_w) if d_w else None, n, C.byref(loss),
                                                 g.ctypes.data_as(C.POINTER(C.c_float)) if want_grads else None))
        return (loss.value, g) if want_grads else loss.value

    def adam_step(self, lr):
        self._check(gpu().fbse_mlp_adam_step(self.m, lr))

    def features_dev(self, nx, ny, nz, origin, dx, atoms_xyz, species, n_species, d_vloc, center, r_scale, d_x):
        xyz = _f64(atoms_xyz).reshape(-1)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        org, cen = _f64(origin), _f64(center)
        self._check(gpu().fbse_mlp_features_dev(self.m, nx, ny, nz, _dp(org), dx, _dp(xyz),
                                                sp.ctypes.data_as(C.POINTER(C.c_int)), len(sp), n_species,
                                                C.c_void_p(d_vloc), _dp(cen), r_scale, C.c_void_p(d_x)))
