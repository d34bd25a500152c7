"""CPU tests of the host-side library (include/fbse_host.h) against the oracle
(bit-exact: integer, index and identically-ordered FP64 work), and of the
C-ABI libraries' exports (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from filterbse_amd import api
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fbse_[A-Za-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("headers,path", [(("fbse_gpu.h", "fbse_mlp.h", "fbse_io.h"), api.GPU_LIB),
                                          (("fbse_host.h",), api.HOST_LIB)])
def test_library_exports_every_declared_symbol(headers, path):
    names = sorted(set(n for h in headers for n in declared(h)))
    for h in headers:
        assert declared(h), h
    lib = ctypes.CDLL(path)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    sigs = api.GPU_SIGNATURES if path == api.GPU_LIB else api.HOST_SIGNATURES
    assert sorted(sigs) == names


def test_gpu_library_fails_loudly_without_device_or_bad_args():
    # fbse_create with a NULL descriptor is an argument error, not a fallback
    out = ctypes.c_void_p()
    assert api.gpu().fbse_create(None, None, -1, ctypes.byref(out)) == -1


@pytest.mark.parametrize("M", [64, 512, 2048])
def test_leja_bit_exact(M):
    assert np.array_equal(api.leja_points(M), O.leja_points(M))


@pytest.mark.parametrize("nspin,cplx", [(1, 0), (2, 1), (1, 1)])
def test_start_states_bit_exact(nspin, cplx):
    a, sa = api.init_filter_states(3, 4, -11, 9, 10, 12, nspin, cplx, 0.27)
    b, sb = O.init_filter_states(3, 4, -11, 9, 10, 12, nspin, cplx, 0.27)
    assert sa == sb and np.array_equal(a, b)


def test_energy_range_start_vectors_follow_fixed_seed():
    p0, p1 = api.energy_range_start_states(6, 7, 8, 2, 1, 0.4)
    seed = -874917403
    ref = []
    for _ in range(2):
        comps = []
        for _s in range(2):
            c, seed = O.init_psi(seed, 6, 7, 8, 1, 0.4)
            comps.append(c.view(np.float64))
        ref.append(np.concatenate(comps))
    assert np.array_equal(p0, ref[0]) and np.array_equal(p1, ref[1])


def test_grid_and_ksqr_bit_exact():
    rng = np.random.default_rng(3)
    R = rng.uniform(-9, 9, size=(40, 3))
    for dx in (0.5, 0.7):
        n1, m1, x1, k1 = api.grid_params(R, (8, 8, 8), (dx, dx, dx))
        n2, m2, x2, k2 = O.init_grid_params(R, (8, 8, 8), (dx, dx, dx))
        assert np.array_equal(n1, n2) and np.array_equal(m1, m2) and np.array_equal(x1, x2) and np.array_equal(k1, k2)
        a = api.grid_ksqr(n1, (dx, dx, dx), m1, k1, 20.0)
        b = O.build_grid_ksqr(n2, (dx, dx, dx), m2, k2, 20.0)
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


def test_newton_coefficients_bit_exact():
    z = api.leja_points(1024)
    t = api.ene_targets(10, 6, -0.5, -0.2, -0.167, -0.1)
    dE = 26.0193
    dt = (1024 / (2.5 * dE)) ** 2
    assert np.array_equal(api.newton_coeff(z, t, -0.756627, dE, dt), O.newton_coeff(z, t, -0.756627, dE, dt))
    assert np.array_equal(t, O.ene_targets(10, 6, -0.5, -0.2, -0.167, -0.1))


def test_ls_bit_exact():
    assert np.array_equal(api.def_LS(), O.def_LS())
