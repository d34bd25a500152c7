"""Wide orthogonalisation and Rayleigh-Ritz: the Gram products split over row
slabs (a remainder slab included) and formed block-upper-triangular for
n >= 512 columns must agree with LAPACK on the same inputs.

Grid 17 x 19 x 22 spinor: 14,212 complex rows, split into three slabs of
4,737 rows and a one-row remainder."""
import numpy as np
import pytest

from oracle import oracle as O
from systems import gpu_context, oracle_ham, small_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wide():
    s = small_system(n=(17, 19, 22), nspin=2, is_complex=1, so=1, natoms=2)
    print("case grid 17x19x22 spinor", flush=True)
    return s, oracle_ham(s), gpu_context(s)


def test_wide_ortho_and_rr(wide):
    s, h, g = wide
    stlen = 2 * s["nspin"] * s["nx"] * s["ny"] * s["nz"]
    n = 560
    rng = np.random.default_rng(4)
    A = rng.standard_normal((n, stlen))
    # two directions far below the rest: one kept (1e-8), one cut (1e-13)
    A[n - 2] = A[0] + 1e-8 * A[n - 2]
    A[n - 1] = A[1] + 1e-13 * A[n - 1]
    psi = A.ravel()
    out, k, sv = g.ortho(psi, n)
    _, sv_ref, k_ref = O.ortho(psi, n, stlen, 1, s["dv"])
    assert k == k_ref == n - 1, (k, k_ref, sv[-3:] / sv[0])
    assert np.allclose(sv[: n - 2], sv_ref[: n - 2], rtol=1e-9, atol=0)
    assert abs(sv[n - 2] / sv_ref[n - 2] - 1.0) < 1e-4
    U = O.as_matrix(out, k, stlen, 1)
    gram = s["dv"] * (U.conj().T @ U)
    assert np.abs(gram - np.eye(k)).max() < 1e-10

    m = 520
    HU = O.as_matrix(h.apply(out[: m * stlen]), m, stlen, 1)
    w_ref, _ = O.rayleigh_ritz(U[:, :m], HU, s["dv"])
    rot, w = g.rayleigh_ritz(out[: m * stlen], m)
    assert np.abs(w - w_ref).max() < 1e-10 * np.abs(w_ref).max()
    V = O.as_matrix(rot, m, stlen, 1)
    HV = O.as_matrix(h.apply(rot), m, stlen, 1)
    Hs = s["dv"] * (V.conj().T @ HV)
    assert np.abs(Hs - np.diag(w)).max() < 1e-9 * np.abs(w).max()


def test_tall_ortho_blocked_solve():
    """A block taller than one row block of the Cholesky re-orthogonalisation's
    triangular solve (262,144 rows per solve): 64 x 64 x 48 spinor, 393,216
    complex rows, two blocks.  Rank and singular values against LAPACK, and an
    orthonormal basis of the same span."""
    s = small_system(n=(64, 64, 48), nspin=2, is_complex=1)
    print("case grid 64x64x48 spinor, 24 columns", flush=True)
    g = gpu_context(s)
    stlen = 2 * s["nspin"] * s["nx"] * s["ny"] * s["nz"]
    n = 24
    rng = np.random.default_rng(8)
    A = rng.standard_normal((n, stlen))
    A[n - 1] = A[2] + 1e-9 * A[n - 1]
    psi = A.ravel()
    out, k, sv = g.ortho(psi, n)
    _, sv_ref, k_ref = O.ortho(psi, n, stlen, 1, s["dv"])
    assert k == k_ref == n
    assert np.allclose(sv[: n - 1], sv_ref[: n - 1], rtol=1e-9, atol=0)
    U = O.as_matrix(out, k, stlen, 1)
    assert np.abs(s["dv"] * (U.conj().T @ U) - np.eye(k)).max() < 1e-10
    # same span: the projection of the inputs onto U reproduces them
    Am = O.as_matrix(psi, n, stlen, 1)
    resid = Am - U @ (s["dv"] * (U.conj().T @ Am))
    assert np.abs(resid).max() < 1e-8 * np.abs(Am).max()
    g.close()
