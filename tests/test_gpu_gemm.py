"""FP64 matrix-core subspace products (gemm_kernels.hip) through the C ABI:
fbse_overlap_dev (dv * A^H B, the dotp / dotpreal matrix of diag_H,
Hmat.c:141-164, 301-338) and fbse_rotate_dev (the change of basis of diag_H,
Hmat.c:230-281), against numpy's products in FP64.  Shapes cover partial
64 x 64 tiles, row counts that are not multiples of the 16-row slices, the
split-K path (many row slabs, partial Grams summed in a fixed order) and real
as well as complex states."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(n, nspin, is_complex, dv=0.3):
    from filterbse_amd import api
    ng = n[0] * n[1] * n[2]
    return api.Context(n[0], n[1], n[2], nspin, is_complex, dv, np.zeros(ng), np.zeros(nspin * ng), None)


def _block(ctx, ncols, seed):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randn(ncols * ctx.stlen, dtype=torch.float64, device="cuda", generator=g)


def _np(ctx, d, ncols):
    a = d.cpu().numpy().reshape(ncols, -1)
    if ctx.is_complex:
        a = a.view(np.complex128)
    return a.T  # column i = state i


@pytest.mark.parametrize("n,nspin,cplx,n1,n2", [
    ((10, 11, 13), 2, 1, 70, 130),     # complex spinors, partial tiles, m = 2860
    ((10, 11, 13), 1, 0, 65, 3),       # real states
    ((40, 40, 41), 2, 1, 129, 64),     # m = 131200: split-K over row slabs
    ((40, 40, 41), 1, 0, 200, 200),    # real, square, split-K
])
def test_overlap(n, nspin, cplx, n1, n2):
    import torch
    ctx = _ctx(n, nspin, cplx)
    a = _block(ctx, n1, 1)
    b = _block(ctx, n2, 2)
    torch.cuda.synchronize()
    s = ctx.overlap_dev(a.data_ptr(), n1, b.data_ptr(), n2)
    A, B = _np(ctx, a, n1), _np(ctx, b, n2)
    ref = ctx.dv * (A.conj().T @ B)
    err = np.abs(s - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err
    ctx.close()


@pytest.mark.parametrize("n,nspin,cplx,nin,nout", [
    ((10, 11, 13), 2, 1, 90, 70),
    ((10, 11, 13), 1, 0, 33, 33),
    ((40, 40, 41), 2, 1, 130, 129),
])
def test_rotate(n, nspin, cplx, nin, nout):
    import torch
    ctx = _ctx(n, nspin, cplx)
    a = _block(ctx, nin, 3)
    A = _np(ctx, a, nin).copy()
    rng = np.random.default_rng(4)
    V = rng.standard_normal((nin, nout))
    if cplx:
        V = V + 1j * rng.standard_normal((nin, nout))
    torch.cuda.synchronize()
    ctx.rotate_dev(a.data_ptr(), V)
    out = _np(ctx, a, nin)
    ref = A @ V
    err = np.abs(out[:, :nout] - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err
    # columns beyond nout are left as they were
    assert np.array_equal(out[:, nout:], A[:, nout:])
    ctx.close()


def test_hermitian_gram_via_ortho():
    """ortho's Grams take the upper-tile path (A^H A): an orthonormal result
    (U^H U dv = I) from random columns, several tile rows of 64."""
    import torch
    ctx = _ctx((24, 24, 25), 2, 1)
    n = 200
    a = _block(ctx, n, 5)
    torch.cuda.synchronize()
    rank, sv = ctx.ortho_dev(a.data_ptr(), n)
    assert rank == n
    s = ctx.overlap_dev(a.data_ptr(), n, a.data_ptr(), n)
    assert np.abs(s - np.eye(n)).max() < 1e-12
    ctx.close()
