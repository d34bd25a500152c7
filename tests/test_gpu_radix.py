"""Plane/slab FFT passes on grids whose radix schedules use every composite
butterfly (6, 9, 10, 12, 15, 16 as well as 2, 3, 4, 5, 8): H from each FP64
plane/slab kernel instance, with z slabs whole or cut into 2 or 3 column
tiles, must equal H from the per-axis passes (which the parity tests hold to
the oracle) to FFT rounding, and the oracle itself.

Schedules (fewest passes, ties to the smaller largest radix):
  12 staged values: 60 -> 12*5, 45 -> 15*3, 80 -> 16*5, 96 -> 12*8,
                    50 -> 5*5*2, 36 -> 6*6
  16 staged values: 60 -> 10*6, 45 -> 9*5, 80 -> 10*8, 96 -> 16*6,
                    50 -> 10*5, 36 -> 6*6
"""
import numpy as np
import pytest

from systems import gpu_context, oracle_ham, random_states, small_system

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("grid", [(60, 45, 80), (96, 50, 36)])
def test_tile_radices(grid):
    s = small_system(n=grid, dx=0.5, nspin=2, is_complex=1)
    print("case grid %s" % (grid,), flush=True)
    g = gpu_context(s)
    psi = random_states(s, 2, seed=5)
    g.set_fft_kernel(0, 0)
    ref = g.apply_h(psi)
    assert rel(ref, oracle_ham(s).apply(psi)) < 1e-11
    for inst in (1, 2):
        g.set_fft_kernel(1, inst)
        for split in (1, 2, 3):  # z-slab tiles of 60, 30, 20 (96: 96, 48, 32) columns
            g.set_slab_split(split)
            out = g.apply_h(psi)
            assert rel(out, ref) < 1e-12, (inst, split)
    g.set_slab_split(0)


@pytest.mark.parametrize("n", [64, 96])
def test_fixed_instances(n):
    """The fixed-schedule persistent instances of cubic 64^3 and 96^3 grids
    (FP64 instance 3, FP32 instance 4) against the generic instance 1: H in
    FP64 to FFT rounding, and the mixed filter (FP32 recurrence through the
    fixed forward, z-slab and combine passes with the projector expansion)
    to FP32 rounding."""
    s = small_system(n=(n, n, n), dx=0.5, nspin=2, is_complex=1, so=1, nl=1, natoms=6,
                     n_so_proj=1, n_nl_proj=1)
    print("case fixed instance %d^3" % n, flush=True)
    g = gpu_context(s)
    psi = random_states(s, 2, seed=7)
    g.set_fft_kernel(1, 1)
    ref = g.apply_h(psi)
    g.set_fft_kernel(1, 3)
    out = g.apply_h(psi)
    assert rel(out, ref) < 1e-12
    # mixed filter: 2 starts x 2 targets, 12 Newton steps
    from test_gpu_parity import filter_inputs
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, 2, 2, 12)
    g.set_fft_kernel(1, 1)
    fref = g.filter(psi_rank, 2, 2, an, zn, Vmin, dE_1, precision=1)
    g.set_fft_kernel(1, 4)
    fout = g.filter(psi_rank, 2, 2, an, zn, Vmin, dE_1, precision=1)
    g.set_fft_kernel(-1, 0)
    assert rel(fout, fref) < 1e-5
    g.close()


@pytest.mark.parametrize("grid", [(256, 160, 12), (160, 12, 256)])
def test_fixed_axis_instances(grid):
    """The fixed-schedule per-axis instances (axis lengths 160 and 256, 16,
    32 or 64 lines per tile: x rows forward and combine, y lines forward and
    backward, z lines with k^2) against the generic per-axis kernel: H in FP64
    to FFT rounding, the mixed filter to FP32 rounding."""
    s = small_system(n=grid, dx=0.5, nspin=2, is_complex=1, so=1, nl=1, natoms=6, n_so_proj=1, n_nl_proj=1)
    print("case fixed axis instances %s" % (grid,), flush=True)
    g = gpu_context(s)
    psi = random_states(s, 2, seed=9)
    g.set_fft_kernel(0, 16)
    ref = g.apply_h(psi)
    assert rel(ref, oracle_ham(s).apply(psi)) < 1e-11
    from test_gpu_parity import filter_inputs
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, 2, 2, 12)
    fref = g.filter(psi_rank, 2, 2, an, zn, Vmin, dE_1, precision=1)
    # refinement: residuals, the k-space preconditioner pass (z lines), Grams, RR
    u, _, _ = g.ortho(random_states(s, 4, seed=10), 4)
    u, ev0 = g.rayleigh_ritz(u, 4)
    sg0, _ = g.sigma(u, 4)
    _, _, rref, _ = g.refine(u, 4, ev0, sg0, -1e3, 1e3, sigma_max=1e3)
    for lines in (1016, 1032, 1064):  # 1064: FP32 fixed instance, FP64 generic 64-line tiles
        g.set_fft_kernel(0, lines)
        out = g.apply_h(psi)
        assert rel(out, ref) < 1e-12, lines
        fout = g.filter(psi_rank, 2, 2, an, zn, Vmin, dE_1, precision=1)
        assert rel(fout, fref) < 1e-5, lines
        _, _, r, _ = g.refine(u, 4, ev0, sg0, -1e3, 1e3, sigma_max=1e3)
        assert np.abs(r - rref).max() < 1e-10 * np.abs(rref).max(), lines
    g.set_fft_kernel(-1, 0)
    g.close()


@pytest.mark.parametrize("n", [60, 72, 80, 90, 100, 108, 120])
def test_fixed_instances_fp32_sizes(n):
    """The FP32 fixed-schedule instances of the other cubic sizes (the uneven
    dots' grids: radices 10*6, 9*8, 10*8, 10*9, 10*10, 12*9, 15*8): the mixed
    filter through the fixed forward, z-slab and combine passes (projector
    expansion included) against the generic per-axis passes, to FP32 rounding."""
    s = small_system(n=(n, n, n), dx=0.5, nspin=2, is_complex=1, so=1, nl=1, natoms=6,
                     n_so_proj=1, n_nl_proj=1)
    print("case fixed FP32 instance %d^3" % n, flush=True)
    g = gpu_context(s)
    from test_gpu_parity import filter_inputs
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, 2, 2, 12)
    g.set_fft_kernel(0, 16)
    fref = g.filter(psi_rank, 2, 2, an, zn, Vmin, dE_1, precision=1)
    g.set_fft_kernel(1, 4)
    fout = g.filter(psi_rank, 2, 2, an, zn, Vmin, dE_1, precision=1)
    g.set_fft_kernel(-1, 0)
    assert rel(fout, fref) < 1e-5
    g.close()
