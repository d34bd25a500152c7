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
