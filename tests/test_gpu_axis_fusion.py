"""The fused per-axis combine (fbse_set_axis_fusion): on x axes of length 160
and 256 the Newton filter's x combine pass also adds the projector expansion
of its rows and writes the forward x transform the next step starts from.
With every per-axis pass forced to the fixed 16-line instance, the fused and
the separate passes run the same arithmetic in the same order, so the
filtered states must be bit-identical; the FP64 filter must also match the
oracle.  Windows are forced short so that graph replays carry the chained
forward transform from one window into the next.  Grids: 160 x 32 x 48 (the
combine's row blocks per item, 96, a multiple of 8: XCD tile order on) and
256 x 16 x 20 (20 row blocks: plain order), spinors with SO + NL projectors
and packed real states with NL projectors.  Spinor combine tiles hold 8 rows
of both spins of one state by default (fusion 1) or 16 consecutive rows of
one grid (fusion 16): both must equal the separate passes bit for bit.  On
160^3 the default mode also runs the step's y and z passes in the z-tile
layout (y forward into the step's ring slot, z there in contiguous tiles, y
backward back): again bit for bit.  These comparisons run with the
projections in their own pass (the default); with the projections inside the
spin-paired combine (fbse_set_combine_projection: per-tile sums, reduced per
atom by the next step) only the summation order of the projections changes,
so the filtered states agree with the separate passes to rounding (relative
1e-12 in FP64, 1e-5 in FP32), and the FP64 filter matches the oracle."""
import numpy as np
import pytest

from systems import gpu_context, oracle_ham, small_system
from test_gpu_parity import filter_inputs

pytestmark = pytest.mark.gpu

CASES = {
    "spinor_160x32x48": dict(n=(160, 32, 48), dx=0.5, nspin=2, is_complex=1, so=1, nl=1, natoms=6, n_so_proj=1,
                             n_nl_proj=1),
    "spinor_256x16x20": dict(n=(256, 16, 20), dx=0.4, nspin=2, is_complex=1, so=1, nl=1, natoms=5, n_so_proj=1,
                             n_nl_proj=1),
    "real_256x16x20": dict(n=(256, 16, 20), dx=0.4, nspin=1, is_complex=0, so=0, nl=1, natoms=5, n_nl_proj=1),
    # y and z of length 160: the default mode runs the y and z passes in the
    # z-tile layout (no oracle run at this size: the modes against each other)
    "spinor_160x160x160": dict(n=(160, 160, 160), dx=0.5, nspin=2, is_complex=1, so=1, nl=1, natoms=6, n_so_proj=1,
                               n_nl_proj=1),
}
ORACLE_MAX_POINTS = 64 ** 3


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("precision", [0, 1])
def test_fused_combine_identical(case, precision):
    s = small_system(**CASES[case])
    print("case axis fusion %s precision %d" % (case, precision), flush=True)
    g = gpu_context(s)
    g.set_fft_kernel(0, 1016)
    g.set_window(4)
    n_starts, ms, M = 3, 4, 40
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, n_starts, ms, M)
    g.set_combine_projection(0)
    g.set_axis_fusion(0)
    sep = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    g.set_axis_fusion(1)
    fused = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    assert g.graph_replays() > 0
    assert np.array_equal(fused, sep)
    # spinors: the default tiles pair 8 rows of both spins; 16 consecutive rows of one grid
    g.set_axis_fusion(16)
    rows16 = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    g.set_axis_fusion(1)
    assert np.array_equal(rows16, sep)
    # projections summed inside the spin-paired combine
    g.set_combine_projection(1)
    in_combine = g.filter(psi_rank, n_starts, ms, an, zn, Vmin, dE_1, precision=precision)
    assert g.graph_replays() > 0
    rel = np.abs(in_combine - sep).max() / np.abs(sep).max()
    print("projections in the combine vs their own pass: %.2e" % rel)
    assert rel < (1e-12 if precision == 0 else 1e-5)
    if CASES[case]["nspin"] == 1:
        assert np.array_equal(in_combine, sep)  # no spin-paired tiles: the projection pass runs
    if precision == 0 and np.prod(CASES[case]["n"]) <= ORACLE_MAX_POINTS:
        ref = oracle_ham(s).filter_cycles(psi_rank, n_starts, ms, an, zn, Vmin, dE_1)
        for name, got in (("fused", fused), ("projections in the combine", in_combine)):
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print("%s FP64 filter vs oracle: %.2e" % (name, err))
            assert err < 1e-10
    g.set_combine_projection(0)
    g.close()
