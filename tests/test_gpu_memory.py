"""Device memory around the filter: the ring of Newton terms is returned when
the context's work buffers are released, and the reporting entry points say
what the filter took and what the context holds."""
import pytest

from systems import gpu_context, small_system
from test_gpu_parity import filter_inputs

pytestmark = pytest.mark.gpu


def test_ring_returned_after_graph_filter():
    """Regression: on the per-axis path the projection pass used to be forked
    onto the side stream also inside the window-graph capture, and the
    captured fork kept the filter ring's memory allocated after hipFree (the
    whole ring was lost per graph-replayed filter).  Two graph-replayed mixed
    filters (per-axis passes, one radial projector per channel, window 40 ->
    a ring of 42 slots of 16.8 MB) each followed by fbse_trim leave the
    device's free memory where it was."""
    import torch

    s = small_system(n=(64, 64, 64), dx=0.5, nspin=2, is_complex=1, so=1, nl=1, natoms=6,
                     n_so_proj=1, n_nl_proj=1)
    print("case 64^3 spinor SO+NL, per-axis passes, window 40, M 130", flush=True)
    g = gpu_context(s)
    g.set_fft_kernel(0, 0)
    g.set_window(40)
    psi_rank, stlen, an, zn, Vmin, dE_1 = filter_inputs(s, 4, 2, 130)
    g.filter(psi_rank, 4, 2, an, zn, Vmin, dE_1, precision=1)  # autotuners, tables
    assert g.last_window() == 40
    held = g.held_bytes()
    assert held[0] >= 42 * 4 * 2 * 64 ** 3 * 8  # the ring (+ work slot) in the first work buffer
    g.trim()
    assert g.held_bytes()[:3] == (0, 0, 0)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(2):
        g.filter(psi_rank, 4, 2, an, zn, Vmin, dE_1, precision=1)
        assert g.graph_replays() == 2
        g.trim()
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 >= free0 - (64 << 20), (free0, free1)
    g.set_fft_kernel(-1, 0)
    g.set_window(0)
    g.close()
