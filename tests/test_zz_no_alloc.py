"""fbse_apply_h_dev_noalloc (the H application that allocates nothing, its
work space a static device array) against fbse_apply_h_dev on the same
inputs, through the C ABI: two calls in a row, the second with fewer states
(the static array keeps the first call's contents), for complex spinors with
the fused projector expansion, spinors with five radial projectors, and real
states with nonlocal projectors; and the host-side rejection of a call whose
work space would exceed the array."""
import numpy as np
import pytest

from systems import VARIANTS, gpu_context, random_states, small_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", ["spinor_so_nl_1p", "spinor_so_nl", "real_nl"])
def test_noalloc_matches_apply_h(variant):
    import torch
    s = small_system(**VARIANTS[variant])
    print("case no-alloc H %s" % variant, flush=True)
    g = gpu_context(s)
    stlen = g.stlen
    for n in (5, 2):
        psi = torch.from_numpy(random_states(s, n, seed=10 + n)).cuda()
        ref = torch.empty_like(psi)
        out = torch.full_like(psi, float("nan"))
        torch.cuda.synchronize()
        g.apply_h_dev(psi.data_ptr(), ref.data_ptr(), n)
        g.synchronize()
        g.apply_h_dev_noalloc(psi.data_ptr(), out.data_ptr(), n)
        g.synchronize()
        r, o = ref.cpu().numpy(), out.cpu().numpy()
        assert o.shape == (n * stlen,) and np.isfinite(o).all()
        err = np.abs(o - r).max() / np.abs(r).max()
        assert err < 1e-13, (variant, n, err)
    g.close()


def test_noalloc_rejects_oversize():
    import torch
    from filterbse_amd import api
    s = small_system(**VARIANTS["spinor_vloc"])
    g = gpu_context(s)
    per = s["nspin"] * s["nx"] * s["ny"] * s["nz"] * 16  # bytes of FFT work per state
    n = (640 << 20) // per + 1
    psi = torch.zeros(n * g.stlen, dtype=torch.float64, device="cuda")
    out = torch.zeros_like(psi)
    torch.cuda.synchronize()
    with pytest.raises(api.FbseError) as e:
        g.apply_h_dev_noalloc(psi.data_ptr(), out.data_ptr(), n)
    assert e.value.code == api.FBSE_ERR_ARG
    # the context still works afterwards, with and without allocation
    small = torch.from_numpy(random_states(s, 1, seed=3)).cuda()
    a, b = torch.empty_like(small), torch.empty_like(small)
    torch.cuda.synchronize()
    g.apply_h_dev(small.data_ptr(), a.data_ptr(), 1)
    g.apply_h_dev_noalloc(small.data_ptr(), b.data_ptr(), 1)
    g.synchronize()
    assert (a - b).abs().max().item() <= 1e-13 * a.abs().max().item()
    g.close()
