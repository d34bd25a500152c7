"""State files between HBM and disk (include/fbse_io.h) against the oracle's
restatement of the reference's writer (write.c:289-299, mod_output.c:81-100,
save.c:492) and reader (aux.c:131-156) with the start-state placement the
filter reads (filter.c:253): byte-exact, with staging chunks smaller than a
state so that chunks split states and both buffers cycle."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import io_oracle as IO
from systems import gpu_context, small_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    s = small_system(nspin=2, is_complex=1)
    print("case io spinor 16x18x20", flush=True)
    g = gpu_context(s)
    g.set_io_chunk(64 * 1024)  # a state is 184,320 bytes: three chunks, the last partial
    return s, g, 2 * 2 * 16 * 18 * 20


def test_write_states(ctx, tmp_path):
    import torch
    s, g, stlen = ctx
    rng = np.random.default_rng(3)
    psi = rng.standard_normal((7, stlen))
    d = torch.from_numpy(psi).cuda()
    torch.cuda.synchronize()
    cases = [(None, 1), ([4, 0, 6, 2], 1), ([5, 1, 3], 2), ([], 1)]
    for idx, n_every in cases:
        path = tmp_path / "psi.dat"
        g.write_states(d.data_ptr(), 7, str(path), idx=idx, n_every=n_every)
        assert path.read_bytes() == IO.write_psi_dat(psi, stlen, idx=idx, n_every=n_every), (idx, n_every)


def test_fwrite_states_after_header(ctx, tmp_path):
    """output.dat pattern: a header from fprintf, the psitot block, a trailer
    (save.c:446-503), on one buffered C stream."""
    import torch
    s, g, stlen = ctx
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fputs.argtypes = [C.c_char_p, C.c_void_p]
    libc.fclose.argtypes = [C.c_void_p]
    psi = np.random.default_rng(4).standard_normal((3, stlen))
    d = torch.from_numpy(psi).cuda()
    torch.cuda.synchronize()
    path = tmp_path / "output.dat"
    f = libc.fopen(os.fsencode(str(path)), b"w")
    assert f
    libc.fputs(b"12345\n5760 11520\n", f)
    g.fwrite_states(d.data_ptr(), 3, f)
    libc.fputs(b"\n\nEOF\n", f)
    assert libc.fclose(f) == 0
    assert path.read_bytes() == b"12345\n5760 11520\n" + IO.write_psi_dat(psi, stlen) + b"\n\nEOF\n"


def test_read_start_states(ctx, tmp_path):
    import torch
    from filterbse_amd import api
    s, g, stlen = ctx
    states = np.random.default_rng(5).standard_normal((7, stlen))
    raw = IO.write_psi_dat(states, stlen)
    path = tmp_path / "psi-init.dat"
    path.write_bytes(raw)
    ms = 3
    for start, end, n_every, n_starts in [(0, 6, 2, 5), (2, 6, 1, 3), (1, 5, 2, 4), (0, 7, 1, 2)]:
        d = torch.full((n_starts * ms * stlen,), 7.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        n = g.read_start_states(str(path), start, end, n_every, d.data_ptr(), n_starts, ms)
        sel = IO.read_psi_dat(raw, start, end, stlen, n_every)[:n_starts]
        assert n == len(sel)
        want = IO.start_blocks(sel, n_starts, ms, stlen).reshape(n_starts, ms, stlen)
        got = d.cpu().numpy().reshape(n_starts, ms, stlen)
        assert np.array_equal(got[:n, 0], want[:n, 0]), (start, end, n_every)
        assert (got[:, 1:] == 7.0).all() and (got[n:] == 7.0).all()  # nothing else is touched
    d = torch.zeros(8 * stlen, dtype=torch.float64, device="cuda")
    for start, end, n_every in [(5, 12, 1), (0, 7, 2)]:  # states 5, 6, 7 / 1, 3, 5, 7: the file holds 0..6
        with pytest.raises(api.FbseError) as ei:
            g.read_start_states(str(path), start, end, n_every, d.data_ptr(), 4, 1)
        assert ei.value.code == api.FBSE_ERR_IO
    with pytest.raises(api.FbseError) as ei:
        g.read_start_states(str(tmp_path / "missing.dat"), 0, 2, 1, d.data_ptr(), 1, 1)
    assert ei.value.code == api.FBSE_ERR_IO


def test_written_start_states_feed_the_filter_layout(ctx, tmp_path):
    """States written with n_every = 2 come back, through the spinor reader
    stride, one per start block."""
    import torch
    s, g, stlen = ctx
    psi = np.random.default_rng(6).standard_normal((4, stlen))
    d = torch.from_numpy(psi).cuda()
    torch.cuda.synchronize()
    path = tmp_path / "psi-init.dat"
    g.write_states(d.data_ptr(), 4, str(path), n_every=2)
    ms = 2
    r = torch.zeros(4 * ms * stlen, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert g.read_start_states(str(path), 0, 8, 2, r.data_ptr(), 4, ms) == 4
    got = r.cpu().numpy().reshape(4, ms, stlen)
    assert np.array_equal(got[:, 0], psi) and not got[:, 1].any()
