"""The warm-start network's GPU kernels (BF16 MFMA forward, backward, Adam)
and its grid helpers against the CPU oracle (oracle/mlp_oracle.py)."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from systems import gpu_context, small_system

pytestmark = pytest.mark.gpu


def to_dev_bf16(x, k_in):
    import torch
    xp = np.zeros((x.shape[0], k_in), dtype=np.float32)
    xp[:, :x.shape[1]] = x
    bits = MO.bf16_bits(xp)
    return torch.from_numpy(bits.view(np.int16)).cuda(), MO.bf16(xp)[:, :x.shape[1]]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("width,n", [(128, 1000), (64, 77)])
def test_forward_matches_oracle(width, n):
    import torch
    from filterbse_amd import api
    n_in, n_out = 34, 2
    m = api.Mlp(n_in, width, n_out, seed=3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((n, n_in)).astype(np.float32)
    dx, xq = to_dev_bf16(x, m.k_in)
    y = torch.empty((n, n_out), dtype=torch.float32, device="cuda")
    m.forward_dev(dx.data_ptr(), n, y.data_ptr())
    torch.cuda.synchronize()
    ref = MO.forward(MO.unpack(m.params(), n_in, width, n_out), xq)[0]
    assert rel(y.cpu().numpy(), ref) < 1e-2


def test_gradients_match_oracle():
    import torch
    from filterbse_amd import api
    n_in, width, n_out, n = 34, 128, 2, 3000
    m = api.Mlp(n_in, width, n_out, seed=5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((n, n_in)).astype(np.float32)
    t = rng.standard_normal((n, n_out)).astype(np.float32)
    w = rng.uniform(0.2, 1.0, (n, n_out)).astype(np.float32)
    dx, xq = to_dev_bf16(x, m.k_in)
    dt, dw = torch.from_numpy(t).cuda(), torch.from_numpy(w).cuda()
    loss, g = m.loss_grad_dev(dx.data_ptr(), dt.data_ptr(), dw.data_ptr(), n, want_grads=True)
    layers = MO.unpack(m.params(), n_in, width, n_out)
    loss_ref, gref = MO.loss_grad(layers, xq, t, w)
    assert abs(loss - loss_ref) < 1e-3 * loss_ref
    gl = MO.unpack(g, n_in, width, n_out)
    for l in range(4):
        for a, b in zip(gl[l], gref[l]):
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            assert err < 2e-2, (l, err)


def test_training_reduces_loss():
    import torch
    from filterbse_amd import api
    n_in, n = 34, 8192
    m = api.Mlp(n_in, 128, 2, seed=7)
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (n, n_in)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] ** 2, np.exp(-x[:, 2] ** 2) - 0.5 * x[:, 3]], axis=1)
    dx, _ = to_dev_bf16(x, m.k_in)
    dt = torch.from_numpy(t.astype(np.float32)).cuda()
    first = m.loss_grad_dev(dx.data_ptr(), dt.data_ptr(), None, n)
    for it in range(400):
        loss = m.loss_grad_dev(dx.data_ptr(), dt.data_ptr(), None, n)
        m.adam_step(3e-3)
    last = m.loss_grad_dev(dx.data_ptr(), dt.data_ptr(), None, n)
    print("loss", first, "->", last)
    assert last < 0.05 * first


def test_features_match_oracle():
    import torch
    from filterbse_amd import api
    m = api.Mlp(34, 32, 2, seed=1)
    rng = np.random.default_rng(9)
    n, dx, origin = (16, 18, 20), 0.6, (-4.5, -5.1, -5.7)
    atoms = rng.uniform(-4, 4, (23, 3))
    species = rng.integers(0, 4, 23)
    vloc = rng.standard_normal(n[0] * n[1] * n[2])
    d_v = torch.from_numpy(vloc).cuda()
    ng = n[0] * n[1] * n[2]
    d_x = torch.empty((ng, m.k_in), dtype=torch.int16, device="cuda")
    m.features_dev(*n, origin, dx, atoms, species, 4, d_v.data_ptr(), (0.3, -0.2, 0.1), 5.0, d_x.data_ptr())
    got = (d_x.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    ref = MO.features(n, origin, dx, atoms, species, 4, vloc, (0.3, -0.2, 0.1), 5.0, m.k_in)
    assert np.all(np.abs(got - ref) <= 2 ** -7 * np.abs(ref) + 1e-6)


def test_density_and_warm_start():
    import torch
    from oracle import oracle as O
    s = small_system(nspin=2, is_complex=1)
    g = gpu_context(s)
    ng = s["nx"] * s["ny"] * s["nz"]
    stlen = 2 * 2 * ng
    ns, ms = 3, 2
    psi, _ = O.init_filter_states(ns, ms, -4, s["nx"], s["ny"], s["nz"], 2, 1, s["dv"])
    d = torch.from_numpy(psi.copy()).cuda()
    rho = torch.zeros(ng, dtype=torch.float64, device="cuda")
    g.state_density_dev(d.data_ptr(), [0, 2, 5], rho.data_ptr())
    P = psi.reshape(ns * ms, 2, ng, 2)
    want = sum((P[k] ** 2).sum(axis=(0, 2)) for k in (0, 2, 5))
    assert np.allclose(rho.cpu().numpy(), want, rtol=1e-13, atol=0)
    rng = np.random.default_rng(2)
    y = rng.uniform(-40, 5, (ng, 3)).astype(np.float32)
    dy = torch.from_numpy(y).cuda()
    g.warm_start_dev(dy.data_ptr(), 3, 1, d.data_ptr(), ns, ms)
    out = d.cpu().numpy().reshape(ns, ms, 2, ng, 2)
    env = np.exp(0.5 * np.clip(y[:, 1].astype(np.float64), -30, 30))
    for j in range(ns):
        v = psi.reshape(ns, ms, 2, ng, 2)[j, 0] * env[None, :, None]
        v /= np.sqrt((v ** 2).sum() * s["dv"])
        assert np.allclose(out[j, 0], v, rtol=1e-12, atol=1e-14)
        assert np.array_equal(out[j, 1], psi.reshape(ns, ms, 2, ng, 2)[j, 1])
