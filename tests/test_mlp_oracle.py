"""Pins of the warm-start network's CPU oracle (oracle/mlp_oracle.py), no GPU:
BF16 rounding against a brute-force nearest-representable search, the
backward pass against central finite differences of the loss (float64, no
rounding), and the features against hand-computed values."""
import numpy as np

from oracle import mlp_oracle as MO


def brute_bf16(x):
    """Nearest BF16 value by comparing the two candidates around x (ties to even)."""
    out = np.empty_like(x)
    for i, v in enumerate(x):
        u = int(np.float32(v).view(np.uint32))
        lo = (u >> 16) << 16
        hi = lo + (1 << 16)
        flo = float(np.uint32(lo).view(np.float32))
        fhi = float(np.uint32(hi & 0xFFFFFFFF).view(np.float32))
        dl, dh = abs(float(v) - flo), abs(fhi - float(v))
        if dl < dh or (dl == dh and ((lo >> 16) & 1) == 0):
            out[i] = flo
        else:
            out[i] = fhi
    return out


def test_bf16_rounding():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(2000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 2000),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0])])
    assert np.array_equal(MO.bf16(x), brute_bf16(x))
    assert MO.bf16(np.float32(1.0 + 2 ** -8)) == 1.0            # tie to even (down)
    assert MO.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.015625   # tie to even (up)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    n_in, width, n_out, n = 3, 5, 2, 7
    dims = [(n_in, width), (width, width), (width, width), (width, n_out)]
    layers = [(rng.standard_normal(d) * 0.7, rng.standard_normal(d[1]) * 0.3) for d in dims]
    x = rng.standard_normal((n, n_in))
    t = rng.standard_normal((n, n_out))
    w = rng.uniform(0.5, 2.0, (n, n_out))
    loss, grads = MO.loss_grad(layers, x, t, w, emulate=False)
    p = MO.pack(layers)
    g = MO.pack(grads)
    eps = 1e-6
    num = np.zeros_like(p)
    for i in range(len(p)):
        for sgn in (1, -1):
            q = p.copy()
            q[i] += sgn * eps
            num[i] += sgn * MO.loss_grad(MO.unpack(q, n_in, width, n_out), x, t, w, emulate=False)[0]
        num[i] /= 2 * eps
    assert np.abs(num - g).max() < 1e-6 * max(1.0, np.abs(g).max()), np.abs(num - g).max()
    assert loss > 0


def test_features_hand_values():
    atoms = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [0.0, 0.0, 4.0]])
    species = np.array([0, 1, 0])
    n, origin, dx = (3, 2, 2), (0.0, 0.0, 0.0), 1.0
    vloc = np.arange(12, dtype=float) * 0.1
    f = MO.features(n, origin, dx, atoms, species, 2, vloc, (1.0, 0.0, 0.0), 2.0, 32, nearest=8)
    assert f.shape == (12, 32)
    # point (jx, jy, jz) = (2, 1, 1): index 2 + 3*(1 + 2*1) = 11, at (2, 1, 1)
    p = np.array([2.0, 1.0, 1.0])
    d0 = sorted([np.linalg.norm(p - atoms[0]), np.linalg.norm(p - atoms[2])])
    d1 = np.linalg.norm(p - atoms[1])
    want = np.zeros(32)
    want[0:2] = [1 / (1 + 0.5 * d) for d in d0]
    want[8] = 1 / (1 + 0.5 * d1)
    want[16] = 1.1
    want[17] = np.linalg.norm(p - np.array([1.0, 0.0, 0.0])) / 2.0
    assert np.allclose(f[11], MO.bf16(want.astype(np.float32)))
