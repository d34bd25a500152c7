"""CPU oracle of the warm-start network and its grid features (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernels in filterbse_amd/csrc/mlp_kernels.hip; nothing in the product path
imports it.  The network has no counterpart in the reference (FilterBSE
starts its filter from ran2 noise, init.c:1207-1305); this file states the
intended arithmetic:

    a0 = x (BF16 values)
    z_l = a_{l-1} W_l + b_l          (products of BF16 values, accumulated exactly here)
    a_l = bf16(silu(z_l))            l = 1, 2, 3
    y   = a_3 W_4 + b_4              (linear output)
    loss = sum w (y - t)^2 / sum w

and the backward pass with the GPU's rounding points: every dZ is rounded to
BF16 before it enters a product (dZ_4 = bf16(2 w (y - t) / sum w),
dZ_l = bf16((dZ_{l+1} W_{l+1}^T) * silu'(z_l))); weight gradients are
a_{l-1}^T dZ_l, bias gradients the column sums of dZ_l.  With emulate=False
every rounding is skipped (plain float64), which is what the finite-difference
tests check.
"""
import numpy as np


def bf16(x):
    """Round to the nearest BF16 value, ties to even (finite inputs)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def bf16_bits(x):
    """BF16 bit patterns (uint16) of x, rounded as bf16()."""
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(p, n_in, width, n_out):
    dims = [(n_in, width), (width, width), (width, width), (width, n_out)]
    out, o = [], 0
    for k, n in dims:
        w = np.asarray(p[o:o + k * n], dtype=np.float64).reshape(k, n)
        o += k * n
        b = np.asarray(p[o:o + n], dtype=np.float64)
        o += n
        out.append((w, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])


def forward(layers, x, emulate=True):
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    a = np.asarray(r(x), dtype=np.float64)
    acts, zs = [a], []
    for l, (w, b) in enumerate(layers):
        wq = np.asarray(r(w), dtype=np.float64)
        z = a @ wq + b
        zs.append(z)
        if l < 3:
            a = np.asarray(r(silu(z)), dtype=np.float64)
            acts.append(a)
        else:
            a = z
    return a, acts, zs


def loss_grad(layers, x, t, w=None, emulate=True):
    """(loss, [(gW, gb)] per layer)."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    y, acts, zs = forward(layers, x, emulate)
    t = np.asarray(t, dtype=np.float64)
    wt = np.ones_like(t) if w is None else np.asarray(w, dtype=np.float64)
    sw = wt.sum()
    loss = float((wt * (y - t) ** 2).sum() / sw)
    dz = np.asarray(r(2.0 * wt * (y - t) / sw), dtype=np.float64)
    grads = [None] * 4
    for l in range(3, -1, -1):
        wq = np.asarray(r(layers[l][0]), dtype=np.float64)
        grads[l] = (acts[l].T @ dz, dz.sum(axis=0))
        if l > 0:
            da = dz @ wq.T
            dz = np.asarray(r(da * silu_grad(zs[l - 1])), dtype=np.float64)
    return loss, grads


def features(n, origin, dx, atoms_xyz, species, n_species, vloc, center, r_scale, k_in, nearest=8):
    """Per grid point (x fastest): for each species the `nearest` smallest
    distances as 1/(1 + d/2) ascending (0 when fewer atoms), V_loc, |r-c|/r_scale,
    zero padding to k_in; BF16-rounded like the GPU's output."""
    nx, ny, nz = n
    jz, jy, jx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([origin[0] + dx * jx.ravel(), origin[1] + dx * jy.ravel(), origin[2] + dx * jz.ravel()], axis=1)
    out = np.zeros((len(pts), k_in))
    at = np.asarray(atoms_xyz, dtype=np.float64).reshape(-1, 3)
    sp = np.asarray(species)
    c = 0
    for s in range(n_species):
        a_s = at[sp == s]
        if len(a_s):
            d = np.sqrt(((pts[:, None, :] - a_s[None, :, :]) ** 2).sum(axis=2))
            d.sort(axis=1)
            m = min(nearest, d.shape[1])
            out[:, c:c + m] = 1.0 / (1.0 + 0.5 * d[:, :m])
        c += nearest
    out[:, c] = np.asarray(vloc)[:len(pts)]
    out[:, c + 1] = np.sqrt(((pts - np.asarray(center)) ** 2).sum(axis=1)) / r_scale
    return bf16(out.astype(np.float32))
