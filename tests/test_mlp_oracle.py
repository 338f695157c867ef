"""Pins of the learned-warm-start oracle (CPU only).

The MLP and its features have no reference counterpart; the oracle
(oracle/mlp_oracle.py) is pinned by what fixes each piece independently:
backpropagation by central finite differences, bf16 rounding, GELU and Adam by
torch's own implementations, the features by brute force over mesh edges and by
closed forms (planar bed -> its slope at every node; a linear field -> the
gradient dotted with the mean neighbour displacement).
"""
import numpy as np

import adh_kraken_amd as A
import mlp_oracle as M
from helpers import small_flume


def random_model(rng, scale=0.3):
    p = rng.uniform(-scale, scale, M.NPARAM)
    norm = np.zeros(36)
    norm[0:12] = rng.uniform(-0.5, 0.5, 12)
    norm[12:24] = rng.uniform(0.5, 2.0, 12)
    norm[24:27] = rng.uniform(-0.1, 0.1, 3)
    norm[27:30] = rng.uniform(0.5, 2.0, 3)
    norm[30:33] = 1.0 / norm[27:30]
    return p, norm


def test_backprop_vs_finite_differences():
    rng = np.random.default_rng(0)
    p, norm = random_model(rng)
    x = rng.normal(size=(17, 12))
    t = rng.normal(size=(17, 3))
    loss, g = M.loss_grad(p, norm, x, t, bf16=False)
    idx = np.concatenate([rng.choice(M.NPARAM, 300, replace=False), np.arange(M.OFF_W3, M.NPARAM)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, norm, x, t, bf16=False)[0] - M.loss_grad(pm, norm, x, t, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), i


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(1)
    w = np.concatenate([rng.normal(size=5000).astype(np.float32),
                        np.array([0.0, -0.0, 1.0, 3.0e38, 1e-40, 1.00390625, 1.01171875], np.float32)])
    ref = torch.from_numpy(w).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16_round(w), ref)


def test_gelu_matches_torch():
    import torch
    x = torch.linspace(-8, 8, 2001, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(x, approximate="none")
    y.sum().backward()
    xn = x.detach().numpy()
    assert np.abs(M.gelu(xn) - y.detach().numpy()).max() <= 1e-15
    assert np.abs(M.gelu_d(xn) - x.grad.numpy()).max() <= 1e-14


def test_adam_matches_torch():
    import torch
    rng = np.random.default_rng(2)
    p0 = rng.normal(size=50)
    tp = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for step in range(1, 6):
        g = rng.normal(size=50)
        opt.zero_grad()
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = M.adam(p, m, v, g, step, 3e-3)
    assert np.abs(p - tp.detach().numpy()).max() <= 1e-12


def test_features_brute_force_and_closed_forms():
    spec, (mesh, sol0) = small_flume(npx=13, npy=7, nbumps=0)
    rng = np.random.default_rng(3)
    N = mesh.nnodes
    # planar bed: every node's area-weighted slope is the plane's
    a, b = 0.013, -0.007
    m2 = A.Mesh(mesh.x, mesh.y, a * mesh.x + b * mesh.y + 0.4, mesh.tri, mesh.seg, mesh.seg_kind, mesh.seg_val,
                np.full(mesh.nelems2d, 0.031))
    s, s1 = rng.uniform(0.5, 1.5, 3 * N), rng.uniform(0.5, 1.5, 3 * N)
    f = M.features(m2, s, s1)
    assert np.allclose(f[:, 9], a, rtol=0, atol=1e-15) and np.allclose(f[:, 10], b, rtol=0, atol=1e-15)
    assert np.allclose(f[:, 11], 0.031, rtol=0, atol=1e-15)
    # neighbour means by brute force over the mesh edges
    edges = set()
    for tr in mesh.tri:
        for i in range(3):
            for j in range(3):
                if i != j:
                    edges.add((int(tr[i]), int(tr[j])))
    h, u, v = s[0::3], s[1::3], s[2::3]
    for n in range(N):
        nb = [q for (p_, q) in edges if p_ == n]
        assert abs(f[n, 6] - (np.mean(h[nb]) - h[n])) <= 1e-14
        assert abs(f[n, 7] - (np.mean(h[nb] * u[nb]) - h[n] * u[n])) <= 1e-14
        assert abs(f[n, 8] - (np.mean(h[nb] * v[nb]) - h[n] * v[n])) <= 1e-14
    assert np.array_equal(f[:, 3], s[0::3] - s1[0::3])
    # a linear field: the neighbour-mean offset is the gradient dotted with the mean neighbour displacement
    lin = np.zeros(3 * N)
    lin[0::3] = 1.0 + 0.2 * mesh.x - 0.1 * mesh.y
    lin[1::3] = 1.0
    fl = M.features(mesh, lin, lin)
    for n in range(N):
        nb = [q for (p_, q) in edges if p_ == n]
        want = 0.2 * np.mean(mesh.x[nb] - mesh.x[n]) - 0.1 * np.mean(mesh.y[nb] - mesh.y[n])
        assert abs(fl[n, 6] - want) <= 1e-13 and abs(fl[n, 7] - want) <= 1e-13
    assert np.abs(fl[:, 3:6]).max() == 0.0


def test_forward_loss_consistent():
    rng = np.random.default_rng(4)
    p, norm = random_model(rng)
    x = rng.normal(size=(9, 12))
    y = M.forward(p, norm, x)
    t = y.copy()  # targets equal to the prediction -> zero loss and zero gradient
    loss, g = M.loss_grad(p, norm, x, t)
    assert loss <= 1e-28 and np.abs(g).max() <= 1e-13


def test_matrix_core_rounding_is_bf16_sized():
    """The BF16 roundings of the hidden layer's products (mfma) move the forward by no more than BF16's
    relative step on each product's terms, and the rounded gradient stays close to the exact one (the
    finite-difference-pinned path above) -- a rounding at a wrong point (an output, a weight gradient)
    or none at all fails one of these."""
    rng = np.random.default_rng(7)
    p, norm = random_model(rng)
    x = rng.normal(size=(64, 12))
    t = rng.normal(size=(64, 3))
    xn = M.normalise_x(x, norm)
    y0, (z1, a1, z2, a2) = M.forward_norm(p, xn, bf16=True, mfma=False)
    y1, (_, a1r, z2r, _) = M.forward_norm(p, xn, bf16=True, mfma=True)
    W1, b1, W2, b2, W3, b3 = M.unpack(M.bf16_round(p).astype(np.float64))
    assert np.array_equal(a1r, M.bf16_round(a1).astype(np.float64))
    bound = (np.abs(a1) @ np.abs(W2).T) * 2.0 ** -8  # |z2 change| <= sum_j |W2_ij| |a1_j| 2^-8
    assert np.all(np.abs(z2r - z2) <= bound + 1e-15) and np.abs(z2r - z2).max() > 0
    l0, g0 = M.loss_grad(p, norm, x, t, bf16=True, mfma=False)
    l1, g1 = M.loss_grad(p, norm, x, t, bf16=True, mfma=True)
    assert abs(l1 - l0) <= 2e-2 * l0
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_W2, M.OFF_B2), (M.OFF_W3, M.OFF_B3)):
        assert 0 < np.abs(g1[a:b] - g0[a:b]).max() <= 3e-2 * np.abs(g0[a:b]).max()
    # the bias gradient of the hidden layer sums the unrounded dz2
    dy = 2.0 * (y1 - (t - norm[24:27]) * norm[30:33]) / (64 * 3)
    dz2 = (dy @ W3) * M.gelu_d(z2r)
    assert np.allclose(g1[M.OFF_B2:M.OFF_W3], dz2.sum(0), rtol=1e-13, atol=1e-16)
