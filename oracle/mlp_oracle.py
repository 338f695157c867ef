"""CPU oracle of the learned Newton warm start (numpy, float64 arithmetic).

TEST INFRASTRUCTURE ONLY: the checker of libadhk.so's MLP kernels.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.

There is no reference counterpart: the model is this project's extension (the
reference starts each Newton solve from the previous time level,
smodel_super_prep_sol.c:2-8).  What it restates is the model as the library
defines it:

* features(): 12 per node -- h, hu, hv at t; their change since t-1; the 1-ring
  neighbour means of h, hu, hv at t minus the node's own; the bed slope (x, y)
  and Manning n, area-weighted over the node's triangles;
* forward() / loss_grad(): 12 -> 64 -> 64 -> 3, GELU (erf form), weights rounded
  to BF16 (round to nearest even), the mean squared error on normalised outputs
  and its gradient by backpropagation;
* with the BF16 weights come the roundings of the matrix-core products of the
  64 x 64 layer (mfma=True): the layer's input activations a1 are rounded to BF16
  where they enter the forward product and the weight-gradient product, and its
  output gradients dz2 where they enter the weight-gradient and input-gradient
  products; sums, biases, the other layers and the bias gradients stay unrounded;
* adam(): the optimiser step.

Pinned by tests/test_mlp_oracle.py: the gradient against central finite
differences, bf16 rounding and GELU against torch, features against brute force
and a planar-bed closed form.
"""
import numpy as np
from scipy.special import erf

NIN, HID, NOUT = 12, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * NIN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + NOUT * HID
NPARAM = OFF_B3 + NOUT


def bf16_round(w):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(HID, NIN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(NOUT, HID), p[OFF_B3:NPARAM])


def _weights(params, bf16):
    return bf16_round(params).astype(np.float64) if bf16 else np.asarray(params, np.float64)


def normalise_x(x, norm):
    return (np.asarray(x, np.float64) - norm[0:12]) * norm[12:24]


def forward_norm(params, xn, bf16=True, mfma=None):
    """normalised inputs (n, 12) -> normalised outputs (n, 3) and the activations; mfma (default: with the
    BF16 weights) rounds a1 to BF16 where it enters the 64 x 64 product"""
    mfma = bf16 if mfma is None else mfma
    W1, b1, W2, b2, W3, b3 = unpack(_weights(params, bf16))
    z1 = xn @ W1.T + b1
    a1 = gelu(z1)
    if mfma:
        a1 = bf16_round(a1).astype(np.float64)
    z2 = a1 @ W2.T + b2
    a2 = gelu(z2)
    y = a2 @ W3.T + b3
    return y, (z1, a1, z2, a2)


def forward(params, norm, x, bf16=True, mfma=None):
    """raw features (n, 12) -> raw outputs (n, 3): the change of h, u, v"""
    norm = np.asarray(norm, np.float64)
    y, _ = forward_norm(params, normalise_x(x, norm), bf16, mfma)
    return y * norm[27:30] + norm[24:27]


def loss_grad(params, norm, x, t, bf16=True, mfma=None):
    """mean squared error over n samples and 3 normalised outputs, and its gradient (NPARAM,); with mfma
    the products of the 64 x 64 layer see BF16 a1 and dz2 (the gradient of the rounded forward, the
    roundings taken as the identity)"""
    mfma = bf16 if mfma is None else mfma
    norm = np.asarray(norm, np.float64)
    xn = normalise_x(x, norm)
    tn = (np.asarray(t, np.float64) - norm[24:27]) * norm[30:33]
    W1, b1, W2, b2, W3, b3 = unpack(_weights(params, bf16))
    y, (z1, a1, z2, a2) = forward_norm(params, xn, bf16, mfma)
    n = len(xn)
    d = y - tn
    loss = float(np.sum(d * d) / (n * NOUT))
    dy = 2.0 * d / (n * NOUT)
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = (dy.T @ a2).ravel()
    g[OFF_B3:NPARAM] = dy.sum(0)
    dz2 = (dy @ W3) * gelu_d(z2)
    dz2m = bf16_round(dz2).astype(np.float64) if mfma else dz2  # a1 is already rounded by forward_norm
    g[OFF_W2:OFF_B2] = (dz2m.T @ a1).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2m @ W2) * gelu_d(z1)
    g[OFF_W1:OFF_B1] = (dz1.T @ xn).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g


def adam(params, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = params - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v


def node_neighbours(tri, nnodes):
    nb = [set() for _ in range(nnodes)]
    for a, b, c in np.asarray(tri):
        nb[a] |= {b, c}
        nb[b] |= {a, c}
        nb[c] |= {a, b}
    return [sorted(s) for s in nb]


def node_static(x, y, z, tri, mann):
    """area-weighted bed slope (x, y) and Manning n over each node's triangles"""
    N = len(x)
    t = np.asarray(tri)
    x1, x2, x3 = x[t[:, 0]], x[t[:, 1]], x[t[:, 2]]
    y1, y2, y3 = y[t[:, 0]], y[t[:, 1]], y[t[:, 2]]
    area2 = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)
    gzx = ((y2 - y3) * z[t[:, 0]] + (y3 - y1) * z[t[:, 1]] + (y1 - y2) * z[t[:, 2]]) / area2
    gzy = ((x3 - x2) * z[t[:, 0]] + (x1 - x3) * z[t[:, 1]] + (x2 - x1) * z[t[:, 2]]) / area2
    w = 0.5 * area2
    out = np.zeros((N, 3))
    sw = np.zeros(N)
    for a in range(3):
        np.add.at(out[:, 0], t[:, a], w * gzx)
        np.add.at(out[:, 1], t[:, a], w * gzy)
        np.add.at(out[:, 2], t[:, a], w * mann)
        np.add.at(sw, t[:, a], w)
    return out / np.where(sw > 0, sw, 1.0)[:, None]


def features(mesh, sol_t, sol_tm1):
    """(nnodes, 12) features of the state at t and t-1"""
    N = mesh.nnodes
    h, u, v = sol_t[0::3], sol_t[1::3], sol_t[2::3]
    h1, u1, v1 = sol_tm1[0::3], sol_tm1[1::3], sol_tm1[2::3]
    hu, hv, hu1, hv1 = h * u, h * v, h1 * u1, h1 * v1
    f = np.zeros((N, NIN))
    f[:, 0], f[:, 1], f[:, 2] = h, hu, hv
    f[:, 3], f[:, 4], f[:, 5] = h - h1, hu - hu1, hv - hv1
    for n, nb in enumerate(node_neighbours(mesh.tri, N)):
        if nb:
            f[n, 6] = h[nb].mean() - h[n]
            f[n, 7] = hu[nb].mean() - hu[n]
            f[n, 8] = hv[nb].mean() - hv[n]
    mann = mesh.mann if mesh.mann is not None else np.zeros(mesh.nelems2d)
    f[:, 9:12] = node_static(mesh.x, mesh.y, mesh.z, mesh.tri, mann)
    return f
