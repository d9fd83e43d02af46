"""CPU restatement of the learned-closure MLP (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the HIP
kernels in ``csrc/closure.hip``; the product package never imports it.

The closure has no counterpart in the reference (it corrects the penalty
force of fsi/coupling.py:333-406); what pins this restatement:

* the float64 mode (``bf16_round=False``) against central finite differences of
  its own loss (``tests/test_closure_oracle.py``), so a dropped term, a
  wrong transpose or a wrong GELU derivative fails there;
* GELU against the closed form at known points (GELU(0) = 0,
  GELU'(0) = 1/2, GELU(x) - GELU(-x) = x);
* Adam against a hand-evaluated first step (|update| = lr for any nonzero
  gradient at step 1);
* ``closure_features`` against hand-built fields with known answers (a
  uniform velocity gives u - v and a zero gradient; a linear one its
  constant gradient, which quadratic B-spline interpolation reproduces
  exactly; a planar occupancy ramp its normal).

With ``bf16_round=True`` it rounds exactly where the kernels round (weights,
inputs, hidden activations and the backward deltas to BF16, round to
nearest even) and accumulates in float64, so the kernels' FP32 accumulation
is the only remaining difference.

Parameter layout (one flat vector, as the kernels): W1[64][32] (input
columns 18..31 zero), b1[64], W2[64][64], b2[64], W3[16][64] (rows 3..15
zero), b3[16]; layer l computes z = W h + b.
"""

from __future__ import annotations

import math

import numpy as np

IN, IN_PAD, HID, OUT, OUT_PAD = 18, 32, 64, 3, 16
O_W1 = 0
O_B1 = O_W1 + HID * IN_PAD
O_W2 = O_B1 + HID
O_B2 = O_W2 + HID * HID
O_W3 = O_B2 + HID
O_B3 = O_W3 + OUT_PAD * HID
N_PARAMS = O_B3 + OUT_PAD

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[O_W1:O_B1].reshape(HID, IN_PAD), p[O_B1:O_W2], p[O_W2:O_B2].reshape(HID, HID), p[O_B2:O_W3],
            p[O_W3:O_B3].reshape(OUT_PAD, HID), p[O_B3:N_PARAMS])


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, zero padding (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    w1 = np.zeros((HID, IN_PAD), np.float32)
    w1[:, :IN] = rng.uniform(-1, 1, (HID, IN)) * math.sqrt(6.0 / IN)
    w3 = np.zeros((OUT_PAD, HID), np.float32)
    w3[:OUT] = rng.uniform(-1, 1, (OUT, HID)) * math.sqrt(6.0 / HID) * 0.1
    p[O_W1:O_B1] = w1.reshape(-1)
    p[O_W2:O_B2] = (rng.uniform(-1, 1, (HID, HID)) * math.sqrt(6.0 / HID)).reshape(-1)
    p[O_W3:O_B3] = w3.reshape(-1)
    return p


def forward(p, X, bf16_round=True):
    """Outputs (n x 3) and the cached activations of rows X (n x 18)."""
    r = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.zeros((len(X), IN_PAD))
    x[:, :IN] = np.asarray(X, np.float64)
    x = r(x)
    z1 = x @ r(W1).T + b1
    h1 = r(gelu(z1))
    z2 = h1 @ r(W2).T + b2
    h2 = r(gelu(z2))
    y = h2 @ r(W3).T + b3
    return y[:, :OUT], (x, z1, h1, z2, h2)


def loss_and_grad(p, X, T, bf16_round=True):
    """Sum of squared errors, and d(mean squared error)/dp (flat)."""
    r = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (x, z1, h1, z2, h2) = forward(p, X, bf16_round)
    n = len(X)
    e = y - np.asarray(T, np.float64)
    dyf = np.zeros((n, OUT_PAD))
    dyf[:, :OUT] = 2.0 * e / (3.0 * n)
    dy = r(dyf)
    dz2 = (dy @ r(W3)) * gelu_grad(z2)
    dz2r = r(dz2)
    dz1 = (dz2r @ r(W2)) * gelu_grad(z1)
    dz1r = r(dz1)
    g = np.zeros(N_PARAMS)
    g[O_W1:O_B1] = (dz1r.T @ x).reshape(-1)
    g[O_B1:O_W2] = dz1.sum(0)
    g[O_W2:O_B2] = (dz2r.T @ h1).reshape(-1)
    g[O_B2:O_W3] = dz2.sum(0)
    g[O_W3:O_B3] = (dy.T @ h2).reshape(-1)
    g[O_B3:N_PARAMS] = dyf.sum(0)
    return float((e * e).sum()), g


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba, ICLR 2015); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def closure_features(x, v, u, solid, phi, dx=1.0):
    """The 18 closure features per particle (float64; the kernel computes in
    the particle precision and stores float32). x, v: (n, 3); u: (nx, ny,
    nz, 3) lattice velocity; solid: (nx, ny, nz) mask; phi: (nx, ny, nz)
    occupancy. Quadratic B-spline weights and their x-derivatives, support
    and velocity over in-lattice fluid nodes, phi over in-lattice nodes."""
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    nx, ny, nz = solid.shape
    inv_dx = 1.0 / dx
    out = np.zeros((len(x), IN))
    for p in range(len(x)):
        gp = x[p] * inv_dx
        base = np.trunc(gp - 0.5).astype(int)
        fx = gp - base
        w = np.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2])
        dw = np.stack([(fx - 1.5), -2.0 * (fx - 1.0), (fx - 0.5)]) * inv_dx
        W = 0.0
        ui = np.zeros(3)
        G = np.zeros((3, 3))
        ph = 0.0
        gph = np.zeros(3)
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    n = base + (i, j, k)
                    if not (0 <= n[0] < nx and 0 <= n[1] < ny and 0 <= n[2] < nz):
                        continue
                    wt = w[i, 0] * w[j, 1] * w[k, 2]
                    gw = np.array([dw[i, 0] * w[j, 1] * w[k, 2], w[i, 0] * dw[j, 1] * w[k, 2],
                                   w[i, 0] * w[j, 1] * dw[k, 2]])
                    f = phi[n[0], n[1], n[2]]
                    ph += wt * f
                    gph += gw * f
                    if solid[n[0], n[1], n[2]] != 0:
                        continue
                    un = u[n[0], n[1], n[2]]
                    W += wt
                    ui += wt * un
                    G += np.outer(un, gw)
        rel = (ui / W if W > 0 else np.zeros(3)) - v[p]
        gn = np.linalg.norm(gph)
        out[p, 0:3] = rel
        out[p, 3:12] = G.reshape(-1)
        out[p, 12] = ph
        out[p, 13:16] = -gph / gn if gn > 1e-6 else 0.0
        out[p, 16] = W
        out[p, 17] = np.linalg.norm(rel)
    return out
