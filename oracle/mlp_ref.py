"""CPU reference of the warm-start MLP (44 -> 64 -> 64 -> 3, ReLU, mean squared error).

TEST INFRASTRUCTURE ONLY (tests/ use it as the checker of the GPU kernels).

Parameters are the flat FP32 vector of include/ncx.h: W1 [64][48], b1 [64], W2 [64][64],
b2 [64], W3 [16][64], b3 [16] (rows/columns past 44 inputs / 3 outputs are zero).

`emulated=True` rounds to float16 exactly where the GPU does (features, weights, post-ReLU
activations, output and hidden gradients before each matrix product) and accumulates in float32,
so GPU and CPU differ only by float32 summation order.  `emulated=False` is the plain float64
network used for the finite-difference check of the gradient itself.
"""
import numpy as np

IN, INP, H, OUT, OUTP = 44, 48, 64, 3, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * INP
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + OUTP * H
NP = OFF_B3 + OUTP


def unpack(p):
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(H, INP), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(OUTP, H), p[OFF_B3:NP])


def _h(x, emulated):
    return x.astype(np.float16).astype(np.float32) if emulated else x


def forward(p, X, emulated=True):
    """X: (n, 48).  Returns (Z3 (n, 16), cache)."""
    dt = np.float32 if emulated else np.float64
    W1, b1, W2, b2, W3, b3 = [a.astype(dt) for a in unpack(p)]
    X = _h(np.asarray(X, dt), emulated)
    W1h, W2h, W3h = _h(W1, emulated), _h(W2, emulated), _h(W3, emulated)
    Z1 = X @ W1h.T + b1
    H1 = _h(np.maximum(Z1, 0), emulated)
    Z2 = H1 @ W2h.T + b2
    H2 = _h(np.maximum(Z2, 0), emulated)
    Z3 = H2 @ W3h.T + b3
    return Z3, (X, Z1, H1, Z2, H2, W1h, W2h, W3h)


def loss_and_grad(p, X, Y, emulated=True):
    n = X.shape[0]
    Z3, (Xh, Z1, H1, Z2, H2, W1h, W2h, W3h) = forward(p, X, emulated)
    dt = Z3.dtype
    Yp = np.zeros((n, OUTP), dt)
    Yp[:, :OUT] = Y
    E = Z3 - Yp
    E[:, OUT:] = 0
    loss = np.sum(E * E) / n
    D3 = 2 * E
    D3h = _h(D3, emulated)
    dH2 = D3h @ W3h
    D2 = dH2 * (Z2 > 0)
    D2h = _h(D2, emulated)
    dH1 = D2h @ W2h
    D1 = dH1 * (Z1 > 0)
    D1h = _h(D1, emulated)
    g = np.zeros(NP, dt)
    g[OFF_W1:OFF_B1] = (D1h.T @ Xh).ravel() / n
    g[OFF_B1:OFF_W2] = D1.sum(0) / n
    g[OFF_W2:OFF_B2] = (D2h.T @ H1).ravel() / n
    g[OFF_B2:OFF_W3] = D2.sum(0) / n
    g[OFF_W3:OFF_B3] = (D3h.T @ H2).ravel() / n
    g[OFF_B3:NP] = D3.sum(0) / n
    return loss, g


def valid_mask():
    m = np.zeros(NP, bool)
    m[OFF_W1:OFF_B1] = (np.arange(H * INP) % INP) < IN
    m[OFF_B1:OFF_W2] = True
    m[OFF_W2:OFF_B2] = True
    m[OFF_B2:OFF_W3] = True
    m[OFF_W3:OFF_B3] = (np.arange(OUTP * H) // H) < OUT
    m[OFF_B3:NP] = np.arange(OUTP) < OUT
    return m
