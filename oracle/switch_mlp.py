"""CPU restatement of the learned switch's network (TEST INFRASTRUCTURE ONLY).

The network has no counterpart in the reference: it stands in for the hybrid
sensor (rhs_hybrid_high_llf.m4:239-256).  This module restates what the HIP
kernels compute, in float64 numpy, so that the device forward pass, loss and
gradients can be checked:
  features  asinh(d_rho / sqrt(eps)), asinh(d_p / (c^2 sqrt(eps))) of the x
            and y forward differences on a 5x5 window, rounded to BF16;
  network   100-64-32-1, ReLU, sigmoid, parameters and the first hidden
            layer rounded to BF16;
  loss      weighted mean binary cross-entropy, sum(w_b * l_b) / sum(w_b).
The gradients are pinned by central finite differences of this loss
(tests/test_switch_oracle.py), independent of the backward formulas here.
"""
import numpy as np

N_IN, N_H1, N_H2 = 100, 64, 32
N_PARAM = N_H1 * N_IN + N_H1 + N_H2 * N_H1 + N_H2 + N_H2 + 1


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    W1 = p[o:o + N_H1 * N_IN].reshape(N_H1, N_IN); o += N_H1 * N_IN
    b1 = p[o:o + N_H1]; o += N_H1
    W2 = p[o:o + N_H2 * N_H1].reshape(N_H2, N_H1); o += N_H2 * N_H1
    b2 = p[o:o + N_H2]; o += N_H2
    w3 = p[o:o + N_H2]; o += N_H2
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [b3]])


def forward(params, X, round_h1=True):
    """Logits and hidden activations for feature rows X (already BF16 values);
    the first hidden layer is rounded to BF16 as the device feeds it to the
    second."""
    W1, b1, W2, b2, w3, b3 = unpack(bf16(params))
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    if round_h1:
        h1 = bf16(h1)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3 + b3
    return z, h1, h2


def loss(params, X, y, w, round_h1=True):
    z = forward(params, X, round_h1)[0]
    l = np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z)))
    return float(np.sum(w * l) / np.sum(w))


def loss_exact_params(pvals, X, y, w):
    """The loss with the parameters taken as given (no BF16 rounding): the
    function whose finite differences pin the gradient."""
    W1, b1, W2, b2, w3, b3 = unpack(pvals)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    z = h2 @ w3 + b3
    l = np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z)))
    return float(np.sum(w * l) / np.sum(w))


def gradients(params, X, y, w, round_h1=True):
    """d loss / d params, evaluated at the BF16-rounded parameters (the
    rounding of the first hidden layer passes gradients straight through)."""
    W1, b1, W2, b2, w3, b3 = unpack(bf16(params))
    z, h1, h2 = forward(params, X, round_h1)
    p = 1.0 / (1.0 + np.exp(-z))
    dz = w * (p - y) / np.sum(w)
    gw3 = h2.T @ dz
    gb3 = dz.sum()
    dh2 = np.outer(dz, w3) * (h2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    return pack(gW1, gb1, gW2, gb2, gw3, gb3)


def features(rho, p, cells, nx, g, eps, c):
    """The 100 BF16 features of interior cells (i + j*nx) of ghosted rho and p
    arrays (shape (Ny, Nx))."""
    out = np.empty((len(cells), N_IN))
    # as the device: the FP64 difference and the scale in float32, asinh of
    # the float32 product
    ir = np.float32(1.0 / np.sqrt(eps))
    ip = np.float32(1.0 / (c * c * np.sqrt(eps)))
    for q, cell in enumerate(cells):
        i, j = cell % nx + g, cell // nx + g
        k = 0
        for dj in range(-2, 3):
            for di in range(-2, 3):
                jj, ii = j + dj, i + di
                vals = (np.float32(rho[jj, ii + 1] - rho[jj, ii]) * ir, np.float32(rho[jj + 1, ii] - rho[jj, ii]) * ir,
                        np.float32(p[jj, ii + 1] - p[jj, ii]) * ip, np.float32(p[jj + 1, ii] - p[jj, ii]) * ip)
                for v in vals:
                    out[q, k] = np.float32(np.arcsinh(np.float64(v)))
                    k += 1
    return bf16(out)
