"""The oracle's absorbing envelopes (oracle.absorb, enlightning.C:630-786)
pinned by what the envelope's definition fixes, independent of its code:

* the ambient state (rho0, 0, 0, 0, T0*rho0; p0; T0) is a fixed point;
* momenta scale by tmp = 1 - 2^(-(d/W)^2): by 1/2 at distance d = W from the
  absorbing edge, by 1 - 2^-4 at d = 2W;
* the damped band ends where 1 - tmp first drops to 1e-8, i.e. at
  d = W sqrt(log2(1e8)) (the first cell past it is the last one processed),
  and cells farther in are untouched;
* a state mirrored in x (y) stays mirrored under left + right (bottom + top)
  envelopes;
* the pre-loop test (enlightning.C:675-676): a patch whose edge cell is
  already past the band leaves its first row alone on the left / right sides
  and does nothing on the bottom / top sides.
"""
import math

import numpy as np

import oracle
from enlightning_amd.gas import Gas


def state(nx, ny, g, gas, rng=None, ambient=False):
    N0, N1 = nx + 2 * g, ny + 2 * g
    w = np.empty((5, N1, N0))
    w[0] = gas.rho0
    w[1:4] = 0.0
    w[4] = gas.T0 * gas.rho0
    p = np.full((N1, N0), gas.p0)
    T = np.full((N1, N0), gas.T0)
    if not ambient:
        w[0] += 0.05 * rng.standard_normal((N1, N0))
        w[1:4] = rng.standard_normal((3, N1, N0))
        w[4] += 3.0 * rng.standard_normal((N1, N0))
        p += 50.0 * rng.standard_normal((N1, N0))
        T += rng.standard_normal((N1, N0))
    return w, p, T


def run(box, g, pxlo, dx, gxlo, gxhi, sides, width, gas, w, p, T):
    w, p, T = w.copy(), p.copy(), T.copy()
    oracle.absorb(box, g, pxlo, dx, gxlo, gxhi, sides, width, gas, w, p, T)
    return w, p, T


def test_ambient_is_a_fixed_point():
    gas = Gas()
    n, g = 64, 4
    w, p, T = state(n, n, g, gas, ambient=True)
    w2, p2, T2 = run((0, n - 1, 0, n - 1), g, (0.0, 0.0), (1 / n, 1 / n), (0.0, 0.0), (1.0, 1.0), (1, 1, 1, 1),
                     0.1, gas, w, p, T)
    assert np.array_equal(w2, w) and np.array_equal(p2, p) and np.array_equal(T2, T)


def test_momentum_profile_and_band_end():
    gas = Gas()
    n, g = 400, 4
    dx = 1.0 / n
    W = 8 * dx  # cell centres at d = (i + 1/2) dx: d = W at i = 7.5 -> use the exact points below
    w, p, T = state(n, 4, g, gas, ambient=True)
    w[1] = 1.0  # unit x momentum everywhere
    # left side only, over a 400 x 4 patch of a [0, 1] x [0, 4/400] domain
    w2, _, _ = run((0, n - 1, 0, 3), g, (0.0, 0.0), (dx, dx), (0.0, 0.0), (1.0, 4 * dx), (1, 0, 0, 0), W, gas, w,
                   p, T)
    row = w2[1, g + 1, g:-g]  # second row (the first row follows the same rule here)
    d = (np.arange(n) + 0.5) * dx
    # closed form 1 - 2^(-(d/W)^2) at the cells, to rounding
    assert np.allclose(row[:20], 1 - 2.0 ** (-(d[:20] / W) ** 2), rtol=0, atol=1e-14)
    # band end: cells up to the first with 1 - tmp <= 1e-8 are damped, the
    # rest carry exactly 1
    d_end = W * math.sqrt(math.log2(1e8))
    k_end = int(np.argmax(d >= d_end))  # first cell at or beyond d_end
    assert np.all(row[:k_end] < 1.0)
    assert np.all(row[k_end + 1:] == 1.0)
    # half damping at d = W: interpolate between the two neighbouring cells
    assert abs(np.interp(W, d, row) - 0.5) < 0.02


def test_mirror_symmetry():
    gas = Gas()
    rng = np.random.default_rng(5)
    nx, ny, g = 60, 40, 3
    w, p, T = state(nx, ny, g, gas, rng)
    # make the state mirror-symmetric in x (x momentum antisymmetric)
    w = 0.5 * (w + w[:, :, ::-1])
    w[1] = 0.0
    p = 0.5 * (p + p[:, ::-1])
    T = 0.5 * (T + T[:, ::-1])
    dx = (1.0 / nx, 1.0 / ny)
    w2, p2, T2 = run((0, nx - 1, 0, ny - 1), g, (0.0, 0.0), dx, (0.0, 0.0), (1.0, 1.0), (1, 1, 0, 0), 0.05, gas,
                     w, p, T)
    assert np.allclose(w2, w2[:, :, ::-1], rtol=0, atol=1e-12)
    assert np.allclose(p2, p2[:, ::-1], rtol=0, atol=1e-9)
    # and the same in y with bottom + top
    w = 0.5 * (w + w[:, ::-1, :])
    w[2] = 0.0
    p = 0.5 * (p + p[::-1, :])
    T = 0.5 * (T + T[::-1, :])
    w2, p2, T2 = run((0, nx - 1, 0, ny - 1), g, (0.0, 0.0), dx, (0.0, 0.0), (1.0, 1.0), (0, 0, 1, 1), 0.05, gas,
                     w, p, T)
    assert np.allclose(w2, w2[:, ::-1, :], rtol=0, atol=1e-12)
    assert np.allclose(T2, T2[::-1, :], rtol=0, atol=1e-12)


def test_patch_past_the_band():
    """A patch far from every edge: its first row is untouched on the left and
    right sides and nothing happens on the bottom and top sides; every later
    row gets its edge cell processed with tmp = 1 (a no-op for exactly
    representable deviations)."""
    gas = Gas()
    rng = np.random.default_rng(9)
    n, g = 32, 3
    dx = 1.0 / 256
    w, p, T = state(n, n, g, gas, rng)
    pxlo = (100 * dx, 100 * dx)  # d >= 100 dx >> 5.2 W
    w2, p2, T2 = run((100, 100 + n - 1, 100, 100 + n - 1), g, pxlo, (dx, dx), (0.0, 0.0), (1.0, 1.0), (1, 1, 1, 1),
                     2 * dx, gas, w, p, T)
    # momenta scale by exactly 1, rho and the rest go through (v - v0) * 1 + v0
    assert np.array_equal(w2[1:4], w[1:4])
    changed = (w2[0] != w[0]) | (p2 != p) | (T2 != T)
    rows, cols = np.nonzero(changed)
    # only edge cells of rows after the first, on the left and right edges
    assert np.all((rows >= g + 1) & (rows <= g + n - 1))
    assert np.all((cols == g) | (cols == g + n - 1))
    rel = np.abs(w2[0] - w[0]).max() / gas.rho0
    assert rel <= 2.3e-16
