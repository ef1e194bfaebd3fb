"""Seeded test inputs in the SAMRAI CellData layout (shared by CPU and GPU tests)."""
import numpy as np


def patch_arrays(nx, ny, g):
    Nx, Ny = nx + 2 * g, ny + 2 * g
    return (np.zeros((5, Ny, Nx)), np.zeros((Ny, Nx)), np.zeros((Ny, Nx)), np.zeros((ny, nx)))


def random_state(nx, ny, g, seed, kind="mixed", gas=None):
    """A physically plausible state (rho ~ 1.2, |u| ~ 10 m/s, p ~ 1e5 Pa,
    T ~ 294 K) with smooth waves, optional discontinuities and noise, so both
    the DRP and the WENO branch of the hybrid scheme are exercised."""
    rng = np.random.default_rng(seed)
    Nx, Ny = nx + 2 * g, ny + 2 * g
    x = np.linspace(0, 1, Nx)[None, :]
    y = np.linspace(0, 1, Ny)[:, None]

    def wave(a):
        kx, ky, ph = rng.uniform(1, 4), rng.uniform(1, 4), rng.uniform(0, 6.28)
        return a * np.sin(2 * np.pi * (kx * x + ky * y) + ph)

    # density waves gentle enough for the hybrid sensor (tolerance 5e-3) to
    # keep the smooth regions on DRP; the steps below trip it
    rho = 1.2 + wave(0.002) + wave(0.001)
    u = wave(8.0) + wave(3.0)
    v = wave(8.0) + wave(3.0)
    s = wave(2.0)
    Tn = 294.0 + wave(5.0)
    p = 101325.0 + wave(3000.0)
    T = 294.0 + wave(4.0)
    if kind in ("mixed", "shock"):
        # a few steps and a kink so that the sensor tags cells
        # (tiny patches: the step sits at the middle column / row)
        cx = int(rng.integers(g + 2, Nx - g - 2)) if Nx - g - 2 > g + 2 else Nx // 2
        cy = int(rng.integers(g + 2, Ny - g - 2)) if Ny - g - 2 > g + 2 else Ny // 2
        rho = rho + 0.15 * (x > x[0, cx]) - 0.08 * (y > y[cy, 0])
        p = p + 9000.0 * (x > x[0, cx])
        u = u + 20.0 * (x > x[0, cx])
        T = T + 3.0 * (y > y[cy, 0])
    if kind in ("mixed", "noise"):
        rho = rho + rng.uniform(-2e-5, 2e-5, size=rho.shape)
        p = p + rng.uniform(-10, 10, size=p.shape)
    w = np.empty((5, Ny, Nx))
    w[0] = rho
    w[1] = rho * u
    w[2] = rho * v
    w[3] = rho * s
    w[4] = rho * Tn
    src = np.zeros((ny, nx))
    if kind == "mixed":
        src = rng.uniform(-1.0, 1.0, size=(ny, nx)) * 1e-2
    return np.ascontiguousarray(w), np.ascontiguousarray(p), np.ascontiguousarray(T), src


def interior(a, g):
    return a[..., g:-g, g:-g]
