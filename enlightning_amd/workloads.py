"""Synthetic inputs of the measured workloads.

A workload is one square patch of n x n interior cells with g ghost cells on
[0, L]^2, reflecting walls on all sides, a superposition of Gaussian
overpressure pulses dp(x) = A*p0*exp(-|x-x_c|^2 / (2 sigma^2)) and optional
uniform noise of amplitude noise*p0.  The reference has no overpressure
initial condition (it only injects a mass source, enlightning.C:2176-2192),
so the pulse is realised as a density perturbation at zero entropy: from the
second-order EOS of update_state.m4:90-94 with s = 0,
    dp = c^2 (rho' + a rho'^2),  a = (gamma-1)/(2 rho0)
    => rho' = (-1 + sqrt(1 + 4 a dp / c^2)) / (2 a),
with rho*u = rho*v = rho*s = 0 and rho*T_n = rho*T0 (enlightning.C:418-424).

Arrays use the SAMRAI CellData layout: numpy C-order shape (5, Ny, Nx) for
w_n (component slowest, i fastest) and (Ny, Nx) for scalars.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class Workload:
    name: str
    n: int              # interior cells per side
    g: int = 4          # ghost width
    L: float = 1.0      # domain side, m
    steps: int = 100
    cfl: float = 0.5
    variant: str = "hybrid_high_llf"
    seed: int = 0
    npulses: int = 1
    noise: float = 0.0  # uniform noise amplitude / p0
    # single-pulse parameters (npulses == 1 and fixed == True)
    fixed: bool = False
    amp: float = 0.5
    sigma: float = 0.02
    center: tuple = (0.5, 0.5)
    amp_range: tuple = (0.05, 0.8)
    sigma_range: tuple = (0.005, 0.03)
    # hierarchy workloads: base-grid steps taken before the tags are formed
    presteps: int = 0

    @property
    def dx(self):
        return self.L / self.n

    def dt(self, gas):
        # enlightning.C:446-447: dt = cfl * min(dx, dy) / c
        return self.cfl * (self.dx / gas.c)

    def pulses(self):
        """(centers [m], amplitudes [dp/p0], sigmas [m]) of the pulses."""
        if self.fixed:
            return (np.array([self.center], float), np.array([self.amp]), np.array([self.sigma]))
        rng = np.random.default_rng(self.seed)
        centers = rng.uniform(0.0, self.L, size=(self.npulses, 2))
        amps = rng.uniform(*self.amp_range, size=self.npulses)
        sigmas = rng.uniform(*self.sigma_range, size=self.npulses) * self.L
        return centers, amps, sigmas


# The plain-method workloads: a 256^2 CPU-runnable case and the 4096^2 case
# the single-GPU throughput is quoted on.
SMALL_256 = Workload("pulse_256_single", n=256, steps=100, seed=0, fixed=True, amp=0.5, sigma=0.02,
                     center=(0.5, 0.5), noise=1e-6)
PULSES_4096 = Workload("pulses_4096_8", n=4096, steps=500, seed=1, npulses=8)


def pulses_at(wl, x, y):
    """Sum of the workload's pulses (dp/p0) on the grid x (row) times y
    (column) of cell centres; shape (len(y), len(x))."""
    X = np.asarray(x)[None, :]
    Y = np.asarray(y)[:, None]
    dp = np.zeros((Y.shape[0], X.shape[1]))
    centers, amps, sigmas = wl.pulses()
    for (cx, cy), a, s in zip(centers, amps, sigmas):
        # separable evaluation keeps this cheap at 4096^2
        ex = np.exp(-((X - cx) ** 2) / (2 * s * s))
        ey = np.exp(-((Y - cy) ** 2) / (2 * s * s))
        dp += a * (ey * ex)
    return dp


def overpressure(wl):
    """dp [Pa / p0 units] on the ghosted cell centres, shape (Ny, Nx)."""
    n, g, dx = wl.n, wl.g, wl.dx
    xc = (np.arange(-g, n + g) + 0.5) * dx
    dp = pulses_at(wl, xc, xc)
    if wl.noise:
        rng = np.random.default_rng(wl.seed)
        dp[g:g + n, g:g + n] += rng.uniform(-1.0, 1.0, size=(n, n)) * wl.noise
    return dp


def state_from_overpressure(dp_over_p0, gas):
    """w (5, Ny, Nx) at zero entropy and rest for an overpressure field."""
    dp = dp_over_p0 * gas.p0
    a = (gas.gamma1 - 1.0) / (2.0 * gas.rho0)
    rho_p = (-1.0 + np.sqrt(1.0 + 4.0 * a * dp / gas.c ** 2)) / (2.0 * a)
    rho = gas.rho0 + rho_p
    w = np.zeros((5,) + rho.shape)
    w[0] = rho
    w[4] = rho * gas.T0
    return w


def initial_fields(wl, gas):
    """Host arrays (w, p, T, source) of the initial state.  p and T hold the
    ambient values; the caller derives them from w with update_state and
    fills the ghosts with the physical boundary routine."""
    w = state_from_overpressure(overpressure(wl), gas)
    N = wl.n + 2 * wl.g
    p = np.full((N, N), gas.p0)
    T = np.full((N, N), gas.T0)
    src = np.zeros((N, N))
    return w, p, T, src


# The batched-hierarchy workload: base 1024^2 on [0,1]^2 m in 64^2 patches,
# eight seeded pulses as in PULSES_4096 (seed 2), advanced 500 steps on the
# uniform base grid so that they have spread into interacting rings; then
# one finer level (ratio 4) over tag_cells' tags (565 patches of 8..64 cells
# a side, 1.2 M fine cells), initialised by linear refinement.
AMR_1024 = Workload("amr_1024_r4_8pulses", n=1024, steps=100, seed=2, npulses=8, presteps=500)


# ---- the precision form's weak-acoustic workload -----------------------------
# An exponentially stratified ambient at rest and uniform pressure p0:
#   rho_a(y) = rho0 exp(-y / H),  rho*s_a = -c_p ((rho_a - rho0) + a (rho_a - rho0)^2)
# (the EOS of update_state.m4:90-94 then gives p = p0 at c_lr = 0, so the
# ambient is in mechanical equilibrium; there is no gravity in the model),
# rho*T_n in equilibrium with T (update_state.m4:103-105).  Pulses are
# isentropic (specific entropy of the ambient row kept) with dp/p0 drawn in
# [amp_range]; rho from the EOS quadratic at the cell's s.


@dataclass
class StratifiedWorkload(Workload):
    scale_height: float = 8000.0


WEAK_2KM = StratifiedWorkload("weak_acoustic_2km_16pulses", n=1024, L=2000.0, steps=2000, seed=3, npulses=16,
                              amp_range=(1e-5, 1e-3), sigma_range=(0.005, 0.03), scale_height=8000.0)


def ambient_rows(wl, gas, y):
    """(rho, rho*s, rho*T_n) of the stratified ambient at heights y."""
    y = np.asarray(y, dtype=np.float64)
    a = (gas.gamma1 - 1.0) / (2.0 * gas.rho0)
    rho = gas.rho0 * np.exp(-y / wl.scale_height)
    D = rho - gas.rho0
    rhos = -gas.c_p * (D + a * D * D)
    s = rhos / rho
    T = gas.T0 * np.exp((s - gas.R * np.log(gas.rho0 / rho)) / gas.c_v)
    return rho, rhos, rho * T


def stratified_state(wl, gas, x, y):
    """w (5, len(y), len(x)) on cell centres: ambient + isentropic pulses."""
    rho_a, rhos_a, _ = ambient_rows(wl, gas, y)
    s = (rhos_a / rho_a)[:, None] * np.ones((1, len(x)))
    dp = pulses_at(wl, x, y) * gas.p0
    a = (gas.gamma1 - 1.0) / (2.0 * gas.rho0)
    # c^2 (D + a D^2 + (rho0 + D) s / c_p) = dp, D = rho - rho0 (c_lr = 0)
    b = 1.0 + s / gas.c_p
    cc = gas.rho0 * s / gas.c_p - dp / gas.c ** 2
    D = (-b + np.sqrt(b * b - 4.0 * a * cc)) / (2.0 * a)
    # the ambient rows' own root, so that dp = 0 reproduces them exactly
    D = np.where(dp == 0.0, (rho_a - gas.rho0)[:, None], D)
    rho = gas.rho0 + D
    T = gas.T0 * np.exp((s - gas.R * np.log(gas.rho0 / rho)) / gas.c_v)
    w = np.zeros((5,) + rho.shape)
    w[0] = rho
    w[3] = rho * s
    w[4] = rho * T
    return w
