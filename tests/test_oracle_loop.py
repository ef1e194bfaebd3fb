"""Pin of the oracle's MOL stage loop (orc_advance_patch).

SAMRAI's MethodOfLinesIntegrator is not in the reference tree, so the stage
loop (copy current -> scratch, then per stage: ghost fill, rhs, runge_kutta,
update_state; enlightning.C:459-583) has no reference test.  It is pinned here
by a closed form.

A uniform patch at rest stays uniform: every stencil difference is zero, the
fluxes' divergence vanishes, and the rhs reduces to the local source terms of
rhs_hybrid_high_llf.m4:159-237 at zero gradients,

    rho'   = H0                                    (mass source, :159)
    (rho s)'  = (rho/T) A_n M_n - (rho/T_n) c_vn M_n   (:200-234)
    (rho T_n)' = (rho/tau_n) (T - T_n)               (:237)

with T from update_state.m4:103-105 at the current state, T_n = rhoT_n/rho,
M_n = (T - T_n)/tau_n, c_vn = 0.78 R (T*/T_n)^2 exp(-T*/T_n) and A_n =
(T/T_n - 1) c_vn.  The loop must then advance (rho, rho s, rho T_n) exactly as
the Runge-Kutta method advances this ODE.  The expected values come from the
Butcher form of each method (k_i = f(y + dt sum_j a_ij k_j), y + dt sum b_i
k_i), not from the Shu-Osher form the loop uses, so a wrong coefficient, a
stage fed the wrong state, the step-start copy taken at the wrong time or a
stale T (update_state skipped between stages) each move the result by
O(dt/tau_n)^2 or more of the step's change, far above the 1e-12 bar.
"""
import types

import numpy as np
import pytest

from enlightning_amd.gas import Gas

F078 = float(np.float32(0.78))


def _ode_rhs(y, gas, H0):
    rho, rhos, rhoTn = y
    T = gas.T0 * np.exp((rhos / rho - gas.R * np.log(gas.rho0 / rho)) / gas.c_v)
    Tn = rhoTn / rho
    Mn = (T - Tn) / gas.tau_n
    cvn = F078 * gas.R * (gas.T_star_n / Tn) ** 2 * np.exp(-gas.T_star_n / Tn)
    An = (T / Tn - 1) * cvn
    return np.array([H0, (rho / T) * An * Mn - (rho / Tn) * cvn * Mn, (rho / gas.tau_n) * (T - Tn)])


# Butcher tableaux (A, b) of the methods the MOL coefficients encode.
SSPRK3_BUTCHER = (np.array([[0, 0, 0], [1, 0, 0], [0.25, 0.25, 0]]), np.array([1 / 6, 1 / 6, 2 / 3]))
HEUN_BUTCHER = (np.array([[0, 0], [1, 0]]), np.array([0.5, 0.5]))


def _butcher_step(y, dt, tableau, f):
    A, b = tableau
    ks = []
    for i in range(len(b)):
        ks.append(f(y + dt * sum(A[i, j] * ks[j] for j in range(i))))
    return y + dt * sum(bi * ki for bi, ki in zip(b, ks))


# Shu-Osher (alpha_1, alpha_2, beta) of the same methods: SSP-RK3 with the
# exact thirds (matlab/hybrid.m:132-145) and the active RK2 of input.txt:135-138.
RK3 = ((1.0, 0.75, 1.0 / 3.0), (0.0, 0.25, 2.0 / 3.0), (1.0, 0.25, 2.0 / 3.0))
RK2 = ((1.0, 0.5), (0.0, 0.5), (1.0, 0.5))


@pytest.mark.parametrize("rk,tableau", [(RK3, SSPRK3_BUTCHER), (RK2, HEUN_BUTCHER)])
@pytest.mark.parametrize("variant,level", [("hybrid_high_llf", 0), ("hybrid_high_llf", 1), ("drp", 0),
                                           ("weno_low_llf", 0)])
def test_mol_loop_uniform_relaxation(orc, rk, tableau, variant, level):
    gas = Gas()
    n, g = 8, 4
    N = n + 2 * g
    rho, s, Tn = 1.25, 3.0, gas.T0 - 40.0
    H0 = 4.0
    dt = 0.3 * gas.tau_n
    nsteps = 4
    # 1 m cells: c*dt/dx = 0.12, so rounding residuals at the walls stay small
    wl = types.SimpleNamespace(n=n, g=g, dx=1.0, variant=variant)
    w = np.zeros((5, N, N))
    w[0], w[3], w[4] = rho, rho * s, rho * Tn
    p, T = np.zeros((N, N)), np.zeros((N, N))
    src = np.full((N, N), H0)
    h = orc.HostPatch(wl, gas)
    h.init_state(w, p, T, src)
    h.advance(nsteps, dt, rk, level=level)

    f = lambda y: _ode_rhs(y, gas, H0)
    y = np.array([rho, rho * s, rho * Tn])
    start = y.copy()
    for _ in range(nsteps):
        y = _butcher_step(y, dt, tableau, f)
    got = np.array([h.w[k][g:-g, g:-g] for k in (0, 3, 4)])
    change = np.abs(y - start)
    assert (change > 1e-3 * np.abs(y)).all(), change  # source and relaxation move every field
    for c, k in enumerate((0, 3, 4)):
        assert np.abs(got[c] - y[c]).max() <= 1e-12 * np.abs(y[c]), (k, got[c].ravel()[:2], y[c])
    # the state stays at rest and uniform
    assert np.abs(h.w[1]).max() < 1e-9 and np.abs(h.w[2]).max() < 1e-9
    for k in (0, 3, 4):
        assert np.ptp(h.w[k][g:-g, g:-g]) <= 1e-13 * np.abs(h.w[k]).max()
