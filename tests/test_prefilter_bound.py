"""The learned switch's prefilter against the oracle's density sensor.

The prefilter calls a cell DRP without the network when each of its six
sensor pairs (t1, t2) has t1^2 + t2^2 or (t1 - t2)^2 below 5e-5 eps.  The
sensor (rhs_hybrid_high_llf.m4:239-256) picks WENO where r = (2|t1 t2| +
eps) / (t1^2 + t2^2 + eps) < 0.9999 for some pair; since 1 - r =
(|t1| - |t2|)^2 / (t1^2 + t2^2 + eps), either bound keeps 1 - r below 5e-5.
Checked here against the oracle's switch maps (not a restatement of r):
* on steep straight density profiles with bounded curvature (a stratified
  ambient is one) no cell is WENO, although t1^2 + t2^2 is far above the
  bound there: the second form is what keeps them off the network;
* on fields with kinks of every size around the bound, no cell the
  prefilter calls DRP is a sensor-WENO cell, and cells just outside the
  bound are sensor-WENO (the bound is not vacuous).
"""
import numpy as np
import pytest

import oracle
from enlightning_amd.gas import Gas, sensor_eps
from tests.fields import random_state

NX, NY, G = 40, 36, 4


def sensor_map(rho, seed):
    """The oracle's level-0 switch map of a state whose density is rho
    (ghosted), other fields plausible."""
    gas = Gas()
    w, p, T, src = random_state(NX, NY, G, seed=seed, kind="smooth")
    w = w.copy()
    u, v = w[1] / w[0], w[2] / w[0]
    s, tn = w[3] / w[0], w[4] / w[0]
    w[0] = rho
    w[1], w[2], w[3], w[4] = rho * u, rho * v, rho * s, rho * tn
    KK = np.zeros((5, NY, NX))
    sm = np.zeros((NY, NX), np.int32)
    oracle.rhs("hybrid_high_llf", (0, NX - 1, 0, NY - 1), G, (0.0, 0.0), (1.0 / 512, 1.0 / 384),
               np.ascontiguousarray(w), src, KK, p, T, gas, 0, sm)
    return sm


def pair_margin(rho):
    """Per interior cell, the largest over its six pairs of min(t1^2 + t2^2,
    (t1 - t2)^2): the cell is provably DRP when this is below 5e-5 eps."""
    worst = np.zeros((NY, NX))
    J = slice(G, G + NY)
    I = slice(G, G + NX)
    for q in (-1, 0, 1):
        for axis in (1, 0):
            def at(d):
                return np.roll(rho, -d, axis=axis)[J, I]
            t1 = at(q) - at(q - 1)
            t2 = at(q + 1) - at(q)
            worst = np.maximum(worst, np.minimum(t1 * t1 + t2 * t2, (t1 - t2) * (t1 - t2)))
    return worst


def prefilter_drp(rho, thr):
    """Cells (interior) whose six pairs are all provably smooth."""
    return pair_margin(rho) < thr


def thr():
    return 5e-5 * sensor_eps(Gas().tag_tolerance[2])


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("frac", [0.5, 0.99])
def test_steep_straight_profiles_stay_drp(axis, frac):
    """rho = rho0 + s*x + a random second difference of at most
    sqrt(frac * thr) per cell, with s = 0.3 kg/m^3 per cell: t1^2 + t2^2 ~
    0.18 >> thr, (t1 - t2)^2 < thr everywhere."""
    rng = np.random.default_rng(17 + axis)
    n = (NY if axis == 0 else NX) + 2 * G
    c = np.sqrt(frac * thr())
    d2 = rng.uniform(-c, c, size=n)  # second differences
    d1 = 0.3 + np.concatenate([[0.0], np.cumsum(d2[:-1])])
    prof = 1.2 + np.concatenate([[0.0], np.cumsum(d1[:-1])])
    prof -= prof.mean() - 1.2
    rho = np.empty((NY + 2 * G, NX + 2 * G))
    if axis == 0:
        rho[:] = prof[:, None]
    else:
        rho[:] = prof[None, :]
    assert prefilter_drp(rho, thr()).all()
    sm = sensor_map(rho, seed=3)
    assert sm.sum() == 0


def test_kinks_around_the_bound():
    """A smooth field with kinks (jumps of the slope) of size k * sqrt(thr),
    k from 0.3 to 30: every sensor-WENO cell is a candidate, and the kinks
    just outside the bound already make sensor-WENO cells."""
    rng = np.random.default_rng(5)
    t = thr()
    Ny, Nx = NY + 2 * G, NX + 2 * G
    x = np.arange(Nx)[None, :]
    y = np.arange(Ny)[:, None]
    rho = 1.2 + 1e-3 * np.sin(0.1 * x) * np.cos(0.13 * y) + 0.02 * y / Ny
    for m in range(12):
        k = 10 ** rng.uniform(np.log10(0.3), np.log10(30))
        ci, cj = rng.integers(G + 3, Nx - G - 3), rng.integers(G + 3, Ny - G - 3)
        amp = k * np.sqrt(t)
        if m % 2:
            rho = rho + amp * np.maximum(x - ci, 0)
        else:
            rho = rho + amp * np.maximum(y - cj, 0)
    sm = sensor_map(rho, seed=4)
    drp = prefilter_drp(rho, t)
    assert not (drp & (sm == 1)).any()
    assert (sm == 1).sum() > 0 and drp.sum() > 0
    # the bound is near tight: some sensor-WENO cell lies within 10x of it
    # (the sensor needs (|t1| - |t2|)^2 >= ~1e-4 eps = 2x the bound)
    closest = pair_margin(rho)[sm == 1].min() / t
    print("closest sensor-WENO cell at %.2f x the bound" % closest)
    assert closest < 10.0
    # the prefilter leaves most smooth cells to nobody: far more DRP by the
    # second form than by the first alone
    first = np.ones_like(drp)
    J, I = slice(G, G + NY), slice(G, G + NX)
    for q in (-1, 0, 1):
        for axis in (1, 0):
            def at(d):
                return np.roll(rho, -d, axis=axis)[J, I]
            t1, t2 = at(q) - at(q - 1), at(q + 1) - at(q)
            first &= t1 * t1 + t2 * t2 < t
    assert drp.sum() > first.sum()
