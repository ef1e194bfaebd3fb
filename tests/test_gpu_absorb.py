"""Absorbing envelopes on the device (enl_absorb, enlightning.C:630-786).

* enl_absorb against the oracle, bit for bit (the tmp table is built on the
  host with the reference's expressions; the device applies it with the
  same unfused arithmetic), over patch shapes that exercise the selection:
  a whole-domain patch with all four sides absorbing, a patch touching two
  edges, a patch inside the domain (edge cells only, with tmp = 1) and a
  band narrower than one cell (the pre-loop test skips the first row).
* The patch loops with ABSORB sides (the reference input's left, right and
  top) against the oracle's loop: 1e-10 per field after 20 RK3 steps; the
  fused loop bit-identical to the plain one.
"""
import numpy as np
import pytest

import oracle
from enlightning_amd.gas import Gas, RK3
from enlightning_amd.workloads import Workload, initial_fields
from tests import devops as D

pytestmark = pytest.mark.gpu


def random_state(nx, ny, g, gas, seed):
    rng = np.random.default_rng(seed)
    N0, N1 = nx + 2 * g, ny + 2 * g
    w = np.empty((5, N1, N0))
    w[0] = gas.rho0 * (1 + 0.05 * rng.standard_normal((N1, N0)))
    w[1:4] = rng.standard_normal((3, N1, N0))
    w[4] = gas.T0 * gas.rho0 + 3.0 * rng.standard_normal((N1, N0))
    p = gas.p0 + 50.0 * rng.standard_normal((N1, N0))
    T = gas.T0 + rng.standard_normal((N1, N0))
    return w, p, T


CASES = {
    # box, ghosts, patch xlo, dx, domain lo, domain hi, sides, width
    "whole_domain": ((0, 63, 0, 47), 4, (0.0, 0.0), (1 / 64, 1 / 48), (0.0, 0.0), (1.0, 1.0), (1, 1, 1, 1), 0.06),
    "two_edges": ((0, 39, 0, 29), 3, (0.0, 0.0), (1 / 80, 1 / 60), (0.0, 0.0), (1.0, 1.0), (1, 0, 1, 0), 0.05),
    "far_right_top": ((50, 79, 30, 59), 4, (50 / 80, 30 / 60), (1 / 80, 1 / 60), (0.0, 0.0), (1.0, 1.0),
                      (1, 1, 1, 1), 0.04),
    "inside": ((100, 131, 100, 131), 3, (100 / 256, 100 / 256), (1 / 256, 1 / 256), (0.0, 0.0), (1.0, 1.0),
               (1, 1, 1, 1), 2 / 256),
    "sub_cell_band": ((0, 31, 0, 31), 3, (0.0, 0.0), (1 / 32, 1 / 32), (0.0, 0.0), (1.0, 1.0), (1, 1, 1, 1), 1e-5),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_absorb_matches_oracle(case):
    box, g, pxlo, dx, glo, ghi, sides, width = CASES[case]
    gas = Gas()
    nx, ny = box[1] - box[0] + 1, box[3] - box[2] + 1
    w, p, T = random_state(nx, ny, g, gas, seed=len(case))
    wh, ph, Th = w.copy(), p.copy(), T.copy()
    oracle.absorb(box, g, pxlo, dx, glo, ghi, sides, width, gas, wh, ph, Th)
    wd, pd, Td = D.dev(w), D.dev(p), D.dev(T)
    D.P.absorb(box, g, pxlo, dx, glo, ghi, sides, width, gas, wd, pd, Td)
    wd, pd, Td = D.host(wd), D.host(pd), D.host(Td)
    assert np.array_equal(wd, wh) and np.array_equal(pd, ph) and np.array_equal(Td, Th), case
    # the envelope did something where the band covers cells ("inside" and
    # "sub_cell_band" only touch edge cells with tmp = 1)
    if case not in ("inside", "sub_cell_band"):
        assert not np.array_equal(wh, w)


def test_patch_loops_with_absorbing_sides():
    if D.DRY:
        pytest.skip("the CPU stand-in advances in 64-cell tiles, each with its own envelope")
    gas = Gas()
    wl = Workload("absorb_96", n=96, g=4, seed=7, npulses=3, amp_range=(0.05, 0.2), sigma_range=(0.03, 0.06))
    w, p, T, s = initial_fields(wl, gas)
    reflect = (0, 0, 1, 0)  # input.txt:14-17: ABSORB left, right, top; REFLECT bottom
    width = 0.1
    hp = oracle.HostPatch(wl, gas)
    hp.reflect = reflect
    hp.absorb_width = width
    hp.init_state(w, p, T, s)
    a = D.DevicePatch(wl, gas)
    b = D.DevicePatch(wl, gas)
    for d in (a, b):
        d.reflect = reflect
        d.absorb_width = width
        d.init_state(w, p, T, s)
    dt = wl.dt(gas)
    g = wl.g
    for _ in range(4):
        hp.advance(5, dt, RK3, tile=0)
        a.advance(5, dt)
        b.advance_fused(5, dt)
        wa, pa, Ta, _ = a.to_host()
        wb, pb, Tb, _ = b.to_host()
        assert np.array_equal(wa[:, g:-g, g:-g], wb[:, g:-g, g:-g])
        assert np.array_equal(pa[g:-g, g:-g], pb[g:-g, g:-g]) and np.array_equal(Ta[g:-g, g:-g], Tb[g:-g, g:-g])
        for k in range(5):
            ref = hp.w[k, g:-g, g:-g]
            scale = max(np.abs(ref - ref.mean()).max(), 1e-300) if k in (0, 4) else max(np.abs(ref).max(), 1e-300)
            assert np.abs(wa[k, g:-g, g:-g] - ref).max() <= 1e-10 * scale, k
        assert np.abs(pa[g:-g, g:-g] - hp.p[g:-g, g:-g]).max() <= 1e-9 * np.abs(hp.p - gas.p0).max()
    # the envelopes damp: the left edge column is closer to ambient than the middle
    assert np.abs(wa[1, g:-g, g]).max() < np.abs(wa[1, g:-g, g + 48]).max()
