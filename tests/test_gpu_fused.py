"""The fused stage path (enl_advance_patch_fused) against the plain patch loop
(enl_advance_patch): same steps, bit-identical interior state.

The fused kernels run the same per-cell arithmetic in the same order as the
rhs, runge_kutta and update_state kernels (which match the oracle to <= 1e-12);
only the storage differs (no rhs array for DRP cells, ping-pong states, no
current -> scratch copy), so the results must agree to the last bit.
"""
import numpy as np
import pytest

from enlightning_amd.gas import Gas
from enlightning_amd.workloads import SMALL_256, Workload, initial_fields
from tests import devops as D

pytestmark = pytest.mark.gpu

RK2 = ((1.0, 0.5), (0.0, 0.5), (1.0, 0.5))  # input.txt:135-138
EULER = ((1.0,), (0.0,), (1.0,))
RK3 = ((1.0, 0.75, 1.0 / 3.0), (0.0, 0.25, 2.0 / 3.0), (1.0, 0.25, 2.0 / 3.0))


@pytest.mark.parametrize("variant,level,rk", [
    ("hybrid_high_llf", 0, RK3),   # split kernel + compacted WENO list
    ("hybrid_high_llf", 1, RK3),   # sensor off: dense WENO
    ("weno_low", 0, RK2),
    ("drp", 0, RK3),
    ("hybrid_low_llf", 0, EULER),  # one-stage scheme goes through the copy-back
])
def test_fused_matches_plain(variant, level, rk):
    gas = Gas()
    wl = SMALL_256
    w, p, T, s = initial_fields(wl, gas)
    a = D.DevicePatch(wl, gas)
    b = D.DevicePatch(wl, gas)
    a.init_state(w, p, T, s)
    b.init_state(w, p, T, s)
    dt = wl.dt(gas)
    for _ in range(2):
        a.advance(4, dt, variant=variant, level=level, rk=rk)
        b.advance_fused(4, dt, variant=variant, level=level, rk=rk)
        wa, pa, Ta, _ = a.to_host()
        wb, pb, Tb, _ = b.to_host()
        g = wl.g
        assert np.array_equal(wa[:, g:-g, g:-g], wb[:, g:-g, g:-g])
        assert np.array_equal(pa[g:-g, g:-g], pb[g:-g, g:-g])
        assert np.array_equal(Ta[g:-g, g:-g], Tb[g:-g, g:-g])


def test_fused_odd_patch():
    gas = Gas(c_lr=-3.9e-3)
    wl = Workload("odd_77", n=77, g=3, seed=4, npulses=3, amp_range=(0.3, 0.8), sigma_range=(0.02, 0.05))
    w, p, T, s = initial_fields(wl, gas)
    a = D.DevicePatch(wl, gas)
    b = D.DevicePatch(wl, gas)
    a.init_state(w, p, T, s)
    b.init_state(w, p, T, s)
    a.advance(5, wl.dt(gas))
    b.advance_fused(5, wl.dt(gas))
    wa = a.to_host()[0]
    wb = b.to_host()[0]
    assert np.array_equal(wa[:, 3:-3, 3:-3], wb[:, 3:-3, 3:-3])


def _rect_patch(wl, gas, nx, ny, reflect, seed):
    """Host fields and a DevicePatch over an nx x ny box (nx != ny, so a
    swapped extent shows): the workload's initial fields cut to the box,
    with +-3 % noise on rho in alternate 16-cell stretches of the two cell
    rows along each side, which sends those boundary cells to WENO on level
    0 and leaves the others DRP."""
    g = wl.g
    w, p, T, s = (np.ascontiguousarray(a[..., :ny + 2 * g, :nx + 2 * g]) for a in initial_fields(wl, gas))
    rng = np.random.default_rng(seed)
    edge = np.zeros((ny, nx), bool)
    edge[:2] = edge[-2:] = True
    edge[:, :2] = edge[:, -2:] = True
    edge &= (np.add.outer(np.arange(ny), np.arange(nx)) // 16) % 2 == 0
    w[0, g:-g, g:-g] *= np.where(edge, 1 + 0.03 * rng.choice([-1.0, 1.0], size=(ny, nx)), 1.0)
    d = D.DevicePatch(wl, gas)
    d.box = (0, nx - 1, 0, ny - 1)
    d.reflect = reflect
    d.w = D.zeros((5, ny + 2 * g, nx + 2 * g))
    d.p, d.T, d.src = (D.zeros((ny + 2 * g, nx + 2 * g)) for _ in range(3))
    d.ww, d.KK = D.zeros((5, ny, nx)), D.zeros((5, ny, nx))
    d.init_state(w, p, T, s)
    return (w, p, T, s), d


@pytest.mark.skipif(D.DRY, reason="the CPU stand-in refreshes ghosts before each stage, not after")
@pytest.mark.parametrize("variant,level,rk,reflect", [
    ("hybrid_high_llf", 0, RK3, (1, 1, 1, 1)),    # split kernel + WENO list at the sides
    ("hybrid_high_llf", 0, RK3, (0, 1, 1, 0)),
    ("hybrid_high_llf", 1, RK3, (1, 0, 0, 1)),    # dense WENO kernel
    ("drp", 0, RK2, (0, 0, 1, 1)),
    ("hybrid_low_llf", 0, EULER, (1, 0, 1, 0)),   # one stage: the folded state is copied back
])
def test_fused_stages_write_the_physical_ghosts(variant, level, rk, reflect):
    """Without an absorbing envelope the fused stage kernels write their
    output's physical boundary ghosts themselves (the boundary launch runs
    once per call): the interiors stay bit-identical to the plain loop,
    which fills the ghosts before every stage, and the returned ghosts are
    the oracle's boundary fill of the returned interior, bit for bit."""
    import oracle
    gas = Gas()
    wl = Workload("rect", n=160, g=4, seed=9, npulses=4, amp_range=(0.1, 0.4), sigma_range=(0.03, 0.06))
    nx, ny, g = 150, 97, wl.g
    fields, a = _rect_patch(wl, gas, nx, ny, reflect, seed=3)
    _, b = _rect_patch(wl, gas, nx, ny, reflect, seed=3)
    if level == 0 and variant.startswith("hybrid"):
        # the oracle's switch sends some cells on every side to WENO, and
        # keeps others DRP
        w, p, T, s = (x.copy() for x in fields)
        oracle.fill_physical_bc(a.box, g, (1, 1, 1, 1), reflect, w, p, T, s)
        oracle.update_state(a.box, g, a.xlo, a.dx, w, p, T, gas)
        oracle.fill_physical_bc(a.box, g, (1, 1, 1, 1), reflect, w, p, T, s)
        sw = np.zeros((ny, nx), np.int32)
        oracle.rhs(variant, a.box, g, a.xlo, a.dx, w, s, np.zeros((5, ny, nx)), p, T, gas, level, sw)
        for side in (sw[0], sw[-1], sw[:, 0], sw[:, -1]):
            assert side.any() and not side.all()
        assert 0 < sw[0, 0] + sw[0, -1] + sw[-1, 0] + sw[-1, -1] < 4
    dt = wl.dt(gas)
    for _ in range(2):
        a.advance(3, dt, variant=variant, level=level, rk=rk)
        b.advance_fused(3, dt, variant=variant, level=level, rk=rk)
        wa, pa, Ta, _ = a.to_host()
        wb, pb, Tb, _ = b.to_host()
        assert np.array_equal(wa[:, g:-g, g:-g], wb[:, g:-g, g:-g])
        assert np.array_equal(pa[g:-g, g:-g], pb[g:-g, g:-g]) and np.array_equal(Ta[g:-g, g:-g], Tb[g:-g, g:-g])
        wc, pc, Tc = wb.copy(), pb.copy(), Tb.copy()
        oracle.fill_physical_bc(b.box, g, (1, 1, 1, 1), reflect, wc, pc, Tc)
        assert np.array_equal(wb, wc) and np.array_equal(np.signbit(wb), np.signbit(wc))
        assert np.array_equal(pb, pc) and np.array_equal(Tb, Tc)


@pytest.mark.skipif(D.DRY, reason="the CPU stand-in refreshes ghosts before each stage, not after")
@pytest.mark.parametrize("nx,ny,variant,reflect", [
    (1, 40, "drp", (1, 0, 0, 1)),               # one column: a cell on both x sides, its corners on both
    (37, 1, "hybrid_high_llf", (0, 1, 1, 1)),  # one row
    (2, 3, "weno_low", (1, 1, 0, 1)),
])
def test_fused_ghosts_on_thin_patches(nx, ny, variant, reflect):
    """The boundary fold where one cell borders two opposite sides: the
    interiors match the plain loop and the ghosts the oracle's fill."""
    import oracle
    gas = Gas()
    wl = Workload("thin", n=64, g=4, seed=5, npulses=3, amp_range=(0.1, 0.4), sigma_range=(0.05, 0.1))
    g = wl.g
    _, a = _rect_patch(wl, gas, nx, ny, reflect, seed=4)
    _, b = _rect_patch(wl, gas, nx, ny, reflect, seed=4)
    dt = wl.dt(gas)
    a.advance(4, dt, variant=variant)
    b.advance_fused(4, dt, variant=variant)
    wa, pa, Ta, _ = a.to_host()
    wb, pb, Tb, _ = b.to_host()
    assert np.array_equal(wa[:, g:-g, g:-g], wb[:, g:-g, g:-g])
    assert np.array_equal(pa[g:-g, g:-g], pb[g:-g, g:-g]) and np.array_equal(Ta[g:-g, g:-g], Tb[g:-g, g:-g])
    wc, pc, Tc = wb.copy(), pb.copy(), Tb.copy()
    oracle.fill_physical_bc(b.box, g, (1, 1, 1, 1), reflect, wc, pc, Tc)
    assert np.array_equal(wb, wc) and np.array_equal(pb, pc) and np.array_equal(Tb, Tc)
