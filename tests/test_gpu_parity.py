"""Parity of the HIP kernels (called through the C ABI) with the CPU oracle.

Bar (per field, relative L-infinity against the field's max |value|):
  rhs            <= 1e-12 (the kernels keep the reference operation order and
                 IEEE division; only exp() in the N2 relaxation term may round
                 differently), DRP/WENO switch map bit-exact;
  runge_kutta    bit-exact;
  update_state   pressure bit-exact, temperature <= 1e-14 (exp/log);
  tag_cells      bit-exact;
  physical BC    bit-exact.
"""
import math

import numpy as np
import pytest

import oracle
from enlightning_amd.gas import Gas, RK3
from tests import devops as D
from tests.devops import dev, host
from tests.fields import random_state

pytestmark = pytest.mark.gpu

ALL = ["hybrid_high_llf", "hybrid_high", "hybrid_low_llf", "hybrid_low",
       "weno_high_llf", "weno_high", "weno_low_llf", "weno_low", "drp"]
RHS_TOL = 1e-12


def relerr(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


@pytest.fixture(scope="module", autouse=True)
def _stream():
    D.use_stream()
    yield


def gpu_rhs(variant, box, g, xlo, dx, w, p, T, src, gas, level):
    ny, nx = box[3] - box[2] + 1, box[1] - box[0] + 1
    KK = D.zeros((5, ny, nx))
    sm = D.zeros((ny, nx), np.int32)
    D.P.rhs(variant, box, g, xlo, dx, dev(w), dev(src), KK, dev(p), dev(T), gas, level, sm)
    return host(KK), host(sm)


def cpu_rhs(variant, box, g, xlo, dx, w, p, T, src, gas, level):
    ny, nx = box[3] - box[2] + 1, box[1] - box[0] + 1
    KK = np.zeros((5, ny, nx))
    sm = np.zeros((ny, nx), np.int32)
    oracle.rhs(variant, box, g, xlo, dx, w, src, KK, p, T, gas, level, sm)
    return KK, sm


def assert_rhs_parity(Kg, smg, Kc, smc, tag=""):
    assert np.array_equal(smg, smc), tag
    for k in range(5):
        e = relerr(Kg[k], Kc[k])
        assert e <= RHS_TOL, (tag, k, e)


@pytest.mark.parametrize("variant", ALL)
@pytest.mark.parametrize("level", [0, 2])
def test_rhs_parity(variant, level):
    gas = Gas()
    nx, ny, g = 77, 45, 4
    w, p, T, src = random_state(nx, ny, g, seed=100 + ALL.index(variant))
    box, xlo, dx = (0, nx - 1, 0, ny - 1), (0.0, 0.0), (1.0 / 512, 1.0 / 384)
    Kg, smg = gpu_rhs(variant, box, g, xlo, dx, w, p, T, src, gas, level)
    Kc, smc = cpu_rhs(variant, box, g, xlo, dx, w, p, T, src, gas, level)
    if level == 0 and variant.startswith("hybrid"):
        assert 0 < smc.sum() < smc.size  # both branches exercised
    assert_rhs_parity(Kg, smg, Kc, smc, variant)


@pytest.mark.parametrize("shape,g,box0,variant", [
    ((1, 1), 3, (0, 0), "hybrid_high_llf"),
    ((33, 1), 4, (5, -3), "hybrid_high_llf"),
    ((1, 40), 3, (-8, 100), "weno_high_llf"),
    ((65, 9), 2, (7, 7), "weno_low_llf"),
    ((64, 64), 4, (128, 64), "hybrid_low"),
    ((131, 17), 5, (-40, 12), "drp"),
])
def test_rhs_ragged_and_offset_boxes(shape, g, box0, variant):
    gas = Gas(c_lr=-3.9e-3)  # exercise the LLF height term
    nx, ny = shape
    w, p, T, src = random_state(nx, ny, g, seed=nx * 1000 + ny)
    box = (box0[0], box0[0] + nx - 1, box0[1], box0[1] + ny - 1)
    xlo, dx = (12.5, -300.0), (0.25, 0.5)
    Kg, smg = gpu_rhs(variant, box, g, xlo, dx, w, p, T, src, gas, 0)
    Kc, smc = cpu_rhs(variant, box, g, xlo, dx, w, p, T, src, gas, 0)
    assert_rhs_parity(Kg, smg, Kc, smc, str(shape))


def _step_ratio(k, ep):
    """The sensor's ratio for the triple (0, dd) of a density step from 1 to
    1 + k 2^-52 (dd = k 2^-52 exactly), in the reference's double
    operations: (2|0*dd| + ep) / (0^2 + dd^2 + ep)."""
    dd = np.float64(k) * np.float64(2.0) ** -52
    return float(np.float64(ep) / (np.float64(0.0) * np.float64(0.0) + dd * dd + np.float64(ep)))


def _threshold_cases(ep):
    """Step sizes k (in units of 2^-52) whose ratio r lands (1) on each of the
    double values next to float(0.9999), where only the exact quotient can
    decide, and (2) across the +-1e-6 relative band around it in which the
    HIP sensor leaves its single-precision estimate for the exact quotient,
    and just outside that band."""
    thr = float(np.float32(0.9999))
    d0 = math.sqrt(ep / thr - ep)
    k0 = int(round(d0 / 2.0 ** -52))
    # r decreases with k: find the last k with r >= thr
    k = k0 - 64
    while _step_ratio(k + 1, ep) >= thr:
        k += 1
    assert _step_ratio(k, ep) >= thr > _step_ratio(k + 1, ep)
    ks = list(range(k - 3, k + 5))
    for rel in (-2e-6, -1.1e-6, -9e-7, -5e-7, -1e-7, 1e-7, 5e-7, 9e-7, 1.1e-6, 2e-6):
        ks.append(int(round(math.sqrt(ep / (thr * (1 + rel)) - ep) / 2.0 ** -52)))
    return thr, sorted(set(ks))


@pytest.mark.parametrize("axis", [0, 1])
def test_sensor_near_threshold_bit_exact(axis):
    """Switch maps at level 0 from the HIP split kernel against the oracle
    (and the closed form: the step is flagged on the four cells whose triples
    hold it iff r < float(0.9999), rhs_hybrid_high_llf.m4:244-254) for steps
    whose ratio sits within a few ulps of the threshold and inside the band
    where the device decides by the exact quotient."""
    gas = Gas()
    ep = oracle.sensor_eps(gas.tag_tolerance[2])
    thr, ks = _threshold_cases(ep)
    n, g, J = 16, 4, 8
    below = 0
    for k in ks:
        r = _step_ratio(k, ep)
        w, p, T, src = random_state(n, n, g, seed=40)
        hi = 1.0 + k * 2.0 ** -52
        prof = np.where(np.arange(n + 2 * g) - g >= J, hi, 1.0)
        w[0] = prof[None, :] if axis == 0 else prof[:, None]
        w[1] = w[2] = 0.0
        box, xlo, dx = (0, n - 1, 0, n - 1), (0.0, 0.0), (1.0 / 64, 1.0 / 64)
        Kg, smg = gpu_rhs("hybrid_high_llf", box, g, xlo, dx, w, p, T, src, gas, 0)
        Kc, smc = cpu_rhs("hybrid_high_llf", box, g, xlo, dx, w, p, T, src, gas, 0)
        expect = np.zeros((n, n), np.int32)
        if r < thr:
            if axis == 0:
                expect[:, J - 2:J + 2] = 1
            else:
                expect[J - 2:J + 2, :] = 1
        assert np.array_equal(smc, expect), (k, r)
        assert np.array_equal(smg, smc), (k, r, r / thr - 1)
        assert_rhs_parity(Kg, smg, Kc, smc, str(k))
        below += r < thr
    assert 0 < below < len(ks)  # both decisions, on either side of the threshold


def test_rhs_host_pointer_staging_matches_device_path():
    gas = Gas()
    nx, ny, g = 40, 30, 4
    w, p, T, src = random_state(nx, ny, g, seed=7)
    box, xlo, dx = (0, nx - 1, 0, ny - 1), (0.0, 0.0), (0.01, 0.01)
    KK_h = np.zeros((5, ny, nx))
    sm_h = np.zeros((ny, nx), np.int32)
    D.P.rhs("hybrid_high_llf", box, g, xlo, dx, w, src, KK_h, p, T, gas, 0, sm_h)  # host arrays
    Kg, smg = gpu_rhs("hybrid_high_llf", box, g, xlo, dx, w, p, T, src, gas, 0)
    assert np.array_equal(KK_h, Kg) and np.array_equal(sm_h, smg)


@pytest.mark.parametrize("nx,g", [(70, 4), (71, 3), (64, 2)])  # 16-byte and scalar paths
def test_runge_kutta_bit_exact(nx, g):
    ny = 33
    rng = np.random.default_rng(1)
    w = rng.normal(size=(5, ny + 2 * g, nx + 2 * g))
    ww = rng.normal(size=(5, ny, nx))
    KK = rng.normal(size=(5, ny, nx)) * 1e3
    box = (3, 3 + nx - 1, -2, -2 + ny - 1)
    for a1, a2, b in zip(*RK3):
        wc = w.copy()
        oracle.runge_kutta(box, g, 3.7e-7, a1, a2, b, wc, ww, KK)
        wd = dev(w)
        D.P.runge_kutta(box, g, 3.7e-7, a1, a2, b, wd, dev(ww), dev(KK))
        assert np.array_equal(host(wd), wc)


@pytest.mark.parametrize("nx,g", [(90, 4), (91, 4), (90, 3)])  # 16-byte and scalar paths
def test_update_state_parity(nx, g):
    gas = Gas(c_lr=-3.9e-3)
    ny = 21
    w, p, T, _ = random_state(nx, ny, g, seed=3)
    box, xlo, dx = (0, nx - 1, 4, 4 + ny - 1), (0.0, 50.0), (0.5, 2.0)
    pc, Tc = p.copy(), T.copy()
    oracle.update_state(box, g, xlo, dx, w, pc, Tc, gas)
    pd, Td = dev(p), dev(T)
    D.P.update_state(box, g, xlo, dx, dev(w), pd, Td, gas)
    pg, Tg = host(pd), host(Td)
    assert np.array_equal(pg, pc)
    assert relerr(Tg, Tc) <= 1e-14
    assert np.array_equal(Tg[:g], T[:g])  # ghosts untouched


@pytest.mark.parametrize("tol,val", [((1.5, 1e-2, 5e-3), 1), ((0.2, 1e-3, 5e-3), 3), ((0.0, 0.0, 0.0), 1)])
def test_tag_cells_bit_exact(tol, val):
    nx, ny, g = 60, 50, 2
    w, _, _, _ = random_state(nx, ny, g, seed=21)
    rng = np.random.default_rng(22)
    H = np.zeros((ny + 2 * g, nx + 2 * g))
    H[10:30, 15:40] = rng.uniform(0, 3, size=(20, 25))
    tc = np.full((ny, nx), -1, np.int32)
    box = (0, nx - 1, 0, ny - 1)
    oracle.tag_cells(box, g, tc, w, H, val, tol)
    td = D.full((ny, nx), -1, np.int32)
    D.P.tag_cells(box, g, td, dev(w), dev(H), val, tol)
    tg = host(td)
    assert np.array_equal(tg, tc)
    if tol[0] == 1.5:
        assert 0 < (tc != 0).sum() < tc.size


@pytest.mark.parametrize("touch,reflect", [((1, 1, 1, 1), (1, 1, 1, 1)), ((1, 1, 1, 1), (0, 1, 0, 1)),
                                           ((1, 0, 0, 1), (1, 1, 1, 1))])
def test_physical_bc_bit_exact(touch, reflect):
    nx, ny, g = 19, 13, 4
    rng = np.random.default_rng(5)
    w = rng.normal(size=(5, ny + 2 * g, nx + 2 * g))
    p, T, s = (rng.normal(size=(ny + 2 * g, nx + 2 * g)) for _ in range(3))
    box = (0, nx - 1, 0, ny - 1)
    wc, pc, Tc, sc = w.copy(), p.copy(), T.copy(), s.copy()
    oracle.fill_physical_bc(box, g, touch, reflect, wc, pc, Tc, sc)
    wd, pd, Td, sd = dev(w), dev(p), dev(T), dev(s)
    D.P.fill_physical_bc(box, g, touch, reflect, wd, pd, Td, sd)
    assert np.array_equal(host(wd), wc)
    assert np.array_equal(host(pd), pc) and np.array_equal(host(Td), Tc) and np.array_equal(host(sd), sc)


@pytest.mark.skipif(D.DRY, reason="exercises the HIP library's error channel")
def test_error_side_channel():
    from enlightning_amd import _lib
    gas = Gas()
    w, p, T, src = random_state(8, 8, 2, seed=1)
    with pytest.raises(RuntimeError, match="ghost width"):
        gpu_rhs("hybrid_high_llf", (0, 7, 0, 7), 2, (0, 0), (1, 1), w, p, T, src, gas, 0)
    assert _lib.last_error()[0] == 0  # cleared after being reported
