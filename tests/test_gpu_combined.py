"""The learned switch inside the batched FP32-perturbation hierarchy.

Before level 0's rhs the prefilter and the network run over every level-0
patch of the hierarchy at once and write a decision map per patch, which the
FP32 stage kernel reads in place of the density sensor
(rhs_hybrid_high_llf.m4:239-256, level 0 only).  Checked here:
* with every candidate left to the sensor (lo = 0, hi = 1, decisions every
  stage) the combined form is bit-identical to the FP32 form with the sensor:
  the batched prefilter marks DRP only where the FP32 sensor says DRP;
* on a uniform ambient (where the perturbations' differences are the
  field's) the batched decisions on the FP32 perturbations agree with the
  single-patch switch's decisions on the FP64 state of each level-0 patch;
* with the trained network, weak acoustic pulses on a stratified ambient
  stay as close to the FP64 form as the FP32 form with the sensor.  Cells
  that are candidates only through the ambient (next to the reflecting
  walls) are left to the sensor; the network's DRP calls are switched off
  (lo = 0): next to the walls the ambient's own adjustment makes
  sensor-WENO cells unlike any the network was trained on (shocks of the
  channel workload on a uniform ambient), and it calls some of them DRP.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from enlightning_amd import amr
from enlightning_amd.gas import Gas
from enlightning_amd.hierarchy import level_tags
from enlightning_amd.workloads import StratifiedWorkload, Workload, ambient_rows, stratified_state
from tests import devops as D

pytestmark = pytest.mark.gpu

if D.DRY:
    pytest.skip("the combined form has no CPU stand-in", allow_module_level=True)

from enlightning_amd import switch as S  # noqa: E402

TAGS_WL = Workload("amr_test_64", n=64, seed=7, npulses=4, amp_range=(0.3, 0.8), sigma_range=(0.03, 0.06))
# strong pulses on a stratified ambient: the sensor flags WENO cells on level 0
STRONG = StratifiedWorkload("strong_test_64", n=64, L=100.0, seed=11, npulses=4, amp_range=(0.3, 0.8),
                            sigma_range=(0.03, 0.06), scale_height=400.0)
# the same pulses on a uniform ambient
STRONG_UNIFORM = StratifiedWorkload("strong_uniform_test_64", n=64, L=100.0, seed=11, npulses=4,
                                    amp_range=(0.3, 0.8), sigma_range=(0.03, 0.06), scale_height=float("inf"))
WEAK = StratifiedWorkload("weak_test_64", n=64, L=100.0, seed=5, npulses=4, amp_range=(1e-4, 1e-3),
                          sigma_range=(0.05, 0.1), scale_height=400.0)


def setup(wl):
    gas = Gas()
    tags = level_tags(gas, TAGS_WL, oracle.tag_cells)
    h = amr.build(64, 4, 4, wl.L, tags, tile0=16)
    assert h.nlevels == 2
    states = []
    for p in h.patches:
        dx = h.dx(p.level)
        x = (np.arange(p.box[0] - p.g, p.box[1] + p.g + 1) + 0.5) * dx
        y = (np.arange(p.box[2] - p.g, p.box[3] + p.g + 1) + 0.5) * dx
        states.append(stratified_state(wl, gas, x, y))

    def ambient(lev):
        n, g = h.ncells(lev), h.g
        y = (np.arange(-g, n + g) + 0.5) * h.dx(lev)
        return (-g,) + ambient_rows(wl, gas, y)

    return gas, h, states, ambient


def fp32_run(d, wl, states, ambient, steps, dt, sw=None, lo=0.0, hi=1.0, every_stage=True, totals=None):
    d.load_initial(wl, states)
    d.set_ambient(ambient)
    d.fp32_set_switch(sw, lo, hi, every_stage, totals)
    d.fp32_prepare()
    d.advance(steps, dt, mode="fp32_graph")
    d.fp32_store()
    return d.states()


def test_undecided_switch_equals_sensor():
    gas, h, states, ambient = setup(STRONG)
    dt = 0.5 * h.dx(1) / gas.c
    d = D.Hierarchy(h, gas)
    d.load_initial(STRONG, states)
    d0 = d.states()  # ghosts filled
    ref = fp32_run(d, STRONG, states, ambient, 4, dt)
    sw = S.Switch()
    totals = torch.zeros(3, dtype=torch.int64, device="cuda")
    got = fp32_run(d, STRONG, states, ambient, 4, dt, sw, 0.0, 1.0, True, totals)
    cand, weno, undecided = (int(v) for v in totals.cpu())
    d.close()
    sw.close()
    # the oracle's sensor flags WENO cells of level 0 on the initial state,
    # so the identity covers both schemes
    flagged = 0
    for p, (w, pr, T) in zip(h.patches, d0):
        if p.level == 0:
            sm = np.zeros((p.ny, p.nx), np.int32)
            KK = np.zeros((5, p.ny, p.nx))
            src = np.zeros((p.ny, p.nx))
            oracle.rhs("hybrid_high_llf", p.box, p.g, (0.0, 0.0), (h.dx(0), h.dx(0)), w, src, KK, pr, T, gas, 0,
                       switch_map=sm)
            flagged += int(sm.sum())
    assert flagged > 0
    assert cand > 0 and weno == 0 and undecided == cand, (cand, weno, undecided)
    for a, b in zip(ref, got):
        for k in range(3):
            assert np.array_equal(a[k], b[k])


def _single_patch_decisions(sw, p, gas, w, pr, lo, hi):
    wt = torch.from_numpy(np.ascontiguousarray(w)).cuda()
    pt = torch.from_numpy(np.ascontiguousarray(pr)).cuda()
    dec = torch.empty((p.ny, p.nx), dtype=torch.uint8, device="cuda")
    counts = (ctypes.c_uint * 3)()
    gs = gas.to_struct()
    torch.cuda.synchronize()
    rc = sw.lib.enl_switch_decide(sw.handle, p.box[0], p.box[1], p.box[2], p.box[3], p.g,
                                  ctypes.c_void_p(wt.data_ptr()), ctypes.c_void_p(pt.data_ptr()), ctypes.byref(gs),
                                  float(lo), float(hi), ctypes.c_void_p(dec.data_ptr()), counts)
    assert rc == 0
    torch.cuda.synchronize()
    return dec.cpu().numpy()


def test_batched_decisions_match_single_patch_switch():
    params = S.load_params()
    if params is None:
        pytest.skip("no trained parameters")
    lo, hi = 0.2, 0.8
    wl = STRONG_UNIFORM
    gas, h, states, ambient = setup(wl)
    dt = 0.5 * h.dx(1) / gas.c
    d = D.Hierarchy(h, gas)
    sw = S.Switch(params)
    d.load_initial(wl, states)
    s0 = d.states()  # ghosts filled
    # one step with decisions kept from its first stage: the maps left are
    # those of the initial state
    fp32_run(d, wl, states, ambient, 1, dt, sw, lo, hi, every_stage=False)
    total = differ = cand = 0
    for pid, p in enumerate(h.patches):
        if p.level != 0:
            continue
        batched = d.fp32_decisions(pid)
        single = _single_patch_decisions(sw, p, gas, s0[pid][0], s0[pid][1], lo, hi)
        total += batched.size
        cand += int((single != 0).sum())
        differ += int((batched != single).sum())
    d.close()
    sw.close()
    print("level-0 cells %d, single-patch candidates %d, decisions that differ %d" % (total, cand, differ))
    assert cand > 0
    # the FP32 perturbations round the differences the features are made of
    # (BF16-rounded afterwards) and the prefilter's sums near its threshold
    assert differ <= max(2, 2e-3 * cand), (differ, cand)


def test_trained_switch_weak_pulses_error():
    params = S.load_params()
    if params is None:
        pytest.skip("no trained parameters")
    gas, h, states, ambient = setup(WEAK)
    dt = 0.5 * h.dx(1) / gas.c
    d = D.Hierarchy(h, gas)
    d.load_initial(WEAK, states)
    d.advance(20, dt, mode="graph")
    s64 = d.states()
    s32 = fp32_run(d, WEAK, states, ambient, 20, dt)
    sw = S.Switch(params)
    totals = torch.zeros(3, dtype=torch.int64, device="cuda")
    sc = fp32_run(d, WEAK, states, ambient, 20, dt, sw, 0.0, 0.8, every_stage=True, totals=totals)
    cand, weno, undecided = (int(v) for v in totals.cpu())
    print("candidates %d, network WENO %d, left to the sensor %d" % (cand, weno, undecided))
    assert cand > 0 and weno == 0
    d.close()
    sw.close()
    for f_i in range(6):
        num32 = numc = den = 0.0
        for p, a, b, c in zip(h.patches, s64, s32, sc):
            g = p.g
            pick = (lambda st: st[0][f_i]) if f_i < 5 else (lambda st: st[1])
            x, y, z = (pick(s)[g:-g, g:-g] for s in (a, b, c))
            dx = h.dx(p.level)
            yy = (np.arange(p.box[2], p.box[3] + 1) + 0.5) * dx
            amb = ambient_rows(WEAK, gas, yy)
            ref = x - amb[(0, None, None, 1, 2)[f_i]][:, None] if f_i in (0, 3, 4) else (
                x if f_i in (1, 2) else x - gas.p0)
            num32 += float(np.sum((x - y) ** 2))
            numc += float(np.sum((x - z) ** 2))
            den += float(np.sum(ref ** 2))
        e32, ec = np.sqrt(num32 / den), np.sqrt(numc / den)
        print("field %d: FP32 sensor %.3e, FP32 + switch %.3e" % (f_i, e32, ec))
        assert ec <= max(2 * e32, 1e-6), (f_i, e32, ec)
