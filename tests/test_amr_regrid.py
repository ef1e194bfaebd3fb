"""Regrid transfer (main.C:489-502) on the CPU: the schedule
(enlightning_amd.amr.make_transfer) and the oracle's transfer
(oracle.amr.HostHierarchy.transfer_from).

SAMRAI's regrid is not vendored, so the transfer is pinned by what its
operators must do, independent of the code under test:

* every interior cell of every new patch is listed exactly once (whole
  patches whose box did not change, else cell by cell); a copy names the
  same global cell on an old patch of the same level; a refine names the
  coarse cell (i // r, j // r) with the fine centre's offset; the level-0
  fill schedule lent by the previous hierarchy equals a fresh one;
* copies are exact, and a transfer onto the same boxes is the identity;
* linear refinement reproduces linear fields (cells whose coarse stencil
  leaves the domain excepted: the wall ghosts are copies, not linear);
* the r x r fine cells refined from one coarse cell average back to it
  (centred slopes: the offsets sum to zero), for arbitrary data.
"""
import numpy as np
import pytest

import oracle
from enlightning_amd import amr
from enlightning_amd.gas import Gas
from enlightning_amd.hierarchy import level_tags
from enlightning_amd.workloads import Workload
from oracle.amr import HostHierarchy

WL = Workload("amr_test_64", n=64, seed=7, npulses=4, amp_range=(0.3, 0.8), sigma_range=(0.03, 0.06))


@pytest.fixture(scope="module")
def pair():
    gas = Gas()
    tags = level_tags(gas, WL, oracle.tag_cells)
    old = amr.build(WL.n, 4, 4, 1.0, tags, tile0=16)
    shifted = np.roll(tags, (6, -5), axis=(0, 1))
    new = amr.build(WL.n, 4, 4, 1.0, shifted, tile0=16)
    return old, new


def global_of(p, local):
    return local % p.Nx - p.g + p.box[0], local // p.Nx - p.g + p.box[2]


def test_schedule_covers_each_interior_cell_once(pair):
    old, new = pair
    tr = amr.make_transfer(new, old)
    r = new.ratio
    dp, dl, sp, sl = tr["copy"]
    ip, il, isp, isl, fx, fy = tr["interp"]
    assert len(ip) > 0 and len(dp) > 0
    sn, so = tr["same"]
    seen = {p.pid: np.zeros(p.Nx * p.Ny, np.int32) for p in new.patches}
    # unchanged patches: every level-0 patch, and fine boxes that survived
    assert len(sn) > len(new.level_patches(0))
    for a, b in zip(sn, so):
        A, B = new.patches[a], old.patches[b]
        assert (A.level, A.box) == (B.level, B.box)
        m = np.zeros((A.Ny, A.Nx), np.int32)
        m[A.g:-A.g, A.g:-A.g] = 1
        seen[a] += m.ravel()
    for q in range(len(dp)):
        seen[dp[q]][dl[q]] += 1
    for q in range(len(ip)):
        seen[ip[q]][il[q]] += 1
    for p in new.patches:
        want = np.zeros((p.Ny, p.Nx), np.int32)
        want[p.g:-p.g, p.g:-p.g] = 1
        assert np.array_equal(seen[p.pid], want.ravel()), p.pid
    # copies: same level, same global cell
    newp, oldp = new.patches, old.patches
    for q in range(len(dp)):
        D, S = newp[dp[q]], oldp[sp[q]]
        assert D.level == S.level
        assert global_of(D, dl[q]) == global_of(S, sl[q])
    # refines: only where no old patch of the level held the cell, from the
    # coarse cell (i // r, j // r) one level down
    old_fine = np.zeros((old.ncells(1),) * 2, bool)
    for p in old.level_patches(1):
        old_fine[p.box[2]:p.box[3] + 1, p.box[0]:p.box[1] + 1] = True
    for q in range(len(ip)):
        D, C = newp[ip[q]], oldp[isp[q]]
        assert D.level == 1 and C.level == 0
        i, j = global_of(D, il[q])
        assert not old_fine[j, i]
        assert global_of(C, isl[q]) == (i // r, j // r)
        assert fx[q] == ((i % r) + 0.5) / r - 0.5 and fy[q] == ((j % r) + 0.5) / r - 0.5
    # level 0 is unchanged: whole-patch pairs only
    assert all(newp[p].level == 1 for p in ip) and all(newp[p].level == 1 for p in dp)


def test_level0_fill_lent_by_the_previous_hierarchy(pair):
    old, new = pair
    tags = np.zeros((WL.n, WL.n), np.int32)
    tags[10:14, 20:30] = 1
    fresh = amr.build(WL.n, 4, 4, 1.0, tags, tile0=16)
    lent = amr.build(WL.n, 4, 4, 1.0, tags, tile0=16, previous=old)
    assert lent.fill[0] is old.fill[0]
    for key in fresh.fill[0]:
        for x, y in zip(fresh.fill[0][key], lent.fill[0][key]):
            assert np.array_equal(x, y), key
    for key in fresh.fill[1]:
        for x, y in zip(fresh.fill[1][key], lent.fill[1][key]):
            assert np.array_equal(x, y), key
    # a different base tiling is not lent
    other = amr.build(WL.n, 4, 4, 1.0, tags, tile0=32, previous=old)
    assert other.fill[0] is not old.fill[0]


def test_identity_on_the_same_boxes(pair):
    old, _ = pair
    gas = Gas()
    a = HostHierarchy(old, gas)
    rng = np.random.default_rng(1)
    a.pool[:] = rng.standard_normal(a.pool.size)
    for s in a.src:
        s[...] = rng.standard_normal(s.shape)
    b = HostHierarchy(old, gas)
    tr = amr.make_transfer(old, old)
    b.transfer_from(a, tr)
    assert len(tr["interp"][0]) == 0 and len(tr["copy"][0]) == 0 and len(tr["same"][0]) == len(old.patches)
    for k, p in enumerate(old.patches):
        g = p.g
        wa, pa, Ta = a.view(k)
        wb, pb, Tb = b.view(k)
        assert np.array_equal(wa[:, g:-g, g:-g], wb[:, g:-g, g:-g])
        assert np.array_equal(pa[g:-g, g:-g], pb[g:-g, g:-g]) and np.array_equal(Ta[g:-g, g:-g], Tb[g:-g, g:-g])
        assert np.array_equal(a.src[k][g:-g, g:-g], b.src[k][g:-g, g:-g])


def _linear(h, coef):
    """Per patch: 8 components (w_n, p, T, source) a + bx*x + by*y on every
    ghosted cell centre."""
    out = []
    for p in h.patches:
        dx = h.dx(p.level)
        X, Y = np.meshgrid((np.arange(p.box[0] - p.g, p.box[1] + p.g + 1) + 0.5) * dx,
                           (np.arange(p.box[2] - p.g, p.box[3] + p.g + 1) + 0.5) * dx)
        out.append(np.stack([a + bx * X + by * Y for a, bx, by in coef]))
    return out


def test_linear_fields_and_exact_copies(pair):
    old, new = pair
    gas = Gas()
    coef = [(1.2, 0.1, -0.05), (0.3, 2.0, 1.0), (-0.2, 0.5, -1.5), (0.05, 0.01, 0.02), (350.0, 3.0, -2.0),
            (101325.0, 100.0, 50.0), (294.0, 1.0, -0.5), (0.7, -0.3, 0.2)]
    a = HostHierarchy(old, gas)
    for k, f in enumerate(_linear(old, coef)):
        w, pr, T = a.view(k)
        w[...], pr[...], T[...] = f[:5], f[5], f[6]
        a.src[k][...] = f[7]
    before = [tuple(x.copy() for x in a.view(k)) for k in range(len(old.patches))]
    b = HostHierarchy(new, gas)
    tr = amr.make_transfer(new, old)
    b.transfer_from(a, tr)
    # the old interiors are untouched (only its ghosts were refilled)
    for k, p in enumerate(old.patches):
        g = p.g
        for x, y in zip(before[k], a.view(k)):
            assert np.array_equal(x[..., g:-g, g:-g], y[..., g:-g, g:-g])
    want = _linear(new, coef)
    n0, r = new.ncells(0), new.ratio
    checked = 0
    for k, p in enumerate(new.patches):
        g = p.g
        w, pr, T = b.view(k)
        got = np.concatenate([w, pr[None], T[None], b.src[k][None]])[:, g:-g, g:-g]
        exp = want[k][:, g:-g, g:-g]
        I, J = np.meshgrid(np.arange(p.box[0], p.box[1] + 1), np.arange(p.box[2], p.box[3] + 1))
        if p.level == 0:
            ok = np.ones(I.shape, bool)
        else:
            ic, jc = I // r, J // r
            ok = (ic > 0) & (ic < n0 - 1) & (jc > 0) & (jc < n0 - 1)
        for c in range(8):
            scale = abs(coef[c][0]) + abs(coef[c][1]) + abs(coef[c][2])
            assert np.abs(got[c][ok] - exp[c][ok]).max() <= 1e-13 * scale, (k, c)
        checked += ok.sum()
    assert checked > 0.9 * sum(p.nx * p.ny for p in new.patches)
    # copies carry the old values bit for bit
    for x, y in zip(*tr["same"]):
        g = new.patches[x].g
        for u, v in zip(b.view(x), a.view(y)):
            assert np.array_equal(u[..., g:-g, g:-g], v[..., g:-g, g:-g])
    dp, dl, sp, sl = tr["copy"]
    for q in range(0, len(dp), 7):
        D, S = new.patches[dp[q]], old.patches[sp[q]]
        wd = b.view(dp[q])[0].reshape(5, -1)
        ws = a.view(sp[q])[0].reshape(5, -1)
        assert np.array_equal(wd[:, dl[q]], ws[:, sl[q]])
        assert b.src[dp[q]].ravel()[dl[q]] == a.src[sp[q]].ravel()[sl[q]]


def test_refined_cells_average_back_to_the_coarse_cell(pair):
    old, new = pair
    gas = Gas()
    a = HostHierarchy(old, gas)
    rng = np.random.default_rng(3)
    a.pool[:] = 1.0 + rng.random(a.pool.size)
    b = HostHierarchy(new, gas)
    tr = amr.make_transfer(new, old)
    b.transfer_from(a, tr)
    r = new.ratio
    # coarse cells whose r x r fine cells were all refined
    ip, il, isp, isl, _, _ = tr["interp"]
    groups = {}
    for q in range(len(ip)):
        groups.setdefault((int(ip[q]), int(isp[q]), int(isl[q])), []).append(int(il[q]))
    full = [(k, v) for k, v in groups.items() if len(v) == r * r]
    assert len(full) > 20
    for (pid, cid, cl), fine in full:
        wf = b.view(pid)[0].reshape(5, -1)[:, fine]
        wc = a.view(cid)[0].reshape(5, -1)[:, cl]
        assert np.abs(wf.mean(axis=1) - wc).max() <= 1e-15 * np.abs(wc).max() * r * r
