"""The planner's rectangle schedules (enl_rect_t: fill, coarsen and regrid
transfer) against the numpy builders' per-cell lists of
enlightning_amd/amr.py.  Each rectangle is expanded here, on the CPU, by the
rule the public header states for enl_rect_t, and the expansion must name
exactly the numpy lists' items: the same destinations, each with the same
source and the same refinement offsets.  The device expands the same
rectangles into its lists (enl_batch_set_fill_rects) and moves the transfer
from them; tests/test_gpu_regrid.py holds those against the CPU twin."""
import numpy as np
import pytest

from enlightning_amd import amr
from tests.test_amr_native import blobs


def _sy(h):
    return np.array([p.Nx for p in h.patches], np.int64)


def _cells(rect):
    dp, sp, dl, sl, ni, nj, kind, phase = (int(v) for v in rect)
    b, a = np.divmod(np.arange(ni * nj), ni)
    return dp, sp, dl, sl, kind, phase, a, b


def expand_fill(h, level):
    rects, n1, n2 = h.plan.fill_rects(level)
    assert len(rects) == n1 + n2
    sy, r = _sy(h), h.ratio
    out = {"copy": {}, "interp": {}, "edge": {}, "corner": {}}
    for q, rect in enumerate(rects):
        dp, sp, dl, sl, kind, phase, a, b = _cells(rect)
        dst = dl + a + b * sy[dp]
        if kind == 0:
            assert q < n1
            src = sl + a + b * sy[sp]
            out["copy"].update({(dp, int(d)): (sp, int(s)) for d, s in zip(dst, src)})
        elif kind == 1:
            assert q < n1
            pa, pb = (phase & 255) + a, (phase >> 8) + b
            ia, ja = pa // r, pb // r
            src = sl + ia + ja * sy[sp]
            fx = ((pa - ia * r) + 0.5) / r - 0.5
            fy = ((pb - ja * r) + 0.5) / r - 0.5
            out["interp"].update({(dp, int(d)): (sp, int(s), float(x), float(y))
                                  for d, s, x, y in zip(dst, src, fx, fy)})
        else:
            side = kind - 1
            assert 1 <= side <= 4 and sp == dp
            src = sl + b * sy[dp] if side <= 2 else sl + a
            key = "edge" if q < n1 else "corner"
            for d, s in zip(dst, src):
                assert (dp, int(d)) not in out[key]
                out[key][(dp, int(d))] = (int(s), side)
    return out


def lists_fill(h, level):
    f = h.fill[level]
    cp, cd, cs, cl = f["copy"]
    ip, idd, ic, il, fx, fy = f["interp"]
    ep, ed, es, en = f["edge"]
    kp, kd, ks, kn = f["corner"]
    return {"copy": {(int(p), int(d)): (int(s), int(l)) for p, d, s, l in zip(cp, cd, cs, cl)},
            "interp": {(int(p), int(d)): (int(c), int(l), float(x), float(y))
                       for p, d, c, l, x, y in zip(ip, idd, ic, il, fx, fy)},
            "edge": {(int(p), int(d)): (int(s), int(n)) for p, d, s, n in zip(ep, ed, es, en)},
            "corner": {(int(p), int(d)): (int(s), int(n)) for p, d, s, n in zip(kp, kd, ks, kn)}}


@pytest.mark.parametrize("n0,tile0,seed,edge", [(64, 16, 1, False), (64, 16, 2, True), (96, 32, 3, True),
                                                 (100, 32, 4, True), (33, 16, 6, True)])
def test_fill_and_coarsen_rects_expand_to_the_numpy_lists(n0, tile0, seed, edge):
    tags = blobs(n0, seed, edge=edge)
    a = amr.build(n0, 4, 4, 1.0, tags, tile0=tile0, native=False)
    b = amr.build(n0, 4, 4, 1.0, tags, tile0=tile0, native=True)
    assert [(p.level, p.box) for p in a.patches] == [(p.level, p.box) for p in b.patches]
    for lev in range(a.nlevels):
        want, got = lists_fill(a, lev), expand_fill(b, lev)
        for kind in want:
            assert got[kind] == want[kind], (lev, kind)
        # and far fewer rectangles than cells
        rects, n1, n2 = b.plan.fill_rects(lev)
        assert len(rects) * 8 <= sum(len(v) for v in want.values())
    sy, r = _sy(b), b.ratio
    got = {}
    for rect in b.plan.coarsen_rects():
        cp, fp, cl, fl, kind, phase, x, y = _cells(rect)
        assert kind == 0 and phase == 0
        for c, f in zip(cl + x + y * sy[cp], fl + x * r + y * r * sy[fp]):
            assert (cp, int(c)) not in got
            got[(cp, int(c))] = (fp, int(f))
    cp, cl, fp, fl = a.coarsen
    assert got == {(int(p), int(c)): (int(q), int(f)) for p, c, q, f in zip(cp, cl, fp, fl)}


def test_lazy_cell_lists_equal_eager_ones():
    """The native plan builds its per-cell lists on request; they are the
    numpy builder's, list by list (also for a level-0 schedule lent by the
    previous plan)."""
    from tests.test_amr_native import assert_same_hierarchy
    t1, t2 = blobs(96, 11, edge=True), blobs(96, 12)
    old = amr.build(96, 4, 4, 1.0, t1, tile0=32, native=True)
    new = amr.build(96, 4, 4, 1.0, t2, tile0=32, previous=old, native=True)
    assert_same_hierarchy(amr.build(96, 4, 4, 1.0, t2, tile0=32, native=False), new)
    rects0, _, _ = new.plan.fill_rects(0)
    assert np.array_equal(rects0, old.plan.fill_rects(0)[0])


@pytest.mark.parametrize("n0,seed", [(64, 21), (96, 22), (100, 23)])
def test_transfer_rects_expand_to_the_numpy_lists(n0, seed):
    t1 = blobs(n0, seed, edge=True)
    t2 = np.roll(t1, (5, -7), axis=(0, 1)) | blobs(n0, seed + 100)
    old_a = amr.build(n0, 4, 4, 1.0, t1, tile0=32, native=False)
    new_a = amr.build(n0, 4, 4, 1.0, t2, tile0=32, previous=old_a, native=False)
    old_b = amr.build(n0, 4, 4, 1.0, t1, tile0=32, native=True)
    new_b = amr.build(n0, 4, 4, 1.0, t2, tile0=32, previous=old_b, native=True)
    want = amr.make_transfer(new_a, old_a, native=False)
    got = amr.transfer_rects(new_b, old_b)
    for x, y in zip(want["same"], got["same"]):
        assert np.array_equal(x, y)
    assert len(want["same"][0]) > 0
    nsy, osy, r = _sy(new_b), _sy(old_b), new_b.ratio
    copy, interp = {}, {}
    for rect in got["rects"]:
        dp, sp, dl, sl, kind, phase, a, b = _cells(rect)
        assert rect[4] * rect[5] <= 1024
        dst = dl + a + b * nsy[dp]
        if kind == 0:
            for d, s in zip(dst, sl + a + b * osy[sp]):
                assert (dp, int(d)) not in copy
                copy[(dp, int(d))] = (sp, int(s))
        else:
            assert kind == 1
            pa, pb = (phase & 255) + a, (phase >> 8) + b
            ia, ja = pa // r, pb // r
            for d, s, x, y in zip(dst, sl + ia + ja * osy[sp], ((pa - ia * r) + 0.5) / r - 0.5,
                                  ((pb - ja * r) + 0.5) / r - 0.5):
                interp[(dp, int(d))] = (sp, int(s), float(x), float(y))
    cdp, cdl, csp, csl = want["copy"]
    idp, idl, isp, isl, ifx, ify = want["interp"]
    assert copy == {(int(p), int(d)): (int(s), int(l)) for p, d, s, l in zip(cdp, cdl, csp, csl)}
    assert interp == {(int(p), int(d)): (int(s), int(l), float(x), float(y))
                      for p, d, s, l, x, y in zip(idp, idl, isp, isl, ifx, ify)}
    assert len(interp) > 0 and len(copy) > 0
