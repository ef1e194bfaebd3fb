"""Checkpoint / resume of the device hierarchy (enlightning_amd.output;
enlightning.C:1859-2060 and SAMRAI's patch-data restart).

A run of 3 + 2 RK3 steps and a run of 3 steps, checkpoint, a fresh
hierarchy loaded from the checkpoint and 2 more steps end in bit-identical
interiors: the reference's restart resumes the same trajectory, with
absorbing sides and a nonzero source.  The plot quantities read back from
the device equal the interiors of the state.  The one-launch gather and
scatter of every patch's SOLN planes (enl_batch_gather / _scatter: plot
quantities, checkpoints) equal the per-patch copies.
"""
import numpy as np
import pytest

import oracle
from enlightning_amd import amr, output
from enlightning_amd.gas import Gas
from enlightning_amd.hierarchy import level_tags, patch_initial_state
from enlightning_amd.workloads import Workload
from tests import devops as D

pytestmark = pytest.mark.gpu

WL64 = Workload("amr_test_64", n=64, seed=7, npulses=4, amp_range=(0.3, 0.8), sigma_range=(0.03, 0.06))


def test_resume_from_checkpoint_is_bit_identical(tmp_path):
    gas = Gas()
    h = amr.build(WL64.n, 4, 4, 1.0, level_tags(gas, WL64, oracle.tag_cells), tile0=16)
    dt = 0.5 * h.dx(1) / gas.c
    rng = np.random.default_rng(4)
    states = [patch_initial_state(h, p, WL64, gas) for p in h.patches]
    sources = [1e-3 * rng.standard_normal((p.Ny, p.Nx)) for p in h.patches]
    a = D.Hierarchy(h, gas, reflect=(0, 0, 1, 0))
    a.set_absorb(0.1)
    a.load_initial(WL64, states=states, sources=sources)
    a.advance(3, dt, mode="graph")
    path = tmp_path / "chk.npz"
    output.save_restart(path, a, iteration=3, loop_time=3 * dt)
    a.advance(2, dt, mode="graph")
    b, d = output.load_restart(path, factory=D.Hierarchy)
    assert d["p_iteration_number"] == 3 and b.reflect == (0, 0, 1, 0)
    b.advance(2, dt, mode="graph")
    for (wa, pa, Ta), (wb, pb, Tb), p in zip(a.states(), b.states(), h.patches):
        g = p.g
        assert np.array_equal(wa[:, g:-g, g:-g], wb[:, g:-g, g:-g]), p.pid
        assert np.array_equal(pa[g:-g, g:-g], pb[g:-g, g:-g]) and np.array_equal(Ta[g:-g, g:-g], Tb[g:-g, g:-g])
    q = output.plot_quantities(b, viz_w_n=1, viz_source=1)
    for k, ((w, pr, _), s, p) in enumerate(zip(b.states(), b.sources(), h.patches)):
        g = p.g
        assert np.array_equal(q["w_n #0"][k], w[0, g:-g, g:-g]) and np.array_equal(q["pressure"][k], pr[g:-g, g:-g])
        assert np.array_equal(q["source"][k], s[g:-g, g:-g])
    a.close()
    b.close()


def per_patch_copies(dev, h):
    """Every patch's ghosted w_n, pres, temp and source through the
    per-patch copy (enl_batch_copy_state), the path the gather replaces."""
    from enlightning_amd import _lib
    from enlightning_amd.hierarchy import _p
    out = []
    for k, p in enumerate(h.patches):
        w, pr, T, s = np.empty((5, p.Ny, p.Nx)), np.empty((p.Ny, p.Nx)), np.empty((p.Ny, p.Nx)), np.empty((p.Ny, p.Nx))
        _lib.check(dev.lib.enl_batch_copy_state(dev.handle, k, 0, _p(w), _p(pr), _p(T), _p(s)))
        out.append(np.concatenate([w, pr[None], T[None], s[None]]))
    return out


def test_gather_and_scatter_match_per_patch_copies():
    """enl_batch_gather (one launch over every patch) returns, plane by plane,
    the values the per-patch copies read, over the ghosted boxes and over the
    interiors, for every plane subset tried; enl_batch_scatter of a gathered
    state into a fresh hierarchy reproduces it bit for bit; a buffer of the
    wrong size is refused."""
    gas = Gas()
    h = amr.build(WL64.n, 4, 4, 1.0, level_tags(gas, WL64, oracle.tag_cells), tile0=16)
    a = D.Hierarchy(h, gas)
    if not hasattr(a, "gather"):
        pytest.skip("the gather is the C ABI's")
    rng = np.random.default_rng(11)
    a.load_initial(WL64, sources=[rng.standard_normal((p.Ny, p.Nx)) for p in h.patches])
    a.advance(2, 0.5 * h.dx(1) / gas.c, mode="graph")
    ref = per_patch_copies(a, h)
    for planes, ghosted in ((0xFF, True), (0x7F, True), (0x80, True), (0x20, False), (0x9F, False), (0x41, False)):
        ks = [k for k in range(8) if (planes >> k) & 1]
        for v, r, p in zip(a.gather(planes, ghosted), ref, h.patches):
            g = p.g
            want = r[ks] if ghosted else r[ks][:, g:-g, g:-g]
            assert np.array_equal(v, want), (planes, ghosted, p.pid)
    b = D.Hierarchy(h, gas)
    b.scatter(0xFF, a.gather(0xFF, True))
    for x, y in zip(per_patch_copies(b, h), ref):
        assert np.array_equal(x, y)
    from enlightning_amd import _lib
    from enlightning_amd.hierarchy import _p
    small = np.empty(10)
    with pytest.raises(RuntimeError, match="holds 10 values"):
        _lib.check(a.lib.enl_batch_gather(a.handle, 0xFF, 1, _p(small), 10))
    a.close()
    b.close()
