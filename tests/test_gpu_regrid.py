"""Regrid on the device (enl_batch_transfer, DeviceHierarchy.level0_tags /
regrid; main.C:489-502).

* The device transfer equals the oracle's (oracle.amr.HostHierarchy.
  transfer_from) bit for bit on every ghosted array of the new hierarchy
  (w_n, p, T and the source), from the same old state: copies and the
  unfused linear refinement are the same arithmetic.
* level0_tags (tag_cells_ on the device data of every level-0 patch) equals
  the oracle's tag_cells on the same filled states.
* enl_batch_tag_cells into device memory equals the host path; a transfer
  onto the same boxes (whole-patch pairs only) is the identity.
* Regrid inside a run: two RK3 steps, tags of the evolved state (plus the
  same tags moved, so the boxes change), regrid, two more steps; the device
  follows its CPU twin to 1e-10 per field, as without the regrid (with a
  floor of 16 ulps of the field's natural scale on quiet patches); and at
  the reference's cadence, a regrid every two steps for three cycles.
"""
import numpy as np
import pytest

import oracle
from enlightning_amd import amr
from enlightning_amd.gas import Gas, RK3
from enlightning_amd.hierarchy import level_tags, patch_initial_state
from enlightning_amd.workloads import Workload
from oracle.amr import HostHierarchy
from tests import devops as D

pytestmark = pytest.mark.gpu

WL64 = Workload("amr_test_64", n=64, seed=7, npulses=4, amp_range=(0.3, 0.8), sigma_range=(0.03, 0.06))


def hierarchies():
    gas = Gas()
    tags = level_tags(gas, WL64, oracle.tag_cells)
    old = amr.build(WL64.n, 4, 4, 1.0, tags, tile0=16)
    new = amr.build(WL64.n, 4, 4, 1.0, np.roll(tags, (6, -5), axis=(0, 1)), tile0=16)
    assert old.nlevels == 2 and new.nlevels == 2
    return gas, old, new


def random_sources(h, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((p.Ny, p.Nx)) for p in h.patches]


def twin_of(dev, h, gas):
    """A CPU twin holding the device hierarchy's exact current arrays."""
    cpu = HostHierarchy(h, gas, dev.reflect)
    for k, (w, pr, T) in enumerate(dev.states()):
        a, b, c = cpu.view(k)
        a[...], b[...], c[...] = w, pr, T
    for k, s in enumerate(dev.sources()):
        cpu.src[k][...] = s
    return cpu


def test_transfer_matches_oracle_bit_for_bit():
    gas, old, new = hierarchies()
    states = [patch_initial_state(old, p, WL64, gas) for p in old.patches]
    dev = D.Hierarchy(old, gas)
    dev.load_initial(WL64, states=states, sources=random_sources(old, 5))
    cpu = twin_of(dev, old, gas)
    tr = amr.make_transfer(new, old)
    print("transfer: %d copies, %d refinements" % (len(tr["copy"][0]), len(tr["interp"][0])), flush=True)
    dnew = dev.regrid(new, tr)
    cnew = HostHierarchy(new, gas, dev.reflect)
    cnew.transfer_from(cpu, tr)
    for k, ((wd, pd, Td), (wc, pc, Tc), sd, sc) in enumerate(zip(dnew.states(), cnew.states(), dnew.sources(),
                                                              cnew.src)):
        assert np.array_equal(wd, wc), k
        assert np.array_equal(pd, pc) and np.array_equal(Td, Tc), k
        assert np.array_equal(sd, sc), k
    # the old hierarchy's interiors are as they were
    for (wa, _, _), (wb, _, _), p in zip(dev.states(), cpu.states(), old.patches):
        g = p.g
        assert np.array_equal(wa[:, g:-g, g:-g], wb[:, g:-g, g:-g])
    dnew.close()
    dev.close()


def test_transfer_rects_match_oracle_bit_for_bit():
    """The regrid's default path: the planner's rectangles
    (amr.transfer_rects, enl_batch_transfer_rects) move the same values as
    the oracle's per-cell transfer, bit for bit, into every ghosted array."""
    gas, old, new = hierarchies()
    states = [patch_initial_state(old, p, WL64, gas) for p in old.patches]
    dev = D.Hierarchy(old, gas)
    dev.load_initial(WL64, states=states, sources=random_sources(old, 6))
    cpu = twin_of(dev, old, gas)
    if not D.DRY:
        rt = amr.transfer_rects(new, old)
        print("transfer: %d rectangles, %d whole patches" % (len(rt["rects"]), len(rt["same"][0])), flush=True)
    dnew = dev.regrid(new)
    cnew = HostHierarchy(new, gas, dev.reflect)
    cnew.transfer_from(cpu, amr.make_transfer(new, old, native=False))
    for k, ((wd, pd, Td), (wc, pc, Tc), sd, sc) in enumerate(zip(dnew.states(), cnew.states(), dnew.sources(),
                                                              cnew.src)):
        assert np.array_equal(wd, wc), k
        assert np.array_equal(pd, pc) and np.array_equal(Td, Tc), k
        assert np.array_equal(sd, sc), k
    dnew.close()
    dev.close()


def test_rectangle_schedules_equal_list_schedules():
    """A hierarchy whose fill and coarsen lists the device expanded from the
    planner's rectangles (a native plan) steps bit-identically to the same
    hierarchy given the numpy builders' per-cell lists."""
    gas = Gas()
    tags = level_tags(gas, WL64, oracle.tag_cells)
    hr = amr.build(WL64.n, 4, 4, 1.0, tags, tile0=16)
    hl = amr.build(WL64.n, 4, 4, 1.0, tags, tile0=16, native=False)
    assert hr.plan is not None and hl.plan is None
    states = [patch_initial_state(hr, p, WL64, gas) for p in hr.patches]
    srcs = random_sources(hr, 8)
    out = []
    for h in (hr, hl):
        dev = D.Hierarchy(h, gas)
        dev.load_initial(WL64, states=states, sources=srcs)
        dev.advance(3, 0.5 * h.dx(1) / gas.c, mode="graph")
        out.append(dev.states())
        dev.close()
    for k, (a, b) in enumerate(zip(*out)):
        for x, y in zip(a, b):
            assert np.array_equal(x, y), k


def test_transfer_rects_reject_out_of_range():
    gas, old, new = hierarchies()
    if D.DRY:
        pytest.skip("the range check is the C ABI's")
    dev = D.Hierarchy(old, gas)
    dev.load_initial(WL64)
    rt = amr.transfer_rects(new, old)
    refine = np.nonzero(rt["rects"][:, 6] == 1)[0]
    assert len(refine)
    bad = {"same": rt["same"], "rects": rt["rects"].copy()}
    bad["rects"][refine[0], 3] = 0  # the refinement stencil would leave the coarse patch's ghosted box
    with pytest.raises(RuntimeError, match="out of range"):
        dev.regrid(new, bad)
    bad = {"same": rt["same"], "rects": rt["rects"].copy()}
    bad["rects"][0, 4] = 10 ** 6  # a rectangle wider than its patch
    with pytest.raises(RuntimeError, match="out of range"):
        dev.regrid(new, bad)
    dev.close()


def test_transfer_rejects_out_of_range_items():
    gas, old, new = hierarchies()
    dev = D.Hierarchy(old, gas)
    dev.load_initial(WL64)
    tr = amr.make_transfer(new, old)
    dp, dl, sp, sl, fx, fy = (a.copy() for a in tr["interp"])
    sl[0] = 0  # the refinement stencil would leave the coarse patch's ghosted box
    bad = {"same": tr["same"], "copy": tr["copy"], "interp": (dp, dl, sp, sl, fx, fy)}
    if D.DRY:
        pytest.skip("the range check is the C ABI's")
    with pytest.raises(RuntimeError, match="out of range"):
        dev.regrid(new, bad)
    # whole-patch pairs whose boxes differ: another shape, and the same
    # shape at another place
    fine_new = [p.pid for p in new.level_patches(1)]
    A = new.patches[fine_new[0]]
    other_shape = [p.pid for p in old.level_patches(1) if (p.nx, p.ny) != (A.nx, A.ny)]
    same_shape = [p.pid for p in old.level_patches(1) if (p.nx, p.ny) == (A.nx, A.ny) and p.box != A.box]
    for q in other_shape[:1] + same_shape[:1]:
        sn, so = (a.copy() for a in tr["same"])
        sn[0], so[0] = A.pid, q
        bad = {"same": (sn, so), "copy": tr["copy"], "interp": tr["interp"]}
        with pytest.raises(RuntimeError, match="differs in level or box"):
            dev.regrid(new, bad)
    dev.close()


def test_level0_tags_match_oracle():
    gas, old, _ = hierarchies()
    dev = D.Hierarchy(old, gas)
    dev.load_initial(WL64, sources=random_sources(old, 9))
    dev.advance(2, 0.5 * old.dx(1) / gas.c, mode="graph")
    tags = dev.level0_tags()
    cpu = twin_of(dev, old, gas)
    want = np.zeros_like(tags)
    for p in old.level_patches(0):
        w, _, _ = cpu.view(p.pid)
        t = np.zeros((p.ny, p.nx), np.int32)
        oracle.tag_cells(p.box, p.g, t, np.ascontiguousarray(w), cpu.src[p.pid], 1, gas.tag_tolerance)
        want[p.box[2]:p.box[3] + 1, p.box[0]:p.box[1] + 1] = t
    assert tags.any()
    assert np.array_equal(tags, want)
    dev.close()


def test_regrid_inside_a_run_follows_cpu_twin():
    gas, old, _ = hierarchies()
    dt = 0.5 * old.dx(1) / gas.c
    states = [patch_initial_state(old, p, WL64, gas) for p in old.patches]
    cpu = HostHierarchy(old, gas)
    cpu.load_initial(states)
    dev = D.Hierarchy(old, gas)
    dev.load_initial(WL64, states=states)
    cpu.advance(2, dt, RK3)
    dev.advance(2, dt, mode="graph")
    tags = dev.level0_tags()
    tags = tags | np.roll(tags, (3, -2), axis=(0, 1))
    new = amr.build(WL64.n, 4, 4, 1.0, tags, tile0=16, previous=old)
    assert [p.box for p in new.level_patches(1)] != [p.box for p in old.level_patches(1)]
    tr = amr.make_transfer(new, old)
    dnew = dev.regrid(new, tr)
    cnew = HostHierarchy(new, gas)
    cnew.transfer_from(cpu, tr)
    cnew.advance(2, dt, RK3)
    dnew.advance(2, dt, mode="graph")
    # 1e-10 of each field's size on the patch; the moved boxes also refine
    # quiet regions, where the momenta are rounding of ambient-scale terms
    # (~1e-9): there the floor is 16 ulps of the field's natural scale
    natural = (gas.rho0, gas.rho0 * gas.c, gas.rho0 * gas.c, gas.rho0 * gas.R, gas.rho0 * gas.T0)
    ulp = np.finfo(np.float64).eps
    for (wc, pc, Tc), (wd, pd, Td), p in zip(cnew.states(), dnew.states(), new.patches):
        g = p.g
        for k in range(5):
            a, b = wd[k, g:-g, g:-g], wc[k, g:-g, g:-g]
            assert np.abs(a - b).max() <= max(1e-10 * np.abs(b).max(), 16 * ulp * natural[k]), (p.pid, k)
        assert np.abs(pd[g:-g, g:-g] - pc[g:-g, g:-g]).max() <= 1e-10 * np.abs(pc[g:-g, g:-g] - gas.p0).max() + 1e-9
    dnew.close()
    dev.close()


def test_tags_into_device_memory_and_identity_transfer():
    """enl_batch_tag_cells into a device buffer equals the host path; a
    transfer onto the same boxes (whole-patch pairs only) is the identity."""
    if D.DRY:
        pytest.skip("device buffers and the C ABI's pair path")
    import ctypes

    import torch

    from enlightning_amd import _lib
    gas, old, _ = hierarchies()
    dev = D.Hierarchy(old, gas)
    dev.load_initial(WL64, sources=random_sources(old, 11))
    dev.advance(2, 0.5 * old.dx(1) / gas.c, mode="graph")
    host_tags = dev.level0_tags()
    pats = old.level_patches(0)
    buf = torch.zeros(sum(p.nx * p.ny for p in pats), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()  # the fill runs on torch's stream, the tags on the library's
    tol = (ctypes.c_double * 3)(*gas.tag_tolerance)
    refl = (ctypes.c_int * 4)(*dev.reflect)
    _lib.check(dev.lib.enl_batch_tag_cells(dev.handle, 0, refl, 1, tol, ctypes.c_void_p(buf.data_ptr())))
    _lib.check(dev.lib.enl_sync())
    flat = buf.cpu().numpy()
    off = 0
    for p in pats:
        t = flat[off:off + p.nx * p.ny].reshape(p.ny, p.nx)
        assert np.array_equal(t, host_tags[p.box[2]:p.box[3] + 1, p.box[0]:p.box[1] + 1]), p.pid
        off += p.nx * p.ny
    tr = amr.make_transfer(old, old)
    assert len(tr["copy"][0]) == 0 and len(tr["interp"][0]) == 0
    same = dev.regrid(old, tr)
    for (wa, pa, Ta), (wb, pb, Tb), sa, sb, p in zip(dev.states(), same.states(), dev.sources(), same.sources(),
                                                      old.patches):
        assert np.array_equal(wa, wb) and np.array_equal(pa, pb) and np.array_equal(Ta, Tb), p.pid
        assert np.array_equal(sa, sb), p.pid
    same.close()
    dev.close()


def test_regrid_every_two_steps_follows_cpu_twin():
    """The reference's cadence (regrid_step = 2, input.txt:7) with its
    absorbing sides: three cycles of two RK3 steps and a regrid on the
    evolved state's device tags (moved by a few cells as well, so that boxes
    change); the device follows its CPU twin to 1e-10 per field after every
    cycle, and the envelopes and tile shapes carry over."""
    gas, h, _ = hierarchies()
    dt = 0.5 * h.dx(1) / gas.c
    states = [patch_initial_state(h, p, WL64, gas) for p in h.patches]
    reflect, width = (0, 0, 1, 0), 0.1  # input.txt:14-17: ABSORB left, right, top
    cpu = HostHierarchy(h, gas, reflect)
    cpu.absorb_width = width
    cpu.load_initial(states)
    dev = D.Hierarchy(h, gas, reflect)
    dev.set_absorb(width)  # carried over by every regrid
    dev.load_initial(WL64, states=states)
    dev.set_tiles(16, 8, level=1)  # carried over by every regrid
    natural = (gas.rho0, gas.rho0 * gas.c, gas.rho0 * gas.c, gas.rho0 * gas.R, gas.rho0 * gas.T0)
    ulp = np.finfo(np.float64).eps
    for cycle in range(3):
        cpu.advance(2, dt, RK3)
        dev.advance(2, dt, mode="graph")
        # the evolved state's tags plus the same moved by a few cells, so
        # that boxes change at every cycle
        tags = dev.level0_tags()
        tags = tags | np.roll(tags, (2 * cycle + 1, -cycle - 1), axis=(0, 1))
        new = amr.build(WL64.n, 4, 4, 1.0, tags, tile0=16, previous=h)
        tr = amr.make_transfer(new, h)
        assert len(tr["copy"][0]) + len(tr["interp"][0]) > 0
        print("cycle %d: %d fine patches, %d pairs, %d copies, %d refinements" % (
            cycle, len(new.level_patches(1)), len(tr["same"][0]), len(tr["copy"][0]), len(tr["interp"][0])),
            flush=True)
        dnew = dev.regrid(new, tr)
        cnew = HostHierarchy(new, gas, reflect)
        cnew.absorb_width = width
        cnew.transfer_from(cpu, tr)
        dev.close()
        dev, cpu, h = dnew, cnew, new
        assert dev.absorb_width == width
        if not D.DRY:
            assert dev.tiles == {1: (16, 8)}
        for (wc, pc, _), (wd, pd, _), p in zip(cpu.states(), dev.states(), h.patches):
            g = p.g
            for k in range(5):
                a, b = wd[k, g:-g, g:-g], wc[k, g:-g, g:-g]
                assert np.abs(a - b).max() <= max(1e-10 * np.abs(b).max(), 16 * ulp * natural[k]), (cycle, p.pid, k)
            assert np.abs(pd[g:-g, g:-g] - pc[g:-g, g:-g]).max() <= \
                1e-10 * np.abs(pc[g:-g, g:-g] - gas.p0).max() + 1e-9
    dev.close()
