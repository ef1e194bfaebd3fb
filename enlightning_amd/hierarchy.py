"""A two-level hierarchy resident in HBM, driven through the batch C ABI
(enl_batch_*): build from an amr.Hierarchy, load the initial state, advance
in plain (drop-in per patch), fused or graph mode, read states back, tag the
base level and regrid onto new boxes (main.C:489-502)."""
import ctypes

import numpy as np

from . import _lib, amr
from .gas import RK3
from .workloads import pulses_at, state_from_overpressure

MODES = {"plain": 0, "fused": 1, "graph": 2, "fp32": 3, "fp32_graph": 4}


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def level_tags(gas, wl, tag_fn):
    """Tags of the workload's initial state on the base grid (one patch
    covering the domain), from tag_fn(box, g, tags, w, source, val, tol)."""
    n, g = wl.n, wl.g
    xc = (np.arange(-g, n + g) + 0.5) * wl.dx
    w = np.ascontiguousarray(state_from_overpressure(pulses_at(wl, xc, xc), gas))
    src = np.zeros((n + 2 * g, n + 2 * g))
    tags = np.zeros((n, n), np.int32)
    tag_fn((0, n - 1, 0, n - 1), g, tags, w, src, 1, gas.tag_tolerance)
    return tags


def patch_initial_state(h, patch, wl, gas):
    """w on the patch's ghosted cell centres (analytic initial condition)."""
    dx = h.dx(patch.level)
    i0, i1, j0, j1 = patch.box
    g = patch.g
    xc = (np.arange(i0 - g, i1 + g + 1) + 0.5) * dx
    yc = (np.arange(j0 - g, j1 + g + 1) + 0.5) * dx
    return np.ascontiguousarray(state_from_overpressure(pulses_at(wl, xc, yc), gas))


def refine_from_base(h, patch, wb):
    """w on a patch's ghosted cells from a ghosted base-grid state wb
    (5, n0+2gb, n0+2gb): level-0 patches take their window of wb; level-1
    patches the linear refinement of the ghost fill,
        c + fx*(c[i+1]-c[i-1])/2 + fy*(c[j+1]-c[j-1])/2,
    (fx, fy) the fine centre's offset from its coarse centre in coarse
    widths.  Ghost values outside the domain are placeholders (the first
    ghost fill replaces them)."""
    n0 = h.n0
    gb = (wb.shape[-1] - n0) // 2
    i0, i1, j0, j1 = patch.box
    g = patch.g
    if patch.level == 0:
        return np.ascontiguousarray(wb[:, j0 - g + gb:j1 + g + 1 + gb, i0 - g + gb:i1 + g + 1 + gb])
    if patch.level != 1:
        raise ValueError("refine_from_base covers levels 0 and 1")
    r = h.ratio
    I = np.arange(i0 - g, i1 + g + 1)
    J = np.arange(j0 - g, j1 + g + 1)
    ic = np.clip(np.floor_divide(I, r), -gb + 1, n0 + gb - 2)
    jc = np.clip(np.floor_divide(J, r), -gb + 1, n0 + gb - 2)
    fx = (I + 0.5) / r - (ic + 0.5)
    fy = (J + 0.5) / r - (jc + 0.5)
    C = lambda dj, di: wb[:, (jc + dj + gb)[:, None], (ic + di + gb)[None, :]]
    out = C(0, 0) + fx[None, None, :] * (0.5 * (C(0, 1) - C(0, -1))) + fy[None, :, None] * (0.5 * (C(1, 0) - C(-1, 0)))
    return np.ascontiguousarray(out)


def evolved_amr_workload(wl, gas, tile0=64, ratio=4):
    """The two-level hierarchy of an AMR workload on the device: the base
    grid (one patch) advanced wl.presteps fused steps from the analytic
    initial state, tag_cells on that state, one level refined by `ratio`
    over the clustered tags, and every patch's state from the evolved base
    (refine_from_base).  Returns (h, states, tags)."""
    import torch
    from . import patch as P
    from .workloads import initial_fields
    n, g = wl.n, wl.g
    base = P.DevicePatch(wl, gas)
    base.init_state(*initial_fields(wl, gas))
    if wl.presteps:
        base.advance_fused(wl.presteps, wl.dt(gas))
    P.fill_physical_bc(base.box, g, (1, 1, 1, 1), base.reflect, base.w, base.p, base.T, base.src)
    torch.cuda.synchronize()
    wb = base.w.cpu().numpy()
    del base
    tags = np.zeros((n, n), np.int32)
    P.tag_cells((0, n - 1, 0, n - 1), g, tags, wb, np.zeros((n + 2 * g, n + 2 * g)), 1, gas.tag_tolerance)
    h = amr.build(n, ratio, g, 1.0, tags, tile0=tile0)
    states = [refine_from_base(h, p, wb) for p in h.patches]
    return h, states, tags


def weak_acoustic_workload(wl, gas, tags, tile0=64, ratio=4):
    """The FP32-perturbation workload on an AMR workload's boxes: the
    hierarchy clustered from `tags` on the wl.L domain, per patch the
    stratified state carrying wl's pulses and its ambient rows (rho, rho s,
    rho T_n at each ghosted row), and the per-level ambient function that
    DeviceHierarchy.set_ambient takes.  Returns (h, states, ambients, fn)."""
    from .workloads import ambient_rows, stratified_state
    h = amr.build(wl.n, ratio, wl.g, wl.L, tags, tile0=tile0)
    states, ambs = [], []
    for p in h.patches:
        dx = h.dx(p.level)
        x = (np.arange(p.box[0] - p.g, p.box[1] + p.g + 1) + 0.5) * dx
        y = (np.arange(p.box[2] - p.g, p.box[3] + p.g + 1) + 0.5) * dx
        states.append(stratified_state(wl, gas, x, y))
        ambs.append(ambient_rows(wl, gas, y))

    def ambient(lev):
        n, g = h.ncells(lev), h.g
        y = (np.arange(-g, n + g) + 0.5) * h.dx(lev)
        return (-g,) + ambient_rows(wl, gas, y)

    return h, states, ambs, ambient


PERTURBATION_FIELDS = ("rho", "rho_u", "rho_v", "rho_s", "rho_Tn", "p", "T")


def perturbation_errors(h, ref, test, ambients, gas):
    """Per field of two per-patch state lists [(w, p, T), ...] over every
    patch interior: rel_l2 and rel_linf of `test` against `ref`, and
    rel_l2_of_perturbation, the L2 error over the L2 norm of ref's
    perturbation (rho, rho s, rho T_n less their ambient rows, p less p0,
    the momenta themselves; none for T)."""
    err = {}
    for f_i, name in enumerate(PERTURBATION_FIELDS):
        num = den = dev = linf = amax = 0.0
        for p, a, b, am in zip(h.patches, ref, test, ambients):
            g = p.g
            pick = (lambda st: st[0][f_i]) if f_i < 5 else (lambda st: st[f_i - 4])
            x, y = pick(a)[g:-g, g:-g], pick(b)[g:-g, g:-g]
            if f_i in (0, 3, 4):
                pert = x - am[(0, None, None, 1, 2)[f_i]][g:-g, None]
            elif f_i in (1, 2):
                pert = x
            elif f_i == 5:
                pert = x - gas.p0
            else:
                pert = None
            num += float(np.sum((x - y) ** 2))
            den += float(np.sum(x ** 2))
            if pert is not None:
                dev += float(np.sum(pert ** 2))
            linf = max(linf, float(np.abs(x - y).max()))
            amax = max(amax, float(np.abs(x).max()))
        err[name] = {"rel_l2": float(np.sqrt(num / den)), "rel_linf": linf / amax,
                     "rel_l2_of_perturbation": float(np.sqrt(num / dev)) if dev > 0 else None}
    return err


_MAPS = {}


def _pinned_map(n):
    """An (n, n) int32 array in pinned host memory (kept for reuse): the
    device writes a tag map into it directly."""
    import torch
    if n not in _MAPS:
        _MAPS[n] = torch.empty((n, n), dtype=torch.int32, pin_memory=True)
    return _MAPS[n].numpy()


class DeviceHierarchy:
    def __init__(self, h, gas, reflect=(1, 1, 1, 1)):
        self.h, self.gas, self.reflect = h, gas, tuple(reflect)
        self.absorb_width = 0.0
        self.tiles = {}  # level -> (bx, by) set by set_tiles / autotune; None: every level
        self.packing = {}  # level -> (mode, waves per SIMD) set by set_packing / autotune
        self.lib = _lib.load()
        n = len(h.patches)
        # the patch table (enl_patch_desc_t rows) built with numpy
        desc = np.zeros(n, dtype=_lib.PATCH_DESC_DTYPE)
        if n:
            lev = np.array([p.level for p in h.patches], np.int32)
            box = np.array([p.box for p in h.patches], np.int32).reshape(n, 4)
            dx = h.L / (h.n0 * np.power(float(h.ratio), lev))
            desc["level"], desc["box"] = lev, box
            desc["xlo"] = np.stack([box[:, 0] * dx, box[:, 2] * dx], axis=1)
            desc["dx"] = np.stack([dx, dx], axis=1)
        handle = ctypes.c_void_p()
        _lib.check(self.lib.enl_batch_create(n, desc.ctypes.data_as(ctypes.POINTER(_lib.EnlPatchDesc)), h.g, h.ratio,
                                             ctypes.byref(handle)))
        self.handle = handle
        self._keep = []
        if h.plan is not None:
            # the planner's rectangles, expanded into per-cell lists on the device
            for lev in range(h.nlevels):
                rects, n1, n2 = h.plan.fill_rects(lev)
                _lib.check(self.lib.enl_batch_set_fill_rects(self.handle, lev, n1, n2, _p(rects)))
            if h.nlevels > 1:
                rects = h.plan.coarsen_rects()
                _lib.check(self.lib.enl_batch_set_coarsen_rects(self.handle, len(rects), _p(rects)))
            return
        for lev, f in enumerate(h.fill):
            c, it, e, k = f["copy"], f["interp"], f["edge"], f["corner"]
            arrs = [np.ascontiguousarray(a) for a in c + it + e + k]
            self._keep += arrs
            P = [_p(a) for a in arrs]
            _lib.check(self.lib.enl_batch_set_fill(
                self.handle, lev, len(c[0]), *P[0:4], len(it[0]), *P[4:10], len(e[0]), *P[10:14],
                len(k[0]), *P[14:18]))
        if h.coarsen is not None:
            arrs = [np.ascontiguousarray(a) for a in h.coarsen]
            self._keep += arrs
            _lib.check(self.lib.enl_batch_set_coarsen(self.handle, len(arrs[0]), *[_p(a) for a in arrs]))

    def _reflect(self):
        return (ctypes.c_int * 4)(*self.reflect)

    TILE_SHAPES = ((32, 8), (16, 16), (16, 8), (8, 8))

    def set_absorb(self, width):
        """Absorbing envelopes (enl_batch_set_absorb) on the sides whose reflect
        flag is 0; the domain is the level-0 index space at the level-0 dx."""
        n0, dx0 = self.h.ncells(0), self.h.dx(0)
        D = ctypes.c_double * 2
        _lib.check(self.lib.enl_batch_set_absorb(self.handle, D(0.0, 0.0), D(n0 * dx0, n0 * dx0), float(width)))
        self.absorb_width = float(width)

    # how a level's tiles are packed into the stage launch's workgroups
    # (enl_batch_set_level_packing): (mode, waves per SIMD of the persistent grid)
    PACKINGS = ((_lib.PACK_SINGLE, 0), (_lib.PACK_STRIDED, 4), (_lib.PACK_STRIDED, 8), (_lib.PACK_CHUNKED, 4),
                (_lib.PACK_CHUNKED, 8))
    PACK_NAMES = {_lib.PACK_SINGLE: "single", _lib.PACK_STRIDED: "strided", _lib.PACK_CHUNKED: "chunked"}

    def set_packing(self, level, mode, wpc=4):
        """Packing of a level's tiles into workgroups (ENL_PACK_*; wpc: waves
        per SIMD of the persistent grid)."""
        _lib.check(self.lib.enl_batch_set_level_packing(self.handle, level, int(mode), max(1, int(wpc))))
        self.packing[level] = (int(mode), int(wpc))

    def packing_name(self, level):
        mode, wpc = self.packing.get(level, (_lib.PACK_SINGLE, 0))
        return self.PACK_NAMES[mode] + ("" if mode == _lib.PACK_SINGLE else "/%d waves per SIMD" % wpc)

    def set_tiles(self, bx, by, level=None):
        if level is None:
            _lib.check(self.lib.enl_batch_set_tiles(self.handle, bx, by))
            self.tiles = {None: (bx, by)}
        else:
            _lib.check(self.lib.enl_batch_set_level_tiles(self.handle, level, bx, by))
            self.tiles[level] = (bx, by)

    def autotune(self, dt, states, wl, steps=2, variant="hybrid_high_llf", rk=RK3, mode="graph"):
        """Tile shape and packing per level by coordinate descent over
        measured steps of `mode` ("graph" or "fp32_graph"): each level in
        turn tries every tile shape with the others fixed and keeps the
        fastest, then every packing of its tiles into workgroups (one tile
        per workgroup, or persistent packs: strided or cost-balanced chunks
        at 4 or 8 waves per SIMD) at that shape.  Reloads `states` before
        each trial; returns {level: (bx, by)} (the packings are in
        self.packing) and the per-trial times (ms per step)."""
        import time

        def reload():
            self.load_initial(wl, states)
            if mode.startswith("fp32"):
                self.fp32_load()

        def measure():
            reload()
            self.advance(1, dt, mode=mode, variant=variant, rk=rk)  # capture
            reload()
            self.sync()
            t0 = time.perf_counter()
            self.advance(steps, dt, mode=mode, variant=variant, rk=rk)
            self.sync()
            return (time.perf_counter() - t0) * 1e3 / steps

        best, trials = {}, {}
        for lev in range(self.h.nlevels):
            times = {}
            for shape in self.TILE_SHAPES:
                self.set_tiles(*shape, level=lev)
                times[shape] = measure()
                trials["L%d:%dx%d:%s" % ((lev,) + shape + (self.packing_name(lev),))] = times[shape]
            best[lev] = min(times, key=times.get)
            self.set_tiles(*best[lev], level=lev)
            ptimes = {self.packing.get(lev, self.PACKINGS[0]): times[best[lev]]}
            for pk in self.PACKINGS:
                if pk in ptimes:
                    continue
                self.set_packing(lev, *pk)
                ptimes[pk] = measure()
                trials["L%d:%dx%d:%s" % ((lev,) + best[lev] + (self.packing_name(lev),))] = ptimes[pk]
            self.set_packing(lev, *min(ptimes, key=ptimes.get))
        return best, trials

    def load_initial(self, wl, states=None, sources=None):
        """Upload the analytic initial state (or given per-patch w arrays)
        and the ghosted sources (zero unless given), fill ghosts, derive p, T
        and fill again."""
        gas = self.gas
        for k, p in enumerate(self.h.patches):
            w = patch_initial_state(self.h, p, wl, gas) if states is None else np.ascontiguousarray(states[k])
            src = np.zeros((p.Ny, p.Nx)) if sources is None else np.ascontiguousarray(sources[k], dtype=np.float64)
            _lib.check(self.lib.enl_batch_copy_state(self.handle, k, 1, _p(w), None, None, _p(src)))
        refl = self._reflect()
        _lib.check(self.lib.enl_batch_prepare(self.handle, refl))
        _lib.check(self.lib.enl_batch_fill(self.handle, refl))
        gs = gas.to_struct()
        _lib.check(self.lib.enl_batch_update_state(self.handle, ctypes.byref(gs)))
        _lib.check(self.lib.enl_batch_fill(self.handle, refl))

    def advance(self, nsteps, dt, mode="graph", variant="hybrid_high_llf", rk=RK3):
        vid = _lib.VARIANT_ID[variant] if isinstance(variant, str) else int(variant)
        a1, a2, b = rk
        D = ctypes.c_double * len(a1)
        gs = self.gas.to_struct()
        _lib.check(self.lib.enl_batch_advance(self.handle, MODES[mode], vid, ctypes.byref(gs), self._reflect(),
                                              len(a1), D(*a1), D(*a2), D(*b), float(dt), int(nsteps)))

    def states(self):
        """Per patch (w, p, T) host copies of the current state (ghosted), from
        one enl_batch_gather."""
        return [(v[:5], v[5], v[6]) for v in self.gather(0x7F, ghosted=True)]

    def sources(self):
        """Per patch host copies of the ghosted source."""
        return [v[0] for v in self.gather(0x80, ghosted=True)]

    PLANES = ("w_n #0", "w_n #1", "w_n #2", "w_n #3", "w_n #4", "pressure", "temperature", "source")

    def _plane_layout(self, planes, ghosted):
        n = bin(planes).count("1")
        shapes = [(n, p.Ny, p.Nx) if ghosted else (n, p.ny, p.nx) for p in self.h.patches]
        return shapes, int(sum(np.prod(sh) for sh in shapes))

    def gather(self, planes=0xFF, ghosted=True):
        """Per patch an array (planes, rows, columns) of the selected SOLN
        planes (bit k: w_n #k for k < 5, 5 pressure, 6 temperature, 7 source)
        over the ghosted box or the interior, all from one device gather
        (enl_batch_gather) into one host buffer."""
        shapes, total = self._plane_layout(planes, ghosted)
        buf = np.empty(total)
        _lib.check(self.lib.enl_batch_gather(self.handle, int(planes), int(bool(ghosted)), _p(buf), total))
        out, off = [], 0
        for sh in shapes:
            n = int(np.prod(sh))
            out.append(buf[off:off + n].reshape(sh))
            off += n
        return out

    def scatter(self, planes, arrays):
        """The inverse of gather over the ghosted boxes: per patch an array
        (planes, Ny, Nx) into the device (enl_batch_scatter), no ghost fill."""
        shapes, total = self._plane_layout(planes, True)
        buf = np.empty(total)
        off = 0
        for sh, a in zip(shapes, arrays):
            n = int(np.prod(sh))
            buf[off:off + n] = np.asarray(a, dtype=np.float64).reshape(-1)
            off += n
        _lib.check(self.lib.enl_batch_scatter(self.handle, int(planes), _p(buf), total))

    def set_states(self, states):
        """Upload per patch (w, p, T) ghosted host arrays as the current state
        (no ghost fill, no update_state)."""
        self.scatter(0x7F, [np.concatenate([w.reshape(5, *pr.shape), pr[None], T[None]]) for w, pr, T in states])

    def upload_all(self, states, sources):
        """A saved state: per patch (w, p, T) and the ghosted source, then the
        source's and the state's ghost fills (p, T are taken as saved)."""
        self.scatter(0xFF, [np.concatenate([w.reshape(5, *pr.shape), pr[None], T[None], s[None]])
                            for (w, pr, T), s in zip(states, sources)])
        refl = self._reflect()
        _lib.check(self.lib.enl_batch_prepare(self.handle, refl))
        _lib.check(self.lib.enl_batch_fill(self.handle, refl))

    def add_source(self, src, iteration, pulse_time0, dt):
        """enlightning::addSource (enlightning.C:2140-2307) over the hierarchy:
        enl_add_source writes every patch's source (src: a source.Source),
        then enl_batch_prepare refills the source's ghosts and the interior
        copy the drop-in mode reads.  Call it before each step, as the
        reference does for time-dependent sources.

        The reference's order per iteration (main.C:424-480): if
        source.should_add_source(...) then add_source, advance one step,
        then zero_source; a PULSE source is therefore felt for one step only
        and a LIGHTNING channel merges into a zeroed source each time."""
        st = src.to_struct()
        D2 = ctypes.c_double * 2
        for p in self.h.patches:
            base = ctypes.c_void_p()
            _lib.check(self.lib.enl_batch_patch_state(self.handle, p.pid, None, None, None, ctypes.byref(base)))
            interior = ctypes.c_void_p(base.value + 8 * (p.g + p.g * p.Nx))
            dx = self.h.dx(p.level)
            _lib.check(self.lib.enl_add_source(p.box[0], p.box[1], p.box[2], p.box[3], D2(p.box[0] * dx, p.box[2] * dx),
                                               D2(dx, dx), ctypes.byref(st), int(iteration), int(pulse_time0),
                                               float(dt), interior, p.Nx))
        _lib.check(self.lib.enl_batch_prepare(self.handle, self._reflect()))

    def zero_source(self):
        """enlightning::setSourceToZero (enlightning.C:2310-2329): every
        patch's source, ghosts and interior copy, set to 0 on the device
        (enl_batch_zero_source), after a step that added a source."""
        _lib.check(self.lib.enl_batch_zero_source(self.handle))

    def sample_mics(self, mic_x, mic_y):
        """recordAudio's pressure reads (enlightning.C:2392-2506) over the
        hierarchy: for every level in order and every patch of it whose
        bounds hold a microphone (inclusive, as there), the pressure of the
        cell containing it (enl_sample_mics on the pool).  Returns, per
        microphone, the (level, patch id, value) entries in the order the
        reference appends them to its mic-*.txt file."""
        import torch
        mx_h = np.asarray(mic_x, dtype=np.float64)
        my_h = np.asarray(mic_y, dtype=np.float64)
        nm = len(mx_h)
        D2 = ctypes.c_double * 2
        todo = []
        for lev in range(self.h.nlevels):
            dx = self.h.dx(lev)
            for p in self.h.level_patches(lev):
                lo = (p.box[0] * dx, p.box[2] * dx)
                hi = ((p.box[1] + 1) * dx, (p.box[3] + 1) * dx)
                if np.any((mx_h >= lo[0]) & (mx_h <= hi[0]) & (my_h >= lo[1]) & (my_h <= hi[1])):
                    todo.append((lev, p, lo, hi, dx))
        res = [[] for _ in range(nm)]
        if not todo or nm == 0:
            return res
        mx = torch.as_tensor(mx_h, device="cuda")
        my = torch.as_tensor(my_h, device="cuda")
        out = torch.zeros((len(todo), nm), dtype=torch.float64, device="cuda")
        hit = torch.zeros((len(todo), nm), dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()  # the inputs are written on torch's stream
        for t, (lev, p, lo, hi, dx) in enumerate(todo):
            pr = ctypes.c_void_p()
            _lib.check(self.lib.enl_batch_patch_state(self.handle, p.pid, None, ctypes.byref(pr), None, None))
            interior = ctypes.c_void_p(pr.value + 8 * (p.g + p.g * p.Nx))
            _lib.check(self.lib.enl_sample_mics(p.box[0], p.box[1], p.box[2], p.box[3], D2(*lo), D2(*hi), D2(dx, dx),
                                                nm, ctypes.c_void_p(mx.data_ptr()), ctypes.c_void_p(my.data_ptr()),
                                                interior, p.Nx, ctypes.c_void_p(out[t].data_ptr()),
                                                ctypes.c_void_p(hit[t].data_ptr())))
        _lib.check(self.lib.enl_sync())
        out_h, hit_h = out.cpu().numpy(), hit.cpu().numpy()
        for t, (lev, p, _, _, _) in enumerate(todo):
            for k in np.nonzero(hit_h[t])[0]:
                res[k].append((lev, p.pid, float(out_h[t, k])))
        return res

    # ---- regrid -------------------------------------------------------------
    def level0_tags(self, refine_tag_val=1):
        """Tags of the current state on the base grid, (n0, n0) [j, i]: the
        state's ghosts are filled, then tag_cells runs on every level-0
        patch's device data in one call (enl_batch_tag_cells;
        enlightning::tagGradientDetectorCells, enlightning.C:586-628), the
        input of the next regrid's clustering."""
        n0 = self.h.ncells(0)
        tags = _pinned_map(n0)
        tol = (ctypes.c_double * len(self.gas.tag_tolerance))(*self.gas.tag_tolerance)
        _lib.check(self.lib.enl_batch_tag_map(self.handle, 0, self._reflect(), int(refine_tag_val), tol, _p(tags), n0))
        return tags.copy()

    def regrid(self, new_h, transfer=None):
        """A DeviceHierarchy on new_h's boxes holding this hierarchy's state,
        moved over on the device by enl_batch_transfer (schedule
        amr.make_transfer: unchanged patches whole, copies from the old
        patches of each level, linear refinement from the old coarser level
        elsewhere).  This hierarchy
        keeps its interiors and stays usable; absorbing envelopes, tile
        shapes and packings carry over.  The FP64 state moves: after FP32-form steps call
        fp32_store() first."""
        if transfer is None:
            tr = amr.transfer_rects(new_h, self.h) if new_h.plan is not None else amr.make_transfer(new_h, self.h)
        else:
            tr = transfer
        new = DeviceHierarchy(new_h, self.gas, self.reflect)
        try:
            if "rects" in tr:
                sm, rects = tr["same"], np.ascontiguousarray(tr["rects"], dtype=np.int32)
                a, b = np.ascontiguousarray(sm[0], np.int32), np.ascontiguousarray(sm[1], np.int32)
                _lib.check(self.lib.enl_batch_transfer_rects(new.handle, self.handle, self._reflect(), len(a), _p(a),
                                                             _p(b), len(rects), _p(rects)))
            else:
                sm, c, it = tr["same"], tr["copy"], tr["interp"]
                arrs = [np.ascontiguousarray(a) for a in sm + c + it]
                P = [_p(a) for a in arrs]
                _lib.check(self.lib.enl_batch_transfer(new.handle, self.handle, self._reflect(), len(sm[0]),
                                                       *P[0:2], len(c[0]), *P[2:6], len(it[0]), *P[6:12]))
            if self.absorb_width > 0:
                new.set_absorb(self.absorb_width)
            # the tile shapes chosen for this hierarchy (results do not
            # depend on them)
            for lev, (bx, by) in sorted(self.tiles.items(), key=lambda kv: -1 if kv[0] is None else kv[0]):
                if lev is None or lev < new_h.nlevels:
                    new.set_tiles(bx, by, level=lev)
            for lev, pk in sorted(self.packing.items()):
                if lev < new_h.nlevels:
                    new.set_packing(lev, *pk)
        except Exception:
            new.close()
            raise
        return new

    # ---- FP32-perturbation form ------------------------------------------
    def set_ambient(self, ambient):
        """ambient(level) -> (row_lo, rho, rho*s, rho*T_n) arrays over level
        rows row_lo .. row_lo+len-1 covering every patch's ghosted rows."""
        gs = self.gas.to_struct()
        for lev in range(self.h.nlevels):
            row_lo, rho, rhos, rhotn = ambient(lev)
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, rhos, rhotn)]
            _lib.check(self.lib.enl_batch_set_ambient(self.handle, lev, int(row_lo), len(arrs[0]), *[_p(a) for a in arrs],
                                                      ctypes.byref(gs)))

    def fp32_prepare(self, variant="hybrid_high_llf"):
        """Ambient correction and conversion of the current FP64 state."""
        vid = _lib.VARIANT_ID[variant] if isinstance(variant, str) else int(variant)
        gs = self.gas.to_struct()
        _lib.check(self.lib.enl_batch_fp32_prepare(self.handle, vid, ctypes.byref(gs), self._reflect()))

    def fp32_set_switch(self, sw, lo=0.2, hi=0.8, every_stage=True, totals=None):
        """The learned switch (a switch.Switch, or None for the sensor) on
        level 0 of the FP32 form: its decisions are made over all level-0
        patches before level 0's rhs, every stage or once per step
        (every_stage=False).  totals: a device int64 tensor of 3 that gains
        (candidates, WENO, undecided) per decision, or None."""
        if totals is not None:
            assert totals.is_cuda and totals.dtype.itemsize == 8 and totals.numel() >= 3
        self._switch, self._switch_totals = sw, totals  # both must outlive their use
        _lib.check(self.lib.enl_batch_fp32_set_switch(
            self.handle, sw.handle if sw is not None else None, float(lo), float(hi), int(bool(every_stage)),
            ctypes.c_void_p(totals.data_ptr()) if totals is not None else None))

    def fp32_decisions(self, pid):
        """The switch's last decision map of level-0 patch pid (ny x nx uint8:
        0 DRP, 1 WENO, 2 the sensor decides)."""
        p = self.h.patches[pid]
        out = np.empty((p.ny, p.nx), np.uint8)
        _lib.check(self.lib.enl_batch_fp32_decisions(self.handle, int(pid), _p(out)))
        return out

    def fp32_load(self):
        _lib.check(self.lib.enl_batch_fp32_load(self.handle))

    def fp32_store(self):
        _lib.check(self.lib.enl_batch_fp32_store(self.handle))

    def sync(self):
        _lib.check(self.lib.enl_sync())

    def close(self):
        if self.handle:
            self.lib.enl_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
