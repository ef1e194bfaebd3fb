"""CPU twin of the batched hierarchy (TEST INFRASTRUCTURE ONLY).

Steps an enlightning_amd.amr.Hierarchy the reference's way: per RK stage and
level, fill the level's ghosts from the schedule (numpy, same arithmetic as
the device fill kernels), call the oracle's rhs / runge_kutta / update_state
on each patch (enlightning.C:459-583), refill the level's ghosts for the next
finer level's coarse-to-fine refine; after the step, average fine data onto
the coarse level.  At a regrid (main.C:489-502) transfer_from moves the SOLN
variables into a new hierarchy.  The ghost-fill, coarsen and transfer
semantics restate SAMRAI's operators, which are not vendored: parity with
the reference is unpinned there (closed forms pin them: copies are exact,
linear refinement reproduces linear fields and conserves coarse averages);
the per-patch kernels are the oracle's.
"""
import numpy as np

import oracle


class HostHierarchy:
    def __init__(self, h, gas, reflect=(1, 1, 1, 1)):
        self.h, self.gas, self.reflect = h, gas, tuple(reflect)
        self.absorb_width = 0.0  # envelopes on the sides whose reflect flag is 0, when > 0
        self.off, self.sk, self.sy = [], [], []
        off = 0
        for p in h.patches:
            self.off.append(off)
            self.sk.append(p.Nx * p.Ny)
            self.sy.append(p.Nx)
            off += 7 * p.Nx * p.Ny
        self.pool = np.zeros(off)
        self.off = np.array(self.off, np.int64)
        self.sk = np.array(self.sk, np.int64)
        self.sy = np.array(self.sy, np.int64)
        self.src = [np.zeros((p.Ny, p.Nx)) for p in h.patches]

    def view(self, k):
        p = self.h.patches[k]
        o, sk = self.off[k], self.sk[k]
        w = self.pool[o:o + 5 * sk].reshape(5, p.Ny, p.Nx)
        pr = self.pool[o + 5 * sk:o + 6 * sk].reshape(p.Ny, p.Nx)
        T = self.pool[o + 6 * sk:o + 7 * sk].reshape(p.Ny, p.Nx)
        return w, pr, T

    # ---- ghost fill (mirrors fill_copy / fill_interp / fill_bc kernels) ----
    def fill_level(self, lev):
        f = self.h.fill[lev]
        P = self.pool
        dp, dl, sp, sl = f["copy"]
        if len(dp):
            d0, s0 = self.off[dp] + dl, self.off[sp] + sl
            for k in range(7):
                P[d0 + k * self.sk[dp]] = P[s0 + k * self.sk[sp]]
        dp, dl, sp, sl, fx, fy = f["interp"]
        if len(dp):
            d0, c0 = self.off[dp] + dl, self.off[sp] + sl
            sy = self.sy[sp]
            for k in range(7):
                c = c0 + k * self.sk[sp]
                P[d0 + k * self.sk[dp]] = P[c] + fx * (0.5 * (P[c + 1] - P[c - 1])) + fy * (0.5 * (P[c + sy] - P[c - sy]))
        for key in ("edge", "corner"):
            pp, dl, sl, side = f[key]
            if not len(pp):
                continue
            d0, s0 = self.off[pp] + dl, self.off[pp] + sl
            refl = np.array(self.reflect)[side - 1] != 0
            negk = np.where(refl, np.where(side <= 2, 1, 2), -1)
            for k in range(7):
                v = P[s0 + k * self.sk[pp]]
                P[d0 + k * self.sk[pp]] = np.where(negk == k, -v, v)

    def fill_all(self):
        for lev in range(self.h.nlevels):
            self.fill_level(lev)

    def _src_pool(self):
        """The per-patch ghosted sources as one flat array (offsets, strides)."""
        off = np.concatenate([[0], np.cumsum([s.size for s in self.src])[:-1]]).astype(np.int64)
        return np.concatenate([s.ravel() for s in self.src]), off

    def _set_src_pool(self, flat, off):
        for k, s in enumerate(self.src):
            s[...] = flat[off[k]:off[k] + s.size].reshape(s.shape)

    def zero_source(self):
        """enlightning::setSourceToZero (enlightning.C:2310-2329): the whole
        source of every patch set to 0."""
        for s in self.src:
            s[...] = 0.0

    def fill_source(self):
        """Ghost fill of the source (one component, nothing negated at the
        walls), level by level, with the state's schedule."""
        S, off = self._src_pool()
        sy = self.sy
        for lev in range(self.h.nlevels):
            f = self.h.fill[lev]
            dp, dl, sp, sl = f["copy"]
            S[off[dp] + dl] = S[off[sp] + sl]
            dp, dl, sp, sl, fx, fy = f["interp"]
            c = off[sp] + sl
            S[off[dp] + dl] = S[c] + fx * (0.5 * (S[c + 1] - S[c - 1])) + fy * (0.5 * (S[c + sy[sp]] - S[c - sy[sp]]))
            for key in ("edge", "corner"):
                pp, dl, sl, _ = f[key]
                S[off[pp] + dl] = S[off[pp] + sl]
        self._set_src_pool(S, off)

    def transfer_from(self, old, tr):
        """Regrid transfer from HostHierarchy `old` (schedule
        enlightning_amd.amr.make_transfer): old ghosts refilled (source, then
        state), whole interiors of unchanged patches, copies and linear
        refinements of w_n, p, T and the source into this hierarchy's
        interiors, then this hierarchy's source and
        state ghost fills.  Same arithmetic as the device transfer kernel."""
        old.fill_source()
        old.fill_all()
        P, Q = self.pool, old.pool
        for dk, sk in zip(*tr["same"]):
            a, b = self.view(int(dk)), old.view(int(sk))
            g = self.h.patches[dk].g
            for x, y in zip(a, b):
                x[..., g:-g, g:-g] = y[..., g:-g, g:-g]
            self.src[dk][g:-g, g:-g] = old.src[sk][g:-g, g:-g]
        S, soff = self._src_pool()
        OS, osoff = old._src_pool()
        dp, dl, sp, sl = tr["copy"]
        if len(dp):
            d0, s0 = self.off[dp] + dl, old.off[sp] + sl
            for k in range(7):
                P[d0 + k * self.sk[dp]] = Q[s0 + k * old.sk[sp]]
            S[soff[dp] + dl] = OS[osoff[sp] + sl]
        dp, dl, sp, sl, fx, fy = tr["interp"]
        if len(dp):
            d0 = self.off[dp] + dl
            sy = old.sy[sp]

            def refine(A, c):
                return A[c] + fx * (0.5 * (A[c + 1] - A[c - 1])) + fy * (0.5 * (A[c + sy] - A[c - sy]))

            for k in range(7):
                P[d0 + k * self.sk[dp]] = refine(Q, old.off[sp] + sl + k * old.sk[sp])
            S[soff[dp] + dl] = refine(OS, osoff[sp] + sl)
        self._set_src_pool(S, soff)
        self.fill_source()
        self.fill_all()

    def coarsen(self):
        if self.h.coarsen is None:
            return
        cp, cl, fp, fl = self.h.coarsen
        r = self.h.ratio
        P = self.pool
        c0, f0 = self.off[cp] + cl, self.off[fp] + fl
        for k in range(7):
            s = np.zeros(len(cp))
            for fj in range(r):
                for fi in range(r):
                    s = s + P[f0 + fi + fj * self.sy[fp] + k * self.sk[fp]]
            P[c0 + k * self.sk[cp]] = s * (1.0 / (r * r))

    # ---- state --------------------------------------------------------------
    def load_initial(self, states, sources=None):
        for k, w in enumerate(states):
            self.view(k)[0][...] = w
            if sources is not None:
                self.src[k][...] = sources[k]
        self.fill_source()
        self.fill_all()
        for k, p in enumerate(self.h.patches):
            w, pr, T = self.view(k)
            dx = self.h.dx(p.level)
            oracle.update_state(p.box, p.g, (p.box[0] * dx, p.box[2] * dx), (dx, dx), w, pr, T, self.gas)
        self.fill_all()

    def advance(self, nsteps, dt, rk, variant="hybrid_high_llf"):
        h = self.h
        a1, a2, b = rk
        for _ in range(nsteps):
            ww = [np.ascontiguousarray(self.view(k)[0][:, p.g:-p.g, p.g:-p.g]) for k, p in enumerate(h.patches)]
            for st in range(len(a1)):
                for lev in range(h.nlevels):
                    self.fill_level(lev)
                    for k, p in enumerate(h.patches):
                        if p.level != lev:
                            continue
                        w, pr, T = self.view(k)
                        dx = h.dx(p.level)
                        xlo = (p.box[0] * dx, p.box[2] * dx)
                        KK = np.zeros((5, p.ny, p.nx))
                        src = np.ascontiguousarray(self.src[k][p.g:-p.g, p.g:-p.g])
                        oracle.rhs(variant, p.box, p.g, xlo, (dx, dx), w, src, KK, pr, T, self.gas, p.level)
                        oracle.runge_kutta(p.box, p.g, dt, a1[st], a2[st], b[st], w, ww[k], KK)
                        oracle.update_state(p.box, p.g, xlo, (dx, dx), w, pr, T, self.gas)
                        if self.absorb_width > 0:
                            n0 = h.ncells(0) * h.dx(0)
                            sides = tuple(int(not r) for r in self.reflect)
                            oracle.absorb(p.box, p.g, xlo, (dx, dx), (0.0, 0.0), (n0, n0), sides,
                                          self.absorb_width, self.gas, w, pr, T)
                    if lev + 1 < h.nlevels:
                        self.fill_level(lev)
            self.coarsen()

    def states(self):
        return [tuple(a.copy() for a in self.view(k)) for k in range(len(self.h.patches))]
