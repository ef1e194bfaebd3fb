"""Device operations used by the GPU tests.

Normally this is the HIP library through enlightning_amd.patch with torch
device memory.  With ENL_TEST_DRY_RUN=1 (pytest --dry-gpu) the same tests run
on the CPU with the oracle standing in for the device, which checks the
tests' own logic and inputs without a GPU; it proves nothing about the
kernels.
"""
import os

import numpy as np

DRY = os.environ.get("ENL_TEST_DRY_RUN") == "1"

if not DRY:
    import torch

    from enlightning_amd import patch as P

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a)).cuda()

    def host(t):
        torch.cuda.synchronize()
        return t.cpu().numpy()

    def zeros(shape, dtype=np.float64):
        return torch.zeros(shape, dtype=torch.int32 if dtype == np.int32 else torch.float64, device="cuda")

    def full(shape, value, dtype=np.float64):
        return torch.full(shape, value, dtype=torch.int32 if dtype == np.int32 else torch.float64, device="cuda")

    def sync():
        torch.cuda.synchronize()

    def interior_copy(t, g):
        return t[g:-g, g:-g].contiguous()

    def all_finite(t):
        return bool(torch.isfinite(t).all().item())

    DevicePatch = P.DevicePatch
    use_stream = P.use_torch_stream
else:
    import oracle as P  # noqa: N812  (the oracle stands in for the device)
    from enlightning_amd.gas import RK3

    def dev(a):
        return np.array(a, copy=True, order="C")

    def host(t):
        return t

    def zeros(shape, dtype=np.float64):
        return np.zeros(shape, dtype)

    def full(shape, value, dtype=np.float64):
        return np.full(shape, value, dtype)

    def sync():
        pass

    def interior_copy(t, g):
        return np.ascontiguousarray(t[g:-g, g:-g])

    def all_finite(t):
        return bool(np.isfinite(t).all())

    class DevicePatch(P.HostPatch):
        def advance(self, nsteps, dt, variant=None, level=0, rk=RK3):
            super().advance(nsteps, dt, rk, variant=variant, level=level, tile=64, nthreads=8)

        def advance_fused(self, nsteps, dt, variant=None, level=0, rk=RK3):
            self.advance(nsteps, dt, variant=variant, level=level, rk=rk)

        def to_host(self):
            return self.w.copy(), self.p.copy(), self.T.copy(), self.src.copy()

    def use_stream():
        pass


# ---- hierarchies ----------------------------------------------------------
if not DRY:
    from enlightning_amd.hierarchy import DeviceHierarchy as Hierarchy
else:
    from oracle.amr import HostHierarchy as _HostHierarchy

    class Hierarchy(_HostHierarchy):
        """The CPU twin behind DeviceHierarchy's interface."""

        def load_initial(self, wl, states=None, sources=None):
            from enlightning_amd.hierarchy import patch_initial_state
            if states is None:
                states = [patch_initial_state(self.h, p, wl, self.gas) for p in self.h.patches]
            super().load_initial(states, sources)

        def sources(self):
            return [s.copy() for s in self.src]

        def set_states(self, states):
            for k, (w, pr, T) in enumerate(states):
                a, b, c = self.view(k)
                a[...], b[...], c[...] = w, pr, T

        def upload_all(self, states, sources):
            self.set_states(states)
            for k, src in enumerate(sources):
                self.src[k][...] = src
            self.fill_source()
            self.fill_all()

        def advance(self, nsteps, dt, mode="graph", variant="hybrid_high_llf", rk=RK3):
            super().advance(nsteps, dt, rk, variant=variant)

        def set_tiles(self, bx, by, level=None):
            pass

        def set_packing(self, level, mode, wpc=4):
            pass

        def set_absorb(self, width):
            self.absorb_width = width

        def add_source(self, src, iteration, pulse_time0, dt):
            for p in self.h.patches:
                dx = self.h.dx(p.level)
                tmp = np.zeros((p.ny, p.nx))
                P.add_source(p.box, (p.box[0] * dx, p.box[2] * dx), (dx, dx), src, iteration, pulse_time0, dt, tmp)
                self.src[p.pid][p.g:-p.g, p.g:-p.g] = tmp
            self.fill_source()

        def sample_mics(self, mic_x, mic_y):
            res = [[] for _ in mic_x]
            for lev in range(self.h.nlevels):
                dx = self.h.dx(lev)
                for p in self.h.level_patches(lev):
                    lo = (p.box[0] * dx, p.box[2] * dx)
                    hi = ((p.box[1] + 1) * dx, (p.box[3] + 1) * dx)
                    pr = np.ascontiguousarray(self.view(p.pid)[1][p.g:-p.g, p.g:-p.g])
                    v, hit = P.sample_mics(p.box, lo, hi, (dx, dx), mic_x, mic_y, pr)
                    for k in np.nonzero(hit)[0]:
                        res[k].append((lev, p.pid, float(v[k])))
            return res

        def level0_tags(self, refine_tag_val=1):
            self.fill_all()
            n0 = self.h.ncells(0)
            tags = np.zeros((n0, n0), np.int32)
            for p in self.h.level_patches(0):
                w, _, _ = self.view(p.pid)
                t = np.zeros((p.ny, p.nx), np.int32)
                P.tag_cells(p.box, p.g, t, np.ascontiguousarray(w), self.src[p.pid], refine_tag_val,
                                 self.gas.tag_tolerance)
                tags[p.box[2]:p.box[3] + 1, p.box[0]:p.box[1] + 1] = t
            return tags

        def regrid(self, new_h, transfer=None):
            from enlightning_amd import amr
            tr = amr.make_transfer(new_h, self.h) if transfer is None else transfer
            new = Hierarchy(new_h, self.gas, self.reflect)
            new.absorb_width = self.absorb_width
            new.transfer_from(self, tr)
            return new

        def close(self):
            pass
