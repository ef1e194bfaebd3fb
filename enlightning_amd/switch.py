"""The learned WENO/DRP switch: its C ABI (enl_switch_*), the tortuous
line-source workload it is trained and judged on, the training-set builder
(plain-sensor maps of seeded channels) and the training loop, all on the
device through the library's own kernels.

The network stands in for the hybrid sensor (rhs_hybrid_high_llf.m4:239-256)
on level 0: p(WENO) > hi -> WENO, < lo -> DRP, otherwise the sensor decides.
"""
import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from . import patch as P
from .gas import RK3
from .workloads import Workload, state_from_overpressure

N_PARAM = 8577
PARAMS_FILE = os.path.join(os.path.dirname(__file__), "data", "switch_params.npy")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def init_params(seed=0):
    """He-initialised parameters, packed as W1[64][100], b1, W2[32][64], b2, w3, b3."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(scale=math.sqrt(2.0 / 100), size=(64, 100))
    W2 = rng.normal(scale=math.sqrt(2.0 / 64), size=(32, 64))
    w3 = rng.normal(scale=math.sqrt(1.0 / 32), size=32)
    return np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(32), w3, [0.0]]).astype(np.float32)


class Switch:
    def __init__(self, params=None):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.enl_switch_create(ctypes.byref(h)))
        self.handle = h
        self.set_params(init_params() if params is None else params)

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == N_PARAM
        _lib.check(self.lib.enl_switch_set_params(self.handle, ctypes.c_void_p(p.ctypes.data)))

    def params(self):
        p = np.empty(N_PARAM, np.float32)
        _lib.check(self.lib.enl_switch_get_params(self.handle, ctypes.c_void_p(p.ctypes.data)))
        return p

    def features(self, dp, cells):
        """BF16 feature rows (int16 tensor, n x 100) of interior cells of a
        DevicePatch state."""
        cells = cells.to(dtype=torch.int32, device=dp.w.device).contiguous()
        X = torch.empty((cells.numel(), 100), dtype=torch.int16, device=dp.w.device)
        gs = dp.gas.to_struct()
        P.use_torch_stream()
        _lib.check(self.lib.enl_switch_features(dp.box[0], dp.box[1], dp.box[2], dp.box[3], dp.g, _ptr(dp.w),
                                                _ptr(dp.p), ctypes.byref(gs), _ptr(cells), cells.numel(), _ptr(X)))
        return X

    def forward(self, X):
        p = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        P.use_torch_stream()
        _lib.check(self.lib.enl_switch_forward(self.handle, _ptr(X), X.shape[0], _ptr(p)))
        return p

    def gradients(self, X, y, w):
        g = torch.empty(N_PARAM, dtype=torch.float32, device=X.device)
        loss = ctypes.c_float()
        P.use_torch_stream()
        _lib.check(self.lib.enl_switch_gradients(self.handle, _ptr(X), _ptr(y), _ptr(w), X.shape[0],
                                                 float(w.sum().item()), _ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, X, y, w, lr, wsum=None):
        loss = ctypes.c_float()
        P.use_torch_stream()
        ws = float(w.sum().item()) if wsum is None else wsum
        _lib.check(self.lib.enl_switch_train_step(self.handle, _ptr(X), _ptr(y), _ptr(w), X.shape[0], ws,
                                                  float(lr), ctypes.byref(loss)))
        return loss.value

    def decide(self, dp, lo, hi):
        """Decision map (uint8 tensor ny x nx: 0 DRP, 1 WENO, 2 sensor) of a
        DevicePatch state, and the counts (candidates, WENO, undecided)."""
        d = torch.empty((dp.n, dp.n), dtype=torch.uint8, device=dp.w.device)
        counts = (ctypes.c_uint * 3)()
        gs = dp.gas.to_struct()
        P.use_torch_stream()
        _lib.check(self.lib.enl_switch_decide(self.handle, dp.box[0], dp.box[1], dp.box[2], dp.box[3], dp.g,
                                              _ptr(dp.w), _ptr(dp.p), ctypes.byref(gs), float(lo), float(hi),
                                              _ptr(d), counts))
        return d, tuple(counts)

    def advance(self, dp, nsteps, dt, lo, hi, variant=None, rk=RK3, every_stage=True):
        """enl_advance_patch_fused with this switch deciding level 0's cells
        before every stage, or once per step (every_stage=False); returns the
        summed counts (candidates, WENO, undecided) over the decisions."""
        lib = self.lib
        P.use_torch_stream()
        variant = dp.wl.variant if variant is None else variant
        vid = _lib.VARIANT_ID[variant] if isinstance(variant, str) else int(variant)
        nbytes = lib.enl_fused_scratch_bytes(dp.box[0], dp.box[1], dp.box[2], dp.box[3], dp.g)
        if getattr(dp, "_scratch", None) is None or dp._scratch.numel() < nbytes:
            dp._scratch = torch.empty(nbytes, dtype=torch.uint8, device=dp.w.device)
        if getattr(dp, "_decide", None) is None:
            dp._decide = torch.empty(dp.n * dp.n, dtype=torch.uint8, device=dp.w.device)
        totals = torch.zeros(3, dtype=torch.int64, device=dp.w.device)
        a1, a2, b = rk
        D = ctypes.c_double * len(a1)
        gs = dp.gas.to_struct()
        _lib.check(lib.enl_advance_patch_learned(
            self.handle, float(lo), float(hi), int(bool(every_stage)), _ptr(dp._decide), _ptr(totals), vid, dp.box[0], dp.box[1],
            dp.box[2], dp.box[3], dp.g, D(*dp.xlo), D(*dp.dx), ctypes.byref(gs), (ctypes.c_int * 4)(*dp.reflect),
            len(a1), D(*a1), D(*a2), D(*b), float(dt), int(nsteps), _ptr(dp.w), _ptr(dp.p), _ptr(dp.T),
            _ptr(dp.src), _ptr(dp._scratch), nbytes, float(getattr(dp, "absorb_width", 0.0))))
        return tuple(int(v) for v in totals.cpu())

    def close(self):
        if self.handle:
            self.lib.enl_switch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- the tortuous line-source workload --------------------------------------

@dataclass
class ChannelWorkload(Workload):
    segments: int = 200
    seg_len: float = 10.0           # m
    angle_sd_deg: float = 25.0      # random-walk turn per segment
    width: float = 0.5              # Gaussian cross-section sigma, m
    dp_range: tuple = (2.0, 10.0)   # dp/p0 on the channel axis


# 4096^2 base at 0.25 m; CFL 0.2 (the strong channel shocks travel at up to
# ~3 c, beyond what the reference's c-based dt allows at 0.5).
CHANNEL_4096 = ChannelWorkload("channel_4096_025m", n=4096, L=1024.0, steps=100, seed=1000, cfl=0.2)


def channel_path(wl, seed):
    """Polyline of the random-walk channel (segments+1 points, metres),
    reflected at a 5 % margin inside the domain."""
    rng = np.random.default_rng(seed)
    lo, hi = 0.05 * wl.L, 0.95 * wl.L
    pts = [rng.uniform(0.3 * wl.L, 0.7 * wl.L, size=2)]
    ang = rng.uniform(0.0, 2 * np.pi)
    for _ in range(wl.segments):
        ang += np.deg2rad(rng.normal(0.0, wl.angle_sd_deg))
        q = pts[-1] + wl.seg_len * np.array([np.cos(ang), np.sin(ang)])
        for k in range(2):
            if q[k] < lo or q[k] > hi:
                q[k] = 2 * (lo if q[k] < lo else hi) - q[k]
                ang = np.pi - ang if k == 0 else -ang
        pts.append(q)
    return np.array(pts), float(rng.uniform(*wl.dp_range))


def channel_overpressure(wl, seed):
    """dp/p0 on the ghosted cell centres: amp * exp(-d^2 / (2 w^2)), d the
    distance to the channel polyline (evaluated within 5 w of each segment)."""
    n, g, dx = wl.n, wl.g, wl.dx
    N = n + 2 * g
    pts, amp = channel_path(wl, seed)
    best = np.zeros((N, N))
    reach = 5 * wl.width
    for a, b in zip(pts[:-1], pts[1:]):
        i0 = max(int((min(a[0], b[0]) - reach) / dx) + g, 0)
        i1 = min(int((max(a[0], b[0]) + reach) / dx) + g + 1, N)
        j0 = max(int((min(a[1], b[1]) - reach) / dx) + g, 0)
        j1 = min(int((max(a[1], b[1]) + reach) / dx) + g + 1, N)
        x = (np.arange(i0, i1) - g + 0.5) * dx
        y = (np.arange(j0, j1) - g + 0.5) * dx
        X, Y = np.meshgrid(x, y)
        ab = b - a
        t = np.clip(((X - a[0]) * ab[0] + (Y - a[1]) * ab[1]) / (ab @ ab), 0.0, 1.0)
        d2 = (X - a[0] - t * ab[0]) ** 2 + (Y - a[1] - t * ab[1]) ** 2
        np.maximum(best[j0:j1, i0:i1], np.exp(-d2 / (2 * wl.width ** 2)), out=best[j0:j1, i0:i1])
    return amp * best


def channel_patch(wl, gas, seed, device="cuda"):
    """A DevicePatch holding the channel's initial state."""
    dp = P.DevicePatch(wl, gas, device=device)
    w = state_from_overpressure(channel_overpressure(wl, seed), gas)
    N = wl.n + 2 * wl.g
    dp.init_state(w, np.full((N, N), gas.p0), np.full((N, N), gas.T0), np.zeros((N, N)))
    return dp


# ---- training data and training -----------------------------------------------

def sensor_map(dp):
    """The plain sensor's choice (int32 ny x nx, 1 = WENO) on a DevicePatch
    state, from the rhs drop-in's switch-map output."""
    KK = torch.empty((5, dp.n, dp.n), dtype=torch.float64, device=dp.w.device)
    sw = torch.empty((dp.n, dp.n), dtype=torch.int32, device=dp.w.device)
    P.fill_physical_bc(dp.box, dp.g, (1, 1, 1, 1), dp.reflect, dp.w, dp.p, dp.T)
    src = dp.src[dp.g:-dp.g, dp.g:-dp.g].contiguous()
    P.rhs(dp.wl.variant, dp.box, dp.g, dp.xlo, dp.dx, dp.w, src, KK, dp.p, dp.T, dp.gas, 0, switch_map=sw)
    return sw


def candidates(dp):
    """The prefilter's candidates (bool ny x nx): some sensor pair with
    both t1^2 + t2^2 and (t1 - t2)^2 at or above 5e-5 eps (either below it
    proves the sensor's DRP)."""
    from .gas import sensor_eps
    g, n = dp.g, dp.n
    r = dp.w[0]
    thr = 5e-5 * sensor_eps(dp.gas.tag_tolerance[2])
    dxr = r[:, 1:] - r[:, :-1]   # dxr[j, i] = r[j, i+1] - r[j, i]
    dyr = r[1:, :] - r[:-1, :]
    cand = torch.zeros((n, n), dtype=torch.bool, device=r.device)
    J = slice(g, g + n)
    for q in (-1, 0, 1):
        t1 = dxr[J, g - 1 + q - 1 + 1:g - 1 + q + n]          # r[i+q] - r[i-1+q]
        t2 = dxr[J, g + q:g + q + n]                          # r[i+1+q] - r[i+q]
        cand |= (t1 * t1 + t2 * t2 >= thr) & ((t1 - t2) * (t1 - t2) >= thr)
        t1 = dyr[g - 1 + q:g - 1 + q + n, g:g + n]
        t2 = dyr[g + q:g + q + n, g:g + n]
        cand |= (t1 * t1 + t2 * t2 >= thr) & ((t1 - t2) * (t1 - t2) >= thr)
    return cand


def build_samples(sw, wl, gas, seeds, snapshots=(0, 20, 60), dt=None, per_snapshot=40000, rng_seed=0,
                  weno_share=0.75):
    """Features, labels and weights from plain-sensor maps: for each channel
    seed and snapshot step, every sensor-WENO candidate (up to
    per_snapshot/2) and as many sensor-DRP candidates; labels are the
    sensor's choice.  The WENO class carries weno_share of the total weight,
    so that calling a sensor-WENO cell DRP (a missed shock) costs more than
    the opposite error."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(rng_seed)
    Xs, ys = [], []
    dt = wl.dt(gas) if dt is None else dt
    for seed in seeds:
        dp = channel_patch(wl, gas, seed)
        done = 0
        for s in snapshots:
            if s > done:
                dp.advance_fused(s - done, dt)
                done = s
            lab = sensor_map(dp).flatten()
            cand = candidates(dp).flatten()
            pos = torch.nonzero(cand & (lab == 1)).flatten()
            neg = torch.nonzero(cand & (lab == 0)).flatten()
            k = per_snapshot // 2
            if pos.numel() > k:
                pos = pos[torch.randperm(pos.numel(), generator=gen, device=pos.device)[:k]]
            kn = min(neg.numel(), max(k, per_snapshot - pos.numel()))
            neg = neg[torch.randperm(neg.numel(), generator=gen, device=neg.device)[:kn]]
            cells = torch.cat([pos, neg])
            Xs.append(sw.features(dp, cells))
            ys.append(torch.cat([torch.ones(pos.numel()), torch.zeros(neg.numel())]).to(device="cuda"))
        del dp
        torch.cuda.empty_cache()
    X = torch.cat(Xs)
    y = torch.cat(ys).float()
    npos = float(y.sum().item())
    nneg = float(y.numel() - npos)
    w = torch.where(y > 0.5, weno_share / max(npos, 1.0), (1.0 - weno_share) / max(nneg, 1.0)).float() * y.numel()
    return X, y, w


def train(sw, X, y, w, epochs=4, batch=8192, lr=2e-3, seed=0, log=None):
    """Adam over shuffled mini-batches; returns the per-epoch mean losses."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    n = X.shape[0]
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen, device=X.device)
        tot, cnt = 0.0, 0
        for b0 in range(0, n - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            Xb, yb, wb = X[idx].contiguous(), y[idx].contiguous(), w[idx].contiguous()
            tot += sw.train_step(Xb, yb, wb, lr * (0.5 ** ep))
            cnt += 1
        losses.append(tot / max(cnt, 1))
        if log:
            log("epoch %d loss %.5f" % (ep, losses[-1]))
    return losses


def evaluate(sw, X, y, lo=0.2, hi=0.8):
    """Decision statistics of the network on labelled rows: accuracy at 0.5,
    the fractions decided WENO / DRP / left to the sensor, and the sensor-WENO
    rows the network confidently called DRP (missed shocks)."""
    p = sw.forward(X)
    pos = y > 0.5
    weno = p > hi
    drp = p < lo
    return {
        "rows": int(y.numel()),
        "accuracy": float(((p > 0.5) == pos).float().mean().item()),
        "decided_weno": float(weno.float().mean().item()),
        "decided_drp": float(drp.float().mean().item()),
        "to_sensor": float((~weno & ~drp).float().mean().item()),
        "missed_shocks": int((pos & drp).sum().item()),
        "false_weno": int((~pos & weno).sum().item()),
    }


def channel_probabilities(sw, wl, gas, seeds, snapshots=(0, 20, 60, 100), dt=None):
    """p(WENO) of the network on every prefilter candidate of the given
    channels at the given steps of the plain FP64 fused advance, split by the
    sensor's own choice there: (p of sensor-WENO cells, p of sensor-DRP
    candidates), float32 host arrays, and the level-0 cell count visited."""
    dt = wl.dt(gas) if dt is None else dt
    pw, pd = [], []
    visited = 0
    for seed in seeds:
        dp = channel_patch(wl, gas, seed)
        done = 0
        for st in snapshots:
            if st > done:
                dp.advance_fused(st - done, dt)
                done = st
            lab = sensor_map(dp).flatten()
            cand = candidates(dp).flatten()
            assert not bool((lab.bool() & ~cand).any().item()), "a sensor-WENO cell outside the prefilter's candidates"
            for sel, out in ((cand & (lab == 1), pw), (cand & (lab == 0), pd)):
                cells = torch.nonzero(sel).flatten()
                for c0 in range(0, cells.numel(), 1 << 20):
                    X = sw.features(dp, cells[c0:c0 + (1 << 20)])
                    out.append(sw.forward(X).cpu().numpy())
            visited += wl.n * wl.n
        del dp
        torch.cuda.empty_cache()
    cat = lambda a: np.concatenate(a) if a else np.zeros(0, np.float32)
    return cat(pw), cat(pd), visited


def calibrate_lo(p_weno, margin=0.5):
    """The DRP threshold that calls none of the given sensor-WENO cells DRP:
    margin times the smallest p among them (the matrix-core inference may
    round p differently from the forward kernel by far less than that)."""
    return float(margin * p_weno.min()) if p_weno.size else 0.0


def decision_shares(p_weno, p_drp, lo, hi):
    """Counts of the network's calls on candidates at thresholds (lo, hi):
    missed shocks (sensor WENO, p < lo), network DRP, network WENO, left to
    the sensor."""
    allp = np.concatenate([p_weno, p_drp])
    return {"candidates": int(allp.size), "sensor_weno": int(p_weno.size),
            "missed_shocks": int((p_weno < lo).sum()), "network_drp": int((allp < lo).sum()),
            "network_weno": int((allp > hi).sum()), "to_sensor": int(((allp >= lo) & (allp <= hi)).sum()),
            "false_weno": int((p_drp > hi).sum())}


THRESHOLDS_FILE = os.path.join(os.path.dirname(__file__), "data", "switch_thresholds.json")


def load_thresholds(default=(0.2, 0.8)):
    """(lo, hi) calibrated for the shipped parameters (train_switch
    --calibrate), or `default` if no calibration is shipped."""
    import json
    if os.path.exists(THRESHOLDS_FILE):
        with open(THRESHOLDS_FILE) as f:
            t = json.load(f)
        return float(t["lo"]), float(t["hi"])
    return default


def load_params():
    """The trained parameters shipped with the package (None if absent)."""
    if os.path.exists(PARAMS_FILE):
        return np.load(PARAMS_FILE).astype(np.float32)
    return None
