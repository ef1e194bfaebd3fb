"""Throughput of the per-patch update path on one MI355X.

Workload (one step = one SSP-RK3 step of the whole patch):
  pulses_4096_8 -- one 4096 x 4096 patch on [0,1]^2 m, 4 ghost cells,
  reflecting walls, ideal gas gamma = 1.4, rho0 = 1.2, p0 = 101325, eight
  seeded Gaussian overpressure pulses (dp/p0 in [0.05, 0.8], sigma in
  [0.005, 0.03] m), hybrid WENO5/DRP with LLF splitting (rhs_hybrid_high_llf),
  CFL 0.5, FP64: the plain drop-in (rhs -> runge_kutta -> update_state per
  stage, SAMRAI-layout arrays resident in HBM).

Both forms run from the same initial state for the same steps: the plain
drop-in (per stage rhs_, runge_kutta_, update_state_ plus SAMRAI's copy and
boundary fill) and the tuned fused form (one kernel per stage, bit-identical
results, checked at the end of the run).

Prints one JSON line.  value = interior cell-updates per second of the tuned
form over the timed steps (inputs resident in HBM); plain_fp64_dropin holds
the plain form's figures.  roofline: the dominant kernel's algorithmic HBM
bytes over its average launch duration, timed with HIP events on its launch
stream during the timed region (fused stage kernel: 146.7 B per cell on
average; plain rhs: 104 B per cell).  cpu_baseline: the CPU restatement of
the same kernels (oracle/) on this host's cores, over a bounded sample.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
RHS_BYTES_PER_CELL = 104.0  # 13 doubles: w_n(5), p, T, source read; rhs(5) written
STEP_BYTES_PER_CELL = 888.0  # 3 stages x (rhs 104 + runge_kutta 160 + update_state 32)
FUSED_STAGE_BYTES_PER_CELL = (120.0 + 160.0 + 160.0) / 3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096, help="interior cells per side")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-amr", action="store_true")
    ap.add_argument("--amr-steps", type=int, default=40)
    ap.add_argument("--amr-plain-steps", type=int, default=4)
    ap.add_argument("--precision-steps", type=int, default=2000)
    ap.add_argument("--switch-steps", type=int, default=100)
    return ap.parse_args()


TRAFFIC_FILES = {"rhs": "rhs_hybrid_high_llf_4096_traffic.txt", "stage_fused": "stage_fused_4096_traffic.txt"}


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (the line `traffic_bytes_per_launch: <bytes>`), or None."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILES[kernel])
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def cpu_baseline(wl, gas, steps):
    import numpy as np  # noqa: F401
    import oracle
    from enlightning_amd.gas import RK3
    from enlightning_amd.workloads import initial_fields
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(16, cores))
    w, p, T, s = initial_fields(wl, gas)
    hp = oracle.HostPatch(wl, gas)
    hp.init_state(w, p, T, s)
    del w, p, T, s
    t0 = time.perf_counter()
    hp.advance(steps, wl.dt(gas), RK3, tile=64, nthreads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": wl.n * wl.n * steps / dt,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d RK3 step(s) of %s from its initial state: the C restatement of the reference kernels "
                  "(oracle/enl_oracle.c, gcc -O2, no FMA), 64x64 tiles (the largest level-0 patch of "
                  "input.txt:114-116) over %d OpenMP threads; %.2f s wall" % (steps, wl.name, threads, dt),
    }


def amr_bench(args):
    """The batched two-level hierarchy: base 1024^2 in 64^2 patches, advanced
    on the uniform grid until the pulses have spread, then one level refined
    by 4 over tag_cells' tags.  The drop-in entry points patch by patch
    (plain) against the batched fused stages replayed from a hipGraph, with
    the stage kernels' tile shape picked by timing each candidate; then one
    regrid of the evolved state (device tags, host clustering and schedules,
    the new batch and the transfer on the device), timed by part."""
    import numpy as np
    import torch
    from enlightning_amd import amr
    from enlightning_amd import patch as P
    from enlightning_amd.gas import Gas
    from enlightning_amd.hierarchy import DeviceHierarchy, refine_from_base
    from enlightning_amd.workloads import AMR_1024, initial_fields

    gas = Gas()
    wl = AMR_1024
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
    h = amr.build(n, 4, 4, 1.0, tags, tile0=64)
    states = [refine_from_base(h, p, wb) for p in h.patches]
    dt = 0.5 * h.dx(h.nlevels - 1) / gas.c  # one dt for all levels (enlightning.C:446-447)
    cells = h.interior_cells()
    d = DeviceHierarchy(h, gas)

    def run(mode, steps):
        d.load_initial(wl, states)
        d.advance(1, dt, mode=mode)  # warmup (and graph capture)
        d.load_initial(wl, states)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.advance(steps, dt, mode=mode)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    best, tune = d.autotune(dt, states, wl, steps=3)
    t_graph = run("graph", args.amr_steps)
    t_plain = run("plain", args.amr_plain_steps)
    plain_states = d.states()
    d.load_initial(wl, states)
    d.advance(args.amr_plain_steps, dt, mode="graph")
    graph_states = d.states()
    same = all(np.array_equal(a[k][..., p.g:-p.g, p.g:-p.g], b[k][..., p.g:-p.g, p.g:-p.g])
               for a, b, p in zip(plain_states, graph_states, h.patches) for k in range(3))
    # one regrid of the evolved state (input.txt:7 regrids every 2 steps):
    # tags on the device, clustering and schedules on the host, the new
    # batch and the transfer on the device
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rtags = d.level0_tags()
    t1 = time.perf_counter()
    h2 = amr.build(n, 4, 4, 1.0, rtags, tile0=64, previous=h)
    tr = amr.make_transfer(h2, h)
    t2 = time.perf_counter()
    d2 = d.regrid(h2, tr)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    d2.close()
    regrid = {"tags_device_ms": (t1 - t0) * 1e3, "cluster_and_schedules_host_ms": (t2 - t1) * 1e3,
              "new_batch_and_transfer_device_ms": (t3 - t2) * 1e3, "total_ms": (t3 - t0) * 1e3,
              "fine_patches_after": len(h2.level_patches(1)), "copied_cells": int(len(tr["copy"][0])),
              "refined_cells": int(len(tr["interp"][0]))}
    d.close()
    sizes = [max(p.nx, p.ny) for p in h.level_patches(1)]
    return {
        "workload": wl.name,
        "base_steps_before_tagging": wl.presteps,
        "tagged_base_cells": int(tags.sum()),
        "patches_per_level": [len(h.level_patches(l)) for l in range(h.nlevels)],
        "cells_per_level": [h.interior_cells(l) for l in range(h.nlevels)],
        "fine_patch_side_min_max": [int(min(sizes)), int(max(sizes))] if sizes else None,
        "autotune_ms_per_step": tune,
        "tiles_per_level": ["%dx%d" % best[l] for l in sorted(best)],
        "graph": {"value": cells * args.amr_steps / t_graph, "ms_per_step": t_graph * 1e3 / args.amr_steps,
                  "steps": args.amr_steps},
        "plain_dropin_per_patch": {"value": cells * args.amr_plain_steps / t_plain,
                                   "ms_per_step": t_plain * 1e3 / args.amr_plain_steps,
                                   "steps": args.amr_plain_steps},
        "speedup_graph_vs_plain": (t_plain / args.amr_plain_steps) / (t_graph / args.amr_steps),
        "graph_bit_identical_to_plain": same,
        "regrid": regrid,
        "unit": "cell-updates/s",
    }, tags


def precision_bench(args, tags):
    """The FP32-perturbation form on the same hierarchy (the boxes of the
    AMR workload, on a 2 km x 2 km domain): 16 weak isentropic pulses on an
    exponentially stratified ambient (scale height 8 km), FP64 batched graph
    against FP32 batched graph from the same state, per-field error."""
    import numpy as np
    import torch
    from enlightning_amd import amr
    from enlightning_amd.gas import Gas
    from enlightning_amd.hierarchy import DeviceHierarchy
    from enlightning_amd.workloads import WEAK_2KM, ambient_rows, stratified_state

    gas = Gas()
    wl = WEAK_2KM
    h = amr.build(wl.n, 4, 4, wl.L, tags, tile0=64)
    states, ambs = [], []
    for p in h.patches:
        dx = h.dx(p.level)
        x = (np.arange(p.box[0] - p.g, p.box[1] + p.g + 1) + 0.5) * dx
        y = (np.arange(p.box[2] - p.g, p.box[3] + p.g + 1) + 0.5) * dx
        states.append(stratified_state(wl, gas, x, y))
        rho, rhos, rhotn = ambient_rows(wl, gas, y)
        ambs.append((rho, rhos, rhotn))

    def ambient(lev):
        n, g = h.ncells(lev), h.g
        y = (np.arange(-g, n + g) + 0.5) * h.dx(lev)
        return (-g,) + ambient_rows(wl, gas, y)

    steps = args.precision_steps
    dt = 0.5 * h.dx(h.nlevels - 1) / gas.c
    cells = h.interior_cells()
    d = DeviceHierarchy(h, gas)
    tiles64, _ = d.autotune(dt, states, wl, steps=3)
    d.load_initial(wl, states)
    d.advance(1, dt, mode="graph")  # capture
    d.load_initial(wl, states)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.advance(steps, dt, mode="graph")
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    s64 = d.states()
    d.load_initial(wl, states)
    d.set_ambient(ambient)
    d.fp32_prepare()
    tiles32, _ = d.autotune(dt, states, wl, steps=3, mode="fp32_graph")
    d.load_initial(wl, states)
    d.fp32_load()
    d.advance(1, dt, mode="fp32_graph")  # capture
    d.load_initial(wl, states)
    d.fp32_load()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.advance(steps, dt, mode="fp32_graph")
    torch.cuda.synchronize()
    t32 = time.perf_counter() - t0
    d.fp32_store()
    s32 = d.states()
    d.close()
    names = ("rho", "rho_u", "rho_v", "rho_s", "rho_Tn", "p", "T")
    err = {}
    for f_i, name in enumerate(names):
        num = den = dev = linf = amax = 0.0
        for p, a, b, am in zip(h.patches, s64, s32, ambs):
            g = p.g
            pick = (lambda st: st[0][f_i]) if f_i < 5 else (lambda st: st[f_i - 4])
            x, y = pick(a)[g:-g, g:-g], pick(b)[g:-g, g:-g]
            if f_i in (0, 3, 4):
                ref = x - am[(0, None, None, 1, 2)[f_i]][g:-g, None]
            elif f_i in (1, 2):
                ref = x
            elif f_i == 5:
                ref = x - gas.p0
            else:
                ref = None
            num += float(np.sum((x - y) ** 2))
            den += float(np.sum(x ** 2))
            if ref is not None:
                dev += float(np.sum(ref ** 2))
            linf = max(linf, float(np.abs(x - y).max()))
            amax = max(amax, float(np.abs(x).max()))
        err[name] = {"rel_l2": float(np.sqrt(num / den)), "rel_linf": linf / amax,
                     "rel_l2_of_perturbation": float(np.sqrt(num / dev)) if dev > 0 else None}
    return {
        "workload": wl.name,
        "hierarchy": "the amr workload's boxes on a %g m domain" % wl.L,
        "steps": steps,
        "fp64_graph": {"value": cells * steps / t64, "ms_per_step": t64 * 1e3 / steps,
                       "tiles_per_level": ["%dx%d" % tiles64[l] for l in sorted(tiles64)]},
        "fp32_perturbation_graph": {"value": cells * steps / t32, "ms_per_step": t32 * 1e3 / steps,
                                    "tiles_per_level": ["%dx%d" % tiles32[l] for l in sorted(tiles32)]},
        "speedup_fp32_vs_fp64": t64 / t32,
        "error_fp32_vs_fp64": err,
        "unit": "cell-updates/s",
    }


def switch_bench(args):
    """The learned switch on the tortuous line source (4096^2 base at 0.25 m,
    held-out channel seed): the plain hybrid (density sensor) against the
    hybrid with the trained network deciding level 0's cells (p > 0.8 WENO,
    p < 0.2 DRP, else the sensor), both fused FP64 from the same state; the
    plain solution judges the learned one's error, and the network's calls are
    compared with the sensor's on the same states (missed shocks: sensor WENO,
    network confident DRP)."""
    import numpy as np
    import torch
    from enlightning_amd import switch as S
    from enlightning_amd.gas import Gas

    params = S.load_params()
    if params is None:
        return {"skipped": "no trained parameters (python -m enlightning_amd.train_switch)"}
    gas = Gas()
    wl = S.CHANNEL_4096
    dt = wl.dt(gas)
    steps = args.switch_steps
    sw = S.Switch(params)
    lo, hi = 0.2, 0.8
    from enlightning_amd import patch as P
    from enlightning_amd.workloads import state_from_overpressure
    N = wl.n + 2 * wl.g
    w0 = state_from_overpressure(S.channel_overpressure(wl, wl.seed), gas)
    ic = (w0, np.full((N, N), gas.p0), np.full((N, N), gas.T0), np.zeros((N, N)))
    plain = P.DevicePatch(wl, gas)
    learned = P.DevicePatch(wl, gas)
    plain.init_state(*ic)
    learned.init_state(*ic)
    plain.advance_fused(2, dt)  # warm-up
    sw.advance(learned, 2, dt, lo, hi)
    plain.init_state(*ic)
    learned.init_state(*ic)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.advance_fused(steps, dt)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    cand, weno, undecided = sw.advance(learned, steps, dt, lo, hi)
    torch.cuda.synchronize()
    t_learned = time.perf_counter() - t0
    # without a switch: WENO on every cell (rhs_weno_high_llf), same steps
    allw = P.DevicePatch(wl, gas)
    allw.init_state(*ic)
    allw.advance_fused(1, dt, variant="weno_high_llf")
    allw.init_state(*ic)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    allw.advance_fused(steps, dt, variant="weno_high_llf")
    torch.cuda.synchronize()
    t_allweno = time.perf_counter() - t0
    del allw
    # decisions against the sensor on the final learned state
    S.P.fill_physical_bc(learned.box, learned.g, (1, 1, 1, 1), learned.reflect, learned.w, learned.p, learned.T)
    sens = S.sensor_map(learned)
    dec, _ = sw.decide(learned, lo, hi)
    missed = int(((sens == 1) & (dec == 0)).sum().item())
    false_weno = int(((sens == 0) & (dec == 1)).sum().item())
    sensor_weno = int((sens == 1).sum().item())
    g = wl.g
    err = {}
    names = ("rho", "rho_u", "rho_v", "rho_s", "rho_Tn")
    for k, name in enumerate(names):
        x = plain.w[k, g:-g, g:-g]
        y = learned.w[k, g:-g, g:-g]
        ref = x - x.mean() if k in (0, 3, 4) else x
        err[name] = {"rel_l2": float(torch.linalg.norm(x - y) / torch.linalg.norm(ref)),
                     "linf": float((x - y).abs().max()), "ref_linf": float(ref.abs().max())}
    cells = wl.n * wl.n
    out = {
        "workload": wl.name, "steps": steps, "cfl": wl.cfl, "thresholds": [lo, hi],
        "plain_sensor": {"value": cells * steps / t_plain, "ms_per_step": t_plain * 1e3 / steps},
        "learned_switch": {"value": cells * steps / t_learned, "ms_per_step": t_learned * 1e3 / steps},
        "weno_everywhere_no_switch": {"value": cells * steps / t_allweno, "ms_per_step": t_allweno * 1e3 / steps},
        "speedup_learned_vs_plain": t_plain / t_learned,
        "speedup_learned_vs_no_switch": t_allweno / t_learned,
        "decisions_summed_over_stages": {"candidates": cand, "network_weno": weno, "to_sensor": undecided,
                                         "network_drp": cand - weno - undecided,
                                         "prefilter_drp": 3 * steps * cells - cand},
        "final_state_vs_sensor": {"sensor_weno": sensor_weno, "missed_shocks": missed, "false_weno": false_weno},
        "error_vs_plain_hybrid": err,
        "unit": "cell-updates/s",
    }
    sw.close()
    return out


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from enlightning_amd import _lib
    from enlightning_amd import patch as P
    from enlightning_amd.gas import Gas
    from enlightning_amd.workloads import PULSES_4096, Workload, initial_fields

    gas = Gas()
    wl = PULSES_4096 if args.n == 4096 else Workload("pulses_%d_8" % args.n, n=args.n, seed=1, npulses=8)
    dt = wl.dt(gas)
    w, p, T, s = initial_fields(wl, gas)
    plain = P.DevicePatch(wl, gas)
    fused = P.DevicePatch(wl, gas)
    lib = _lib.load()
    nk = len(_lib.KERNEL_KINDS)

    def timed(patch, run):
        """Restart from the initial state, then time `steps` steps; per-kernel
        device times come from HIP events on the launch stream."""
        patch.init_state(w, p, T, s)
        torch.cuda.synchronize()
        _lib.check(lib.enl_profile_start(args.steps * 16))
        t0 = time.perf_counter()
        run(args.steps, dt)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        ms = (ctypes.c_double * nk)()
        cnt = (ctypes.c_int * nk)()
        _lib.check(lib.enl_profile_stop(ms, cnt, nk))
        per = {k: (ms[i] / cnt[i] if cnt[i] else None) for i, k in enumerate(_lib.KERNEL_KINDS)}
        return elapsed, per

    # warmup of both paths on the initial state
    for patch, run in ((plain, plain.advance), (fused, fused.advance_fused)):
        patch.init_state(w, p, T, s)
        if args.warmup:
            run(args.warmup, dt)
    torch.cuda.synchronize()

    t_plain, k_plain = timed(plain, plain.advance)
    t_fused, k_fused = timed(fused, fused.advance_fused)
    del w, p, T, s

    n, g = wl.n, wl.g
    same = all(torch.equal(a[..., g:-g, g:-g], b[..., g:-g, g:-g])
               for a, b in ((plain.w, fused.w), (plain.p, fused.p), (plain.T, fused.T)))

    # fraction of WENO cells in the final state (level-0 sensor)
    sm = torch.empty((n, n), dtype=torch.int32, device="cuda")
    KK = torch.empty((5, n, n), dtype=torch.float64, device="cuda")
    P.fill_physical_bc(plain.box, g, (1, 1, 1, 1), plain.reflect, plain.w, plain.p, plain.T, plain.src)
    P.rhs("hybrid_high_llf", plain.box, g, plain.xlo, plain.dx, plain.w, plain.src[g:-g, g:-g].contiguous(), KK,
          plain.p, plain.T, gas, 0, sm)
    weno_frac = float(sm.double().mean().item())
    finite = bool(torch.isfinite(plain.w).all().item())
    del KK, sm

    cells = wl.n * wl.n
    # plain drop-in: rhs is the dominant kernel
    rhs_bytes = RHS_BYTES_PER_CELL * cells
    rhs_achieved = rhs_bytes / (k_plain["rhs"] * 1e-3) / 1e9
    # fused: one stage kernel per stage; per cell it reads w_n(5), p, T,
    # source (+ ww(5) in stages 2 and 3, which is w_n itself in stage 1) and
    # writes w_n(5), p, T: 120 / 160 / 160 B, mean 146.7 B per launch-cell
    stage_bytes = FUSED_STAGE_BYTES_PER_CELL * cells
    stage_achieved = stage_bytes / (k_fused["stage_fused"] * 1e-3) / 1e9
    value = cells * args.steps / t_fused
    out = {
        "metric": "RK3 cell-updates/s on 1 MI355X, tuned (fused FP64) vs plain FP64 drop-in; %HBM peak",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_fused * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "grid": [wl.n, wl.n], "ghosts": wl.g, "rhs": "hybrid_high_llf",
                   "time_integrator": "ssp-rk3", "cfl": wl.cfl, "form": "fused_fp64_stage",
                   "walls": "reflecting"},
        "roofline": {"bound": "hbm", "kernel": "fused stage (rhs+runge_kutta+update_state, hybrid_high_llf)",
                     "achieved": stage_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": stage_achieved / HBM_PEAK_GBS, "traffic": read_traffic("stage_fused"),
                     "algorithmic_bytes_per_launch": stage_bytes},
        "speedup_vs_plain": t_plain / t_fused,
        "fused_bit_identical_to_plain": same,
        "kernel_ms_avg": k_fused,
        "plain_fp64_dropin": {
            "value": cells * args.steps / t_plain,
            "ms_per_step": t_plain * 1e3 / args.steps,
            "step_hbm_frac": STEP_BYTES_PER_CELL * cells / (t_plain / args.steps) / 1e9 / HBM_PEAK_GBS,
            "roofline": {"bound": "hbm", "kernel": "rhs_hybrid_high_llf", "achieved": rhs_achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": rhs_achieved / HBM_PEAK_GBS,
                         "traffic": read_traffic("rhs"), "algorithmic_bytes_per_launch": rhs_bytes},
            "kernel_ms_avg": k_plain,
        },
        "weno_fraction_final": weno_frac,
        "state_finite": finite,
        "cpu_baseline": None,
    }
    del plain, fused
    torch.cuda.empty_cache()
    if not args.no_amr:
        out["amr"], tags = amr_bench(args)
        torch.cuda.empty_cache()
        if args.precision_steps > 0:
            out["precision"] = precision_bench(args, tags)
            torch.cuda.empty_cache()
        if args.switch_steps > 0:
            out["learned_switch"] = switch_bench(args)
            torch.cuda.empty_cache()
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, gas, args.cpu_steps)
    print(json.dumps(out, default=lambda o: o.item()), flush=True)


if __name__ == "__main__":
    main()
