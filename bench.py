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
# the plain step as the reference's patch loop runs it: the time
# integrator's copy of the current state into the scratch state (w_n read
# and written, 80 B) and 3 stages x (rhs 104 + runge_kutta 160 +
# update_state 32)
STEP_BYTES_PER_CELL = 80.0 + 3 * (104.0 + 160.0 + 32.0)
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
    ap.add_argument("--regrid-cycles", type=int, default=5, help="regrid-every-2-steps cycles timed on the AMR workload")
    ap.add_argument("--precision-steps", type=int, default=2000)
    ap.add_argument("--switch-steps", type=int, default=100)
    ap.add_argument("--weno-steps", type=int, default=10,
                    help="steps of the all-WENO fused form timed for the FP64 VALU roofline")
    return ap.parse_args()


def valu_block(kernel, flops_per_launch, ms_per_launch, basis):
    """FP64 vector roofline of one kernel: algorithmic FLOPs per launch (the
    reference's arithmetic, enlightning_amd/flops.py) over its average launch
    duration."""
    achieved = flops_per_launch / (ms_per_launch * 1e-3) / 1e12
    return {"bound": "fp64_valu", "kernel": kernel, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "flops_per_launch": flops_per_launch,
            "basis": basis}


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


def cpu_baseline(wl, gas, steps, runs=5):
    import numpy as np
    import oracle
    from enlightning_amd.gas import RK3
    from enlightning_amd.workloads import initial_fields
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    # every core the process may use, within the OpenMP thread count the
    # host sets for this job (its CPU share)
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(cores, omp) if omp > 0 else cores)
    w, p, T, s = initial_fields(wl, gas)
    hp = oracle.HostPatch(wl, gas)
    times = []
    for _ in range(runs):
        hp.init_state(w, p, T, s)
        t0 = time.perf_counter()
        hp.advance(steps, wl.dt(gas), RK3, tile=64, nthreads=threads)
        times.append(time.perf_counter() - t0)
    del w, p, T, s
    dt = float(np.median(times))
    return {
        "value": wl.n * wl.n * steps / dt,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "runs": runs,
        "run_seconds": times,
        "sample": "median of %d runs of %d RK3 step(s) of %s from its initial state: the C restatement of the "
                  "reference kernels (oracle/enl_oracle.c, gcc -O2, no FMA), 64x64 tiles (the largest level-0 "
                  "patch of input.txt:114-116) over %d OpenMP threads (%d cores visible); %.2f s wall per run"
                  % (runs, steps, wl.name, threads, cores, dt),
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
    from enlightning_amd.gas import Gas
    from enlightning_amd.hierarchy import DeviceHierarchy, evolved_amr_workload
    from enlightning_amd.workloads import AMR_1024

    gas = Gas()
    wl = AMR_1024
    n = wl.n
    h, states, tags = evolved_amr_workload(wl, gas)
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
    # batch and the transfer on the device; timed after an untimed one that
    # fills the allocation cache and the pinned tag map
    warm = d.regrid(amr.build(n, 4, 4, 1.0, d.level0_tags(), tile0=64, previous=h))
    warm.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rtags = d.level0_tags()
    t1 = time.perf_counter()
    h2 = amr.build(n, 4, 4, 1.0, rtags, tile0=64, previous=h)
    tr = amr.transfer_rects(h2, h)
    t2 = time.perf_counter()
    d2 = d.regrid(h2, tr)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    d2.close()
    rc = tr["rects"]
    regrid = {"tags_device_ms": (t1 - t0) * 1e3, "cluster_and_schedules_host_ms": (t2 - t1) * 1e3,
              "new_batch_and_transfer_device_ms": (t3 - t2) * 1e3, "total_ms": (t3 - t0) * 1e3,
              "fine_patches_after": len(h2.level_patches(1)), "unchanged_patches": int(len(tr["same"][0])),
              "transfer_rectangles": int(len(rc)),
              "copied_cells": int((rc[:, 4] * rc[:, 5])[rc[:, 6] == 0].sum()),
              "refined_cells": int((rc[:, 4] * rc[:, 5])[rc[:, 6] == 1].sum()),
              "planner": "native (enl_plan_build / enl_plan_transfer_rects), schedules as rectangles"}
    # the reference's cadence (input.txt:7, main.C:489-502): two graph steps,
    # then a regrid (device tags, native plan, new batch, transfer; the new
    # hierarchy's graph is captured on its first step), repeated
    d.load_initial(wl, states)
    cur, hcur, cells_done = d, h, 0

    def cycle(cur, hcur):
        cur.advance(2, dt, mode="graph")
        hn = amr.build(n, 4, 4, 1.0, cur.level0_tags(), tile0=64, previous=hcur)
        nxt = cur.regrid(hn)
        if cur is not d:
            cur.close()
        return nxt, hn

    cur, hcur = cycle(cur, hcur)  # one untimed cycle
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.regrid_cycles):
        cells_done += 2 * hcur.interior_cells()
        cur, hcur = cycle(cur, hcur)
    torch.cuda.synchronize()
    t_cyc = time.perf_counter() - t0
    if cur is not d:
        cur.close()
    regrid_every_2 = {"cycles": args.regrid_cycles, "ms_per_step": t_cyc * 1e3 / (2 * args.regrid_cycles),
                      "value": cells_done / t_cyc, "unit": "cell-updates/s",
                      "basis": "2 graph steps + 1 regrid per cycle (after one untimed cycle), the regrid's time "
                               "included: device tags, native plan, new batch and transfer, the new hierarchy's "
                               "tiles and graph capture on its first step"}
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
        "packing_per_level": [d.packing_name(l) for l in sorted(best)],
        "graph": {"value": cells * args.amr_steps / t_graph, "ms_per_step": t_graph * 1e3 / args.amr_steps,
                  "steps": args.amr_steps},
        "plain_dropin_per_patch": {"value": cells * args.amr_plain_steps / t_plain,
                                   "ms_per_step": t_plain * 1e3 / args.amr_plain_steps,
                                   "steps": args.amr_plain_steps},
        "speedup_graph_vs_plain": (t_plain / args.amr_plain_steps) / (t_graph / args.amr_steps),
        "graph_bit_identical_to_plain": same,
        "regrid": regrid,
        "with_regrid_every_2_steps": regrid_every_2,
        "unit": "cell-updates/s",
    }, tags


def precision_bench(args, tags):
    """The FP32-perturbation form on the same hierarchy (the boxes of the
    AMR workload, on a 2 km x 2 km domain): 16 weak isentropic pulses on an
    exponentially stratified ambient (scale height 8 km), FP64 batched graph
    against FP32 batched graph from the same state, per-field error; then the
    plain FP64 drop-in patch by patch and the tuned form (FP32 perturbations
    with the learned switch on level 0) on the same hierarchy and state."""
    import torch
    from enlightning_amd import switch as S
    from enlightning_amd.gas import Gas
    from enlightning_amd.hierarchy import DeviceHierarchy, perturbation_errors, weak_acoustic_workload
    from enlightning_amd.workloads import WEAK_2KM

    gas = Gas()
    wl = WEAK_2KM
    h, states, ambs, ambient = weak_acoustic_workload(wl, gas, tags)
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
    out = {
        "workload": wl.name,
        "hierarchy": "the amr workload's boxes on a %g m domain" % wl.L,
        "patches_per_level": [len(h.level_patches(l)) for l in range(h.nlevels)],
        "cells_per_level": [h.interior_cells(l) for l in range(h.nlevels)],
        "steps": steps,
        "fp64_graph": {"value": cells * steps / t64, "ms_per_step": t64 * 1e3 / steps,
                       "tiles_per_level": ["%dx%d" % tiles64[l] for l in sorted(tiles64)]},
        "fp32_perturbation_graph": {"value": cells * steps / t32, "ms_per_step": t32 * 1e3 / steps,
                                    "tiles_per_level": ["%dx%d" % tiles32[l] for l in sorted(tiles32)]},
        "speedup_fp32_vs_fp64": t64 / t32,
        "error_fp32_vs_fp64": perturbation_errors(h, s64, s32, ambs, gas),
        "unit": "cell-updates/s",
    }
    # the plain FP64 drop-in, patch by patch, on the same hierarchy and state
    pst = args.amr_plain_steps
    d.load_initial(wl, states)
    d.advance(1, dt, mode="plain")
    d.load_initial(wl, states)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.advance(pst, dt, mode="plain")
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    out["plain_fp64_dropin_per_patch"] = {"value": cells * pst / t_plain, "ms_per_step": t_plain * 1e3 / pst,
                                          "steps": pst}
    # the tuned form: FP32 perturbations + the learned switch on level 0,
    # decisions before every RK stage, graph-replayed
    params = S.load_params()
    if params is None:
        out["fp32_switch_graph"] = {"skipped": "no trained parameters"}
        d.close()
        return out
    sw = S.Switch(params)
    # the network's DRP calls off (lo = 0): next to the reflecting walls the
    # stratified ambient's own adjustment makes sensor-WENO cells unlike the
    # channel shocks it was trained on; its WENO calls stand, every other
    # candidate goes to the sensor
    lo, hi = 0.0, S.load_thresholds()[1]
    totals = torch.zeros(3, dtype=torch.int64, device="cuda")
    d.load_initial(wl, states)
    d.fp32_set_switch(sw, lo, hi, every_stage=True, totals=totals)
    d.fp32_load()
    d.advance(1, dt, mode="fp32_graph")  # capture
    d.load_initial(wl, states)
    d.fp32_load()
    totals.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.advance(steps, dt, mode="fp32_graph")
    torch.cuda.synchronize()
    tc = time.perf_counter() - t0
    d.fp32_store()
    sc = d.states()
    cand, weno, undecided = (int(v) for v in totals.cpu())
    d.fp32_set_switch(None)
    d.close()
    sw.close()
    err = perturbation_errors(h, s64, sc, ambs, gas)
    worst = max(e["rel_l2_of_perturbation"] for e in err.values() if e["rel_l2_of_perturbation"] is not None)
    out["fp32_switch_graph"] = {
        "value": cells * steps / tc, "ms_per_step": tc * 1e3 / steps, "thresholds": [lo, hi],
        "decisions": "before every RK stage",
        "decisions_summed_over_stages": {"level0_cells_per_stage": h.interior_cells(0), "candidates": cand,
                                         "network_weno": weno, "to_sensor": undecided},
        "speedup_vs_plain_dropin": (t_plain / pst) / (tc / steps),
        "speedup_vs_fp64_graph": t64 / tc,
        "error_vs_fp64": err,
        "worst_rel_l2_of_perturbation": worst,
    }
    return out


def switch_bench(args):
    """The learned switch on the tortuous line source (4096^2 base at 0.25 m,
    held-out channel seed): the plain hybrid (density sensor) against the
    hybrid with the trained network deciding level 0's cells (p > hi WENO,
    p < lo DRP, else the sensor; lo calibrated on other held-out channels),
    decisions once per step (and, timed beside it, before every stage), both
    fused FP64 from the same state; the plain solution judges the learned
    one's error, and the network's calls are compared with the sensor's on
    the same states (missed shocks: sensor WENO, network confident DRP)."""
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
    lo, hi = S.load_thresholds()
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
    # decisions before every stage
    learned.init_state(*ic)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cand1, weno1, undecided1 = sw.advance(learned, steps, dt, lo, hi, every_stage=True)
    torch.cuda.synchronize()
    t_every = time.perf_counter() - t0
    plain.init_state(*ic)
    learned.init_state(*ic)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.advance_fused(steps, dt)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    # decisions once per step, kept for its three stages (the reported form)
    t0 = time.perf_counter()
    cand, weno, undecided = sw.advance(learned, steps, dt, lo, hi, every_stage=False)
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
        "learned_switch": {"value": cells * steps / t_learned, "ms_per_step": t_learned * 1e3 / steps,
                           "decisions": "once per step, kept for its three stages"},
        "learned_switch_every_stage": {"value": cells * steps / t_every, "ms_per_step": t_every * 1e3 / steps,
                                       "candidates": cand1, "network_weno": weno1, "to_sensor": undecided1},
        "cost_learned_vs_plain": t_learned / t_plain,
        "weno_everywhere_no_switch": {"value": cells * steps / t_allweno, "ms_per_step": t_allweno * 1e3 / steps},
        "speedup_learned_vs_plain": t_plain / t_learned,
        "speedup_learned_vs_no_switch": t_allweno / t_learned,
        "decisions_summed_over_steps": {"candidates": cand, "network_weno": weno, "to_sensor": undecided,
                                        "network_drp": cand - weno - undecided,
                                        "prefilter_drp": steps * cells - cand},
        "final_state_vs_sensor": {"sensor_weno": sensor_weno, "missed_shocks": missed, "false_weno": false_weno},
        "error_vs_plain_hybrid": err,
        "unit": "cell-updates/s",
    }
    sw.close()
    return out


def channel_amr_bench(args):
    """The tortuous line source on its two-level hierarchy: the 4096^2 base
    at 0.25 m (held-out channel seed), one level refined by 4 over
    tag_cells' tags of the initial state, fine patches 8-64 cells a side.
    The FP64 batched graph with the density sensor against the tuned form
    (FP32 perturbations of the uniform ambient, the learned switch on level
    0, decisions before every stage) from the same state, plus the plain FP64
    drop-in patch by patch; per-field error of the tuned form against FP64,
    and the network's calls against the sensor's on the final level-0
    state."""
    import numpy as np
    import torch
    from enlightning_amd import amr
    from enlightning_amd import patch as P
    from enlightning_amd import switch as S
    from enlightning_amd.gas import Gas
    from enlightning_amd.hierarchy import DeviceHierarchy, perturbation_errors, refine_from_base
    from enlightning_amd.workloads import StratifiedWorkload, ambient_rows, state_from_overpressure

    params = S.load_params()
    if params is None:
        return {"skipped": "no trained parameters"}
    gas = Gas()
    wl = S.CHANNEL_4096
    n, g = wl.n, wl.g
    N = n + 2 * g
    wb = np.ascontiguousarray(state_from_overpressure(S.channel_overpressure(wl, wl.seed), gas))
    tags = np.zeros((n, n), np.int32)
    P.tag_cells((0, n - 1, 0, n - 1), g, tags, wb, np.zeros((N, N)), 1, gas.tag_tolerance)
    h = amr.build(n, 4, g, wl.L, tags, tile0=64)
    states = [refine_from_base(h, p, wb) for p in h.patches]
    del wb
    uniform = StratifiedWorkload("uniform", n=n, L=wl.L, scale_height=float("inf"))

    def ambient(lev):
        nl = h.ncells(lev)
        y = (np.arange(-g, nl + g) + 0.5) * h.dx(lev)
        return (-g,) + ambient_rows(uniform, gas, y)

    ambs = []
    for p in h.patches:
        y = (np.arange(p.box[2] - p.g, p.box[3] + p.g + 1) + 0.5) * h.dx(p.level)
        ambs.append(ambient_rows(uniform, gas, y))
    dt = wl.cfl * h.dx(h.nlevels - 1) / gas.c  # one dt for all levels (enlightning.C:446-447)
    steps = args.switch_steps
    cells = h.interior_cells()
    d = DeviceHierarchy(h, gas)

    def timed(mode, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.advance(k, dt, mode=mode)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    d.load_initial(wl, states)
    d.advance(1, dt, mode="graph")  # capture
    d.load_initial(wl, states)
    t64 = timed("graph", steps)
    s64 = d.states()
    pst = max(1, args.amr_plain_steps // 2)
    d.load_initial(wl, states)
    t_plain = timed("plain", pst)
    sw = S.Switch(params)
    lo, hi = S.load_thresholds()
    totals = torch.zeros(3, dtype=torch.int64, device="cuda")
    d.load_initial(wl, states)
    d.set_ambient(ambient)
    d.fp32_set_switch(sw, lo, hi, every_stage=True, totals=totals)
    d.fp32_prepare()
    d.advance(1, dt, mode="fp32_graph")  # capture
    d.load_initial(wl, states)
    d.fp32_load()
    totals.zero_()
    tc = timed("fp32_graph", steps)
    d.fp32_store()
    sc = d.states()
    cand, weno, undecided = (int(v) for v in totals.cpu())
    d.fp32_set_switch(None)
    d.close()
    del d
    torch.cuda.empty_cache()
    err = perturbation_errors(h, s64, sc, ambs, gas)
    # the network's calls against the sensor's on the final level-0 state,
    # assembled into one base patch
    base = P.DevicePatch(wl, gas)
    wfull = np.zeros((5, N, N))
    for p, st in zip(h.patches, sc):
        if p.level == 0:
            wfull[:, p.box[2] + g:p.box[3] + g + 1, p.box[0] + g:p.box[1] + g + 1] = st[0][:, g:-g, g:-g]
    base.init_state(wfull, np.full((N, N), gas.p0), np.full((N, N), gas.T0), np.zeros((N, N)))
    del wfull
    sens = S.sensor_map(base)
    dec, _ = sw.decide(base, lo, hi)
    missed = int(((sens == 1) & (dec == 0)).sum().item())
    false_weno = int(((sens == 0) & (dec == 1)).sum().item())
    sensor_weno = int((sens == 1).sum().item())
    del base, sens, dec
    sw.close()
    torch.cuda.empty_cache()
    sizes = [max(p.nx, p.ny) for p in h.level_patches(1)]
    worst = max(e["rel_l2_of_perturbation"] for e in err.values() if e["rel_l2_of_perturbation"] is not None)
    return {
        "workload": wl.name + "_2level", "seed": wl.seed, "steps": steps, "dt": dt,
        "patches_per_level": [len(h.level_patches(l)) for l in range(h.nlevels)],
        "cells_per_level": [h.interior_cells(l) for l in range(h.nlevels)],
        "fine_patch_side_min_max": [int(min(sizes)), int(max(sizes))] if sizes else None,
        "fp64_graph_sensor": {"value": cells * steps / t64, "ms_per_step": t64 * 1e3 / steps},
        "plain_fp64_dropin_per_patch": {"value": cells * pst / t_plain, "ms_per_step": t_plain * 1e3 / pst,
                                        "steps": pst},
        "tuned_fp32_switch_graph": {"value": cells * steps / tc, "ms_per_step": tc * 1e3 / steps,
                                    "thresholds": [lo, hi], "decisions": "before every RK stage"},
        "speedup_tuned_vs_plain_dropin": (t_plain / pst) / (tc / steps),
        "speedup_tuned_vs_fp64_graph": t64 / tc,
        "decisions_summed_over_stages": {"level0_cells_per_stage": h.interior_cells(0), "candidates": cand,
                                         "network_weno": weno, "to_sensor": undecided,
                                         "network_drp": cand - weno - undecided},
        "final_level0_state_vs_sensor": {"sensor_weno": sensor_weno, "missed_shocks": missed,
                                         "false_weno": false_weno},
        "error_tuned_vs_fp64": err,
        "worst_rel_l2_of_perturbation": worst,
        "unit": "cell-updates/s",
    }


def weno_dense_bench(args, patch, wl, gas, dt):
    """The WENO path's FP64 roofline: the same patch advanced by the fused
    form of rhs_weno_high_llf (every cell WENO5 + LLF, the dense tile kernel
    with runge_kutta and update_state fused) from the workload's initial
    state; the stage kernel's average duration from HIP events."""
    import torch
    from enlightning_amd import _lib
    from enlightning_amd import flops as FL
    from enlightning_amd.workloads import initial_fields
    lib = _lib.load()
    nk = len(_lib.KERNEL_KINDS)
    patch.init_state(*initial_fields(wl, gas))
    patch.advance_fused(1, dt, variant="weno_high_llf")
    torch.cuda.synchronize()
    _lib.check(lib.enl_profile_start(args.weno_steps * 16))
    t0 = time.perf_counter()
    patch.advance_fused(args.weno_steps, dt, variant="weno_high_llf")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = (ctypes.c_double * nk)()
    cnt = (ctypes.c_int * nk)()
    _lib.check(lib.enl_profile_stop(ms, cnt, nk))
    k = _lib.KERNEL_KINDS.index("stage_fused")
    stage_ms = ms[k] / cnt[k]
    cells = wl.n * wl.n
    return {
        "workload": wl.name, "rhs": "weno_high_llf", "form": "fused_fp64_stage", "steps": args.weno_steps,
        "value": cells * args.weno_steps / elapsed, "unit": "cell-updates/s",
        "ms_per_step": elapsed * 1e3 / args.weno_steps, "stage_kernel_ms": stage_ms,
        "valu": valu_block("fused stage, dense WENO tile kernel", FL.stage_flops("weno_high_llf", 0, 0, cells),
                           stage_ms, "every cell WENO5 + LLF (3740 reference operations) + runge_kutta (30) + "
                                     "update_state (23)"),
        "hbm": {"achieved": FUSED_STAGE_BYTES_PER_CELL * cells / (stage_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s"},
    }


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

    def weno_fraction(patch):
        """Fraction of the patch's cells the level-0 sensor sends to WENO."""
        sm = torch.empty((wl.n, wl.n), dtype=torch.int32, device="cuda")
        KK = torch.empty((5, wl.n, wl.n), dtype=torch.float64, device="cuda")
        g_ = wl.g
        P.fill_physical_bc(patch.box, g_, (1, 1, 1, 1), patch.reflect, patch.w, patch.p, patch.T, patch.src)
        P.rhs("hybrid_high_llf", patch.box, g_, patch.xlo, patch.dx, patch.w, patch.src[g_:-g_, g_:-g_].contiguous(),
              KK, patch.p, patch.T, gas, 0, sm)
        return float(sm.double().mean().item())

    plain.init_state(w, p, T, s)
    weno_start = weno_fraction(plain)

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
    weno_frac = weno_fraction(plain)
    finite = bool(torch.isfinite(plain.w).all().item())

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
    # FP64 work per launch: the reference's operations per cell, with the
    # WENO share taken as the mean of the first and last states' sensor maps
    from enlightning_amd import flops as FL
    wf = 0.5 * (weno_start + weno_frac)
    n_weno = wf * cells
    basis = ("hybrid_high_llf level 0, %.3g%% WENO cells (mean of the first and last states' sensor maps); "
             "reference operations per cell from the counting restatement" % (100 * wf))
    valu_stage = valu_block("fused stage", FL.stage_flops("hybrid_high_llf", 0, cells - n_weno, n_weno),
                            k_fused["stage_fused"], basis)
    valu_rhs = valu_block("rhs_hybrid_high_llf", FL.rhs_flops("hybrid_high_llf", 0, cells - n_weno, n_weno),
                          k_plain["rhs"], basis)
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
        "valu": valu_stage,
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
            "valu": valu_rhs,
            "kernel_ms_avg": k_plain,
        },
        "weno_fraction_start": weno_start,
        "weno_fraction_final": weno_frac,
        "state_finite": finite,
        "cpu_baseline": None,
    }
    del plain
    if args.weno_steps > 0:
        out["weno_dense"] = weno_dense_bench(args, fused, wl, gas, dt)
    del fused
    torch.cuda.empty_cache()
    if not args.no_amr:
        out["amr"], tags = amr_bench(args)
        torch.cuda.empty_cache()
        if args.precision_steps > 0:
            out["precision"] = pr = precision_bench(args, tags)
            torch.cuda.empty_cache()
            c = pr.get("fp32_switch_graph", {})
            if "value" in c:
                out["tuned_form_on_amr"] = {
                    "form": "batched FP32-perturbation + learned switch, hipGraph",
                    "workload": pr["workload"], "hierarchy_cells_per_level": pr["cells_per_level"],
                    "value": c["value"], "unit": "cell-updates/s",
                    "plain_fp64_dropin_value": pr["plain_fp64_dropin_per_patch"]["value"],
                    "speedup_vs_plain_dropin": c["speedup_vs_plain_dropin"],
                    "worst_rel_l2_of_perturbation": c["worst_rel_l2_of_perturbation"],
                    "steps": pr["steps"]}
        if args.switch_steps > 0:
            out["learned_switch"] = switch_bench(args)
            torch.cuda.empty_cache()
            out["learned_switch_amr"] = channel_amr_bench(args)
            torch.cuda.empty_cache()
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, gas, args.cpu_steps)
    print(json.dumps(out, default=lambda o: o.item()), flush=True)


if __name__ == "__main__":
    main()
