"""Throughput of the per-patch update path on one MI355X.

Workload (one step = one SSP-RK3 step of the whole patch):
  pulses_4096_8 -- one 4096 x 4096 patch on [0,1]^2 m, 4 ghost cells,
  reflecting walls, ideal gas gamma = 1.4, rho0 = 1.2, p0 = 101325, eight
  seeded Gaussian overpressure pulses (dp/p0 in [0.05, 0.8], sigma in
  [0.005, 0.03] m), hybrid WENO5/DRP with LLF splitting (rhs_hybrid_high_llf),
  CFL 0.5, FP64: the plain drop-in (rhs -> runge_kutta -> update_state per
  stage, SAMRAI-layout arrays resident in HBM).

Prints one JSON line.  value = interior cell-updates per second over the
timed steps (inputs resident in HBM).  roofline: the rhs kernel's algorithmic
HBM bytes (104 B per interior cell: w_n, p, T, source read once, rhs written
once) over its average launch duration, timed with HIP events on its launch
stream during the timed region.  cpu_baseline: the CPU restatement of the
same kernels (oracle/) on this host's cores, over a bounded sample.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096, help="interior cells per side")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    return ap.parse_args()


def read_traffic():
    """HBM bytes per rhs launch from the committed rocprofv3 counter summary
    (the line `traffic_bytes_per_launch: <bytes>`), or None."""
    path = os.path.join(ROOT, "profiles", "rhs_hybrid_high_llf_4096_traffic.txt")
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
    d = P.DevicePatch(wl, gas)

    # warmup on the initial state, then restart from it for the timed steps
    d.init_state(w, p, T, s)
    if args.warmup:
        d.advance(args.warmup, dt)
    d.init_state(w, p, T, s)
    del w, p, T, s
    torch.cuda.synchronize()

    lib = _lib.load()
    nk = len(_lib.KERNEL_KINDS)
    rc = lib.enl_profile_start(args.steps * 16)
    _lib.check(rc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.advance(args.steps, dt)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = (ctypes.c_double * nk)()
    cnt = (ctypes.c_int * nk)()
    rc = lib.enl_profile_stop(ms, cnt, nk)
    _lib.check(rc)
    kernel_ms = {k: (ms[i] / cnt[i] if cnt[i] else None) for i, k in enumerate(_lib.KERNEL_KINDS)}

    # fraction of WENO cells in the final state (level-0 sensor)
    n, g = wl.n, wl.g
    sm = torch.empty((n, n), dtype=torch.int32, device="cuda")
    KK = torch.empty((5, n, n), dtype=torch.float64, device="cuda")
    P.fill_physical_bc(d.box, g, (1, 1, 1, 1), d.reflect, d.w, d.p, d.T, d.src)
    P.rhs("hybrid_high_llf", d.box, g, d.xlo, d.dx, d.w, d.src[g:-g, g:-g].contiguous(), KK, d.p, d.T, gas, 0, sm)
    weno_frac = float(sm.double().mean().item())
    finite = bool(torch.isfinite(d.w).all().item())
    del KK, sm

    cells = wl.n * wl.n
    value = cells * args.steps / elapsed
    rhs_ms = kernel_ms["rhs"]
    rhs_bytes = RHS_BYTES_PER_CELL * cells
    achieved = rhs_bytes / (rhs_ms * 1e-3) / 1e9
    out = {
        "metric": "RK3 cell-updates/s on 1 MI355X, plain FP64 drop-in (hybrid WENO5/DRP LLF, 4096^2)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "grid": [wl.n, wl.n], "ghosts": wl.g, "rhs": "hybrid_high_llf",
                   "time_integrator": "ssp-rk3", "cfl": wl.cfl, "form": "plain_fp64_dropin",
                   "walls": "reflecting"},
        "roofline": {"bound": "hbm", "kernel": "rhs_hybrid_high_llf", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(),
                     "algorithmic_bytes_per_launch": rhs_bytes},
        "step_hbm_frac": STEP_BYTES_PER_CELL * cells / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS,
        "kernel_ms_avg": kernel_ms,
        "weno_fraction_final": weno_frac,
        "state_finite": finite,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        del d
        torch.cuda.empty_cache()
        out["cpu_baseline"] = cpu_baseline(wl, gas, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
