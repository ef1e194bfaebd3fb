"""Trains the learned WENO/DRP switch on the device and writes its
parameters (enlightning_amd/data/switch_params.npy, or --out) and a JSON
report of the training and of a held-out evaluation.

Training set: plain-sensor maps of 50 seeded tortuous channels on the
4096^2 / 0.25 m base (seeds 0..49), at steps 0, 20 and 60 of the plain FP64
fused advance; per snapshot every sensor-WENO candidate cell (up to half the
quota) and as many sensor-DRP candidates, classes weighted equally.
Validation: channels 50..54, same snapshots.

    python -m enlightning_amd.train_switch [--channels 50] [--out PATH]
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch


def main(argv=None):
    from enlightning_amd import switch as S
    from enlightning_amd.gas import Gas

    ap = argparse.ArgumentParser()
    ap.add_argument("--channels", type=int, default=50)
    ap.add_argument("--val-channels", type=int, default=5)
    ap.add_argument("--per-snapshot", type=int, default=40000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=S.PARAMS_FILE)
    ap.add_argument("--report", default=None)
    ap.add_argument("--calibrate", action="store_true",
                    help="only calibrate the DRP threshold of the shipped parameters on held-out channels")
    ap.add_argument("--cal-seeds", default="1001-1032", help="held-out channels the threshold is set on")
    ap.add_argument("--check-seeds", default="1033-1064", help="held-out channels it is then checked on")
    ap.add_argument("--hi", type=float, default=0.8)
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    gas = Gas()
    wl = S.CHANNEL_4096
    if args.calibrate:
        return calibrate(args, S, gas, wl, log)
    sw = S.Switch(S.init_params(0))
    t0 = time.time()
    X, y, w = S.build_samples(sw, wl, gas, seeds=range(args.channels), snapshots=(0, 20, 60),
                              per_snapshot=args.per_snapshot)
    log("training rows %d (WENO %d) in %.1f s" % (y.numel(), int(y.sum().item()), time.time() - t0))
    Xv, yv, wv = S.build_samples(sw, wl, gas, seeds=range(args.channels, args.channels + args.val_channels),
                                 snapshots=(0, 20, 60), per_snapshot=args.per_snapshot, rng_seed=1)
    untrained = S.evaluate(sw, Xv, yv)
    torch.cuda.synchronize()
    t1 = time.time()
    losses = S.train(sw, X, y, w, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    torch.cuda.synchronize()
    train_s = time.time() - t1
    ev = S.evaluate(sw, Xv, yv)
    log("validation: %s" % ev)
    params = sw.params()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, params)
    report = {
        "workload": wl.name, "channels": args.channels, "snapshots": [0, 20, 60],
        "rows": int(y.numel()), "weno_rows": int(y.sum().item()), "epochs": args.epochs, "batch": args.batch,
        "lr": args.lr, "losses": losses, "train_seconds": train_s,
        "train_steps_per_s": (y.numel() // args.batch) * args.epochs / train_s,
        "validation_untrained": untrained, "validation": ev,
    }
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    log(json.dumps(report))
    sw.close()
    return 0


def _seeds(spec):
    a, b = (int(v) for v in spec.split("-"))
    return range(a, b + 1)


def calibrate(args, S, gas, wl, log):
    """lo := half the smallest p(WENO) the shipped network gives any
    sensor-WENO cell of the calibration channels (steps 0, 20, 60, 100), so
    that none of them is called DRP; then the same counts on other held-out
    channels.  Writes enlightning_amd/data/switch_thresholds.json."""
    params = S.load_params()
    if params is None:
        log("no shipped parameters")
        return 1
    sw = S.Switch(params)
    t0 = time.time()
    pw, pd, visited = S.channel_probabilities(sw, wl, gas, _seeds(args.cal_seeds))
    lo = S.calibrate_lo(pw)
    cal = S.decision_shares(pw, pd, lo, args.hi)
    cal.update({"seeds": args.cal_seeds, "level0_cells_visited": visited,
                "p_weno_min": float(pw.min()) if pw.size else None,
                "p_weno_quantiles_0.001_0.01_0.1": [float(q) for q in np.quantile(pw, [1e-3, 1e-2, 1e-1])]})
    log("calibration: %s" % cal)
    pw2, pd2, visited2 = S.channel_probabilities(sw, wl, gas, _seeds(args.check_seeds))
    chk = S.decision_shares(pw2, pd2, lo, args.hi)
    chk.update({"seeds": args.check_seeds, "level0_cells_visited": visited2})
    old = S.decision_shares(pw2, pd2, 0.2, args.hi)
    log("check: %s" % chk)
    out = {"lo": lo, "hi": args.hi, "snapshots": [0, 20, 60, 100], "calibration": cal, "check": chk,
           "check_at_lo_0.2": old, "seconds": time.time() - t0}
    with open(S.THRESHOLDS_FILE, "w") as f:
        json.dump(out, f, indent=1)
    log(json.dumps(out))
    sw.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
