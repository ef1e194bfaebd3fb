"""Restart and plot output of a device hierarchy.

* Restart: the reference's checkpoint is enlightning::putToRestart /
  getFromRestart (enlightning.C:1859-2060) for the model scalars, plus
  SAMRAI's patch data of the SOLN variables (w_n, pressure, temperature,
  source; enlightning.C:259-281).  Here one NumPy .npz archive with no
  pickled objects holds the base grid (n0, ratio, g, L), the boxes as
  (level, ifirst0, ilast0, ifirst1, ilast1) rows in pid order, every patch's
  ghosted w_n, pres, temp and source, and the scalars under the reference's
  restart keys as JSON text.  Resuming rebuilds the schedules from the boxes
  and continues bit for bit (the ghosts are refilled at every stage).
* Plot quantities: the names and order registerModelVariables registers
  with the VisIt writer (enlightning.C:295-378): "w_n #0" .. "w_n #4",
  "pressure", "temperature", "source", each switched by its viz_* flag
  (input.txt:9-12), as per-patch interior arrays.
"""
import json

import numpy as np

from . import amr
from .gas import Gas

# reference restart key -> Gas field (enlightning.C:1913-1933)
GAS_KEYS = {
    "p_p0": "p0", "p_T0": "T0", "p_c_lr": "c_lr", "p_rho0": "rho0", "p_c": "c", "p_gamma1": "gamma1",
    "p_RH": "RH", "p_R": "R", "p_c_p": "c_p", "p_T_star_n": "T_star_n", "p_T_star_o": "T_star_o", "p_mu": "mu",
    "p_muB": "muB", "p_kappa": "kappa", "p_weno_alpha": "weno_alpha", "p_tau_n": "tau_n", "p_tau_o": "tau_o",
    "p_R_tilde": "R_tilde", "p_c_v": "c_v", "p_M": "M",
}
BC_KEYS = ("p_bc_type_L", "p_bc_type_R", "p_bc_type_B", "p_bc_type_T")
BC_INT = {"REFLECT": 10, "ABSORB": 20}  # enlightning.C:64-65


def restart_scalars(gas, reflect, absorb_width, g, iteration=0, loop_time=0.0, **extra):
    """The model scalars under the reference's restart keys (the subset this
    path carries; `extra` adds any other key, e.g. p_cfl or p_regrid_step)."""
    d = {k: float(getattr(gas, a)) for k, a in GAS_KEYS.items()}
    d["p_tag_tolerance"] = [float(t) for t in gas.tag_tolerance]
    for key, r in zip(BC_KEYS, reflect):
        name = "REFLECT" if r else "ABSORB"
        d[key] = name
        d[key + "_int"] = BC_INT[name]
    d["p_bc_absorb_width"] = float(absorb_width)
    d["p_nghosts"] = [int(g), int(g)]
    d["p_iteration_number"] = int(iteration)
    d["p_loop_time"] = float(loop_time)
    d.update(extra)
    return d


def gas_from_scalars(d):
    """The Gas a restart was written with (every constant as saved)."""
    kw = {a: d[k] for k, a in GAS_KEYS.items()}
    return Gas(tag_tolerance=tuple(d["p_tag_tolerance"]), **kw)


def reflect_from_scalars(d):
    return tuple(int(d[k + "_int"] == BC_INT["REFLECT"]) for k in BC_KEYS)


def save_restart(path, hier, iteration=0, loop_time=0.0, **extra):
    """Checkpoint `hier` (a DeviceHierarchy, or anything with .h, .gas,
    .reflect, .absorb_width, states() and sources()) into `path` (.npz).
    The FP64 state is what is saved: after FP32-form steps call
    hier.fp32_store() first."""
    h = hier.h
    arrays = {
        "grid": np.array([h.n0, h.ratio, h.g], np.int64),
        "L": np.array(h.L, np.float64),
        "boxes": np.array([[p.level, *p.box] for p in h.patches], np.int64).reshape(-1, 5),
        "scalars": np.array(json.dumps(restart_scalars(hier.gas, hier.reflect, hier.absorb_width, h.g, iteration,
                                                       loop_time, **extra))),
    }
    for k, ((w, pr, T), src) in enumerate(zip(hier.states(), hier.sources())):
        arrays["w_%d" % k], arrays["p_%d" % k], arrays["T_%d" % k], arrays["s_%d" % k] = w, pr, T, src
    np.savez(path, **arrays)


def load_restart(path, factory=None):
    """(hierarchy, scalars) from a checkpoint: the hierarchy rebuilt from the
    saved boxes by factory(h, gas, reflect) (default DeviceHierarchy), its
    state and source uploaded and ghost-filled, absorbing envelopes set as
    saved."""
    if factory is None:
        from .hierarchy import DeviceHierarchy as factory
    with np.load(path, allow_pickle=False) as z:
        n0, ratio, g = (int(v) for v in z["grid"])
        d = json.loads(str(z["scalars"]))
        h = amr.from_boxes(n0, ratio, g, float(z["L"]), z["boxes"])
        states = [(z["w_%d" % k], z["p_%d" % k], z["T_%d" % k]) for k in range(len(h.patches))]
        sources = [z["s_%d" % k] for k in range(len(h.patches))]
    hier = factory(h, gas_from_scalars(d), reflect_from_scalars(d))
    hier.upload_all(states, sources)
    if d["p_bc_absorb_width"] > 0:
        hier.set_absorb(d["p_bc_absorb_width"])
    return hier, d


def plot_quantities(hier, viz_w_n=0, viz_pressure=1, viz_temperature=0, viz_source=0):
    """{name: [interior array of each patch]} in registerModelVariables'
    order (enlightning.C:295-378); defaults are input.txt:9-12.  A device
    hierarchy gathers every selected plane of every patch in one launch
    (enl_batch_gather)."""
    if hasattr(hier, "gather"):
        planes = (0x1F if viz_w_n else 0) | (0x20 if viz_pressure else 0) | (0x40 if viz_temperature else 0) | (
            0x80 if viz_source else 0)
        out = {}
        if planes:
            names = [n for k, n in enumerate(hier.PLANES) if (planes >> k) & 1]
            per = hier.gather(planes, ghosted=False)
            for q, name in enumerate(names):
                out[name] = [v[q] for v in per]
        return out
    out = {}
    states = hier.states()
    g = [p.g for p in hier.h.patches]
    if viz_w_n:
        for n in range(5):
            out["w_n #%01d" % n] = [w[n, k:-k, k:-k] for (w, _, _), k in zip(states, g)]
    if viz_pressure:
        out["pressure"] = [pr[k:-k, k:-k] for (_, pr, _), k in zip(states, g)]
    if viz_temperature:
        out["temperature"] = [T[k:-k, k:-k] for (_, _, T), k in zip(states, g)]
    if viz_source:
        out["source"] = [s[k:-k, k:-k] for s, k in zip(hier.sources(), g)]
    return out
