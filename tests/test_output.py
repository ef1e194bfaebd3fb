"""Restart and plot output (enlightning_amd.output) on the CPU.

* A hierarchy rebuilt from its saved boxes has the same schedules.
* The restart scalars use the reference's keys (enlightning.C:1861-1933) and
  restore the Gas and the boundary types exactly (REFLECT = 10, ABSORB = 20,
  enlightning.C:64-65).
* save_restart / load_restart round-trip every SOLN array and scalar bit for
  bit (here through the oracle's hierarchy, with the ghost fills it applies
  on load).
* Plot quantities follow registerModelVariables' names and order
  (enlightning.C:295-378) and input.txt's viz flags.
"""
import json

import numpy as np

import oracle
from enlightning_amd import amr, output
from enlightning_amd.gas import Gas
from enlightning_amd.hierarchy import level_tags, patch_initial_state
from enlightning_amd.workloads import Workload
from oracle.amr import HostHierarchy

WL = Workload("amr_test_64", n=64, seed=7, npulses=4, amp_range=(0.3, 0.8), sigma_range=(0.03, 0.06))

# keys of enlightning::putToRestart this path writes (enlightning.C:1861-1933)
REFERENCE_KEYS = {
    "p_nghosts", "p_loop_time", "p_iteration_number", "p_tag_tolerance", "p_bc_type_L", "p_bc_type_L_int",
    "p_bc_type_R", "p_bc_type_R_int", "p_bc_type_B", "p_bc_type_B_int", "p_bc_type_T", "p_bc_type_T_int",
    "p_bc_absorb_width", "p_p0", "p_T0", "p_c_lr", "p_rho0", "p_c", "p_gamma1", "p_RH", "p_R", "p_c_p",
    "p_T_star_n", "p_T_star_o", "p_mu", "p_muB", "p_kappa", "p_weno_alpha", "p_tau_n", "p_tau_o", "p_R_tilde",
    "p_c_v", "p_M",
}


class Host(HostHierarchy):
    def sources(self):
        return [s.copy() for s in self.src]

    def upload_all(self, states, sources):
        for k, ((w, pr, T), s) in enumerate(zip(states, sources)):
            a, b, c = self.view(k)
            a[...], b[...], c[...] = w, pr, T
            self.src[k][...] = s
        self.fill_source()
        self.fill_all()

    def set_absorb(self, width):
        self.absorb_width = width


def hierarchy():
    gas = Gas()
    return gas, amr.build(WL.n, 4, 4, 1.0, level_tags(gas, WL, oracle.tag_cells), tile0=16)


def test_boxes_rebuild_the_same_schedules():
    _, h = hierarchy()
    r = amr.from_boxes(h.n0, h.ratio, h.g, h.L, [[p.level, *p.box] for p in h.patches])
    assert [(p.level, p.box) for p in r.patches] == [(p.level, p.box) for p in h.patches]
    for lev in range(h.nlevels):
        for key in h.fill[lev]:
            for x, y in zip(h.fill[lev][key], r.fill[lev][key]):
                assert np.array_equal(x, y), (lev, key)
    for x, y in zip(h.coarsen, r.coarsen):
        assert np.array_equal(x, y)


def test_scalars_use_reference_keys_and_restore_exactly():
    gas = Gas(p0=95000.0, RH=35.0, tag_tolerance=(1.5, 2e-2, 1e-3))
    d = output.restart_scalars(gas, (0, 1, 1, 0), 0.125, 4, iteration=17, loop_time=0.25)
    assert set(d) == REFERENCE_KEYS
    d = json.loads(json.dumps(d))
    assert output.gas_from_scalars(d) == gas
    assert output.reflect_from_scalars(d) == (0, 1, 1, 0)
    assert (d["p_bc_type_L"], d["p_bc_type_L_int"], d["p_bc_type_R"], d["p_bc_type_R_int"]) == \
        ("ABSORB", 20, "REFLECT", 10)
    assert d["p_nghosts"] == [4, 4] and d["p_iteration_number"] == 17 and d["p_loop_time"] == 0.25


def test_save_load_round_trip(tmp_path):
    gas, h = hierarchy()
    a = Host(h, gas, reflect=(0, 1, 1, 1))
    a.absorb_width = 0.05
    rng = np.random.default_rng(2)
    states = [patch_initial_state(h, p, WL, gas) for p in h.patches]
    a.load_initial(states, [rng.standard_normal((p.Ny, p.Nx)) for p in h.patches])
    path = tmp_path / "chk.npz"
    output.save_restart(path, a, iteration=3, loop_time=1.5e-4, p_cfl=0.5)
    b, d = output.load_restart(path, factory=Host)
    assert d["p_iteration_number"] == 3 and d["p_loop_time"] == 1.5e-4 and d["p_cfl"] == 0.5
    assert b.reflect == (0, 1, 1, 1) and b.absorb_width == 0.05 and b.gas == gas
    for (wa, pa, Ta), (wb, pb, Tb), sa, sb, p in zip(a.states(), b.states(), a.sources(), b.sources(), h.patches):
        g = p.g
        assert np.array_equal(wa[:, g:-g, g:-g], wb[:, g:-g, g:-g])
        assert np.array_equal(pa[g:-g, g:-g], pb[g:-g, g:-g]) and np.array_equal(Ta[g:-g, g:-g], Tb[g:-g, g:-g])
        assert np.array_equal(sa[g:-g, g:-g], sb[g:-g, g:-g])
    # ghosts too: both were filled by the same schedule from the same interiors
    for (wa, _, _), (wb, _, _) in zip(a.states(), b.states()):
        assert np.array_equal(wa, wb)


def test_plot_quantities_names_and_values():
    gas, h = hierarchy()
    a = Host(h, gas)
    a.load_initial([patch_initial_state(h, p, WL, gas) for p in h.patches])
    q = output.plot_quantities(a, viz_w_n=1, viz_pressure=1, viz_temperature=1, viz_source=1)
    assert list(q) == ["w_n #0", "w_n #1", "w_n #2", "w_n #3", "w_n #4", "pressure", "temperature", "source"]
    assert list(output.plot_quantities(a)) == ["pressure"]  # input.txt:9-12
    for k, p in enumerate(h.patches):
        w, pr, T = a.view(k)
        g = p.g
        assert np.array_equal(q["w_n #3"][k], w[3, g:-g, g:-g]) and np.array_equal(q["pressure"][k], pr[g:-g, g:-g])
        assert q["temperature"][k].shape == (p.ny, p.nx)
