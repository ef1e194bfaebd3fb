// enl_switch.h -- the learned switch as the batched FP32-perturbation form
// calls it (enl_switch.hip implements, enl_batch32.hip calls).
#pragma once
#include <hip/hip_runtime.h>

namespace enl {

struct Switch;  // enl_switch.hip

// One level-0 patch of an FP32-perturbation hierarchy as the switch reads
// it: the perturbations rho' and p' of a stage block (ghosted FP32, row
// stride sy, ghost width g), per ghosted row r the ambient density
// difference to the next row, rho_a(r+1) - rho_a(r), at drho[r*rs] (the
// prefilter's pairs are those of the full density), the patch's interior
// decision map (nx*ny bytes: 0 DRP, 1 WENO, 2 sensor) and per interior cell
// the sensor's choice on the ambient alone (1 WENO).
struct SwPatch {
  const float* rho;
  const float* p;
  const float* drho;
  unsigned char* decide;
  const unsigned char* amb_weno;
  int nx, ny, g, rs;
  long sy;
  long foff;  // offset of the patch's ghosted cells in the switch's feature map
};

// The decisions of every patch of the table `patches` (device, npatch
// entries, at most max_side cells a side, `cells` interior cells in all)
// into their maps, on stream s:
// the exact prefilter (every sensor pair of the full density provably
// smooth -> DRP; of the perturbation alone, or the ambient alone sensor-WENO
// -> 2, the sensor decides), then the network on the remaining cells from
// the differences of the perturbations (p > hi WENO, p < lo DRP, else 2).
// inv_r = 1/sqrt(eps), inv_p = 1/(c^2 sqrt(eps)) normalise the features,
// thr = 5e-5 eps.  Counts (candidates, WENO, undecided) of the call stay in
// the switch for switch_accumulate.  Launches only (no allocation) once
// switch_reserve has sized the candidate list for `cells` and the feature
// map for `fcells` ghosted cells (the patches' foff ranges).
int switch_decide_batch(Switch* sw, const SwPatch* patches, int npatch, int max_side, long cells, long fcells,
                        double thr, float inv_r, float inv_p, float lo, float hi, hipStream_t s);
// Candidate list capacity for `cells` level-0 cells and feature map for
// `fcells` ghosted cells (allocates if needed).
int switch_reserve(Switch* sw, long cells, long fcells);
// Adds the last decisions' counts to totals[3] (device), on stream s.
int switch_accumulate(Switch* sw, unsigned long long* totals, hipStream_t s);

}  // namespace enl
