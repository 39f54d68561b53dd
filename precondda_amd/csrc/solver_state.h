// Device-resident per-RHS ("slot") solver state shared by the convolution and Krylov
// kernels. Every Krylov scalar lives here, on the GPU, so a solve enqueues iteration
// after iteration without a host round trip; kernels read the scalars they need and
// skip slots that are not iterating.
#pragma once

#include "pdda_common.h"

namespace pdda {

// Slot flags.
enum : int {
  F_ACTIVE = 1,      // slot is iterating
  F_WAS_RESET = 2,   // previous iteration ended with a reset (zeta-only next step)
  F_RECOMPUTE = 4,   // this iteration's reset recomputes r = b - A M x
  F_IMPROVED = 8,    // this iteration lowered the best residual norm
  F_RESET = 16,      // this iteration ends with a reset
};

// Slot status (reported to the host).
enum : int {
  ST_IDLE = 0,
  ST_RUNNING = 1,
  ST_CONVERGED = 2,
  ST_FAILED = 3,    // NaN residual or stagnation window exhausted
  ST_MAXITER = 4,
};

// Which slots a launch works on.
enum : int {
  NEED_ALL = 0,        // every slot (no state check)
  NEED_ACTIVE = 1,     // F_ACTIVE
  NEED_REPLACE = 2,    // F_ACTIVE and it == 1 (true-residual replacement step)
  NEED_RECOMPUTE = 3,  // F_RECOMPUTE (reset with residual recompute)
};

// Reduction kinds fused into the last pass of a convolution or into a vector update.
enum : int {
  DOT_NONE = 0,
  DOT_ALPHA = 1,  // <rhat, o>            -> alpha = rho1 / sum
  DOT_ZETA = 2,   // <y,y>,<o,o>,<y,o>,<o,y>,<y,s>,<o,s> -> zeta, eta
  DOT_RESID = 3,  // <o,o>, <rhat,o>      -> sums[0], sums[1]
};

struct SlotCtl {
  cd rho0, rho1, alpha, beta, zeta, eta;
  cd sums[6];
  double initial_norm, current_norm, lowest_norm, previous_lowest, first_norm, target;
  double ratio, rhs_scale;
  int it, max_iter, stagnation, last_change;
  int flags, status, stop_it, resets;
};

__device__ inline bool slot_wanted(const SlotCtl* ctl, int slot, int need) {
  if (ctl == nullptr || need == NEED_ALL) return true;
  const SlotCtl& c = ctl[slot];
  switch (need) {
    case NEED_ACTIVE: return (c.flags & F_ACTIVE) != 0;
    case NEED_REPLACE: return (c.flags & F_ACTIVE) != 0 && c.it == 1;
    case NEED_RECOMPUTE: return (c.flags & F_RECOMPUTE) != 0;
    default: return false;
  }
}

}  // namespace pdda
