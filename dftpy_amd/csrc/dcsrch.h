// Moré-Thuente line search state machine (host entry; the algorithm is in
// dcsrch_core.h, shared with the device-driven optimizer iteration).
#pragma once
#include "dcsrch_core.h"

namespace ofdft {

// One reverse-communication step: with task LS_START initialise from
// (f, g) at stp = 0 and the first trial step stp; afterwards pass the value
// and derivative at the last returned stp.  Evaluate again while LS_FG is
// returned.  state: 15 doubles owned by the caller.
LsTask dcsrch(double &stp, double f, double g, double ftol, double gtol, double xtol, double stpmin, double stpmax,
              LsTask task, double *state);

}  // namespace ofdft
