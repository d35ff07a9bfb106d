// Moré-Thuente line search state machine (host).
#pragma once

namespace ofdft {

enum LsTask { LS_START = 0, LS_FG = 1, LS_CONVERGENCE = 2, LS_WARNING = 3, LS_ERROR = 4 };

// One reverse-communication step: with task LS_START initialise from
// (f, g) at stp = 0 and the first trial step stp; afterwards pass the value
// and derivative at the last returned stp.  Evaluate again while LS_FG is
// returned.  state: 15 doubles owned by the caller.
LsTask dcsrch(double &stp, double f, double g, double ftol, double gtol, double xtol, double stpmin, double stpmax,
              LsTask task, double *state);

}  // namespace ofdft
