// Host entry points of the Moré-Thuente line search (dcsrch_core.h): the
// host optimizer loop's dcsrch and the C-ABI ofdft_dcsrch.
#include "dcsrch.h"

namespace ofdft {

LsTask dcsrch(double &stp, double f, double g, double ftol, double gtol, double xtol, double stpmin, double stpmax,
              LsTask task, double *st) {
    return ls::dcsrch_step(stp, f, g, ftol, gtol, xtol, stpmin, stpmax, task, st);
}

}  // namespace ofdft

extern "C" int ofdft_dcsrch(double *stp, double f, double g, double ftol, double gtol, double xtol, double stpmin,
                            double stpmax, int *task, double *state) {
    if (!stp || !task || !state) return 1;
    *task = (int)ofdft::dcsrch(*stp, f, g, ftol, gtol, xtol, stpmin, stpmax, (ofdft::LsTask)*task, state);
    return 0;
}
