// Moré-Thuente line search (the MINPACK-2 dcsrch/dcstep pair the reference
// reaches through scipy, math_utils.py:6-9,76-106), written from
// J. J. Moré and D. J. Thuente, "Line search algorithms with guaranteed
// sufficient decrease", ACM TOMS 20 (1994) 286-307.  FP64, one
// reverse-communication step per call.  Host and device code: the host
// optimizer loop (dcsrch.cpp) and the device-driven TN iteration (opt.hip)
// run the same arithmetic, with floating-point contraction off so both round
// identically.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#if defined(__clang__)
#define OFDFT_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define OFDFT_NO_CONTRACT
#endif
#define OFDFT_LS_FN __host__ __device__ inline

namespace ofdft {

enum LsTask { LS_START = 0, LS_FG = 1, LS_CONVERGENCE = 2, LS_WARNING = 3, LS_ERROR = 4 };

namespace ls {

// std::max / std::min argument order (the first argument on ties)
OFDFT_LS_FN double dmax(double a, double b) { return a < b ? b : a; }
OFDFT_LS_FN double dmin(double a, double b) { return b < a ? b : a; }

constexpr double kXtrapl = 1.1, kXtrapu = 4.0, kP5 = 0.5, kP66 = 0.66;

// One safeguarded step: update the interval [stx, sty] that brackets a
// minimiser and return the next trial step.
OFDFT_LS_FN void dcstep(double &stx, double &fx, double &dx, double &sty, double &fy, double &dy, double &stp,
                   double fp, double dp, bool &brackt, double stpmin, double stpmax) {
    OFDFT_NO_CONTRACT
    const double sgnd = dp * (dx / fabs(dx));
    double stpf;
    if (fp > fx) {
        // higher function value: the minimum is bracketed (cubic or quadratic)
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = dmax(dmax(fabs(theta), fabs(dx)), fabs(dp));
        double gamma = s * sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
        if (stp < stx) gamma = -gamma;
        const double p = (gamma - dx) + theta;
        const double q = ((gamma - dx) + gamma) + dp;
        const double r = p / q;
        const double stpc = stx + r * (stp - stx);
        const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
        stpf = (fabs(stpc - stx) < fabs(stpq - stx)) ? stpc : stpc + (stpq - stpc) / 2.0;
        brackt = true;
    } else if (sgnd < 0.0) {
        // derivatives of opposite sign: bracketed, cubic vs secant
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = dmax(dmax(fabs(theta), fabs(dx)), fabs(dp));
        double gamma = s * sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
        if (stp > stx) gamma = -gamma;
        const double p = (gamma - dp) + theta;
        const double q = ((gamma - dp) + gamma) + dx;
        const double r = p / q;
        const double stpc = stp + r * (stx - stp);
        const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
        stpf = (fabs(stpc - stp) > fabs(stpq - stp)) ? stpc : stpq;
        brackt = true;
    } else if (fabs(dp) < fabs(dx)) {
        // same sign, derivative magnitude decreases
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = dmax(dmax(fabs(theta), fabs(dx)), fabs(dp));
        double gamma = s * sqrt(dmax(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
        if (stp > stx) gamma = -gamma;
        const double p = (gamma - dp) + theta;
        const double q = (gamma + (dx - dp)) + gamma;
        const double r = p / q;
        double stpc;
        if (r < 0.0 && gamma != 0.0)
            stpc = stp + r * (stx - stp);
        else if (stp > stx)
            stpc = stpmax;
        else
            stpc = stpmin;
        const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
        if (brackt) {
            stpf = (fabs(stpc - stp) < fabs(stpq - stp)) ? stpc : stpq;
            if (stp > stx)
                stpf = dmin(stp + kP66 * (sty - stp), stpf);
            else
                stpf = dmax(stp + kP66 * (sty - stp), stpf);
        } else {
            stpf = (fabs(stpc - stp) > fabs(stpq - stp)) ? stpc : stpq;
            stpf = dmin(stpmax, stpf);
            stpf = dmax(stpmin, stpf);
        }
    } else {
        // same sign, derivative magnitude does not decrease
        if (brackt) {
            const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
            const double s = dmax(dmax(fabs(theta), fabs(dy)), fabs(dp));
            double gamma = s * sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
            if (stp > sty) gamma = -gamma;
            const double p = (gamma - dp) + theta;
            const double q = ((gamma - dp) + gamma) + dy;
            const double r = p / q;
            stpf = stp + r * (sty - stp);
        } else if (stp > stx) {
            stpf = stpmax;
        } else {
            stpf = stpmin;
        }
    }
    if (fp > fx) {
        sty = stp;
        fy = fp;
        dy = dp;
    } else {
        if (sgnd < 0.0) {
            sty = stx;
            fy = fx;
            dy = dx;
        }
        stx = stp;
        fx = fp;
        dx = dp;
    }
    stp = stpf;
}

// state layout: [0] brackt [1] stage [2] finit [3] ginit [4] gtest [5] width
// [6] width1 [7] stx [8] fx [9] gx [10] sty [11] fy [12] gy [13] stmin [14] stmax
OFDFT_LS_FN LsTask dcsrch_step(double &stp, double f, double g, double ftol, double gtol, double xtol,
                               double stpmin, double stpmax, LsTask task, double *st) {
    OFDFT_NO_CONTRACT
    if (task == LS_START) {
        if (stp < stpmin || stp > stpmax || g >= 0.0 || ftol < 0.0 || gtol < 0.0 || xtol < 0.0 || stpmin < 0.0 ||
            stpmax < stpmin)
            return LS_ERROR;
        st[0] = 0.0;
        st[1] = 1.0;
        st[2] = f;
        st[3] = g;
        st[4] = ftol * g;
        st[5] = stpmax - stpmin;
        st[6] = (stpmax - stpmin) / kP5;
        st[7] = 0.0;
        st[8] = f;
        st[9] = g;
        st[10] = 0.0;
        st[11] = f;
        st[12] = g;
        st[13] = 0.0;
        st[14] = stp + kXtrapu * stp;
        return LS_FG;
    }
    bool brackt = st[0] != 0.0;
    int stage = (int)st[1];
    const double finit = st[2], ginit = st[3], gtest = st[4];
    double width = st[5], width1 = st[6];
    double stx = st[7], fx = st[8], gx = st[9], sty = st[10], fy = st[11], gy = st[12];
    double stmin = st[13], stmax = st[14];

    const double ftest = finit + stp * gtest;
    if (stage == 1 && f <= ftest && g >= 0.0) stage = 2;
    LsTask out = LS_FG;
    if (brackt && (stp <= stmin || stp >= stmax)) out = LS_WARNING;
    if (brackt && stmax - stmin <= xtol * stmax) out = LS_WARNING;
    if (stp == stpmax && f <= ftest && g <= gtest) out = LS_WARNING;
    if (stp == stpmin && (f > ftest || g >= gtest)) out = LS_WARNING;
    if (f <= ftest && fabs(g) <= gtol * (-ginit)) out = LS_CONVERGENCE;
    if (out != LS_FG) {
        st[1] = stage;
        return out;
    }
    if (stage == 1 && f <= fx && f > ftest) {
        // modified function psi(stp) = f(stp) - f(0) - ftol stp f'(0)
        const double fm = f - stp * gtest;
        double fxm = fx - stx * gtest, fym = fy - sty * gtest;
        const double gm = g - gtest;
        double gxm = gx - gtest, gym = gy - gtest;
        dcstep(stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax);
        fx = fxm + stx * gtest;
        fy = fym + sty * gtest;
        gx = gxm + gtest;
        gy = gym + gtest;
    } else {
        dcstep(stx, fx, gx, sty, fy, gy, stp, f, g, brackt, stmin, stmax);
    }
    if (brackt) {
        if (fabs(sty - stx) >= kP66 * width1) stp = stx + kP5 * (sty - stx);
        width1 = width;
        width = fabs(sty - stx);
        stmin = dmin(stx, sty);
        stmax = dmax(stx, sty);
    } else {
        stmin = stp + kXtrapl * (stp - stx);
        stmax = stp + kXtrapu * (stp - stx);
    }
    stp = dmax(stp, stpmin);
    stp = dmin(stp, stpmax);
    if ((brackt && (stp <= stmin || stp >= stmax)) || (brackt && stmax - stmin <= xtol * stmax)) stp = stx;
    st[0] = brackt ? 1.0 : 0.0;
    st[1] = stage;
    st[5] = width;
    st[6] = width1;
    st[7] = stx;
    st[8] = fx;
    st[9] = gx;
    st[10] = sty;
    st[11] = fy;
    st[12] = gy;
    st[13] = stmin;
    st[14] = stmax;
    return LS_FG;
}

}  // namespace ls
}  // namespace ofdft
