"""Moré-Thuente line search (MINPACK-2 ``dcsrch``/``dcstep``), restated from
Moré & Thuente, "Line search algorithms with guaranteed sufficient
decrease", ACM TOMS 20 (1994) 286-307.  The reference reaches the Fortran
routine through scipy (``math_utils.py:6-9,76-106``).

State is a dict so the reverse-communication protocol of the original is
kept: call with task 'START', evaluate f,g at the returned stp while the task
starts with 'FG'.
"""
import math

XTRAPL = 1.1
XTRAPU = 4.0
P5 = 0.5
P66 = 0.66


def dcstep(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax):
    sgnd = dp * (dx / abs(dx))
    if fp > fx:
        # case 1: higher function value, the minimum is bracketed
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp < stx:
            gamma = -gamma
        p = (gamma - dx) + theta
        q = ((gamma - dx) + gamma) + dp
        r = p / q
        stpc = stx + r * (stp - stx)
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx)
        if abs(stpc - stx) < abs(stpq - stx):
            stpf = stpc
        else:
            stpf = stpc + (stpq - stpc) / 2.0
        brackt = True
    elif sgnd < 0.0:
        # case 2: derivatives of opposite sign, the minimum is bracketed
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = ((gamma - dp) + gamma) + dx
        r = p / q
        stpc = stp + r * (stx - stp)
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        if abs(stpc - stp) > abs(stpq - stp):
            stpf = stpc
        else:
            stpf = stpq
        brackt = True
    elif abs(dp) < abs(dx):
        # case 3: same sign, derivative magnitude decreases
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt(max(0.0, (theta / s) ** 2 - (dx / s) * (dp / s)))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = (gamma + (dx - dp)) + gamma
        r = p / q
        if r < 0.0 and gamma != 0.0:
            stpc = stp + r * (stx - stp)
        elif stp > stx:
            stpc = stpmax
        else:
            stpc = stpmin
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        if brackt:
            stpf = stpc if abs(stpc - stp) < abs(stpq - stp) else stpq
            if stp > stx:
                stpf = min(stp + P66 * (sty - stp), stpf)
            else:
                stpf = max(stp + P66 * (sty - stp), stpf)
        else:
            stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
            stpf = min(stpmax, stpf)
            stpf = max(stpmin, stpf)
    else:
        # case 4: same sign, derivative magnitude does not decrease
        if brackt:
            theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp
            s = max(abs(theta), abs(dy), abs(dp))
            gamma = s * math.sqrt((theta / s) ** 2 - (dy / s) * (dp / s))
            if stp > sty:
                gamma = -gamma
            p = (gamma - dp) + theta
            q = ((gamma - dp) + gamma) + dy
            r = p / q
            stpc = stp + r * (sty - stp)
            stpf = stpc
        elif stp > stx:
            stpf = stpmax
        else:
            stpf = stpmin
    # update the interval which contains a minimizer
    if fp > fx:
        sty, fy, dy = stp, fp, dp
    else:
        if sgnd < 0.0:
            sty, fy, dy = stx, fx, dx
        stx, fx, dx = stp, fp, dp
    return stx, fx, dx, sty, fy, dy, stpf, brackt


def dcsrch(stp, f, g, ftol, gtol, xtol, task, stpmin, stpmax, st):
    """One reverse-communication step; returns (stp, task)."""
    if task == "START":
        if stp < stpmin:
            return stp, "ERROR: STP .LT. STPMIN"
        if stp > stpmax:
            return stp, "ERROR: STP .GT. STPMAX"
        if g >= 0.0:
            return stp, "ERROR: INITIAL G .GE. ZERO"
        if ftol < 0.0:
            return stp, "ERROR: FTOL .LT. ZERO"
        if gtol < 0.0:
            return stp, "ERROR: GTOL .LT. ZERO"
        if xtol < 0.0:
            return stp, "ERROR: XTOL .LT. ZERO"
        if stpmin < 0.0:
            return stp, "ERROR: STPMIN .LT. ZERO"
        if stpmax < stpmin:
            return stp, "ERROR: STPMAX .LT. STPMIN"
        st.clear()
        st.update(brackt=False, stage=1, finit=f, ginit=g, gtest=ftol * g,
                  width=stpmax - stpmin, width1=(stpmax - stpmin) / P5,
                  stx=0.0, fx=f, gx=g, sty=0.0, fy=f, gy=g,
                  stmin=0.0, stmax=stp + XTRAPU * stp)
        return stp, "FG"
    s = st
    ftest = s["finit"] + stp * s["gtest"]
    if s["stage"] == 1 and f <= ftest and g >= 0.0:
        s["stage"] = 2
    task = None
    if s["brackt"] and (stp <= s["stmin"] or stp >= s["stmax"]):
        task = "WARNING: ROUNDING ERRORS PREVENT PROGRESS"
    if s["brackt"] and s["stmax"] - s["stmin"] <= xtol * s["stmax"]:
        task = "WARNING: XTOL TEST SATISFIED"
    if stp == stpmax and f <= ftest and g <= s["gtest"]:
        task = "WARNING: STP = STPMAX"
    if stp == stpmin and (f > ftest or g >= s["gtest"]):
        task = "WARNING: STP = STPMIN"
    if f <= ftest and abs(g) <= gtol * (-s["ginit"]):
        task = "CONVERGENCE"
    if task is not None:
        return stp, task
    gtest = s["gtest"]
    if s["stage"] == 1 and f <= s["fx"] and f > ftest:
        fm = f - stp * gtest
        fxm = s["fx"] - s["stx"] * gtest
        fym = s["fy"] - s["sty"] * gtest
        gm = g - gtest
        gxm = s["gx"] - gtest
        gym = s["gy"] - gtest
        (s["stx"], fxm, gxm, s["sty"], fym, gym, stp, s["brackt"]) = dcstep(
            s["stx"], fxm, gxm, s["sty"], fym, gym, stp, fm, gm, s["brackt"], s["stmin"], s["stmax"])
        s["fx"] = fxm + s["stx"] * gtest
        s["fy"] = fym + s["sty"] * gtest
        s["gx"] = gxm + gtest
        s["gy"] = gym + gtest
    else:
        (s["stx"], s["fx"], s["gx"], s["sty"], s["fy"], s["gy"], stp, s["brackt"]) = dcstep(
            s["stx"], s["fx"], s["gx"], s["sty"], s["fy"], s["gy"], stp, f, g, s["brackt"], s["stmin"], s["stmax"])
    if s["brackt"]:
        if abs(s["sty"] - s["stx"]) >= P66 * s["width1"]:
            stp = s["stx"] + P5 * (s["sty"] - s["stx"])
        s["width1"] = s["width"]
        s["width"] = abs(s["sty"] - s["stx"])
    if s["brackt"]:
        s["stmin"] = min(s["stx"], s["sty"])
        s["stmax"] = max(s["stx"], s["sty"])
    else:
        s["stmin"] = stp + XTRAPL * (stp - s["stx"])
        s["stmax"] = stp + XTRAPU * (stp - s["stx"])
    stp = max(stp, stpmin)
    stp = min(stp, stpmax)
    if (s["brackt"] and (stp <= s["stmin"] or stp >= s["stmax"])) or (
            s["brackt"] and s["stmax"] - s["stmin"] <= xtol * s["stmax"]):
        stp = s["stx"]
    return stp, "FG"


def line_search(func, alpha0, func0, c1=1e-4, c2=0.9, amax=1.0, amin=0.0, xtol=1e-14, maxiter=100):
    """``math_utils.py:76-106`` (LineSearchDcsrch2).  ``func(alpha)`` returns a
    list whose first two entries are value and derivative.  Returns
    (alpha or None, task, number of loop passes, last func result)."""
    st = {}
    task = "START"
    alpha1 = alpha0
    x1, g1 = func0[0], func0[1]
    func1 = func0
    i = 0
    for i in range(maxiter):
        alpha1, task = dcsrch(alpha1, x1, g1, c1, c2, xtol, task, amin, amax, st)
        if task[:2] == "FG":
            func1 = func(alpha1)
            x1, g1 = func1[0], func1[1]
        else:
            break
    else:
        alpha1 = None
    if task[:5] == "ERROR" or task[:4] == "WARN":
        alpha1 = None
    return alpha1, task, i, func1
