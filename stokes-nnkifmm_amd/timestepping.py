"""Time stepping of swimmer suspensions -- the caller of mobilityProblem in
Data Generation/Squiremers.m:47-54 and SquiremersGrav.m:55-67:

    [t, y] = ode45(@(t,x) mobilityProblem(swimmers, ..., t, x, ...), tspan, y0, opts)

ode45 here is the Dormand-Prince 5(4) pair (Dormand & Prince, J. Comput.
Appl. Math. 6, 1980) with local extrapolation, the mixed error test and the
step-size rules of Shampine & Reichelt ("The MATLAB ODE Suite", SIAM J. Sci.
Comput. 18, 1997): err = h * max_i |sum_j E_j k_j|_i / max(|y_i|, |ynew_i|,
AbsTol/RelTol) <= RelTol, and the free 4th-order continuous extension of
DOPRI5 (Hairer, Norsett & Wanner, Solving ODEs I, II.6) for the 'Refine'
output points.  State layout is the reference's: 9 values per body
[x0; b1; b2] (mobilityProblem.m:27-33), the ODE right-hand side returns
[U; Omega x b1; Omega x b2].

Every right-hand side rebuilds the tree, the KIFMM handles and B/BT, as the
reference does (mobilityProblem.m:26-35).  Two things make the rebuild
cheap when the root box does not change: the tree option 'rootCube', [lo hi]
(fixes the root instead of taking the points' bounding cube) and the
operator cache (operatorCache): handles whose levels have the same box
widths share the per-level SVD factors and translation operators instead of
recomputing them.
"""
import ctypes

import numpy as np

from ._lib import check, lib

# Dormand-Prince 5(4) tableau
_C = np.array([0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0])
_A = [
    [],
    [1 / 5],
    [3 / 40, 9 / 40],
    [44 / 45, -56 / 15, 32 / 9],
    [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
    [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
    [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84],
]
_B = np.array([35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0])
# 5th- minus 4th-order weights (the error estimate)
_E = _B - np.array([5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40])
# continuous extension: y(t + s h) = y + h * sum_j k_j * (sum_m _P[j][m] s^(m+1))
_P = np.array([
    [1.0, -8048581381 / 2820520608, 8663915743 / 2820520608, -12715105075 / 11282082432],
    [0.0, 0.0, 0.0, 0.0],
    [0.0, 131558114200 / 32700410799, -68118460800 / 10900136933, 87487479700 / 32700410799],
    [0.0, -1754552775 / 470086768, 14199869525 / 1410260304, -10690763975 / 1880347072],
    [0.0, 127303824393 / 49829197408, -318862633887 / 49829197408, 701980252875 / 199316789632],
    [0.0, -282668133 / 205662961, 2019193451 / 616988883, -1453857185 / 822651844],
    [0.0, 40617522 / 29380423, -110615467 / 29380423, 69997945 / 29380423],
])

_ODE_DEFAULTS = {"RelTol": 1e-3, "AbsTol": 1e-6, "Refine": 4, "OutputFcn": None, "InitialStep": None,
                 "MaxStep": None}


def odeset(*args, **kwargs):
    """Options for ode45 as name/value pairs (case-insensitive names):
    RelTol, AbsTol, Refine, OutputFcn, InitialStep, MaxStep."""
    if len(args) % 2:
        raise ValueError("odeset: name/value pairs expected")
    opts = dict(_ODE_DEFAULTS)
    keys = {k.lower(): k for k in opts}
    for name, val in list(zip(args[0::2], args[1::2])) + list(kwargs.items()):
        k = keys.get(str(name).lower())
        if k is None:
            raise ValueError(f"odeset: unknown option {name!r}")
        opts[k] = val
    return opts


def dp45_step(f, t, y, h, k1):
    """One Dormand-Prince step from (t, y) with f(t, y) = k1: returns
    (ynew, stages k_1..k_7, error vector h * sum E_j k_j).  k_7 = f(t+h, ynew)
    is the next step's k_1."""
    K = np.empty((7, y.size))
    K[0] = k1
    for s in range(1, 6):
        K[s] = f(t + _C[s] * h, y + h * (np.asarray(_A[s]) @ K[:s]))
    ynew = y + h * (_B[:6] @ K[:6])
    K[6] = f(t + h, ynew)
    return ynew, K, h * (_E @ K)


def _interp(y, h, K, s):
    """Continuous extension at fractions s (array) of the step."""
    s = np.atleast_1d(np.asarray(s, float))
    pw = np.stack([s, s ** 2, s ** 3, s ** 4], 1)   # (ns, 4)
    return y[None, :] + h * (pw @ (K.T @ _P).T)


def ode45(odefun, tspan, y0, options=None):
    """[t, y] = ode45(odefun, tspan, y0, options): odefun(t, y) -> dy/dt.
    tspan = [t0 tf] gives every step plus Refine-1 interpolated points per
    step; a longer tspan gives the solution at exactly those times.  y has
    one row per output time.  OutputFcn(t, y, flag) is called with flag
    'init', '' (new points: t 1-D, y (n, k)) and 'done', like MATLAB's; a
    nonzero return stops the integration."""
    o = dict(_ODE_DEFAULTS) if options is None else options
    tspan = np.asarray(tspan, float).ravel()
    if tspan.size < 2 or np.any(np.diff(tspan) == 0):
        raise ValueError("ode45: tspan must hold at least two distinct times")
    t0, tf = float(tspan[0]), float(tspan[-1])
    direction = 1.0 if tf > t0 else -1.0
    if np.any(direction * np.diff(tspan) < 0):
        raise ValueError("ode45: tspan must be monotone")
    y = np.asarray(y0, float).ravel().copy()
    rtol, atol = float(o["RelTol"]), float(o["AbsTol"])
    rtol = max(rtol, 100 * np.finfo(float).eps)
    threshold = atol / rtol
    refine = int(o["Refine"]) if tspan.size == 2 else 1
    hmax = float(o["MaxStep"]) if o["MaxStep"] is not None else 0.1 * abs(tf - t0)
    outfcn = o["OutputFcn"]

    def f(t, x):
        return np.asarray(odefun(t, x), float).ravel()

    tout, yout = [t0], [y.copy()]
    if outfcn is not None:
        outfcn(np.array([t0, tf]), y.copy(), "init")
    t = t0
    k1 = f(t, y)
    # initial step (Shampine & Reichelt)
    htspan = abs(tf - t0)
    if o["InitialStep"] is not None:
        absh = min(hmax, htspan, abs(float(o["InitialStep"])))
    else:
        absh = min(hmax, htspan)
        rh = np.max(np.abs(k1) / np.maximum(np.abs(y), threshold)) / (0.8 * rtol ** 0.2)
        if absh * rh > 1:
            absh = 1.0 / rh
    next_out = 1
    stop = False
    while not stop:
        hmin = 16 * np.spacing(t)
        absh = min(hmax, max(hmin, absh))
        h = direction * absh
        if 1.1 * absh >= abs(tf - t):
            h = tf - t
            absh = abs(h)
            last = True
        else:
            last = False
        failed = False
        while True:
            tnew = tf if last else t + h
            h = tnew - t
            ynew, K, evec = dp45_step(f, t, y, h, k1)
            err = np.max(np.abs(evec) / np.maximum(np.maximum(np.abs(y), np.abs(ynew)), threshold))
            if err > rtol:
                if absh <= hmin:
                    raise RuntimeError(f"ode45: step size below {hmin:.3g} at t = {t:.6g}")
                if failed:
                    absh = max(hmin, 0.5 * absh)
                else:
                    absh = max(hmin, absh * max(0.1, 0.8 * (rtol / err) ** 0.2))
                    failed = True
                h = direction * absh
                last = False
                continue
            break
        # output
        if tspan.size == 2:
            if refine > 1:
                s = np.arange(1, refine) / refine
                ti = t + s * h
                yi = _interp(y, h, K, s)
                newt = list(ti) + [tnew]
                newy = list(yi) + [ynew]
            else:
                newt, newy = [tnew], [ynew]
        else:
            newt, newy = [], []
            while next_out < tspan.size and direction * (tspan[next_out] - tnew) <= 0:
                tq = tspan[next_out]
                newt.append(tq)
                newy.append(ynew.copy() if tq == tnew else _interp(y, h, K, [(tq - t) / h])[0])
                next_out += 1
        tout.extend(newt)
        yout.extend(newy)
        if outfcn is not None and newt:
            if outfcn(np.asarray(newt), np.stack(newy, 1), ""):
                stop = True
        t, y, k1 = tnew, ynew, K[6]
        if t == tf:
            stop = True
        if not failed:
            temp = 1.25 * (err / rtol) ** 0.2
            absh = absh / temp if temp > 0.2 else 5.0 * absh
    if outfcn is not None:
        outfcn(np.array([]), np.array([]), "done")
    return np.asarray(tout), np.stack(yout)


def trajectoryOutput(filename):
    """OutputFcn that appends [t; y] columns to a CSV file as the steps come
    (the save function of SquiremersGrav.m:88-101)."""
    def out(t, y, flag):
        if flag == "init":
            open(filename, "w").close()
        elif flag == "":
            cols = np.vstack([np.atleast_1d(t)[None, :], np.asarray(y).reshape(-1, len(np.atleast_1d(t)))])
            old = np.loadtxt(filename, delimiter=",", ndmin=2) if _nonempty(filename) else None
            data = cols if old is None else np.hstack([old, cols])
            np.savetxt(filename, data, delimiter=",", fmt="%.17g")
        return 0
    return out


def _nonempty(path):
    import os
    return os.path.exists(path) and os.path.getsize(path) > 0


def writeTrajectory(filename, t, y):
    """writematrix([t'; y'], filename) (SquiremersGrav.m:69)."""
    np.savetxt(filename, np.vstack([np.asarray(t)[None, :], np.asarray(y).T]), delimiter=",", fmt="%.17g")


def operatorCache(limitBytes):
    """Keep up to limitBytes of per-level KIFMM operators after their last
    handle is destroyed (0 keeps none; handles alive at the same time always
    share).  Returns the previous statistics (hits, misses, bytes, entries)."""
    st = operatorCacheStats()
    check(lib().nnk_op_cache_limit(int(limitBytes)))
    return st


def operatorCacheStats():
    """(hits, misses, bytes kept, entries kept) of the operator cache."""
    v = [ctypes.c_int64() for _ in range(4)]
    check(lib().nnk_op_cache_stats(*[ctypes.byref(x) for x in v]))
    return tuple(int(x.value) for x in v)
