"""Wavenumber quadrature of the 2.5-D transform (host, once per survey).

The 2.5-D forward sums nky 2-D solves, u = sum_k w_k u_k.  Points and weights
are fitted so that sum_k 2 w_k K0(k r) = 1/r over the electrode distances of
the survey, with non-negative weights (reference dctorch/quadrature.py:38-154).
The fit follows the reference's steps so that it returns the same points and
weights:

1. sample r log-uniformly on [r_min/margin, r_max*margin] (200 samples);
2. for given log10-wavenumbers, the weights g solve a column-normalised
   non-negative least squares on the design r K0(k r) (bounded least squares
   if the active-set solver stalls), w = g/2;
3. the log-wavenumbers minimise the residual with a trust-region
   least-squares (xtol = ftol = 1e-12, 200 evaluations per point);
4. continuation: start with min(4, nky) points spread over
   [-log10 max r, -log10 min r] and add one point at a time in the middle of
   the widest gap (the two ends count as half-unit gaps);
5. points with zero weight are dropped.

`simpeg_quadrature` restates stock SimPEG 0.25's own fit (used only when
``quadrature="simpeg"`` is asked for); it cannot be checked against SimPEG
here.
"""
import warnings

import numpy as np
from scipy.special import k0


def electrode_distance_range(survey):
    """(r_min, r_max) over every current/potential electrode pair (x-z plane)."""
    lo, hi = np.inf, 0.0
    for src in survey.source_list:
        cur = [np.atleast_2d(np.asarray(src.location_a, float))]
        if getattr(src, "location_b", None) is not None:
            cur.append(np.atleast_2d(np.asarray(src.location_b, float)))
        cur = np.vstack(cur)[:, :2]
        cur = cur[np.isfinite(cur).all(axis=1)]
        for rx in src.receiver_list:
            pot = [np.atleast_2d(np.asarray(rx.locations_m, float))]
            if getattr(rx, "locations_n", None) is not None:
                pot.append(np.atleast_2d(np.asarray(rx.locations_n, float)))
            pot = np.vstack(pot)[:, :2]
            pot = pot[np.isfinite(pot).all(axis=1)]
            d = np.sqrt(((cur[:, None, :] - pot[None, :, :]) ** 2).sum(-1)).ravel()
            d = d[d > 1e-9]
            if d.size:
                lo, hi = min(lo, d.min()), max(hi, d.max())
    if not np.isfinite(lo):
        raise ValueError("the survey has no separated current/potential electrode pair")
    return float(lo), float(hi)


def amplification(weights):
    """sum|w| / |sum w|: error magnification of the quadrature (1 if w >= 0)."""
    w = np.asarray(weights, float)
    return float(np.abs(w).sum() / abs(w.sum()))


def _design(r, k):
    return r[:, None] * k0(r[:, None] * k[None, :])


def _nonneg_lsq(D, rhs):
    from scipy.optimize import lsq_linear, nnls
    norms = np.linalg.norm(D, axis=0)
    norms[norms == 0] = 1.0
    Dn = D / norms
    try:
        g = nnls(Dn, rhs, maxiter=50 * D.shape[1])[0]
    except RuntimeError:
        g = lsq_linear(Dn, rhs, bounds=(0.0, np.inf), method="bvls").x
    return g / norms


def transform_error(points, weights, r_min, r_max, n=400):
    """max |sum_k 2 w_k r K0(k r) - 1| on a log grid of [r_min, r_max]."""
    r = np.logspace(np.log10(r_min), np.log10(r_max), n)
    fit = _design(r, np.asarray(points, float)) @ (2.0 * np.asarray(weights, float))
    return float(np.abs(fit - 1.0).max())


def k0_quadrature(r_min, r_max, nky, margin=2.0, n_samples=200, n_start=4):
    """(points, weights, info) fitted on [r_min/margin, r_max*margin]."""
    from scipy.optimize import least_squares

    r = np.logspace(np.log10(r_min / margin), np.log10(r_max * margin), n_samples)
    ones = np.ones_like(r)

    def weights_of(lk):
        return _nonneg_lsq(_design(r, 10.0 ** lk), ones)

    def residual(lk):
        D = _design(r, 10.0 ** lk)
        return D @ _nonneg_lsq(D, ones) - ones

    def optimise(lk0):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            sol = least_squares(residual, lk0, method="trf", xtol=1e-12, ftol=1e-12,
                                max_nfev=200 * lk0.size)
        return np.sort(sol.x), bool(sol.success)

    lo, hi = -np.log10(r.max()), -np.log10(r.min())
    m0 = min(n_start, nky)
    lk, ok = optimise(np.linspace(lo, hi, m0))
    errs = {m0: transform_error(10.0 ** lk, weights_of(lk) / 2.0, r_min, r_max)}
    for m in range(m0 + 1, nky + 1):
        bounds = np.r_[lo - 0.5, lk, hi + 0.5]
        gap = int(np.argmax(np.diff(bounds)))
        lk, ok = optimise(np.sort(np.r_[lk, 0.5 * (bounds[gap] + bounds[gap + 1])]))
        errs[m] = transform_error(10.0 ** lk, weights_of(lk) / 2.0, r_min, r_max)
    pts = 10.0 ** lk
    wts = weights_of(lk) / 2.0
    keep = wts > 0
    pts, wts = pts[keep], wts[keep]
    info = {"success": ok and bool(np.all(np.isfinite(wts))), "max_rel_err": errs[nky],
            "r_range": (float(r_min), float(r_max)), "err_by_n": errs,
            "n_used": int(pts.size), "amplification": amplification(wts)}
    return pts, wts, info


def simpeg_quadrature(mesh, nky):
    """Stock SimPEG 0.25's fit (Simulation2DNodal): points and weights, and
    whether it fell back to the trapezoid rule.  Restated from SimPEG's
    documented method; not verifiable against SimPEG on this host."""
    from scipy.optimize import minimize
    from scipy.special import k1

    hmin = min(np.min(h) for h in mesh.h)
    extent = max(np.sum(h) for h in mesh.h)
    rs = np.logspace(np.log10(hmin / 4), np.log10(extent * 4), 100)
    e = np.ones_like(rs)

    def g_of(k):
        A = rs[:, None] * k0(rs[:, None] * k)
        return np.linalg.solve(A.T @ A, A.T @ e), A

    def phi(x):
        g, A = g_of(10.0 ** x)
        return np.linalg.norm((A @ g - e) / len(rs))

    def dphi(x):
        k = 10.0 ** x
        g, A = g_of(k)
        r = A @ g - e
        dA = -(rs ** 2)[:, None] * k1(rs[:, None] * k)
        # d|r|/dk_i through A only (g is optimal, so its variation drops out)
        nrm = np.linalg.norm(r)
        J = (dA * g[None, :]).T @ r / (nrm * len(rs) + 1e-300)
        return J * k * np.log(10.0)

    x0 = np.linspace(-np.log10(rs.max()), -np.log10(rs.min()), nky)
    out = minimize(phi, x0, jac=dphi, method="L-BFGS-B")
    if not out.success:
        pts = np.logspace(-4, 1, nky)
        # trapezoid on the log grid, scaled by 1/pi
        w = np.zeros(nky)
        dk = np.diff(pts)
        w[:-1] += dk / 2
        w[1:] += dk / 2
        return pts, w / np.pi, True
    pts = 10.0 ** out.x
    g, _ = g_of(pts)
    return pts, g / 2.0, False
