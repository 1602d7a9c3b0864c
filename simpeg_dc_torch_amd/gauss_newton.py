"""Gauss-Newton step algebra on an engine's exact sensitivities (FP64).

The reference's `TorchGaussNewton` (dctorch/optimization.py:440-983) is a
SimPEG `Minimize` subclass; SimPEG is not a dependency of this package, so
the numerical core of one Gauss-Newton step is provided as functions that
take the sensitivities, the data weights and the regularisation Hessian
directly.  Each mirrors the reference routine named beside it:

* `pcg`            truncated preconditioned CG (optimization.py:405-439);
* `HmSolver`       the H_m SPD solve through `make_solver`: one plan per
                   pattern and right-hand-side width, refactorised per step,
                   residual-checked, SuperLU on the host otherwise
                   (optimization.py:727-761);
* `direct_step`    the exact step in data space by Woodbury
                   (optimization.py:703-725);
* `occam_step`     beta chosen inside the step, every candidate from one
                   eigendecomposition (optimization.py:763-855);
* `cg_step`        CG on (J^T D J + beta H_m) p = -g with the Jacobi
                   preconditioner from the exact diagonal (optimization.py:
                   405-439 and the CG branch of `findSearchDirection`);
* `beta_by_eig`    the initial beta from the two Hessians' largest
                   eigenvalues (SimPEG's BetaEstimate_ByEig);
* `gauss_newton_step`  one step of the L2 misfit at a given beta through an
                   engine's `linearize` (or `Jmatrix`), `Linearization2D`
                   supplying J, J v, J^T v and diag(J^T D J).

`tensor_regularization` builds a SimPEG-style regularisation Hessian on a
tensor mesh (smallness plus first differences between neighbouring cells),
H_m = 2 (alpha_s V + sum_axes alpha_a G_a^T diag(w_a) G_a), for callers that
do not bring SimPEG's; it is not a parity claim on SimPEG's own weights.

Conventions (optimization.py:440-470): the misfit has no 1/2, so for the L2
misfit the gradient weight is W (2 r) and the Gauss-Newton weight D = 2 W^2;
g = J^T (W 2 r) + beta g_m with g_m = H_m (m - m_ref).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch


# ---------------------------------------------------------------------------
# truncated preconditioned CG
# ---------------------------------------------------------------------------
def pcg(apply_H, b, Minv, maxiter, rtol):
    """Solve H x = b from x = 0 by CG preconditioned with the diagonal
    inverse `Minv`; stop at relative residual `rtol`, after `maxiter`
    iterations, or on non-positive curvature.  Returns (x, iterations)."""
    x = torch.zeros_like(b)
    bnorm = float(torch.linalg.norm(b))
    if bnorm == 0.0:
        return x, 0
    r = b.clone()
    z = Minv * r
    p = z.clone()
    rz = float(r @ z)
    it = 0
    for it in range(1, maxiter + 1):
        q = apply_H(p)
        curv = float(p @ q)
        if curv <= 0.0:
            break
        a = rz / curv
        x.add_(p, alpha=a)
        r.add_(q, alpha=-a)
        if float(torch.linalg.norm(r)) <= rtol * bnorm:
            break
        z = Minv * r
        rz_next = float(r @ z)
        p = z + (rz_next / rz) * p
        rz = rz_next
    return x, it


# ---------------------------------------------------------------------------
# the regularisation Hessian's SPD solve
# ---------------------------------------------------------------------------
class HmSolver:
    """H_m^{-1} X for a sparse SPD H_m (scipy) and a dense block X (torch).

    On a GPU device: `make_solver` (this package's supernodal solver), planned
    once per (pattern, width) and refactorised per call; the solution is
    accepted when its relative residual is at most `tol`, otherwise (and on
    the cpu) SuperLU on the host solves it (optimization.py:727-761)."""

    def __init__(self, device="cpu", tol=1e-8):
        self.device = torch.device(device)
        self.tol = tol
        self._plan = None
        self.last = None  # "device" or "host": which path gave the last solution

    def _device_solve(self, H, X):
        from .solver import make_solver
        n, k = H.shape[0], int(X.shape[1])
        c = self._plan
        if (c is None or c["k"] != k or c["n"] != n
                or not np.array_equal(c["indptr"], H.indptr)
                or not np.array_equal(c["indices"], H.indices)):
            crow = torch.as_tensor(H.indptr, dtype=torch.int32, device=self.device)
            col = torch.as_tensor(H.indices, dtype=torch.int32, device=self.device)
            c = {"k": k, "n": n, "indptr": H.indptr.copy(), "indices": H.indices.copy(),
                 "slv": make_solver(crow, col, n, k, device=self.device, dtype=torch.float64,
                                    matrix_type="spd"),
                 "crow": crow.long(), "col": col.long()}
            self._plan = c
        vals = torch.as_tensor(H.data, dtype=torch.float64, device=self.device)
        c["slv"].factorize(vals)
        Y = c["slv"].solve(X.to(device=self.device, dtype=torch.float64).contiguous())
        Ht = torch.sparse_csr_tensor(c["crow"], c["col"], vals, (n, n))
        res = torch.linalg.norm(Ht @ Y - X) / torch.linalg.norm(X)
        if not torch.isfinite(res) or float(res) > self.tol:
            return None
        return Y

    def __call__(self, Hm, X):
        H = sp.csr_matrix(Hm)
        H.sort_indices()
        if self.device.type == "cuda":
            Y = self._device_solve(H, X)
            if Y is not None:
                self.last = "device"
                return Y.to(X.dtype)
        try:
            lu = spla.splu(sp.csc_matrix(H))
        except RuntimeError:
            return None  # singular (alpha_s = 0)
        self.last = "host"
        Y = lu.solve(np.ascontiguousarray(X.detach().cpu().numpy()))
        return torch.as_tensor(Y, dtype=X.dtype, device=X.device)


# ---------------------------------------------------------------------------
# the two inner solves
# ---------------------------------------------------------------------------
def direct_step(J, D, g, Hm, beta, hm_solve):
    """Exact Gauss-Newton step p with (J^T D J + beta H_m) p = -g, solved in
    data space (Woodbury; optimization.py:703-725):

        u = B^-1 g,  K = B^-1 J^T,  S = D^-1 + J K,  p = -(u - K S^-1 J u),

    B = beta H_m.  J: (nD, nM) dense; D: (nD,); g: (nM,).  Returns None when
    B or S is not positive definite."""
    rhs = torch.cat([J.T, g.unsqueeze(1)], dim=1)
    sol = hm_solve(beta * Hm, rhs)
    if sol is None or not torch.isfinite(sol).all():
        return None
    K, u = sol[:, :-1], sol[:, -1]
    S = J @ K
    S.diagonal().add_(1.0 / D)
    Ls, info = torch.linalg.cholesky_ex(S)
    if int(info) != 0:
        return None
    z = torch.cholesky_solve((J @ u).unsqueeze(1), Ls).squeeze(1)
    return -(u - K @ z)


def occam_step(J, D, g_d, g_m, r, beta_now, Hm, hm_solve, phi_d_of_r, w_lin,
               chi_target=1.0, beta_aim=0.8, beta_span=(10.0, 0.1), beta_points=9):
    """Occam's inversion inside one step (Constable et al. 1987; the
    reference's combined Marquardt-Occam search, optimization.py:763-855).

    With J fixed and K0 = H_m^-1 J^T, u_d = H_m^-1 g_d, u_m = H_m^-1 g_m, the
    exact step for any beta is p(beta) = -(u - K0 S^-1 J u / beta) with
    u = u_d / beta + u_m and S = D^-1 + J K0 / beta.  One eigendecomposition
    of M = D^1/2 (J K0) D^1/2 = Q diag(lam) Q^T gives every candidate with
    nD x nD products only.  The linearised misfit phi_d(r + w_lin J p(beta))
    predicts each beta's landing; the largest beta whose prediction reaches
    beta_aim * chi_target * nD is taken (bisection in log beta between the
    grid points), else the grid beta with the lowest prediction.

    Returns (beta, p, predicted phi_d), or None when H_m is singular."""
    nD = J.shape[0]
    rhs = torch.cat([J.T, g_d.unsqueeze(1), g_m.unsqueeze(1)], dim=1)
    sol = hm_solve(Hm, rhs)
    if sol is None or not torch.isfinite(sol).all():
        return None
    K0, ud, um = sol[:, :-2], sol[:, -2], sol[:, -1]
    JK0, Jud, Jum = J @ K0, J @ ud, J @ um
    Dh = torch.sqrt(D)
    M = Dh.unsqueeze(1) * (0.5 * (JK0 + JK0.T)) * Dh.unsqueeze(0)
    lam, Q = torch.linalg.eigh(M)
    lam = torch.clamp(lam, min=0.0)
    cd, cm = Q.T @ (Dh * Jud), Q.T @ (Dh * Jum)

    def parts(b):
        return cd / b + cm, 1.0 / (1.0 + lam / b)

    def Jp(b):
        c, f = parts(b)
        return -((Jud / b + Jum) - (Q @ ((lam / b) * f * c)) / Dh)

    def p_of(b):
        c, f = parts(b)
        z = Dh * (Q @ (f * c))
        return -((ud / b + um) - (K0 @ z) / b)

    def predicted(b):
        return float(phi_d_of_r(r + w_lin * Jp(b)))

    target = beta_aim * chi_target * nD
    hi, lo = beta_span
    betas = beta_now * np.logspace(np.log10(hi), np.log10(lo), beta_points)  # descending
    vals = [predicted(b) for b in betas]
    reach = [i for i, v in enumerate(vals) if v <= target]
    if reach and reach[0] > 0:
        a, c = np.log(betas[reach[0] - 1]), np.log(betas[reach[0]])  # a misses, c reaches
        for _ in range(12):
            mid = 0.5 * (a + c)
            if predicted(np.exp(mid)) <= target:
                c = mid
            else:
                a = mid
        beta = float(np.exp(c))
    elif reach:
        beta = float(betas[0])
    else:
        beta = float(betas[int(np.argmin(vals))])
    p = p_of(beta)
    if not torch.isfinite(p).all():
        return None
    return beta, p, predicted(beta)


def cg_step(Jv, Jtv, D, g, Hm_apply, Hm_diag, beta, jtdj_diag, maxiter=40, rtol=1e-2,
            precond=True):
    """Truncated CG on (J^T D J + beta H_m) p = -g (matrix-free: J v and
    J^T v products), Jacobi-preconditioned with the exact diagonal
    diag(J^T D J) + beta diag(H_m).  Returns (p, iterations)."""
    def H(v):
        return Jtv(D * Jv(v)) + beta * Hm_apply(v)
    if precond:
        diag = jtdj_diag + beta * Hm_diag
        Minv = torch.where(diag > 0, 1.0 / diag, torch.ones_like(diag))
    else:
        Minv = torch.ones_like(g)
    return pcg(H, -g, Minv, maxiter, rtol)


def beta_by_eig(Jv, Jtv, w, Hm, n_model, ratio=100.0, n_iter=4, seed=0, device="cpu"):
    """Initial trade-off beta0 = ratio * lam_max(J^T D J) / lam_max(H_m),
    D = 2 W^2, both largest eigenvalues by `n_iter` power iterations from a
    seeded random vector (SimPEG's BetaEstimate_ByEig, which the reference's
    inversions use with beta0_ratio 10-100)."""
    dev = torch.device(device)
    H = sp.csr_matrix(Hm)
    Ht = torch.sparse_csr_tensor(torch.as_tensor(H.indptr, dtype=torch.int64, device=dev),
                                 torch.as_tensor(H.indices, dtype=torch.int64, device=dev),
                                 torch.as_tensor(H.data, dtype=torch.float64, device=dev),
                                 H.shape)
    w = torch.as_tensor(w, dtype=torch.float64, device=dev)
    D = 2.0 * w ** 2
    g = torch.Generator(device="cpu").manual_seed(seed)

    def lam(apply):
        x = torch.randn(n_model, generator=g, dtype=torch.float64).to(dev)
        x /= torch.linalg.norm(x)
        val = 0.0
        for _ in range(n_iter):
            y = apply(x)
            val = float(x @ y)
            x = y / torch.linalg.norm(y)
        return val

    lam_d = lam(lambda x: Jtv(D * Jv(x)))
    lam_m = lam(lambda x: Ht @ x)
    return ratio * lam_d / lam_m


# ---------------------------------------------------------------------------
# regularisation on a tensor mesh
# ---------------------------------------------------------------------------
def tensor_regularization(mesh, alpha_s=1e-4, alphas=None):
    """SimPEG-style H_m on a tensor mesh (cells x fastest):
    2 (alpha_s V + sum_a alpha_a G_a^T diag(w_a) G_a), V the cell volumes,
    G_a the differences between neighbouring cells along axis a divided by
    their centre distance, w_a the shared face's area times that distance.
    alphas: per-axis weights (default 1)."""
    h = [np.asarray(x, float) for x in mesh.h]
    n = tuple(x.size for x in h)
    dim, N = len(n), int(np.prod(n))
    alphas = [1.0] * dim if alphas is None else list(alphas)
    idx = np.arange(N).reshape(n[::-1]).T  # idx[ix, iy(, iz)] = cell number
    pos = np.indices(n)
    vol = np.empty(N)
    vol[idx.ravel()] = np.prod([h[b][pos[b]] for b in range(dim)], axis=0).ravel()
    H = 2.0 * alpha_s * sp.diags(vol)
    for a in range(dim):
        if n[a] < 2:
            continue
        lo, hi = [slice(None)] * dim, [slice(None)] * dim
        lo[a], hi[a] = slice(0, n[a] - 1), slice(1, n[a])
        c0, c1 = idx[tuple(lo)].ravel(), idx[tuple(hi)].ravel()
        q = np.indices(idx[tuple(lo)].shape)
        dist = (0.5 * (h[a][q[a]] + h[a][q[a] + 1])).ravel()
        area = np.prod([h[b][q[b]] for b in range(dim) if b != a] or [np.ones(q[a].shape)],
                       axis=0).ravel()
        e = np.arange(c0.size)
        G = sp.csr_matrix((np.r_[-1.0 / dist, 1.0 / dist], (np.r_[e, e], np.r_[c0, c1])),
                          shape=(c0.size, N))
        H = H + 2.0 * alphas[a] * (G.T @ sp.diags(area * dist) @ G)
    return sp.csr_matrix(H)


# ---------------------------------------------------------------------------
# one step through an engine
# ---------------------------------------------------------------------------
def data_residual(d, dobs, w, log_data=False):
    """(r, dr/dd): r = w (d - dobs), or for log data r = w |dobs| (ln(sign(dobs)
    d) - ln|dobs|) with dr/dd = w |dobs| / d (a sign flip clamps to 1e-300,
    a huge misfit) (dctorch/optimization.py:103-124)."""
    if log_data:
        if bool((dobs == 0).any()):
            raise ValueError("log_data needs non-zero observed data")
        wl = w * dobs.abs()
        sd = torch.clamp(torch.sign(dobs) * d, min=1e-300)
        return wl * (torch.log(sd) - torch.log(dobs.abs())), wl / d
    return w * (d - dobs), w


def student_t_weights(r, w, nu):
    """(phi_d, w dphi/dr, D) of the weighted residual r = w (d - dobs):
    L2 (nu = 0): ||r||^2, 2 w r, 2 w^2; Student-t (nu > 0):
    (nu + 1) sum log1p(r^2 / nu), its gradient weight w 2 (nu + 1) r /
    (nu + r^2) and the IRLS Gauss-Newton weight 2 (nu + 1) w^2 / (nu + r^2)
    (dctorch/optimization.py:126-131, :621-630)."""
    if nu and nu > 0:
        phi = float((nu + 1.0) * torch.log1p(r ** 2 / nu).sum())
        return phi, w * (2.0 * (nu + 1.0) * r / (nu + r ** 2)), \
            2.0 * (nu + 1.0) / (nu + r ** 2) * w ** 2
    return float((r ** 2).sum()), w * (2.0 * r), 2.0 * w ** 2


def gauss_newton_step(eng, m, dobs, w, Hm, beta, m_ref=None, inner="direct", hm_solve=None,
                      cg_maxiter=40, cg_rtol=1e-2, nu=0.0, log_data=False, **occam_kw):
    """One Gauss-Newton step of phi = phi_d + beta phi_m at m (FP64), with J
    from the engine's linearisation; phi_d = ||r||^2 of the weighted residual
    r (`data_residual`: linear or log data), or the Student-t misfit with
    IRLS weights when nu > 0 ("occam" is L2 only).

    inner: "direct" (Woodbury, explicit J), "occam" (beta searched inside the
    step, explicit J), "cg" (matrix-free).  Returns a dict with the step `p`,
    the gradient `g`, `beta` (the searched one for "occam"), the misfit
    `phi_d` at m and, for "occam", the predicted misfit."""
    dev = torch.device(eng.device)
    m = torch.as_tensor(m, dtype=torch.float64, device=dev)
    dobs = torch.as_tensor(dobs, dtype=torch.float64, device=dev)
    w = torch.as_tensor(w, dtype=torch.float64, device=dev)
    m_ref = torch.zeros_like(m) if m_ref is None else torch.as_tensor(m_ref, dtype=torch.float64,
                                                                      device=dev)
    hm_solve = hm_solve or HmSolver(dev)
    lin = eng.linearize(m)
    with torch.no_grad():
        d = eng.dpred(m).to(torch.float64)
    r, lw = data_residual(d, dobs, w, log_data)
    phi_d, wd, D = student_t_weights(r, lw, nu)
    if nu and nu > 0 and inner == "occam":
        raise ValueError("the Occam beta search is defined for the L2 misfit (nu = 0)")
    H = sp.csr_matrix(Hm)
    Ht = torch.sparse_csr_tensor(torch.as_tensor(H.indptr, dtype=torch.int64, device=dev),
                                 torch.as_tensor(H.indices, dtype=torch.int64, device=dev),
                                 torch.as_tensor(H.data, dtype=torch.float64, device=dev),
                                 H.shape)
    g_m = Ht @ (m - m_ref)
    out = {"phi_d": phi_d, "beta": beta}
    if inner == "cg":
        g_d = lin.Jtvec(wd)
        g = g_d + beta * g_m
        Hd = torch.as_tensor(H.diagonal(), dtype=torch.float64, device=dev)
        p, it = cg_step(lin.Jvec, lin.Jtvec, D, g, lambda v: Ht @ v, Hd, beta,
                        lin.diag_JtDJ(D), cg_maxiter, cg_rtol)
        out.update(p=p, g=g, cg_iterations=it)
        return out
    J = lin.Jmatrix().to(torch.float64)
    g_d = J.T @ wd
    g = g_d + beta * g_m
    if inner == "direct":
        out.update(p=direct_step(J, D, g, H, beta, hm_solve), g=g)
    elif inner == "occam":
        res = occam_step(J, D, g_d, g_m, r, beta, H, hm_solve,
                         lambda rr: (rr ** 2).sum(), lw, **occam_kw)
        if res is None:
            out.update(p=None, g=g)
        else:
            b, p, pred = res
            out.update(beta=b, p=p, g=g_d + b * g_m, predicted_phi_d=pred)
    else:
        raise ValueError(f"inner must be direct|occam|cg, got {inner!r}")
    return out
