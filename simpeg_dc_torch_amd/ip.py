"""Induced polarisation on the DC engines.

Linearised IP (Seigel), the convention of stock SimPEG's
``induced_polarization.Simulation2DNodal`` / ``Simulation3DNodal`` that the
reference restates (dctorch/ip2d.py:1-27, dctorch/ip3d.py:1-24):

    d_IP = scale * J_sigma[sigma_bg] (-diag(sigma_bg)) eta,
    scale = 1 / V_dc for receivers of data_type "apparent_chargeability",
            1 otherwise.

With the DC model m = log(rho) (sigma = exp(-m), d sigma / d m = -sigma)
J_sigma (-diag sigma) is exactly the Jacobian J_m of the DC data in m, so

    d_IP = scale * J_m(m_bg) eta,   grad_eta phi = J_m^T (scale * dphi/dd),

one J v and one J^T v of the DC engine frozen at the background.  On a GPU
engine those are `Linearization2D.Jvec` / `Jtvec` (libdcx.so: dcx_assemble,
dcx_spmm, dcx_solve on the background factorisation, dcx_survey_dpred; and
dcx_survey_adjoint_rhs, dcx_solve, dcx_grad_values, dcx_lmap_transpose);
the background is factorised once at construction.  On the cpu the 2.5-D
case runs the same linearisation's CPU path and the 3-D case the
reference's own chain (fixed-factor SciPy solves, torch sparse products).

`TorchDCIP2D` is the joint non-linear two-state forward of
dctorch/dcip2d.py:1-60: V0 = F(sigma), V1 = F(sigma (1 - eta)),
d_IP = ip_scale (V1 - V0) / V1, d_DC = V1 ("total") or V0
("instantaneous"), eta = eta_max sigmoid(m_eta).  The two states run as two
`TorchDC2D` engines (one plan each, m_eff = m_rho - log(1 - eta)), so both
forward solves and both adjoint solves are the HIP eval+grad path and autograd
differentiates both blocks of the model.

`TorchCR2D` is the complex-resistivity forward of dctorch/cr2d.py
(sigma* = sigma (1 - i m), complex-symmetric solves).  On the cpu it runs
the reference's chain (complex SuperLU through the complex-safe autograd
op).  On a GPU it needs no complex factorisation: A(sigma*) = A_r + i A_i
with A_r = A(sigma) real SPD and A_i = A(-sigma m), and the complex system
is solved by block Gauss-Seidel on the real factor of A_r,

    x_r = A_r^{-1} (b_r + s A_i x_i),   x_i = A_r^{-1} (b_i - s A_i x_r)

(s = +1 for A, -1 for its conjugate), whose error contracts by
rho(A_r^{-1} A_i)^2 <= max|m|^2 per pass; the adjoint is the same
iteration on the conjugate system, and the value gradients are four entry
gathers (`_ComplexResistivity`).  Every step is a libdcx.so call
(dcx_assemble, dcx_factorize, dcx_solve, dcx_spmm, dcx_survey_dpred,
dcx_survey_adjoint_rhs, dcx_grad_values, dcx_lmap_transpose).
"""
import numpy as np
import torch

from .linearization2d import Linearization2D


def chargeability_scale(survey, v_dc):
    """Per-datum scale of stock SimPEG's BaseIPSimulation._scale: 1 / V_dc on
    the data of "apparent_chargeability" receivers, 1 elsewhere (data in
    survey order: sources, then their receivers).  Follows
    dctorch/ip2d.py:105-116."""
    scale = torch.ones_like(v_dc)
    i0 = 0
    for src in survey.source_list:
        for rx in src.receiver_list:
            i1 = i0 + int(rx.nD)
            if getattr(rx, "data_type", "volt") == "apparent_chargeability":
                scale[i0:i1] = 1.0 / v_dc[i0:i1]
            i0 = i1
    if i0 != v_dc.numel():
        raise ValueError(f"the survey lists {i0} data, the engine {v_dc.numel()}")
    return scale


class _LinearIP(torch.autograd.Function):
    """eta -> scale * J eta through a frozen linearisation (any device)."""

    @staticmethod
    def forward(ctx, e, ip):
        ctx.ip = ip
        return ip._scale * ip._jvec(e.detach())

    @staticmethod
    def backward(ctx, g):
        ip = ctx.ip
        return ip._jtvec(ip._scale * g.detach()), None


class _FixedSolve(torch.autograd.Function):
    """x = A^{-1} b with A fixed and factored (symmetric): differentiable in b
    only, the backward one solve on the same factorisation
    (dctorch/ip2d.py:32-48)."""

    @staticmethod
    def forward(ctx, b, solver):
        ctx.solver = solver
        return solver.solve(b.detach())

    @staticmethod
    def backward(ctx, g):
        return ctx.solver.solve(g.contiguous()), None


#: entries per gather chunk of the cpu dA U product (about 64 MB of FP64)
_SPMV_CHUNK = 8_000_000


class _SpmvFixedU(torch.autograd.Function):
    """rhs = dA(vals) U with U fixed: rhs[r] = sum_p vals[p] U[col_p] over the
    pattern rows; forward and backward in chunks, so nothing nnz x nS-sized
    is kept (the role of dctorch/ip3d.py:36-65)."""

    @staticmethod
    def forward(ctx, vals, U, row, col):
        k = U.shape[1]
        rhs = torch.zeros_like(U)
        step = max(_SPMV_CHUNK // max(k, 1), 1)
        for p0 in range(0, vals.numel(), step):
            p1 = min(p0 + step, vals.numel())
            rhs.index_add_(0, row[p0:p1], vals[p0:p1].detach().unsqueeze(1) * U[col[p0:p1]])
        ctx.save_for_backward(U, row, col)
        return rhs

    @staticmethod
    def backward(ctx, grad_rhs):
        U, row, col = ctx.saved_tensors
        k = U.shape[1]
        g = torch.empty(row.numel(), dtype=U.dtype, device=U.device)
        step = max(_SPMV_CHUNK // max(k, 1), 1)
        for p0 in range(0, row.numel(), step):
            p1 = min(p0 + step, row.numel())
            g[p0:p1] = (grad_rhs[row[p0:p1]] * U[col[p0:p1]]).sum(1)
        return g, None, None, None


class _IPBase:
    def _finish(self, survey, v_dc):
        self.v_dc = v_dc
        self._scale = chargeability_scale(survey, v_dc)

    def misfit(self, eta, dobs, w):
        """phi_d = ||w (dpred - dobs)||^2 (no 1/2 factor)."""
        r = w * (self.dpred(eta) - dobs)
        return (r ** 2).sum()


class TorchIP2D(_IPBase):
    """Differentiable linearised-IP dpred / misfit for the 2.5-D nodal
    problem on a fixed mesh, survey and background (dctorch/ip2d.py:51-143).

    m_dc: (nC,) log-rho background on the full mesh; active_cells: optional
    mask of the eta model (inactive eta = 0).  Extra keywords as TorchDC2D
    (`precision`, `refine`, `leaf`, `rhs_basis`).  The wrapped engine `dc`
    holds the background factorisation: calling its dpred refactorises it
    (the linearisation restores the background before its next product)."""

    def __init__(self, mesh, survey, m_dc, nky=12, device="cuda", dtype=torch.float64,
                 active_cells=None, quadrature=None, **engine_kw):
        from .simulation2d import TorchDC2D
        dc = TorchDC2D(mesh, survey, nky=nky, device=device, dtype=dtype,
                       quadrature=quadrature, **engine_kw)
        self.dc = dc
        self.device, self.dtype = dc.device, dc.dtype
        m_bg = torch.as_tensor(np.asarray(m_dc, dtype=float).reshape(-1), dtype=torch.float64,
                               device=dc.device)
        if m_bg.numel() != mesh.nC:
            raise ValueError("m_dc must live on the full mesh")
        self._sig_bg = torch.exp(-m_bg)
        self._lin = Linearization2D(dc, m_bg)
        self._act = None
        if active_cells is not None:
            self._act = torch.as_tensor(np.where(np.asarray(active_cells, bool))[0],
                                        device=dc.device)
        self._finish(survey, self._lin.d.to(torch.float64))

    def _jvec(self, e):
        return self._lin.Jvec(e).to(torch.float64)

    def _jtvec(self, g):
        return self._lin.Jtvec(g).to(torch.float64)

    def dpred(self, eta):
        """IP data (nD,) for eta on the full mesh, or on the active cells
        when built with active_cells.  Autograd-traceable; linear in eta."""
        dev0, dt0 = eta.device, eta.dtype
        e = eta.to(device=self.device, dtype=torch.float64).reshape(-1)
        if self._act is not None:
            e = torch.zeros(self.dc.nC, dtype=torch.float64,
                            device=self.device).index_put((self._act,), e)
        return _LinearIP.apply(e, self).to(device=dev0, dtype=dt0)


class TorchIP3D(_IPBase):
    """Linearised IP for the 3-D nodal problem (dctorch/ip3d.py:68-151).

    m_dc: log-rho background on the active cells when active_cells is given
    (inactive cells at val_inactive), else on the full mesh; eta lives on the
    same cells (inactive eta = 0)."""

    def __init__(self, mesh, survey, m_dc, device="cuda", dtype=torch.float64,
                 active_cells=None, val_inactive=None, bc_type="Robin", **engine_kw):
        from .simulation3d import TorchDC3D
        dc = TorchDC3D(mesh, survey, device=device, dtype=dtype, active_cells=active_cells,
                       val_inactive=val_inactive, bc_type=bc_type, **engine_kw)
        self.dc = dc
        self.device, self.dtype = dc.device, dc.dtype
        m_bg = torch.as_tensor(np.asarray(m_dc, dtype=float).reshape(-1), dtype=torch.float64,
                               device=dc.device)
        nP = mesh.nC if dc._P_inj is None else int(dc._act.numel())
        if m_bg.numel() != nP:
            raise ValueError(f"m_dc must have {nP} entries (the engine's model cells)")
        with torch.no_grad():
            self._sig_bg = torch.exp(-dc._model_to_full(m_bg).to(torch.float64))
        if dc._gpu is not None:
            self._lin = Linearization2D(dc, m_bg)
            v_dc = self._lin.d.to(torch.float64)
        else:
            v_dc = self._init_cpu()
        self._finish(survey, v_dc)

    # the reference's CPU chain, with the background factorised once
    def _init_cpu(self):
        dc = self.dc
        with torch.no_grad():
            vals = torch.sparse.mm(dc.L3, self._sig_bg.unsqueeze(1)).squeeze(1)
            dc.solver.factorize(vals)
            self._U = dc.solver.solve(dc.q)
            self._row = dc.solver.row.to(torch.int64)
            self._col = dc.solver.col.to(torch.int64)
            return (dc.Pd @ self._U)[dc.iD, dc.sidx]

    def _dpred_cpu(self, e_full):
        dc = self.dc
        dvals = torch.sparse.mm(dc.L3, (-self._sig_bg * e_full).unsqueeze(1)).squeeze(1)
        rhs = _SpmvFixedU.apply(dvals, self._U, self._row, self._col)
        dU = _FixedSolve.apply(-rhs, dc.solver)
        return (dc.Pd @ dU)[dc.iD, dc.sidx] * self._scale

    def _jvec(self, e):
        return self._lin.Jvec(e).to(torch.float64)

    def _jtvec(self, g):
        return self._lin.Jtvec(g).to(torch.float64)

    def dpred(self, eta):
        """IP data (nD,) for eta on the engine's model cells.  Autograd-
        traceable; linear in eta."""
        dev0, dt0 = eta.device, eta.dtype
        e = eta.to(device=self.device, dtype=torch.float64).reshape(-1)
        if self.dc._gpu is not None:
            d = _LinearIP.apply(e, self)
        else:
            if self.dc._P_inj is not None:
                e = torch.zeros(self.dc.nC, dtype=torch.float64,
                                device=self.device).index_put((self.dc._act,), e)
            d = self._dpred_cpu(e)
        return d.to(device=dev0, dtype=dt0)


class TorchDCIP2D:
    """Joint non-linear (rho, eta) 2.5-D DC-IP forward (dctorch/dcip2d.py:
    62-193): model m = cat(log rho, logit(eta / eta_max)) on the model cells;
    `dpred(m)` -> (d_DC, d_IP); `misfit(m, dobs_dc, w_dc, dobs_ip, w_ip, lam)`.
    Receivers are plain "volt" (the ratio is formed here).  Extra keywords
    go to both TorchDC2D engines."""

    def __init__(self, mesh, survey, nky=12, device="cuda", dtype=torch.float64,
                 active_cells=None, val_inactive=None, idx_dc=None, idx_ip=None, eta_max=1.0,
                 ip_scale=1.0, dc_state="total", quadrature=None, **engine_kw):
        from .simulation2d import TorchDC2D
        if dc_state not in ("total", "instantaneous"):
            raise ValueError(f"dc_state must be total|instantaneous, got {dc_state!r}")
        kw = dict(nky=nky, device=device, dtype=dtype, active_cells=active_cells,
                  val_inactive=val_inactive, quadrature=quadrature, **engine_kw)
        self.dc = TorchDC2D(mesh, survey, **kw)            # instantaneous state sigma
        self.dc_polarized = TorchDC2D(mesh, survey, **kw)  # sigma (1 - eta)
        dc = self.dc
        self.device, self.dtype = dc.device, dc.dtype
        self.nky, self.nD = dc.nky, dc.nD
        self.nP = dc.nC if dc._P_inj is None else int(dc._act.numel())
        self.eta_max = float(eta_max)
        self.ip_scale = float(ip_scale)
        self.dc_state = dc_state

        def idx(a):
            if a is None:
                return torch.arange(dc.nD, dtype=torch.int64, device=self.device)
            return torch.as_tensor(np.asarray(a, dtype=np.int64), device=self.device)

        self.idx_dc, self.idx_ip = idx(idx_dc), idx(idx_ip)
        self.nD_dc, self.nD_ip = int(self.idx_dc.numel()), int(self.idx_ip.numel())

    def split(self, m):
        """(m_rho, m_eta) blocks of the stacked model."""
        m = m.reshape(-1)
        if m.numel() != 2 * self.nP:
            raise ValueError(f"model must be 2*{self.nP} long, got {m.numel()}")
        return m[:self.nP], m[self.nP:]

    def eta_of(self, m_eta):
        """Physical eta from the logit block."""
        return self.eta_max * torch.sigmoid(m_eta)

    def m_eta_of(self, eta):
        """Inverse of eta_of (numpy or tensor)."""
        if not torch.is_tensor(eta):
            e = np.clip(np.asarray(eta, dtype=float) / self.eta_max, 1e-12, 1 - 1e-12)
            return np.log(e / (1 - e))
        e = (eta / self.eta_max).clamp(1e-12, 1 - 1e-12)
        return torch.log(e / (1 - e))

    def dpred_from_fields(self, m_rho, eta):
        """(d_DC, d_IP) from log-rho and physical eta on the model cells."""
        dev0, dt0 = m_rho.device, m_rho.dtype
        mr = m_rho.to(device=self.device, dtype=torch.float64).reshape(-1)
        e = eta.to(device=self.device, dtype=torch.float64).reshape(-1)
        V0 = self.dc.dpred(mr)
        V1 = self.dc_polarized.dpred(mr - torch.log1p(-e))   # rho / (1 - eta)
        d_dc = (V1 if self.dc_state == "total" else V0)[self.idx_dc]
        v1 = V1[self.idx_ip]
        d_ip = self.ip_scale * (v1 - V0[self.idx_ip]) / v1
        return d_dc.to(device=dev0, dtype=dt0), d_ip.to(device=dev0, dtype=dt0)

    def dpred(self, m):
        """(d_DC, d_IP) for the stacked model m = cat(log rho, logit eta)."""
        m_rho, m_eta = self.split(m)
        return self.dpred_from_fields(m_rho, self.eta_of(m_eta))

    def misfit(self, m, dobs_dc, w_dc, dobs_ip, w_ip, lam=1.0):
        """phi_d^DC + lam phi_d^IP, each ||w (dpred - dobs)||^2 (no 1/2)."""
        d_dc, d_ip = self.dpred(m)
        r0 = w_dc * (d_dc - dobs_dc)
        r1 = w_ip * (d_ip - dobs_ip)
        return (r0 ** 2).sum() + lam * (r1 ** 2).sum()


def _cr_solve(g, vim, br, bi, sign, k, tol=1e-13, maxit=60):
    """(x_r, x_i) with (A_r + i sign A_i)(x_r + i x_i) = b_r + i b_i, A_r the
    engine's factorised systems (elimination order, every wavenumber), A_i
    the values vim, by block Gauss-Seidel on A_r (module docstring).  b_r,
    b_i: (n, k) shared or (nbatch, n, k); None is zero.  Stops when both
    parts change by at most tol (relative) in one pass."""
    from . import _lib
    sv = g.solver
    nb, n = sv.nbatch, sv.n

    def full(b):
        if b is None:
            return torch.zeros(nb, n, k, dtype=torch.float64, device=vim.device)
        return b.expand(nb, n, k) if b.dim() == 2 else b

    def spmm(x, alpha):
        y = torch.empty_like(x)
        _lib.check(g.lib.dcx_spmm(sv._h, _lib.ptr(vim), _lib.ptr(x), _lib.ptr(y), int(k),
                                  float(alpha), _lib.DCX_PERMUTED, g.stream()), "spmm")
        return y
    br, bi = full(br), full(bi)
    xi = torch.zeros(nb, n, k, dtype=torch.float64, device=vim.device)
    xr = None
    for it in range(maxit):
        rr = (br + spmm(xi, sign)).contiguous() if it else br.contiguous()
        xr_new = g.solve_perm(rr, k, sparse=False, shared=False)
        xi_new = g.solve_perm((bi + spmm(xr_new, -sign)).contiguous(), k, sparse=False,
                              shared=False)
        if xr is not None:
            dr = float(torch.linalg.norm(xr_new - xr) / torch.linalg.norm(xr_new))
            di = float(torch.linalg.norm(xi_new - xi) / max(float(torch.linalg.norm(xi_new)),
                                                              1e-300))
            if max(dr, di) <= tol:
                return xr_new, xi_new
        xr, xi = xr_new, xi_new
    raise RuntimeError(f"complex-resistivity iteration did not converge in {maxit} passes "
                       "(it contracts by max|m|^2 per pass: |m| must stay below 1)")


class _ComplexResistivity(torch.autograd.Function):
    """(sigma, s2 = -sigma m) -> (Re V*, Im V*) on the GPU: V* = sum_k w_k
    P A_k(sigma*)^{-1} q with A(sigma*) = A(sigma) + i A(s2).  Backward: the
    adjoint fields lam = conj(A)^{-1} (G_r + i G_i), G the projection's
    transpose applied to the data gradients; then per entry
    dL/dA_r = -(lam_r u_r + lam_i u_i), dL/dA_i = lam_r u_i - lam_i u_r
    (summed over sources, times w_k), and the transposed map."""

    @staticmethod
    def forward(ctx, sig, s2, cr):
        g = cr.dc._gpu
        nS, nD = cr.dc.nS, cr.dc.nD
        with torch.cuda.device(cr.device):
            vre, vim = g.assemble(sig), g.assemble(s2)
            g.solver.factorize(vre)
            ur, ui = _cr_solve(g, vim, g.q_perm, None, 1.0, nS)
            Vr = torch.empty(nD, dtype=torch.float64, device=cr.device)
            Vi = torch.empty_like(Vr)
            g.dpred(ur, nS, 0, nS, Vr)
            g.dpred(ui, nS, 0, nS, Vi)
        ctx.cr, ctx.stamp = cr, g.solver.stamp
        ctx.save_for_backward(vre, vim, ur, ui)
        return Vr, Vi

    @staticmethod
    def backward(ctx, gVr, gVi):
        cr = ctx.cr
        g = cr.dc._gpu
        nS = cr.dc.nS
        vre, vim, ur, ui = ctx.saved_tensors
        with torch.cuda.device(cr.device):
            if g.solver.stamp != ctx.stamp:
                g.solver.factorize(vre)

            def rhs(gv):
                return None if gv is None else \
                    g.adjoint_rhs(gv.to(torch.float64).contiguous(), nS, 0, nS)
            lr, li = _cr_solve(g, vim, rhs(gVr), rhs(gVi), -1.0, nS)
            gvv = lambda y, x: g.solver.grad_values(y, x, scale=g.wq, permuted=True)
            gv_re = gvv(lr, ur) + gvv(li, ui)
            gv_im = gvv(li, ur) - gvv(lr, ui)
            minus_one = -torch.ones(cr.dc.nC, dtype=torch.float64, device=cr.device)
            return g.lmap_t(gv_re, minus_one), g.lmap_t(gv_im, minus_one), None


class TorchCR2D:
    """Complex-resistivity 2.5-D nodal forward (dctorch/cr2d.py:13-108):
    sigma* = sigma_DC (1 - i m), one complex-symmetric solve per wavenumber,
    d_rho = |V*| ("magnitude") or Re V* ("real"), d_ip = ip_scale Im V* /
    Re V*; model cat(log rho_DC, m) on the model cells.

    On the cpu the complex-symmetric solves are SciPy SuperLU (as the
    reference's cpu path), through the complex-safe autograd op
    `sparse_solve_batch`.  On a GPU the forward and its gradient are
    `_ComplexResistivity`: block Gauss-Seidel on the real FP64 factor of
    A(sigma) (converges for |m| < 1, by max|m|^2 per pass)."""

    def __init__(self, mesh, survey, nky=12, device="cuda", dtype=torch.complex128,
                 active_cells=None, val_inactive=None, ip_scale=1.0, dc_datum="magnitude",
                 quadrature=None):
        from .simulation2d import TorchDC2D
        from .solver import make_batch_solver
        if dc_datum not in ("magnitude", "real"):
            raise ValueError(f"dc_datum must be magnitude|real, got {dc_datum!r}")
        rdt = torch.float64 if dtype == torch.complex128 else torch.float32
        dc = TorchDC2D(mesh, survey, nky=nky, device=device, dtype=rdt,
                       active_cells=active_cells, val_inactive=val_inactive,
                       quadrature=quadrature)
        self.dc = dc
        self.device, self.dtype, self.rdtype = dc.device, dtype, rdt
        self.nky, self.nD = dc.nky, dc.nD
        self.nP = dc.nC if dc._P_inj is None else int(dc._act.numel())
        self.ip_scale = float(ip_scale)
        self.dc_datum = dc_datum
        if dc._gpu is not None:
            if dtype != torch.complex128:
                raise TypeError("TorchCR2D on a GPU runs in complex128 (FP64 real factor)")
            self.solver = dc.solver
            return
        self.solver = make_batch_solver(dc._crow, dc._ccol, dc.nN, dc.nS, dc.nky,
                                        device=self.device, dtype=dtype,
                                        matrix_type="symmetric")
        self.q = dc.q.to(dtype)

    def _inject_ip(self, m_ip):
        if self.dc._P_inj is None:
            return m_ip
        return torch.zeros(self.dc.nC, dtype=m_ip.dtype,
                           device=self.device).index_put((self.dc._act,), m_ip)

    def split(self, m):
        m = m.reshape(-1)
        if m.numel() != 2 * self.nP:
            raise ValueError(f"model must be 2*{self.nP} long, got {m.numel()}")
        return m[:self.nP], m[self.nP:]

    def dpred_from_fields(self, m_rho, m_ip):
        """(d_rho, d_ip) from log-rho_DC and chargeability m (model cells)."""
        from .solver import sparse_solve_batch
        dc = self.dc
        dev0, dt0 = m_rho.device, m_rho.dtype
        mr = m_rho.to(device=self.device, dtype=self.rdtype).reshape(-1)
        mi = m_ip.to(device=self.device, dtype=self.rdtype).reshape(-1)
        sig = torch.exp(-dc._model_to_full(mr))
        chg = self._inject_ip(mi)
        if dc._gpu is not None:
            Vr, Vi = _ComplexResistivity.apply(sig, -sig * chg, self)
            d_rho = torch.sqrt(Vr ** 2 + Vi ** 2) if self.dc_datum == "magnitude" else Vr
            d_ip = self.ip_scale * Vi / Vr
            return d_rho.to(device=dev0, dtype=dt0), d_ip.to(device=dev0, dtype=dt0)
        vre = torch.sparse.mm(dc.Lstack, sig.unsqueeze(1)).view(dc.nky, dc.nnzA)
        vim = torch.sparse.mm(dc.Lstack, (-sig * chg).unsqueeze(1)).view(dc.nky, dc.nnzA)
        U = sparse_solve_batch(torch.complex(vre, vim), self.q, self.solver)
        Ubar = (dc.wq.to(self.dtype).view(-1, 1, 1) * U).sum(0)
        Vr = torch.sparse.mm(dc.Pd, Ubar.real)[dc.iD, dc.sidx]
        Vi = torch.sparse.mm(dc.Pd, Ubar.imag)[dc.iD, dc.sidx]
        d_rho = torch.sqrt(Vr ** 2 + Vi ** 2) if self.dc_datum == "magnitude" else Vr
        d_ip = self.ip_scale * Vi / Vr
        return d_rho.to(device=dev0, dtype=dt0), d_ip.to(device=dev0, dtype=dt0)

    def dpred(self, m):
        """(d_rho, d_ip) for the stacked model m = cat(log rho_DC, m)."""
        return self.dpred_from_fields(*self.split(m))

    def misfit(self, m, dobs_rho, w_rho, dobs_ip, w_ip, lam=1.0):
        """phi_d^rho + lam phi_d^IP (no 1/2)."""
        d_r, d_i = self.dpred(m)
        return ((w_rho * (d_r - dobs_rho)) ** 2).sum() + lam * ((w_ip * (d_i - dobs_ip)) ** 2).sum()
