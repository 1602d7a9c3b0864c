"""Differentiable 2.5-D DC-resistivity forward, drop-in for dctorch.TorchDC2D.

API (dctorch/simulation2d.py:34-534): ``TorchDC2D(mesh, survey, nky=12,
device, dtype, active_cells, val_inactive, chunk, singularity_removal,
rho_primary, quadrature)`` with ``dpred``, ``misfit``, ``misfit_and_grad``,
``Jtvec``, ``Jmatrix``, ``Jvec``/``linearize`` and the attributes ``nC, nN,
nS, nD, nky, nnzA, kys, wq, q, Pd, sidx, iD, Lstack, solver,
quadrature_info``.  The model is m = log(rho) (sigma = exp(-m)); the misfit
has no 1/2.

On a GPU device one evaluation is, all in libdcx.so kernels:

1. sigma -> A-values of every wavenumber (factored assembly map);
2. numeric multifrontal factorisation of all nky systems (one launch per
   tree level, all wavenumbers in each launch);
3. U_k = A_k^{-1} q for all wavenumbers and sources in one batched solve (the
   forward sweep skips the subtrees holding no electrode);
4. d = (Pd sum_k w_k U_k)[iD, sidx] by a direct gather per datum.

The reverse pass (a torch.autograd.Function) reuses the factor:
G = Pd^T (dL/dd) per source, Lam_k = A_k^{-1} G, then
gv[k, e] = -w_k sum_s Lam_k[row_e, s] U_k[col_e, s] and
dL/dm = -sigma * L^T gv.  The fields are kept in the elimination order
throughout, so no permutation pass touches them.

``precision="mixed"`` factors and sweeps in FP32 and recovers FP64 accuracy
by residual refinement (FP64 residual on the FP64 values); ``"fp64"`` (the
default) is FP64 end to end.  On the cpu device the reference's own chain
runs: torch sparse assembly and SciPy SuperLU (dctorch/solver.py:314-374).
"""
import ctypes as C
import os
import warnings

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .operators2d import Problem2DMaps
from .quadrature import (amplification, electrode_distance_range, k0_quadrature,
                         simpeg_quadrature, transform_error)
from .linearization2d import Linearization2D
from .solver import (HipBatchSolver, make_batch_solver, sparse_solve_batch,
                     sparse_solve_batch_prefactored)


def _s2t(M, device, dtype):
    M = sp.coo_matrix(M)
    return torch.sparse_coo_tensor(
        torch.tensor(np.vstack((M.row, M.col)), dtype=torch.int64, device=device),
        torch.tensor(M.data, dtype=dtype, device=device), M.shape).coalesce()


class _GpuState:
    """Device objects of one engine: plan, assembly map, survey map."""

    def __init__(self, eng, maps, precision, refine, leaf):
        dev = eng.device
        self.solver = HipBatchSolver(
            torch.tensor(maps.crow), torch.tensor(maps.col), maps.nN, maps.nS, eng.nky,
            device=dev, dtype=torch.float64, precision=precision, refine=refine,
            coords=maps.node_coords(), leaf=leaf)
        lib = self.lib = self.solver._lib
        rows = maps.electrode_rows()
        self.solver.set_rhs_support(rows)
        self.perm = torch.tensor(self.solver.perm().astype(np.int64), device=dev)
        # assembly map
        ky2, ky2p = _lib.f64(eng.kys ** 2)
        ptr, ptrp = _lib.i32(maps.map_ptr)
        cell, cellp = _lib.i32(maps.map_cell)
        le, lep = _lib.f64(maps.map_le)
        ln, lnp = _lib.f64(maps.map_ln)
        be, bep = _lib.i32(maps.bc_e)
        bcl, bclp = _lib.i32(maps.bc_cell)
        bco, bcop = _lib.f64(maps.bc_coef)
        h = C.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(lib.dcx_lmap_create(maps.nnz, maps.nC, eng.nky, ky2p, ptrp, cellp, lep,
                                           lnp, int(be.size), bep, bclp, bcop, C.byref(h)),
                       "lmap_create")
        self.lmap = h
        # survey map
        Pd = maps.Pd
        pp, ppp = _lib.i32(Pd.indptr)
        pc, pcp = _lib.i32(Pd.indices)
        pv, pvp = _lib.f64(Pd.data)
        si, sip = _lib.i32(maps.sidx)
        h2 = C.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(lib.dcx_survey_create(self.solver._h, maps.nD, maps.nS, ppp, pcp, pvp,
                                             sip, C.byref(h2)), "survey_create")
        self.survey = h2
        # sources in elimination order, dense (nN x nS)
        q = maps.q.tocsr()[self.solver.perm()].toarray()
        self.q_perm = torch.tensor(q, dtype=torch.float64, device=dev)
        self.wq = torch.tensor(eng.wq_np, dtype=torch.float64, device=dev)
        # mixed precision keeps the eval+grad fields in FP32 (the factor's own
        # precision): sums over data, wavenumbers and sources stay FP64
        self.fields32 = precision == "mixed"
        self._rhs_blocks = {}
        if getattr(eng, "rhs_basis", "sources") == "electrodes":
            self._electrode_basis(maps, dev)

    def _electrode_basis(self, maps, dev):
        """Electrode-basis data: E (elimination order, dense n x ne) as the
        right-hand side, its support rows and weights for E^T G, and the
        source / receiver combinations S (ne x nS) and R (ne x nD)."""
        E, S, R = maps.electrode_basis()
        Ep = E.tocsr()[self.solver.perm()]
        self.ne = E.shape[1]
        self.E_perm = torch.tensor(Ep.toarray(), dtype=torch.float64, device=dev)
        sup = np.unique(Ep.tocoo().row)
        self.e_sup = torch.tensor(sup, dtype=torch.int64, device=dev)
        self.E_supT = torch.tensor(Ep[sup].toarray().T.copy(), dtype=torch.float64, device=dev)
        self.S = torch.tensor(S.toarray(), dtype=torch.float64, device=dev)
        self.R = torch.tensor(R.toarray(), dtype=torch.float64, device=dev)
        self.sidx = torch.tensor(maps.sidx, dtype=torch.int64, device=dev)

    def stream(self):
        return self.solver._stream()

    def assemble(self, sigma):
        vals = torch.empty(self.solver.nbatch, self.solver.col.numel(), dtype=torch.float64,
                           device=sigma.device)
        _lib.check(self.lib.dcx_assemble(self.lmap, _lib.ptr(sigma), _lib.ptr(vals),
                                         self.stream()), "assemble")
        return vals

    def lmap_t(self, gv, sigma):
        out = torch.empty_like(sigma)
        _lib.check(self.lib.dcx_lmap_transpose(self.lmap, _lib.ptr(gv), _lib.ptr(sigma),
                                               _lib.ptr(out), self.stream()), "lmap_transpose")
        return out

    def solve_perm(self, b, k, sparse=True, shared=True, f32=False, defer=False):
        """Solve in elimination order; f32: FP32 solution block (mixed plans);
        defer: the check's verdict is left to solver.check_wait()."""
        sv = self.solver
        x = torch.empty(sv.nbatch, sv.n, k, dtype=torch.float32 if f32 else torch.float64,
                        device=b.device)
        flags = _lib.DCX_PERMUTED | (_lib.DCX_SHARED_RHS if shared else 0) | \
            (_lib.DCX_SPARSE_RHS if sparse else 0) | (_lib.DCX_F32_FIELDS if f32 else 0)
        with sv._lock:
            sv.run_solve(b, x, k, flags, defer=defer)
        return x

    def factor_solve_perm(self, vals, b, k, f32=False, defer=False, split=False):
        """Factor every wavenumber and solve the shared, sparse right-hand
        sides b in elimination order, the two overlapped (dcx_factor_solve)."""
        sv = self.solver
        x = torch.empty(sv.nbatch, sv.n, k, dtype=torch.float32 if f32 else torch.float64,
                        device=b.device)
        flags = _lib.DCX_PERMUTED | _lib.DCX_SHARED_RHS | _lib.DCX_SPARSE_RHS | \
            (_lib.DCX_F32_FIELDS if f32 else 0) | (_lib.DCX_SPLIT_BWD if split else 0)
        v = vals.contiguous()
        with sv._lock:
            sv.run_factor_solve(v, b, x, k, flags, defer=defer)
        return x

    def dpred(self, U, k, s0, s1, d):
        flags = _lib.DCX_PERMUTED | (_lib.DCX_F32_FIELDS if U.dtype == torch.float32 else 0)
        _lib.check(self.lib.dcx_survey_dpred(self.survey, _lib.ptr(U), int(k),
                                             self.solver.nbatch, _lib.ptr(self.wq), int(s0),
                                             int(s1), _lib.ptr(d), flags,
                                             self.stream()), "survey_dpred")

    def adjoint_rhs(self, v, k, s0, s1, reuse=False):
        """G = Pd^T v for the sources [s0, s1) (n x k, elimination order).
        reuse: the engine's block for this window, cleared once and kept --
        every call writes the same target entries and the rest stays zero, so
        the block (0.2 GB on the 25k-cell line) is not cleared on each call;
        the previous G of that window is overwritten."""
        flags = _lib.DCX_PERMUTED
        if reuse:
            key = (int(k), int(s0), int(s1))
            G = self._rhs_blocks.get(key)
            if G is None:
                G = torch.zeros(self.solver.n, k, dtype=torch.float64, device=v.device)
                self._rhs_blocks = {key: G}   # the last window only
            flags |= _lib.DCX_SPARSE_RHS
        else:
            G = torch.empty(self.solver.n, k, dtype=torch.float64, device=v.device)
        _lib.check(self.lib.dcx_survey_adjoint_rhs(self.survey, _lib.ptr(v), _lib.ptr(G),
                                                   int(k), int(s0), int(s1), flags,
                                                   self.stream()),
                   "survey_adjoint_rhs")
        return G

    def free(self):
        if getattr(self, "lmap", None):
            self.lib.dcx_lmap_destroy(self.lmap)
            self.lmap = None
        if getattr(self, "survey", None):
            self.lib.dcx_survey_destroy(self.survey)
            self.survey = None
        self.solver.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _sources_forward(eng, m_full):
    """dpred through the source columns: one factor_solve per wavenumber batch
    with the nS source right-hand sides.  Returns (d, saved)."""
    g = eng._gpu
    sig = torch.exp(-m_full)
    vals = g.assemble(sig)
    U = g.factor_solve_perm(vals, g.q_perm, eng.nS, f32=g.fields32)
    d = torch.empty(eng.nD, dtype=torch.float64, device=eng.device)
    g.dpred(U, eng.nS, 0, eng.nS, d)
    return d, (sig, vals, U)


def _sources_backward(eng, saved, stamp, gd):
    """dphi/dm_full from dphi/dd: one adjoint solve and the entry gathers."""
    sig, vals, U = saved
    g = eng._gpu
    if g.solver.stamp != stamp:      # someone refactorised since
        g.solver.factorize(vals)
    v = gd.to(torch.float64).contiguous()
    G = g.adjoint_rhs(v, eng.nS, 0, eng.nS)
    Lam = g.solve_perm(G, eng.nS, f32=g.fields32, defer=True)
    gv = g.solver.grad_values(Lam, U, scale=g.wq, permuted=True)
    if g.solver.check_wait():        # systems of Lam were solved again in FP64
        gv = g.solver.grad_values(Lam, U, scale=g.wq, permuted=True)
    return g.lmap_t(gv, sig)


def _electrodes_forward(eng, m_full):
    """dpred through the electrode columns (see _ForwardElectrodes2D)."""
    g = eng._gpu
    sig = torch.exp(-m_full)
    vals = g.assemble(sig)
    G = g.factor_solve_perm(vals, g.E_perm, g.ne, f32=g.fields32)
    Ke = torch.matmul(g.E_supT, G[:, g.e_sup, :].to(torch.float64))  # (nky, ne, ne)
    K = (g.wq.view(-1, 1, 1) * Ke).sum(0)
    d = (g.R * (K @ g.S)[:, g.sidx]).sum(0)
    return d, (sig, vals, G)


def _rows_times(g, G, Wt):
    """G_b Wt for every system b (G: nbatch x n x ne, Wt: ne x ne): one
    dcx_rows_mix launch over all rows when ne <= 128, else a batched GEMM."""
    nb, n, ne = G.shape
    W = Wt.to(G.dtype)
    if ne > 128:
        return torch.matmul(G, W.contiguous())
    W = W.unsqueeze(0).expand(nb, ne, ne).contiguous()
    H = torch.empty_like(G)
    seg = np.array([0, n], np.int32)
    flags = _lib.DCX_F32_FIELDS if G.dtype == torch.float32 else 0
    _lib.check(g.lib.dcx_rows_mix(_lib.ptr(G), _lib.ptr(W), _lib.ptr(H), nb, n, ne, 1,
                                  seg.ctypes.data_as(_lib._i32p), flags, g.stream()),
               "rows_mix")
    return H


def _electrodes_backward(eng, saved, stamp, gd):
    sig, vals, G = saved
    g = eng._gpu
    if g.solver.stamp != stamp:      # refactorised since: solve again
        g.solver.factorize(vals)
        G = g.solve_perm(g.E_perm, g.ne, f32=g.fields32)
    v = gd.to(torch.float64)
    T = torch.zeros(g.ne, eng.nS, dtype=torch.float64, device=eng.device)
    T.index_add_(1, g.sidx, g.R * v)
    M = T @ g.S.T
    H = _rows_times(g, G, M.T)
    gv = g.solver.grad_values(G, H, scale=g.wq, permuted=True)
    return g.lmap_t(gv, sig)


def _field_chunk(eng, want_grad):
    """Source columns per pass of the fused eval+grad: all of them when the
    fields of one pass (U and, with the gradient, Lambda: nky x n x nS each)
    fit the budget -- the engine's fields_max_gib, else half of the free
    device memory -- otherwise as many as fit (fields recomputed per chunk
    instead of stored, the reference's memory bound by source chunks,
    dctorch/simulation2d.py:345-375)."""
    g = eng._gpu
    s = 4 if g.fields32 else 8
    per_col = (2 if want_grad else 1) * eng.nky * eng.nN * s + eng.nN * 8
    budget = getattr(eng, "fields_max_gib", None)
    if budget is None:
        free, _ = torch.cuda.mem_get_info(eng.device)
        budget_b = 0.5 * free
    else:
        budget_b = float(budget) * 2 ** 30
    width = int(budget_b // per_col)
    if eng.chunk:
        width = min(width, int(eng.chunk))
    return max(1, min(eng.nS, width))


def _misfit_body_chunked(eng, m_full, dobs, w, want_grad, width):
    """_misfit_body one source chunk of `width` columns at a time: one
    factorisation, then per chunk the forward solve, the projection of its
    data, its adjoint solve and the gather accumulated over the chunks."""
    g = eng._gpu
    sv = g.solver
    sig = torch.exp(-m_full)
    vals = g.assemble(sig)
    sv.factorize(vals)
    d = torch.empty(eng.nD, dtype=torch.float64, device=eng.device)
    gv = torch.zeros(eng.nky, eng.nnzA, dtype=torch.float64, device=eng.device) \
        if want_grad else None
    for s0 in range(0, eng.nS, width):
        s1 = min(s0 + width, eng.nS)
        k = s1 - s0
        U = g.solve_perm(g.q_perm[:, s0:s1].contiguous(), k, f32=g.fields32)
        g.dpred(U, k, s0, s1, d)
        if want_grad:
            rows, _, _ = eng._rows_of(s0, s1)
            v = torch.zeros(eng.nD, dtype=torch.float64, device=eng.device)
            v[rows] = (2.0 * w[rows]) * (w[rows] * (d[rows] - dobs[rows]))
            Lam = g.solve_perm(g.adjoint_rhs(v, k, s0, s1), k, f32=g.fields32)
            gv += sv.grad_values(Lam, U, scale=g.wq, permuted=True)
            del Lam
        del U
    r = w * (d - dobs)
    return (r ** 2).sum(), d, (g.lmap_t(gv, sig) if want_grad else None), r


def _drive(steps):
    """Run a generator of `_misfit_steps` form to its end; its return value."""
    try:
        while True:
            next(steps)
    except StopIteration as e:
        return e.value


def _misfit_body(eng, m_full, dobs, w, want_grad):
    """`_misfit_steps` run to its end: (phi, d, grad_full, r)."""
    return _drive(_misfit_steps(eng, m_full, dobs, w, want_grad))


def _misfit_steps(eng, m_full, dobs, w, want_grad):
    """phi = ||w (dpred - dobs)||^2 and, with want_grad, dphi/dm_full, in one
    pass on the device: the forward solve, the projection, the misfit, the
    adjoint solve and the gathers, with the solves' checks deferred and
    waited for where their results are used (the forward's before the adjoint
    solve, the adjoint's at the end), so the host never waits for the device
    mid-pass; a system refactorised in FP64 by a failing check has everything
    computed from its fields computed again.  A generator: it yields once,
    when every kernel of the pass is queued and the host is about to wait for
    the solves' checks, so that a caller driving several engines can queue
    all of their passes before it waits for any (it resumes each on the
    stream and device it started on; no context is held across the yield).
    Its return value is (phi, d, grad_full, r) with r = w (d - dobs)."""
    g = eng._gpu
    sv = g.solver
    if eng.rhs_basis != "electrodes":
        width = _field_chunk(eng, want_grad)
        if width < eng.nS:
            return _misfit_body_chunked(eng, m_full, dobs, w, want_grad, width)
    sig = torch.exp(-m_full)
    vals = g.assemble(sig)
    if eng.rhs_basis == "electrodes":
        G = g.factor_solve_perm(vals, g.E_perm, g.ne, f32=g.fields32, defer=True)

        def rest():
            Ke = torch.matmul(g.E_supT, G[:, g.e_sup, :].to(torch.float64))  # (nky, ne, ne)
            K = (g.wq.view(-1, 1, 1) * Ke).sum(0)
            d = (g.R * (K @ g.S)[:, g.sidx]).sum(0)
            r = w * (d - dobs)
            if not want_grad:
                return d, r, None
            T = torch.zeros(g.ne, eng.nS, dtype=torch.float64, device=eng.device)
            T.index_add_(1, g.sidx, g.R * ((2.0 * r) * w))
            H = _rows_times(g, G, (T @ g.S.T).T)
            return d, r, sv.grad_values(G, H, scale=g.wq, permuted=True)
        d, r, gv = rest()
        yield
        if sv.check_wait():
            d, r, gv = rest()
    else:
        # the forward solve's backward sweep is split (DCX_SPLIT_BWD): the
        # receivers' supernodes first, which is all the projection reads; the
        # rest overlaps the misfit and the adjoint solve and is joined before
        # the gather.  Both checks are waited for after the gather is queued.
        U = g.factor_solve_perm(vals, g.q_perm, eng.nS, f32=g.fields32, defer=True,
                                split=want_grad)
        d = torch.empty(eng.nD, dtype=torch.float64, device=eng.device)

        def project():
            g.dpred(U, eng.nS, 0, eng.nS, d)
            r = w * (d - dobs)
            G = g.adjoint_rhs(((2.0 * r) * w).contiguous(), eng.nS, 0, eng.nS, reuse=True) \
                if want_grad else None
            return r, G

        def adjoint(G):
            Lam = g.solve_perm(G, eng.nS, f32=g.fields32, defer=True)
            sv.join()
            return Lam, sv.grad_values(Lam, U, scale=g.wq, permuted=True)
        r, G = project()
        gv = None
        if not want_grad:
            yield
            if sv.check_wait():
                r, _ = project()
        else:
            Lam, gv = adjoint(G)
            yield
            if sv.check_wait():          # the forward's: U changed, all again
                sv.check_discard()       # the adjoint's, of a superseded solve
                r, G = project()
                Lam, gv = adjoint(G)
            if sv.check_wait():          # the adjoint's: Lam changed
                gv = sv.grad_values(Lam, U, scale=g.wq, permuted=True)
    phi = (r ** 2).sum()
    return phi, d, (g.lmap_t(gv, sig) if want_grad else None), r


class _Misfit(torch.autograd.Function):
    """(m_full, dobs, w) -> phi on the GPU, the gradient computed with the
    value (one fused eval+grad, `_misfit_body`); backward scales it."""

    @staticmethod
    def forward(ctx, m_full, dobs, w, eng):
        with torch.cuda.device(eng.device):
            phi, d, grad_full, r = _misfit_body(eng, m_full.detach(), dobs.detach(),
                                                w.detach(), True)
        ctx.save_for_backward(grad_full, r, d, dobs, w)
        return phi

    @staticmethod
    def backward(ctx, gphi):
        grad_full, r, d, dobs, w = ctx.saved_tensors
        gm = grad_full * gphi if ctx.needs_input_grad[0] else None
        gdobs = (-2.0 * w * r) * gphi if ctx.needs_input_grad[1] else None
        gw = (2.0 * r * (d - dobs)) * gphi if ctx.needs_input_grad[2] else None
        return gm, gdobs, gw, None


def _gpu_misfit(eng, m, dobs, w):
    """Body of TorchDC2D/TorchDC3D.misfit on a GPU engine: the fused pass,
    through autograd when a gradient is wanted."""
    out_dt = m.dtype
    for t in (dobs, w):
        if torch.is_tensor(t):
            out_dt = torch.promote_types(out_dt, t.dtype)
    m_full = eng._model_to_full(m).to(torch.float64)
    dobs_t = torch.as_tensor(dobs).to(device=eng.device, dtype=torch.float64).reshape(-1)
    w_t = torch.as_tensor(w).to(device=eng.device, dtype=torch.float64).reshape(-1)
    if torch.is_grad_enabled() and (m_full.requires_grad or dobs_t.requires_grad or
                                    w_t.requires_grad):
        phi = _Misfit.apply(m_full, dobs_t, w_t, eng)
    else:
        with torch.no_grad(), torch.cuda.device(eng.device):
            phi = _misfit_body(eng, m_full, dobs_t, w_t, False)[0]
    return phi.to(device=m.device, dtype=out_dt)


def _eval_grad_device(eng, m, dobs, w):
    """Body of TorchDC2D/TorchDC3D.eval_grad_device: the fused pass of
    `misfit` (`_misfit_body`) without autograd."""
    if eng._gpu is None:
        raise RuntimeError("eval_grad_device needs a GPU engine")
    with torch.no_grad(), torch.cuda.device(eng.device):
        return _drive(_eval_grad_device_steps(eng, m, dobs, w))


class PendingEvalGrad:
    """An eval+grad whose kernels are all queued and whose solve checks are
    not yet waited for (`eval_grad_device_begin`); `finish()` waits for the
    checks (recomputing what a failing check's FP64 re-solve changed) and
    returns (phi, grad).  Finish it on the stream it was begun on."""

    def __init__(self, eng, m, dobs, w):
        self._eng = eng
        self._steps = _eval_grad_device_steps(eng, m, dobs, w)
        self._out = None
        self._step()

    def _step(self):
        with torch.no_grad(), torch.cuda.device(self._eng.device):
            try:
                next(self._steps)
            except StopIteration as e:
                self._out, self._steps = e.value, None

    def finish(self):
        while self._steps is not None:
            self._step()
        return self._out


def _eval_grad_device_steps(eng, m, dobs, w):
    """`_eval_grad_device` as a generator of `_misfit_steps` form (one yield,
    before the host waits for the checks); returns (phi, grad).  The caller
    runs each step under torch.no_grad() and the engine's device."""
    m_full = eng._model_to_full(torch.as_tensor(m).detach()).to(torch.float64)
    dobs = torch.as_tensor(dobs, dtype=torch.float64, device=eng.device)
    w = torch.as_tensor(w, dtype=torch.float64, device=eng.device)
    phi, _, grad_full, _ = yield from _misfit_steps(eng, m_full, dobs, w, True)
    return phi, eng._model_from_full(grad_full)


class _Forward2D(torch.autograd.Function):
    """m_full -> dpred on the GPU; backward = one adjoint solve + gathers."""

    @staticmethod
    def forward(ctx, m_full, eng):
        with torch.cuda.device(eng.device):
            d, saved = _sources_forward(eng, m_full.detach())
        ctx.eng = eng
        ctx.stamp = eng._gpu.solver.stamp
        ctx.save_for_backward(*saved)
        return d

    @staticmethod
    def backward(ctx, gd):
        with torch.cuda.device(ctx.eng.device):
            grad_m = _sources_backward(ctx.eng, ctx.saved_tensors, ctx.stamp, gd.detach())
        return grad_m, None


class _ForwardElectrodes2D(torch.autograd.Function):
    """m_full -> dpred on the GPU through the electrode basis.

    Every source and receiver is a signed sum of electrode positions
    (q = E S, Pd = (E R)^T), so one solve per wavenumber with the ne electrode
    columns, G_k = A_k^{-1} E, gives every field: U_k = G_k S and, for the
    adjoint, Lambda_k = G_k T with T[:, s] = sum over the data j of source s of
    v_j R[:, j].  Forward: d_j = R_j^T (sum_k w_k E^T G_k) S_src(j).  Reverse:
    dphi/dA_k = -w_k G_k M G_k^T with M = T S^T (ne x ne), gathered entry-wise
    as -w_k sum_c G_k[row, c] (G_k M^T)[col, c]; no second solve.  The same
    results as the source basis (the reference's formulation), with ne
    right-hand sides instead of nS, and nS instead of 2 nS solved columns."""

    @staticmethod
    def forward(ctx, m_full, eng):
        with torch.cuda.device(eng.device):
            d, saved = _electrodes_forward(eng, m_full.detach())
        ctx.eng = eng
        ctx.stamp = eng._gpu.solver.stamp
        ctx.save_for_backward(*saved)
        return d

    @staticmethod
    def backward(ctx, gd):
        with torch.cuda.device(ctx.eng.device):
            grad_m = _electrodes_backward(ctx.eng, ctx.saved_tensors, ctx.stamp, gd.detach())
        return grad_m, None


class TorchDC2D:
    """Differentiable dpred/misfit for 2.5-D nodal DC on a fixed mesh+survey."""

    def __init__(self, mesh, survey, nky=12, device="cuda", dtype=torch.float64,
                 active_cells=None, val_inactive=None, chunk=None,
                 singularity_removal=False, rho_primary=None, quadrature=None,
                 precision="fp64", refine=1, leaf=None, rhs_basis="sources",
                 fields_max_gib=None):
        if singularity_removal:
            raise NotImplementedError(
                "singularity removal is not part of this engine (off by default in the "
                "reference); use the total-field formulation")
        if quadrature is None:
            quadrature = os.environ.get("DCTORCH_QUADRATURE", "electrodes")
        if quadrature not in ("auto", "simpeg", "electrodes"):
            raise ValueError(f"quadrature must be auto|simpeg|electrodes, got {quadrature!r}")
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        self.chunk = None if chunk is None else int(chunk)
        self._sr = False
        self.precision = precision
        if rhs_basis not in ("sources", "electrodes"):
            raise ValueError(f"rhs_basis must be sources|electrodes, got {rhs_basis!r}")
        self.rhs_basis = rhs_basis
        # device memory budget of the fields one eval+grad keeps (None: half
        # of the free memory); beyond it the sources go in chunks
        self.fields_max_gib = fields_max_gib

        rmin, rmax = electrode_distance_range(survey)
        sp_pts, sp_wts, fell_back = simpeg_quadrature(mesh, nky)
        err_sp = transform_error(sp_pts, sp_wts, rmin, rmax)
        self.quadrature_info = {"source": "simpeg", "simpeg_fell_back": fell_back,
                                "r_range": (rmin, rmax), "max_rel_err": err_sp,
                                "amplification": amplification(sp_wts),
                                "n_used": int(np.size(sp_pts))}
        pts, wts = sp_pts, sp_wts
        if quadrature == "electrodes" or (quadrature == "auto" and fell_back):
            pts, wts, info = k0_quadrature(rmin, rmax, nky)
            self.quadrature_info.update(source="electrodes", max_rel_err=info["max_rel_err"],
                                        amplification=info["amplification"],
                                        n_used=info["n_used"], simpeg_max_rel_err=err_sp)
            if fell_back:
                warnings.warn(
                    f"SimPEG's wavenumber fit failed at nky={nky} and fell back to a "
                    f"trapezoid rule (transform error {err_sp:.1e}); replaced by the "
                    f"electrode fit ({info['max_rel_err']:.1e}).", stacklevel=2)
        elif fell_back:
            warnings.warn(f"SimPEG's wavenumber fit failed at nky={nky}; its trapezoid "
                          f"fallback has transform error {err_sp:.1e}.", stacklevel=2)
        self.kys = np.asarray(pts, float)
        self.wq_np = np.asarray(wts, float)
        maps = Problem2DMaps(mesh, survey, self.kys)
        self._maps = maps
        self.nC, self.nN, self.nS, self.nD = maps.nC, maps.nN, maps.nS, maps.nD
        self.nky = self.kys.size
        self.nnzA = maps.nnz
        dev, dt = self.device, self.dtype
        self.wq = torch.tensor(self.wq_np, dtype=dt, device=dev)
        self.sidx = torch.tensor(maps.sidx, dtype=torch.int64, device=dev)
        self.iD = torch.arange(self.nD, dtype=torch.int64, device=dev)
        self._crow = torch.tensor(maps.crow, dtype=torch.int32)
        self._ccol = torch.tensor(maps.col, dtype=torch.int32)
        self._Pd = None
        self._q = None
        self._Lstack = None
        self._P_inj = None
        if active_cells is not None:
            act = np.asarray(active_cells, dtype=bool)
            ia = np.where(act)[0]
            self._act = torch.tensor(ia, dtype=torch.int64, device=dev)
            v = np.full(self.nC, float(val_inactive))
            v[act] = 0.0
            self._v_inj = torch.tensor(v, dtype=dt, device=dev)
            self._P_inj = True
        self._mesh, self._survey = mesh, survey
        self._solvers = {}
        self._gpu = None
        if self.device.type == "cuda":
            self._gpu = _GpuState(self, maps, precision, refine, leaf)
            self.solver = self._gpu.solver
        else:
            self.solver = self._solver_for(self.nS)

    # ---------------- reference attributes (built on first use) -------------
    @property
    def q(self):
        if self._q is None:
            self._q = torch.tensor(self._maps.q.toarray(), dtype=self.dtype, device=self.device)
        return self._q

    @property
    def Pd(self):
        if self._Pd is None:
            self._Pd = _s2t(self._maps.Pd, self.device, self.dtype)
        return self._Pd

    @property
    def Lstack(self):
        """(nky*nnzA, nC) stacked map, as the reference's COO tensor."""
        if self._Lstack is None:
            m = self._maps
            ent = np.repeat(np.arange(m.nnz), np.diff(m.map_ptr))
            blocks = []
            for ky in self.kys:
                L = sp.csr_matrix((m.map_le + ky ** 2 * m.map_ln, (ent, m.map_cell)),
                                  shape=(m.nnz, m.nC))
                blocks.append(L)
            for k in range(self.nky):
                blocks[k] = blocks[k] + sp.csr_matrix(
                    (m.bc_coef[k], (m.bc_e, m.bc_cell)), shape=(m.nnz, m.nC))
            self._Lstack = _s2t(sp.vstack(blocks), self.device, self.dtype)
        return self._Lstack

    # ---------------- helpers -------------------------------------------------
    def _solver_for(self, n_rhs):
        if n_rhs not in self._solvers:
            self._solvers[n_rhs] = make_batch_solver(self._crow, self._ccol, self.nN, n_rhs,
                                                     self.nky, device=self.device,
                                                     dtype=self.dtype)
        return self._solvers[n_rhs]

    def _source_chunks(self, chunk=None):
        w = self.chunk if chunk is None else chunk
        if not w or w >= self.nS:
            return [(0, self.nS)]
        return [(s, min(s + w, self.nS)) for s in range(0, self.nS, w)]

    def _rows_of(self, s0, s1):
        key = (s0, s1)
        cache = self.__dict__.setdefault("_rowcache", {})
        if key not in cache:
            sel = (self.sidx >= s0) & (self.sidx < s1)
            cache[key] = (torch.nonzero(sel, as_tuple=True)[0], self.iD[sel], self.sidx[sel] - s0)
        return cache[key]

    def _model_to_full(self, m):
        m_loc = m.to(device=self.device, dtype=self.dtype).reshape(-1)
        if self._P_inj is not None:
            full = self._v_inj.clone()
            full = full.index_put((self._act,), m_loc, accumulate=True)
            return full
        return m_loc

    def _model_from_full(self, g_full):
        if self._P_inj is not None:
            return g_full[self._act]
        return g_full

    # ---------------- runtime --------------------------------------------------
    def dpred(self, m):
        """Predicted data (nD,) for m = log(rho).  Autograd-traceable."""
        dev0, dt0 = m.device, m.dtype
        m_full = self._model_to_full(m)
        if self._gpu is not None:
            fn = _ForwardElectrodes2D if self.rhs_basis == "electrodes" else _Forward2D
            d = fn.apply(m_full.to(torch.float64), self)
            return d.to(device=dev0, dtype=dt0)
        if self.chunk and self.chunk < self.nS:
            d = torch.empty(self.nD, device=self.device, dtype=self.dtype)
            for s0, s1 in self._source_chunks():
                rows, _, _ = self._rows_of(s0, s1)
                d[rows] = self._dpred_chunk_cpu(m_full, s0, s1)
            return d.to(device=dev0, dtype=dt0)
        sig = torch.exp(-m_full).unsqueeze(1)
        VALS = torch.sparse.mm(self.Lstack, sig).view(self.nky, self.nnzA)
        U = sparse_solve_batch(VALS, self.q.contiguous(), self.solver)
        Ubar = (self.wq.view(-1, 1, 1) * U).sum(0)
        d = torch.sparse.mm(self.Pd, Ubar)[self.iD, self.sidx]
        return d.to(device=dev0, dtype=dt0)

    def _dpred_chunk_cpu(self, m_full, s0, s1):
        sig = torch.exp(-m_full).unsqueeze(1)
        VALS = torch.sparse.mm(self.Lstack, sig).view(self.nky, self.nnzA)
        U = sparse_solve_batch(VALS, self.q[:, s0:s1].contiguous(), self._solver_for(s1 - s0))
        Ubar = (self.wq.view(-1, 1, 1) * U).sum(0)
        _, gi, li = self._rows_of(s0, s1)
        return torch.sparse.mm(self.Pd, Ubar)[gi, li]

    def misfit(self, m, dobs, w):
        """phi_d = ||w (dpred - dobs)||^2 (no 1/2 factor).  On a GPU engine the
        value and, when a gradient is wanted, dphi/dm (and dphi/ddobs,
        dphi/dw) come from one fused pass; backward() then only scales it."""
        if self._gpu is not None:
            return _gpu_misfit(self, m, dobs, w)
        r = w * (self.dpred(m) - dobs)
        return (r ** 2).sum()

    def eval_grad_device(self, m, dobs, w):
        """(phi_d, dphi_d/dm) as device tensors, with no autograd graph: the
        same fused pass, kernels and order as `misfit(m, dobs, w).backward()`
        (bit-identical results), issued on the calling thread's current
        stream; the host waits only for the solves' checks.  Meant for callers
        that drive many engines from several host threads (torch's autograd
        engine runs every CUDA backward of a device on one worker thread).
        GPU engines only."""
        return _eval_grad_device(self, m, dobs, w)

    def eval_grad_device_begin(self, m, dobs, w):
        """`eval_grad_device` in two halves: every kernel of the pass is
        queued on the current stream and the host has not waited for
        anything yet; the returned PendingEvalGrad's `finish()` waits for the
        solves' checks and gives (phi_d, dphi_d/dm).  A caller driving several
        engines from one thread begins all of them before it finishes any."""
        if self._gpu is None:
            raise RuntimeError("eval_grad_device_begin needs a GPU engine")
        return PendingEvalGrad(self, m, dobs, w)

    # -- chunked reverse passes ---------------------------------------------------
    def _gpu_reverse(self, m, weights_of, chunk):
        """Shared body of misfit_and_grad / Jtvec on the GPU: factor once, then
        per source chunk solve, project, form the adjoint weights (weights_of
        gets the chunk's data and rows), adjoint solve and gather."""
        g = self._gpu
        width = self.chunk if chunk is None else chunk
        width = self.nS if not width else min(int(width), self.nS)
        with torch.no_grad(), torch.cuda.device(self.device):
            m_full = self._model_to_full(m.detach()).to(torch.float64)
            sig = torch.exp(-m_full)
            vals = g.assemble(sig)
            g.solver.factorize(vals)
            gv = torch.zeros(self.nky, self.nnzA, dtype=torch.float64, device=self.device)
            total = 0.0
            d = torch.empty(self.nD, dtype=torch.float64, device=self.device)
            for s0, s1 in self._source_chunks(width):
                k = s1 - s0
                qb = g.q_perm[:, s0:s1].contiguous() if k != self.nS else g.q_perm
                U = g.solve_perm(qb, k, f32=g.fields32)
                g.dpred(U, k, s0, s1, d)
                rows, _, _ = self._rows_of(s0, s1)
                v, f = weights_of(d, rows)
                total += f
                G = g.adjoint_rhs(v, k, s0, s1)
                Lam = g.solve_perm(G, k, f32=g.fields32)
                gv += g.solver.grad_values(Lam, U, scale=g.wq, permuted=True)
                del U, Lam, G
            grad = self._model_from_full(g.lmap_t(gv, sig))
        return total, grad

    def misfit_and_grad(self, m, dobs, w, chunk=None):
        """(phi_d, dphi_d/dm) one source chunk at a time, outside autograd."""
        if self._gpu is not None:
            dobs_t = dobs.to(self.device, torch.float64).reshape(-1)
            w_t = w.to(self.device, torch.float64).reshape(-1)

            def weights_of(d, rows):
                vfull = torch.zeros(self.nD, dtype=torch.float64, device=self.device)
                r = d[rows] - dobs_t[rows]
                vfull[rows] = 2.0 * w_t[rows] ** 2 * r
                return vfull, float(((w_t[rows] * r) ** 2).sum())

            total, grad = self._gpu_reverse(m, weights_of, chunk)
            return total, grad.to(m.device, m.dtype)
        width = self.chunk if chunk is None else chunk
        width = self.nS if not width else min(int(width), self.nS)
        solver = self._solver_for(width)
        with torch.no_grad():
            sig = torch.exp(-self._model_to_full(m)).unsqueeze(1)
            solver.factorize(torch.sparse.mm(self.Lstack, sig).view(self.nky, self.nnzA))
        total, grad = 0.0, torch.zeros_like(m)
        for s0, s1 in self._source_chunks(width):
            rows, gi, li = self._rows_of(s0, s1)
            mc = m.detach().clone().requires_grad_(True)
            sig = torch.exp(-self._model_to_full(mc)).unsqueeze(1)
            VALS = torch.sparse.mm(self.Lstack, sig).view(self.nky, self.nnzA)
            q = self.q[:, s0:s1]
            if width > s1 - s0:
                q = torch.cat([q, q.new_zeros(self.nN, width - (s1 - s0))], dim=1)
            U = sparse_solve_batch_prefactored(VALS, q.contiguous(), solver)
            Ubar = (self.wq.view(-1, 1, 1) * U).sum(0)[:, :s1 - s0]
            d = torch.sparse.mm(self.Pd, Ubar)[gi, li]
            f = ((w[rows] * (d - dobs[rows])) ** 2).sum()
            f.backward()
            total += float(f)
            grad += mc.grad
        return total, grad

    def Jtvec(self, m, v, chunk=None):
        """J^T v without forming J (same reverse pass as the gradient)."""
        if self._gpu is not None:
            v_t = v.to(self.device, torch.float64).reshape(-1)

            def weights_of(d, rows):
                vfull = torch.zeros(self.nD, dtype=torch.float64, device=self.device)
                vfull[rows] = v_t[rows]
                return vfull, 0.0

            _, g = self._gpu_reverse(m, weights_of, chunk)
            return g.to(m.device, m.dtype)
        width = self.chunk if chunk is None else chunk
        width = self.nS if not width else min(int(width), self.nS)
        solver = self._solver_for(width)
        v_t = v.to(device=self.device, dtype=self.dtype).reshape(-1)
        with torch.no_grad():
            sig = torch.exp(-self._model_to_full(m)).unsqueeze(1)
            solver.factorize(torch.sparse.mm(self.Lstack, sig).view(self.nky, self.nnzA))
        out = torch.zeros_like(m)
        for s0, s1 in self._source_chunks(width):
            rows, gi, li = self._rows_of(s0, s1)
            mc = m.detach().clone().requires_grad_(True)
            sig = torch.exp(-self._model_to_full(mc)).unsqueeze(1)
            VALS = torch.sparse.mm(self.Lstack, sig).view(self.nky, self.nnzA)
            q = self.q[:, s0:s1]
            if width > s1 - s0:
                q = torch.cat([q, q.new_zeros(self.nN, width - (s1 - s0))], dim=1)
            U = sparse_solve_batch_prefactored(VALS, q.contiguous(), solver)
            Ubar = (self.wq.view(-1, 1, 1) * U).sum(0)[:, :s1 - s0]
            d = torch.sparse.mm(self.Pd, Ubar)[gi, li]
            d.backward(v_t[rows])
            out = out + mc.grad
        return out

    def Jmatrix(self, m, block=32):
        """The explicit Jacobian (nD, nM) from one factorisation (batched
        adjoint: the adjoint right-hand sides are the columns of Pd^T)."""
        block = max(1, min(int(block), self.nD))
        dev0, dt0 = m.device, m.dtype
        if self._gpu is not None:
            J = Linearization2D(self, m, store_fields=True).Jmatrix(sub=block)
            return J.to(device=dev0, dtype=dt0)
        Lt = self.Lstack.t().coalesce()
        with torch.no_grad():
            m_full = self._model_to_full(m).to(torch.float64)
            sig = torch.exp(-m_full)
            solver = self._solver_for(self.nS)
            VALS = torch.sparse.mm(self.Lstack, sig.unsqueeze(1)).view(self.nky, self.nnzA)
            solver.factorize(VALS.to(self.dtype))
            U = solver.solve(self.q.to(self.dtype)).to(torch.float64)
            PdT = self.Pd.t().coalesce().to_dense().to(torch.float64)
            row, col = solver.row.long(), solver.col.long()
            wq = self.wq.to(torch.float64)
            Jm = torch.empty(self.nC, self.nD, dtype=torch.float64, device=self.device)
            for a in range(0, self.nD, block):
                b = min(a + block, self.nD)
                Y = solver.solve(PdT[:, a:b].contiguous().to(self.dtype)).to(torch.float64)
                cols = self.sidx[a:b]
                per_ky = [(-wq[k]) * Y[k][row] * U[k][:, cols][col] for k in range(self.nky)]
                Jm[:, a:b] = torch.sparse.mm(Lt.to(torch.float64), torch.cat(per_ky, 0))
            Jm = -sig.unsqueeze(1) * Jm
            Jm = self._model_from_full(Jm)
            return Jm.T.contiguous().to(device=dev0, dtype=dt0)

    def linearize(self, m, store_fields=None, fields_max_gib=2.0):
        """The forward frozen at m: one factorisation, the fields kept, and
        J v / J^T v as one batched solve each (Linearization2D)."""
        return Linearization2D(self, m, store_fields=store_fields,
                               fields_max_gib=fields_max_gib)

    def Jvec(self, m, v):
        """J v for J = d(dpred)/dm without forming J: one forward and one extra
        solve.  For many products at one model, `linearize` once."""
        return self.linearize(m).Jvec(v)
