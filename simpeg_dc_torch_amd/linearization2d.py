"""The 2.5-D forward frozen at one model, for Gauss-Newton.

Drop-in for ``dctorch.simulation2d.Linearization2D``
(dctorch/simulation2d.py:537-818): one factorisation at m, the forward fields
U_k kept, and

    J v   :  A_k dU_k = -A_k(dsigma) U_k,   dsigma = -sigma v,
             J v = Pd sum_k w_k dU_k
    J^T v :  A_k Y_k = Pd^T v (the same factorisation, A is symmetric),
             gv[k, e] = -w_k Y_k[row e] . U_k[col e],  J^T v = -sigma L^T gv

each one batched solve over all wavenumbers.  Rows of J (``Jmatrix``,
``diag_JtDJ``) come a block of data at a time from one adjoint solve with the
data's projection rows as right-hand sides, then the pairwise gather
-w_k Y_k[row e, j] U_k[col e, src j] and L^T.

On a GPU engine every step is a libdcx.so call: dcx_assemble (A(sigma),
A(dsigma)), dcx_factorize, dcx_solve, dcx_spmm (the J v right-hand side),
dcx_survey_dpred / dcx_survey_adjoint_rhs, dcx_grad_values +
dcx_lmap_transpose (J^T v), dcx_jacobian_rows (rows of J).  On the cpu the
reference's own chain runs (torch sparse products, SciPy SuperLU).

The solver object is shared with the engine; anything else that factorises it
(a line-search dpred) bumps its ``stamp`` and the factorisation is restored
before the next product, as in the reference.
"""
import numpy as np
import torch

from . import _lib


class Linearization2D:
    """J v, J^T v, diag(J^T D J) and J at one model of a TorchDC2D engine."""

    #: rows of J formed per adjoint solve on the GPU (one lane per datum)
    gpu_block = 64

    def __init__(self, eng, m, store_fields=None, fields_max_gib=2.0):
        if getattr(eng, "_sr", False):
            raise NotImplementedError(
                "Linearization2D assumes the plain formulation; use Jtvec")
        self.eng = eng
        self.m_dtype, self.m_device = m.dtype, m.device
        with torch.no_grad():
            self.sig = torch.exp(-eng._model_to_full(m).to(torch.float64))
        item = 8
        nbytes = eng.nky * eng.nN * eng.nS * item
        if store_fields is None:
            store_fields = nbytes <= fields_max_gib * 2 ** 30
        self._store = bool(store_fields)
        if eng._gpu is not None:
            self._init_gpu()
        else:
            self._init_cpu()

    # ------------------------------------------------------------------ common
    def _to_eng(self, v):
        eng = self.eng
        v = v.to(eng.device, torch.float64).reshape(-1)
        if eng._P_inj is not None:
            full = torch.zeros(eng.nC, dtype=torch.float64, device=eng.device)
            full[eng._act] = v
            return full
        return v

    def _to_model(self, g_full):
        return self.eng._model_from_full(g_full).to(self.m_device, self.m_dtype)

    def _row_blocks(self, sub):
        eng = self.eng
        width = int(sub) if sub else (self.gpu_block if eng._gpu is not None else 16)
        width = max(1, min(width, eng.nD))
        return [(a, min(a + width, eng.nD)) for a in range(0, eng.nD, width)]

    # ===================================================================== GPU
    def _init_gpu(self):
        eng, g = self.eng, self.eng._gpu
        self.g = g
        dev = eng.device
        # engines built with rhs_basis="electrodes": every field from one solve
        # with the electrode columns, G = A^{-1} E (U = G S, the adjoint field
        # of datum j is G R_j), so J, J v and J^T v need no further solve
        self._elec = getattr(eng, "rhs_basis", "sources") == "electrodes"
        self._G = None
        with torch.no_grad(), torch.cuda.device(dev):
            self.vals = g.assemble(self.sig)
            self._factorize()
            if self._elec:
                self._G = g.solve_perm(g.E_perm, g.ne)
            U = self._solve_fields_gpu()
            self._U = U if self._store else None
            d = torch.empty(eng.nD, dtype=torch.float64, device=dev)
            g.dpred(U, eng.nS, 0, eng.nS, d)
        self.d = d.to(self.m_device, self.m_dtype)
        # projection rows in elimination numbering, for the row blocks of J
        Pd = eng._maps.Pd.tocsr()
        pinv = np.empty(eng.nN, np.int64)
        pinv[g.solver.perm()] = np.arange(eng.nN)
        datum = np.repeat(np.arange(eng.nD), np.diff(Pd.indptr))
        self._pd_datum = torch.tensor(datum, device=dev)
        self._pd_row = torch.tensor(pinv[Pd.indices], device=dev)
        self._pd_val = torch.tensor(Pd.data, dtype=torch.float64, device=dev)
        self._src32 = eng.sidx.to(torch.int32)

    def _factorize(self):
        self.g.solver.factorize(self.vals)
        self._stamp = self.g.solver.stamp

    def _ensure(self):
        if self.g.solver.stamp != self._stamp:
            self._factorize()

    def _electrode_fields(self):
        if self.g.solver.stamp != self._stamp:   # refactorised since: solve again
            self._factorize()
            self._G = self.g.solve_perm(self.g.E_perm, self.g.ne)
        return self._G

    def _solve_fields_gpu(self):
        if self._elec:
            return torch.matmul(self._electrode_fields(), self.g.S)
        self._ensure()
        return self.g.solve_perm(self.g.q_perm, self.eng.nS)

    def _fields_gpu(self):
        return self._U if self._U is not None else self._solve_fields_gpu()

    def _jvec_gpu(self, v):
        eng, g = self.eng, self.g
        with torch.no_grad(), torch.cuda.device(eng.device):
            dvals = g.assemble(-self.sig * self._to_eng(v))
            U = self._fields_gpu()
            rhs = torch.empty_like(U)
            _lib.check(g.lib.dcx_spmm(g.solver._h, _lib.ptr(dvals), _lib.ptr(U), _lib.ptr(rhs),
                                      eng.nS, -1.0, _lib.DCX_PERMUTED, g.stream()), "spmm")
            if self._elec:
                # P_j A^{-1} rhs[:, s] = R_j^T (G^T rhs)[:, s], summed over ky
                G = self._electrode_fields()
                Z = (g.wq.view(-1, 1, 1) * torch.matmul(G.transpose(1, 2).contiguous(),
                                                        rhs)).sum(0)
                out = (g.R * Z[:, g.sidx]).sum(0)
                return out.to(self.m_device, self.m_dtype)
            self._ensure()
            dU = g.solve_perm(rhs, eng.nS, sparse=False, shared=False)
            out = torch.empty(eng.nD, dtype=torch.float64, device=eng.device)
            g.dpred(dU, eng.nS, 0, eng.nS, out)
        return out.to(self.m_device, self.m_dtype)

    def _jtvec_gpu(self, v):
        eng, g = self.eng, self.g
        with torch.no_grad(), torch.cuda.device(eng.device):
            vv = v.to(eng.device, torch.float64).reshape(-1).contiguous()
            if self._elec:
                Ge = self._electrode_fields()
                T = torch.zeros(g.ne, eng.nS, dtype=torch.float64, device=eng.device)
                T.index_add_(1, g.sidx, g.R * vv)
                from .simulation2d import _rows_times
                H = _rows_times(g, Ge, (T @ g.S.T).T)
                gv = g.solver.grad_values(Ge, H, scale=g.wq, permuted=True)
                return self._to_model(g.lmap_t(gv, self.sig))
            G = g.adjoint_rhs(vv, eng.nS, 0, eng.nS)
            U = self._fields_gpu()
            self._ensure()
            Y = g.solve_perm(G, eng.nS)
            gv = g.solver.grad_values(Y, U, scale=g.wq, permuted=True)
            out = g.lmap_t(gv, self.sig)
        return self._to_model(out)

    def _rows_gpu(self, sub):
        """Yield (a, b, Jt) with Jt = (rows a..b of J)^T on all cells, (nC, b - a)."""
        eng, g = self.eng, self.g
        with torch.no_grad(), torch.cuda.device(eng.device):
            U = self._fields_gpu()
            for a, b in self._row_blocks(sub):
                r = b - a
                if self._elec:  # adjoint fields of data a..b: G R[:, a:b]
                    Y = torch.matmul(self._electrode_fields(), g.R[:, a:b].contiguous())
                else:
                    sel = (self._pd_datum >= a) & (self._pd_datum < b)
                    E = torch.zeros(eng.nN, r, dtype=torch.float64, device=eng.device)
                    E.index_put_((self._pd_row[sel], self._pd_datum[sel] - a),
                                 self._pd_val[sel], accumulate=True)
                    self._ensure()
                    Y = g.solve_perm(E, r, sparse=True, shared=True)
                src = self._src32[a:b].contiguous()
                Jt = torch.empty(eng.nC, r, dtype=torch.float64, device=eng.device)
                _lib.check(g.lib.dcx_jacobian_rows(
                    g.lmap, g.solver._h, _lib.ptr(Y), r, _lib.ptr(U), eng.nS, _lib.ptr(src),
                    _lib.ptr(g.wq), _lib.ptr(self.sig), _lib.ptr(Jt), _lib.DCX_PERMUTED,
                    g.stream()), "jacobian_rows")
                yield a, b, Jt

    # ===================================================================== CPU
    def _init_cpu(self):
        eng = self.eng
        self.solver = eng._solver_for(eng.nS)
        with torch.no_grad():
            sig = self.sig.to(eng.dtype).unsqueeze(1)
            self.VALS = torch.sparse.mm(eng.Lstack, sig).view(eng.nky, eng.nnzA)
            self._factorize_cpu()
            U = self._solve_cpu(eng.q)
            self._U = U if self._store else None
            Ubar = (eng.wq.view(-1, 1, 1) * U).sum(0)
            d = torch.sparse.mm(eng.Pd, Ubar)[eng.iD, eng.sidx]
        self.d = d.to(self.m_device, self.m_dtype)
        self._row = self.solver.row.long()
        self._col = torch.as_tensor(eng._ccol, dtype=torch.int64)
        self._crow = torch.as_tensor(eng._crow, dtype=torch.int64)

    def _factorize_cpu(self):
        self.solver.factorize(self.VALS)
        self._stamp = self.solver.stamp

    def _solve_cpu(self, b):
        if self.solver.stamp != self._stamp:
            self._factorize_cpu()
        return self.solver.solve(b.contiguous())

    def _fields_cpu(self):
        return self._U if self._U is not None else self._solve_cpu(self.eng.q)

    def _lt_cpu(self, X):
        """L^T X for X (nky * nnzA,) or (nky * nnzA, r), the stacked map."""
        return torch.sparse.mm(self.eng.Lstack.t(), X if X.dim() > 1 else X.unsqueeze(1))

    def _jvec_cpu(self, v):
        eng = self.eng
        dsig = (-self.sig * self._to_eng(v)).to(eng.dtype)
        dvals = torch.sparse.mm(eng.Lstack, dsig.unsqueeze(1)).view(eng.nky, eng.nnzA)
        U = self._fields_cpu()
        rhs = torch.empty_like(U)
        for k in range(eng.nky):
            Ak = torch.sparse_csr_tensor(self._crow, self._col, -dvals[k], (eng.nN, eng.nN))
            rhs[k] = Ak @ U[k]
        dU = self._solve_cpu(rhs)
        dbar = (eng.wq.view(-1, 1, 1) * dU).sum(0)
        return torch.sparse.mm(eng.Pd, dbar)[eng.iD, eng.sidx].to(self.m_device, self.m_dtype)

    def _jtvec_cpu(self, v):
        eng = self.eng
        v = v.to(eng.device, eng.dtype).reshape(-1)
        V = torch.zeros(eng.nD, eng.nS, dtype=eng.dtype)
        V[eng.iD, eng.sidx] = v
        Y = self._solve_cpu(torch.sparse.mm(eng.Pd.t(), V))
        U = self._fields_cpu()
        gv = -(Y[:, self._row] * U[:, self._col]).sum(-1) * eng.wq.view(-1, 1)
        gsig = self._lt_cpu(gv.reshape(-1)).squeeze(1)
        return self._to_model(-self.sig * gsig)

    def _rows_cpu(self, sub):
        eng = self.eng
        U = self._fields_cpu()
        PdT = eng.Pd.t().coalesce().to_dense()
        for a, b in self._row_blocks(sub):
            Y = self._solve_cpu(PdT[:, a:b])
            Us = U[:, :, eng.sidx[a:b]]
            gv = -(Y[:, self._row] * Us[:, self._col]) * eng.wq.view(-1, 1, 1)
            Jt = self._lt_cpu(gv.reshape(eng.nky * eng.nnzA, b - a))
            yield a, b, -self.sig.unsqueeze(1) * Jt

    # ================================================================ products
    @torch.no_grad()
    def Jvec(self, v):
        """J v, (nD,), on the model's device and dtype."""
        return self._jvec_gpu(v) if self.eng._gpu is not None else self._jvec_cpu(v)

    @torch.no_grad()
    def Jtvec(self, v):
        """J^T v, (nM,), on the model's device and dtype."""
        return self._jtvec_gpu(v) if self.eng._gpu is not None else self._jtvec_cpu(v)

    def _rows_of_J(self, sub):
        it = self._rows_gpu(sub) if self.eng._gpu is not None else self._rows_cpu(sub)
        for a, b, Jt in it:
            yield a, b, self.eng._model_from_full(Jt)

    @torch.no_grad()
    def diag_JtDJ(self, dw, sub=None):
        """diag(J^T diag(dw) J), (nM,), exact, without storing J."""
        dw = dw.to(self.eng.device, torch.float64).reshape(-1)
        out = None
        for a, b, Jb in self._rows_of_J(sub):
            part = (Jb.to(torch.float64) ** 2) @ dw[a:b]
            out = part if out is None else out + part
        return out.to(self.m_device, self.m_dtype)

    @torch.no_grad()
    def Jmatrix(self, sub=None):
        """The explicit J (nD, nM) from this linearisation, rows formed a block
        of data at a time."""
        J = None
        for a, b, Jb in self._rows_of_J(sub):
            if J is None:
                J = torch.empty(self.eng.nD, Jb.shape[0], dtype=torch.float64,
                                device=self.eng.device)
            J[a:b] = Jb.T
        return J.to(self.m_device, self.m_dtype)
