"""Differentiable 3-D DC-resistivity forward, drop-in for dctorch.TorchDC3D.

API (dctorch/simulation3d.py:44-227): ``TorchDC3D(mesh, survey,
device="cuda", dtype=torch.float64, active_cells=None, val_inactive=None,
bc_type="Robin")`` with ``dpred``, ``Jtvec``, ``misfit`` and the attributes
``nC, nN, nS, nD, nnzA, L3, L3t, q, Pd, sidx, iD, solver``.  The model is
m = log(rho).  Extra keywords as in TorchDC2D: ``precision`` ("fp64" or
"mixed"), ``refine``, ``leaf``, ``rhs_basis`` ("sources" or "electrodes":
one solve with the distinct electrode positions as right-hand sides, the
adjoint by electrode-space algebra; the same results).

The system is the 2.5-D one with a single "wavenumber" of zero and weight
one, so on a GPU device the same libdcx.so path runs: the factored
sigma -> A-values map (`dcx_assemble`), the supernodal factorisation and
solve of the one system for all sources at once (`dcx_factor_solve`), the
receiver gather (`dcx_survey_dpred`), and in the reverse pass the adjoint
solve, the gather-dot and the transposed map.  On the cpu the reference's
chain runs: a sparse product with L3 and SciPy SuperLU.
"""
import numpy as np
import scipy.sparse as sp
import torch

from .operators3d import Problem3DMaps
from .simulation2d import (PendingEvalGrad, _Forward2D, _ForwardElectrodes2D, _GpuState,
                           _eval_grad_device, _gpu_misfit)
from .solver import make_solver, sparse_solve


def _csr(M, device, dtype):
    M = sp.csr_matrix(M)
    M.sort_indices()
    return torch.sparse_csr_tensor(torch.tensor(M.indptr, dtype=torch.int32, device=device),
                                   torch.tensor(M.indices, dtype=torch.int32, device=device),
                                   torch.tensor(M.data, dtype=dtype, device=device),
                                   size=M.shape, device=device)


class _SpMV(torch.autograd.Function):
    """y = M x with the transpose of the fixed M cached for the backward
    (dctorch/simulation3d.py:25-41)."""

    @staticmethod
    def forward(ctx, x, M, Mt):
        ctx.Mt = Mt
        return (M @ x.unsqueeze(1)).squeeze(1)

    @staticmethod
    def backward(ctx, g):
        return (ctx.Mt @ g.contiguous().unsqueeze(1)).squeeze(1), None, None


class TorchDC3D:
    """Differentiable dpred/misfit for 3-D nodal DC on a fixed mesh+survey."""

    def __init__(self, mesh, survey, device="cuda", dtype=torch.float64, active_cells=None,
                 val_inactive=None, bc_type="Robin", precision="fp64", refine=1, leaf=None,
                 rhs_basis="sources", fields_max_gib=None):
        if bc_type != "Robin":
            raise NotImplementedError("only the Robin boundary condition (the default) is built")
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        self.precision = precision
        if rhs_basis not in ("sources", "electrodes"):
            raise ValueError(f"rhs_basis must be sources|electrodes, got {rhs_basis!r}")
        self.rhs_basis = rhs_basis
        # device memory budget of the fields one eval+grad keeps (None: half
        # of the free memory); beyond it the sources go in chunks
        self.fields_max_gib = fields_max_gib
        self.chunk = None
        maps = Problem3DMaps(mesh, survey)
        self._maps = maps
        self.kys = np.zeros(1)
        self.wq_np = np.ones(1)
        self.nky = 1
        self.nC, self.nN, self.nS, self.nD = maps.nC, maps.nN, maps.nS, maps.nD
        self.nnzA = maps.nnz
        dev, dt = self.device, dtype
        self.sidx = torch.tensor(maps.sidx, dtype=torch.int64, device=dev)
        self.iD = torch.arange(self.nD, dtype=torch.int64, device=dev)
        self._P_inj = None
        if active_cells is not None:
            act = np.asarray(active_cells, dtype=bool)
            self._act = torch.tensor(np.where(act)[0], dtype=torch.int64, device=dev)
            v = np.full(self.nC, float(val_inactive))
            v[act] = 0.0
            self._v_inj = torch.tensor(v, dtype=dt, device=dev)
            self._P_inj = True
        self._L3 = self._L3t = self._q = self._Pd = None
        self._gpu = None
        if self.device.type == "cuda":
            self._gpu = _GpuState(self, maps, precision, refine, leaf)
            self.solver = self._gpu.solver
        else:
            self.solver = make_solver(torch.tensor(maps.crow), torch.tensor(maps.col), self.nN,
                                      self.nS, device=dev, dtype=dt)

    # ---------------- reference attributes (built on first use) -------------
    def _L3_host(self):
        m = self._maps
        ent = np.repeat(np.arange(m.nnz), np.diff(m.map_ptr))
        L = sp.csr_matrix((m.map_le, (ent, m.map_cell)), shape=(m.nnz, m.nC))
        return L + sp.csr_matrix((m.bc_coef[0], (m.bc_e, m.bc_cell)), shape=(m.nnz, m.nC))

    @property
    def L3(self):
        """(nnzA, nC) CSR map: A(sigma).data = L3 sigma."""
        if self._L3 is None:
            L = self._L3_host()
            self._L3 = _csr(L, self.device, self.dtype)
            self._L3t = _csr(L.T, self.device, self.dtype)
        return self._L3

    @property
    def L3t(self):
        self.L3
        return self._L3t

    @property
    def q(self):
        if self._q is None:
            self._q = torch.tensor(self._maps.q.toarray(), dtype=self.dtype, device=self.device)
        return self._q

    @property
    def Pd(self):
        if self._Pd is None:
            self._Pd = _csr(self._maps.Pd, self.device, self.dtype)
        return self._Pd

    # ---------------- helpers -------------------------------------------------
    def _model_to_full(self, m):
        m_loc = m.to(device=self.device, dtype=self.dtype).reshape(-1)
        if self._P_inj is not None:
            full = self._v_inj.clone()
            return full.index_put((self._act,), m_loc, accumulate=True)
        return m_loc

    def _model_from_full(self, g_full):
        return g_full[self._act] if self._P_inj is not None else g_full

    def _rows_of(self, s0, s1):
        """(data rows, their indices, their source offsets) of sources s0..s1."""
        key = (s0, s1)
        cache = self.__dict__.setdefault("_rowcache", {})
        if key not in cache:
            sel = (self.sidx >= s0) & (self.sidx < s1)
            cache[key] = (torch.nonzero(sel, as_tuple=True)[0], self.iD[sel], self.sidx[sel] - s0)
        return cache[key]

    # ---------------- runtime --------------------------------------------------
    def dpred(self, m):
        """Predicted data (nD,) for m = log(rho).  Autograd-traceable."""
        dev0, dt0 = m.device, m.dtype
        m_full = self._model_to_full(m)
        if self._gpu is not None:
            fn = _ForwardElectrodes2D if self.rhs_basis == "electrodes" else _Forward2D
            d = fn.apply(m_full.to(torch.float64), self)
            return d.to(device=dev0, dtype=dt0)
        sig = torch.exp(-m_full)
        VALS = _SpMV.apply(sig, self.L3, self.L3t)
        U = sparse_solve(VALS, self.q, self.solver)
        d = (self.Pd @ U)[self.iD, self.sidx]
        return d.to(device=dev0, dtype=dt0)

    def Jtvec(self, m, v):
        """J^T v without forming J: the reverse pass of dpred with v in place
        of the residual."""
        mc = m.detach().clone().requires_grad_(True)
        d = self.dpred(mc)
        v_t = v.to(device=d.device, dtype=d.dtype).reshape(-1)
        return torch.autograd.grad(d, mc, grad_outputs=v_t)[0]

    def misfit(self, m, dobs, w):
        """phi_d = ||w (dpred - dobs)||^2 (no 1/2 factor); on a GPU engine one
        fused eval+grad pass (see TorchDC2D.misfit)."""
        if self._gpu is not None:
            return _gpu_misfit(self, m, dobs, w)
        r = w * (self.dpred(m) - dobs)
        return (r ** 2).sum()

    def eval_grad_device(self, m, dobs, w):
        """(phi_d, dphi_d/dm) as device tensors without an autograd graph,
        bit-identical to `misfit(...).backward()`, on the calling thread's
        current stream.  GPU engines only."""
        return _eval_grad_device(self, m, dobs, w)

    def eval_grad_device_begin(self, m, dobs, w):
        """`eval_grad_device` in two halves (kernels queued now, the solve
        checks waited for by the returned PendingEvalGrad's `finish()`)."""
        if self._gpu is None:
            raise RuntimeError("eval_grad_device_begin needs a GPU engine")
        return PendingEvalGrad(self, m, dobs, w)

    def misfit_and_grad(self, m, dobs, w):
        """(phi_d, dphi_d/dm) by one reverse pass."""
        mc = m.detach().clone().requires_grad_(True)
        phi = self.misfit(mc, dobs, w)
        phi.backward()
        return phi.detach(), mc.grad
