"""Sparse SPD direct solve as a torch op, with the reference's solver interface.

Mirrors dctorch/solver.py: duck-typed solver objects (``n, n_rhs, nbatch,
device, dtype, crow, col, row, stamp, factorize(values), solve(b), free()``),
the factories `make_solver` / `make_batch_solver` that dispatch on
``torch.device(device).type``, and the autograd ops `sparse_solve`,
`sparse_solve_batch`, `sparse_solve_batch_prefactored`.

On a GPU device ("cuda" is the device type of ROCm torch) the solver is
`HipBatchSolver`: the owned supernodal multifrontal Cholesky of libdcx.so
(nested-dissection symbolic phase on the host once per pattern, numeric
factorisation and multi-RHS sweeps on the GPU).  It raises ImportError, with
the reason, when the library is not built; there is no fallback to the CPU.
On the CPU device the solvers are SciPy SuperLU (`splu(A.tocsc())`, defaults),
as in the reference (dctorch/solver.py:314-374).

Backward (dctorch/solver.py:236-281): the adjoint solve reuses the factor
(A symmetric), and the value gradient is
    grad_vals[b, e] = -sum_s adj[b, row_e, s] x[b, col_e, s]
computed by one gather-dot kernel on the GPU.
"""
import ctypes as C
import os
import threading
import warnings

import numpy as np
import torch

from . import _lib

#: elements per gather chunk of the CPU backward (dctorch/solver.py:72-73)
_GRAD_CHUNK_ELEMS = int(os.environ.get("DCTORCH_GRAD_CHUNK", 8_000_000))
#: leaf size of the nested dissection (dense leaf supernodes)
DEFAULT_LEAF = int(os.environ.get("DCX_LEAF", 48))


def _rows_of(crow, n):
    counts = (crow[1:] - crow[:-1]).long()
    return torch.repeat_interleave(torch.arange(n, device=crow.device), counts)


class HipBatchSolver:
    """Uniform batch of SPD systems sharing one sparsity pattern, on the GPU.

    Parameters follow CuDSSBatchSolver (dctorch/solver.py:181-233).  Extra:
    ``precision`` -- "fp64" (default for float64) or "mixed" (FP32 factor
    and sweeps, FP64 residual refinement; results FP64), ``refine`` -- the
    refinement passes of mixed mode, ``coords`` -- optional (n x dim) node
    coordinates for a geometric nested dissection, ``leaf`` -- leaf size.
    """

    def __init__(self, crow, col, n, n_rhs, nbatch, device="cuda",
                 dtype=torch.float64, matrix_type="spd", precision=None,
                 refine=1, coords=None, leaf=None):
        if matrix_type not in ("spd", None):
            raise NotImplementedError(
                f"matrix_type={matrix_type!r}: the owned solver handles real SPD systems")
        if dtype not in (torch.float64, torch.float32):
            raise TypeError(f"unsupported dtype {dtype}")
        lib = _lib.load()
        self.n, self.n_rhs, self.nbatch = int(n), int(n_rhs), int(nbatch)
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        if precision is None:
            precision = "fp64" if dtype == torch.float64 else "fp32"
        if precision not in ("fp64", "mixed", "fp32"):
            raise ValueError("precision must be 'fp64', 'mixed' or 'fp32'")
        self.precision = precision
        self.crow = crow.to(self.device)
        self.col = col.to(self.device)
        crow_h, crow_p = _lib.i32(crow.cpu().numpy())
        col_h, col_p = _lib.i32(col.cpu().numpy())
        if crow_h.size != self.n + 1:
            raise ValueError("crow must have n + 1 entries")
        co_p, dim = None, 0
        if coords is not None:
            co, co_p = _lib.f64(np.asarray(coords, float).reshape(self.n, -1))
            dim = co.shape[1]
            self._coords = co
        fp = _lib.DCX_FP64 if precision == "fp64" else _lib.DCX_FP32
        self._coords_arg = coords
        self.leaf = int(leaf or DEFAULT_LEAF)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.dcx_plan_create(self.n, crow_p, col_p, self.nbatch, co_p, dim,
                                           self.leaf, fp, C.byref(h)),
                       "plan_create")
        self._h = h
        self._lib = lib
        self.set_refinement(refine if precision == "mixed" else 0)
        self.row = _rows_of(self.crow, self.n)
        self.stamp = 0
        self._lock = threading.Lock()
        self._support = None
        self._twin = None          # FP64 plan of the same pattern (fallback)
        self._twin_stamp = -1
        self.fallbacks = []        # (stamp, systems) of every FP64 fallback
        self._pending = []         # the solves whose checks are deferred, oldest first
        self.set_check()

    # -- info ----------------------------------------------------------------
    def info(self):
        inf = _lib.PlanInfo()
        _lib.check(self._lib.dcx_plan_get_info(self._h, C.byref(inf)), "plan_get_info")
        return {f: getattr(inf, f) for f, _ in _lib.PlanInfo._fields_}

    def perm(self):
        out = np.empty(self.n, np.int32)
        _lib.check(self._lib.dcx_plan_get_perm(self._h, out.ctypes.data_as(_lib._i32p)),
                   "plan_get_perm")
        return out

    def set_refinement(self, passes):
        _lib.check(self._lib.dcx_plan_set_refinement(self._h, int(passes)), "set_refinement")
        self.refine = int(passes)

    def set_fusion(self, rows):
        """LDS row budget of the fused bottom subtrees of the sweeps (0: none)."""
        with torch.cuda.device(self.device):
            _lib.check(self._lib.dcx_plan_set_fusion(self._h, int(rows)), "set_fusion")
        self.fuse_rows = int(rows)

    _TUNE_KEYS = {"sweep_tc": 1, "splitk_below": 2, "wide_below": 3, "panel_width": 4,
                  "rowsplit_below": 5}

    def set_tuning(self, **values):
        """Kernel-variant thresholds of this plan: sweep_tc (16, 32, 64),
        splitk_below, wide_below, panel_width (16, 32) (see dcx_plan_set_tuning)."""
        for name, v in values.items():
            if name not in self._TUNE_KEYS:
                raise ValueError(f"unknown tuning key {name!r}")
            _lib.check(self._lib.dcx_plan_set_tuning(self._h, self._TUNE_KEYS[name], int(v)),
                       "plan_set_tuning")

    def tuning(self):
        out = {}
        for name, key in self._TUNE_KEYS.items():
            v = C.c_int()
            _lib.check(self._lib.dcx_plan_get_tuning(self._h, key, C.byref(v)),
                       "plan_get_tuning")
            out[name] = v.value
        return out

    def set_timing(self, enable=True):
        """Record HIP events around the sweeps of every solve (on its stream)."""
        _lib.check(self._lib.dcx_plan_set_timing(self._h, int(bool(enable))), "set_timing")

    def set_overlap(self, enable=True):
        """The plan's own streams beside the caller's (default on): factor /
        forward-sweep overlap, split backward sweep, deferred checks on a
        check stream.  Off, every kernel goes on the caller's stream in order
        (for many plans driven on streams of their own, which would otherwise
        share the device's few hardware queues with the plans' extra ones)."""
        _lib.check(self._lib.dcx_plan_set_overlap(self._h, int(bool(enable))), "set_overlap")

    def sweep_times(self):
        """(forward ms, backward ms, number of solves) since timing was enabled."""
        f, b, n = C.c_double(), C.c_double(), C.c_int()
        _lib.check(self._lib.dcx_plan_sweep_times(self._h, C.byref(f), C.byref(b), C.byref(n)),
                   "sweep_times")
        return f.value, b.value, n.value

    def set_rhs_support(self, rows):
        r, rp = _lib.i32(np.asarray(rows))
        _lib.check(self._lib.dcx_set_rhs_support(self._h, rp, int(r.size)), "set_rhs_support")
        self._support = r

    # -- solve check and FP64 fallback ----------------------------------------
    def set_check(self, tol=None, max_passes=2, sample_cols=8):
        """Residual check of every solve (dcx_plan_set_check): tol None takes
        the default (relative residual 1e-3 for FP32 factors, 1e-8 for FP64,
        or DCX_CHECK_TOL), tol <= 0 turns the check off.  A mixed-precision
        solve failing it gets up to `max_passes` FP64-residual refinement
        passes, then its failing systems are refactorised and solved in FP64
        (see `_fp64_fallback`); an FP64 solve failing it raises
        `NumericError`."""
        if tol is None:
            env = os.environ.get("DCX_CHECK_TOL")
            tol = float(env) if env else (1e-8 if self.precision == "fp64" else 1e-3)
        _lib.check(self._lib.dcx_plan_set_check(self._h, float(tol), int(max_passes),
                                                int(sample_cols)), "plan_set_check")
        self.check_tol = float(tol)

    def check_report(self):
        """The last solve's check: rho (relative residual per system, the
        tested quantity), eta (normwise backward error), pivot_fail, passes
        (refinement passes run), status."""
        rho = np.zeros(self.nbatch)
        eta = np.zeros(self.nbatch)
        piv = np.zeros(self.nbatch, np.int32)
        passes, status = C.c_int(), C.c_int()
        _lib.check(self._lib.dcx_plan_get_check(
            self._h, rho.ctypes.data_as(_lib._f64p), eta.ctypes.data_as(_lib._f64p),
            piv.ctypes.data_as(_lib._i32p), C.byref(passes), C.byref(status)),
            "plan_get_check")
        return {"rho": rho, "eta": eta, "pivot_fail": piv.astype(bool), "passes": passes.value,
                "status": {0: "none", 1: "passed", 2: "failed", 3: "skipped"}[status.value]}

    def inject_fault(self, batch, rel):
        """Tests: later factorisations of system `batch` factor (1 + rel) A
        (dcx_plan_inject_fault); rel = 0 disarms."""
        _lib.check(self._lib.dcx_plan_inject_fault(self._h, int(batch), float(rel)),
                   "plan_inject_fault")

    def _fp64_twin(self):
        if self._twin is None:
            t = HipBatchSolver(self.crow.cpu(), self.col.cpu(), self.n, self.n_rhs, self.nbatch,
                               device=self.device, precision="fp64",
                               coords=self._coords_arg, leaf=self.leaf)
            if not np.array_equal(t.perm(), self.perm()):
                raise RuntimeError("FP64 fallback plan has another elimination order")
            if self._support is not None:
                t.set_rhs_support(self._support)
            self._twin = t
        return self._twin

    def _fp64_fallback(self, vals, b, x, k, flags, exc):
        """The systems whose FP32 solve failed its check (or met a non-positive
        pivot) are factorised and solved again in FP64 on a second plan of
        the same pattern and order; their rows of x are replaced.  Returns
        those systems; raises NumericError when the FP64 factorisation fails
        too (the matrix is not positive definite)."""
        rep = self.check_report()
        bad = np.nonzero(rep["pivot_fail"] | ~(rep["rho"] <= self.check_tol))[0]
        if bad.size == 0:
            raise exc
        twin = self._fp64_twin()
        st = self._stream()
        flags = int(flags) & ~(_lib.DCX_F32_FIELDS | _lib.DCX_CHECK_DEFER | _lib.DCX_SPLIT_BWD)
        try:
            if self._twin_stamp != self.stamp:
                _lib.check(self._lib.dcx_factorize(twin._h, _lib.ptr(vals), st), "factorize")
                self._twin_stamp = self.stamp
            x64 = torch.empty(self.nbatch, self.n, k, dtype=torch.float64, device=self.device)
            _lib.check(self._lib.dcx_solve(twin._h, _lib.ptr(b), _lib.ptr(x64), int(k), flags,
                                           st), "solve")
        except _lib.NumericError as e2:
            raise _lib.NumericError(
                f"systems {bad.tolist()} failed the FP32 solve check ({exc}) and their FP64 "
                f"factorisation failed too: not positive definite ({e2})") from e2
        idx = torch.as_tensor(bad, device=self.device)
        x[idx] = x64[idx].to(x.dtype)
        self.fallbacks.append((self.stamp, bad.tolist()))
        warnings.warn(f"mixed-precision solve: systems {bad.tolist()} failed the backward-error "
                      f"check (relative residual {rep['rho'][bad].tolist()}) after {rep['passes']} "
                      "refinement pass(es); refactorised and solved in FP64", RuntimeWarning,
                      stacklevel=3)
        return bad.tolist()

    def check_wait(self):
        """Verdict of the oldest deferred check (solves run with
        ``defer=True``; dcx_plan_check_wait): [] when it passed (after any
        refinement passes), else the systems that were refactorised and
        solved in FP64, whose rows of that solve's x were replaced (results
        computed from that x before must be computed again).  Raises
        NumericError for an FP64 plan or a system that is not SPD."""
        if not self._pending:
            return []
        p = self._pending.pop(0)
        try:
            with torch.cuda.device(self.device):
                _lib.check(self._lib.dcx_plan_check_wait(self._h), "check_wait")
        except _lib.NumericError as exc:
            if self.precision == "fp64":
                self._pending.clear()
                self._lib.dcx_plan_check_wait(self._h)   # drop the rest as well
                raise
            with torch.cuda.device(self.device):
                return self._fp64_fallback(*p, exc)
        return []

    def check_discard(self):
        """Complete the oldest deferred check without acting on its verdict
        (its solve was superseded: its results are computed again)."""
        if not self._pending:
            return
        self._pending.pop(0)
        with torch.cuda.device(self.device):
            self._lib.dcx_plan_check_wait(self._h)

    def check_wait_all(self):
        """check_wait for every pending deferred check."""
        out = []
        while self._pending:
            out.append(self.check_wait())
        return out

    def join(self):
        """Order the current stream after the second half of the last split
        backward sweep (dcx_plan_join): every row of that solve's x is then
        final."""
        _lib.check(self._lib.dcx_plan_join(self._h, self._stream()), "plan_join")

    def run_solve(self, b, x, k, flags, defer=False):
        """dcx_solve on tensors (b FP64, x FP64 or FP32 with DCX_F32_FIELDS),
        with the FP64 fallback of failing systems; with defer, the check's
        verdict is left to `check_wait` (b and x must stay as they are until
        then; at most two are pending, a third completes the oldest).  The
        caller holds the lock and the device."""
        if self.refine > 0 and self.precision != "fp64":
            self.check_wait_all()        # the library would complete them first
        elif defer and len(self._pending) >= 2:
            self.check_wait()
        fl = int(flags) | (_lib.DCX_CHECK_DEFER if defer else 0)
        try:
            _lib.check(self._lib.dcx_solve(self._h, _lib.ptr(b), _lib.ptr(x), int(k), fl,
                                           self._stream()), "solve")
        except _lib.NumericError as exc:
            if self.precision == "fp64":
                raise
            self._fp64_fallback(self._vals_keep, b, x, k, flags, exc)
            return
        if defer:
            self._pending.append((self._vals_keep, b, x, k, flags))

    def run_factor_solve(self, vals, b, x, k, flags, defer=False):
        """dcx_factor_solve on tensors (vals FP64 (nbatch, nnz)), with the FP64
        fallback of failing systems (defer: as `run_solve`); bumps the stamp.
        The caller holds the lock and the device."""
        self.check_wait_all()            # the factor they refer to is replaced
        self._vals_keep = vals
        self.stamp += 1
        fl = int(flags) | (_lib.DCX_CHECK_DEFER if defer else 0)
        try:
            _lib.check(self._lib.dcx_factor_solve(self._h, _lib.ptr(vals), _lib.ptr(b),
                                                  _lib.ptr(x), int(k), fl, self._stream()),
                       "factor_solve")
        except _lib.NumericError as exc:
            if self.precision == "fp64":
                raise
            self._fp64_fallback(vals, b, x, k, flags & ~_lib.DCX_SPLIT_BWD, exc)
            return
        if defer:
            self._pending.append((vals, b, x, k, flags & ~_lib.DCX_SPLIT_BWD))

    # -- numeric -------------------------------------------------------------
    def _stream(self):
        return _lib.stream_of(self.device)

    def factorize(self, values):
        """values: (nbatch, nnz) -- refactorise every system, same pattern."""
        v = values.detach().to(device=self.device, dtype=torch.float64).contiguous()
        if v.shape != (self.nbatch, self.col.numel()):
            v = v.reshape(self.nbatch, self.col.numel())
        with self._lock, torch.cuda.device(self.device):
            self.check_wait_all()
            _lib.check(self._lib.dcx_factorize(self._h, _lib.ptr(v), self._stream()),
                       "factorize")
            self._vals_keep = v
        self.stamp += 1

    def solve(self, b, flags=0):
        """Solve all systems. b: (nbatch, n, k) or (n, k) shared -> (nbatch, n, k)."""
        shared = b.dim() == 2
        bb = b.detach().to(device=self.device, dtype=torch.float64).contiguous()
        k = bb.shape[-1] if bb.dim() >= 2 else 1
        if shared:
            bb = bb.reshape(self.n, k)
            flags |= _lib.DCX_SHARED_RHS
        else:
            bb = bb.reshape(self.nbatch, self.n, k)
        x = torch.empty(self.nbatch, self.n, k, dtype=torch.float64, device=self.device)
        with self._lock, torch.cuda.device(self.device):
            self.run_solve(bb, x, k, flags)
        return x.to(self.dtype)

    def factor_solve(self, values, b, flags=0):
        """factorize(values) then solve(b), with the factorisation and the
        solve's forward sweep overlapped on the device (dcx_factor_solve)."""
        v = values.detach().to(device=self.device, dtype=torch.float64).contiguous()
        if v.shape != (self.nbatch, self.col.numel()):
            v = v.reshape(self.nbatch, self.col.numel())
        shared = b.dim() == 2
        bb = b.detach().to(device=self.device, dtype=torch.float64).contiguous()
        k = bb.shape[-1] if bb.dim() >= 2 else 1
        if shared:
            bb = bb.reshape(self.n, k)
            flags |= _lib.DCX_SHARED_RHS
        else:
            bb = bb.reshape(self.nbatch, self.n, k)
        x = torch.empty(self.nbatch, self.n, k, dtype=torch.float64, device=self.device)
        with self._lock, torch.cuda.device(self.device):
            self.run_factor_solve(v, bb, x, k, flags)
        return x.to(self.dtype)

    def grad_values(self, adj, x, scale=None, permuted=False):
        """-scale[b] sum_s adj[b, row_e, s] x[b, col_e, s] -> (nbatch, nnz) FP64.
        Float32 adj and x are read as they are (FP64 sums)."""
        f32 = adj.dtype == torch.float32 and x.dtype == torch.float32
        dt = torch.float32 if f32 else torch.float64
        a = adj.detach().to(device=self.device, dtype=dt).contiguous()
        xx = x.detach().to(device=self.device, dtype=dt).contiguous()
        k = a.shape[-1]
        gv = torch.empty(self.nbatch, self.col.numel(), dtype=torch.float64, device=self.device)
        sc = None if scale is None else scale.to(self.device, torch.float64).contiguous()
        with self._lock, torch.cuda.device(self.device):
            _lib.check(self._lib.dcx_grad_values(
                self._h, _lib.ptr(a), _lib.ptr(xx), int(k), _lib.ptr(sc), _lib.ptr(gv),
                (_lib.DCX_PERMUTED if permuted else 0) | (_lib.DCX_F32_FIELDS if f32 else 0),
                self._stream()), "grad_values")
        return gv

    def free(self):
        if getattr(self, "_twin", None) is not None:
            self._twin.free()
            self._twin = None
        if getattr(self, "_h", None):
            self._lib.dcx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HipSolver(HipBatchSolver):
    """Single SPD system (nbatch = 1), the CuDSSSolver interface
    (dctorch/solver.py:76-138): factorize(values (nnz,)), solve(b (n,)|(n,k))."""

    def __init__(self, crow, col, n, n_rhs, device="cuda", dtype=torch.float64,
                 matrix_type="spd", **kw):
        super().__init__(crow, col, n, n_rhs, 1, device=device, dtype=dtype,
                         matrix_type=matrix_type, **kw)

    def factorize(self, values):
        super().factorize(values.reshape(1, -1))

    def solve(self, b):
        shape = b.shape
        x = super().solve(b.reshape(self.n, -1))
        return x[0].reshape(shape)

    def factor_solve(self, values, b, flags=0):
        shape = b.shape
        x = super().factor_solve(values.reshape(1, -1), b.reshape(self.n, -1), flags)
        return x[0].reshape(shape)


# ---------------------------------------------------------------------------
# CPU solvers: SciPy SuperLU with SciPy defaults (dctorch/solver.py:314-374)
# ---------------------------------------------------------------------------
class SciPySolver:
    def __init__(self, crow, col, n, n_rhs, device="cpu", dtype=torch.float64):
        if torch.device(device).type != "cpu":
            raise ValueError("SciPySolver runs on the cpu device")
        self.n, self.n_rhs = n, n_rhs
        self.device, self.dtype = torch.device("cpu"), dtype
        self.crow = crow.to("cpu")
        self.col = col.to("cpu")
        self._indptr = self.crow.numpy()
        self._indices = self.col.numpy()
        self.row = _rows_of(self.crow, n)
        self._lu = None
        self.stamp = 0

    def factorize(self, values):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        A = sp.csr_matrix((values.detach().cpu().numpy(), self._indices, self._indptr),
                          shape=(self.n, self.n))
        self._lu = spla.splu(A.tocsc())
        self.stamp += 1

    def solve(self, b):
        x = self._lu.solve(np.ascontiguousarray(b.detach().cpu().numpy().reshape(self.n, -1)))
        return torch.as_tensor(x, dtype=self.dtype).reshape(b.shape)

    def free(self):
        self._lu = None


class SciPyBatchSolver:
    def __init__(self, crow, col, n, n_rhs, nbatch, device="cpu", dtype=torch.float64):
        self.n, self.n_rhs, self.nbatch = n, n_rhs, nbatch
        self.device, self.dtype = torch.device("cpu"), dtype
        self.crow = crow.to("cpu")
        self.col = col.to("cpu")
        self._single = [SciPySolver(crow, col, n, n_rhs, dtype=dtype) for _ in range(nbatch)]
        self.row = _rows_of(self.crow, n)
        self.stamp = 0

    def factorize(self, values):
        for i in range(self.nbatch):
            self._single[i].factorize(values[i])
        self.stamp += 1

    def solve(self, b):
        shared = b.dim() == 2
        return torch.stack([s.solve(b if shared else b[i]) for i, s in enumerate(self._single)])

    def free(self):
        for s in self._single:
            s.free()


# ---------------------------------------------------------------------------
# autograd ops
# ---------------------------------------------------------------------------
def _gather_grad(solver, adj, x):
    """-sum_s adj[..., row_e, s] x[..., col_e, s] per stored entry."""
    if isinstance(solver, HipBatchSolver):
        a = adj if adj.dim() == 3 else adj.reshape(1, solver.n, -1)
        xx = x if x.dim() == 3 else x.reshape(1, solver.n, -1)
        return solver.grad_values(a, xx)
    row, col = solver.row, solver.col.long()
    batched = x.dim() == 3
    xm = x if batched else x.reshape(1, solver.n, -1)
    am = adj if batched else adj.reshape(1, solver.n, -1)
    nb, k = xm.shape[0], xm.shape[-1]
    nnz = col.numel()
    out = torch.empty(nb, nnz, dtype=xm.dtype, device=xm.device)
    step = max(_GRAD_CHUNK_ELEMS // max(nb * k, 1), 1)
    for p0 in range(0, nnz, step):
        p1 = min(p0 + step, nnz)
        out[:, p0:p1] = -(am[:, row[p0:p1]] * xm[:, col[p0:p1]]).sum(-1)
    return out


class _SparseSolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, solver):
        if hasattr(solver, "factor_solve"):
            x = solver.factor_solve(values.detach(), b.detach())
        else:
            solver.factorize(values.detach())
            x = solver.solve(b.detach())
        ctx.solver = solver
        ctx.dt_vals, ctx.dt_b = values.dtype, b.dtype
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        # complex-symmetric A (A^T = A): torch passes dL/d conj(x), so
        # adj = A^{-1} conj(grad_x), dL/d conj(b) = conj(adj) and
        # dL/d conj(A_ij) = -conj(adj_i x_j) (solver.py:236-281); every conj
        # is a no-op on real data
        (x,) = ctx.saved_tensors
        solver = ctx.solver
        adj = solver.solve(grad_x.conj().resolve_conj().contiguous())
        grad_vals = None
        if ctx.needs_input_grad[0]:
            g = _gather_grad(solver, adj.reshape(solver.n, -1), x.reshape(solver.n, -1))
            grad_vals = g.reshape(-1).conj().resolve_conj().to(ctx.dt_vals)
        grad_b = adj.conj().resolve_conj().to(ctx.dt_b) if ctx.needs_input_grad[1] else None
        return grad_vals, grad_b, None


def sparse_solve(values, b, solver):
    """Differentiable sparse SPD solve (dctorch/solver.py:176-178)."""
    return _SparseSolveFn.apply(values, b, solver)


class _SparseSolveBatchFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, solver):
        if hasattr(solver, "factor_solve"):
            x = solver.factor_solve(values.detach(), b.detach())
        else:
            solver.factorize(values.detach())
            x = solver.solve(b.detach())
        ctx.solver = solver
        ctx.b_shared = b.dim() == 2
        ctx.dt_vals, ctx.dt_b = values.dtype, b.dtype
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        # complex-symmetric A: see _SparseSolveFn.backward
        (x,) = ctx.saved_tensors
        solver = ctx.solver
        adj = solver.solve(grad_x.conj().resolve_conj().contiguous())
        grad_vals = None
        if ctx.needs_input_grad[0]:
            grad_vals = _gather_grad(solver, adj, x).conj().resolve_conj().to(ctx.dt_vals)
        grad_b = None
        if ctx.needs_input_grad[1]:
            grad_b = (adj.sum(0) if ctx.b_shared else adj).conj().resolve_conj().to(ctx.dt_b)
        return grad_vals, grad_b, None


class _SparseSolveBatchPreFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, solver):
        x = solver.solve(b.detach())
        ctx.solver = solver
        ctx.b_shared = b.dim() == 2
        ctx.dt_vals, ctx.dt_b = values.dtype, b.dtype
        ctx.save_for_backward(x)
        return x

    backward = _SparseSolveBatchFn.backward


def sparse_solve_batch(values, b, solver):
    """Differentiable batched SPD solve on a shared pattern
    (dctorch/solver.py:307-311).  values (nbatch, nnz); b (nbatch, n, k) or
    (n, k) shared."""
    return _SparseSolveBatchFn.apply(values, b, solver)


def sparse_solve_batch_prefactored(values, b, solver):
    """Same solve and gradients on an already-factorised A (solver.py:302-304)."""
    return _SparseSolveBatchPreFn.apply(values, b, solver)


# ---------------------------------------------------------------------------
# factories (dctorch/solver.py:383-399)
# ---------------------------------------------------------------------------
def _default_mtype(dtype, matrix_type):
    if matrix_type is not None:
        return matrix_type
    return "symmetric" if dtype.is_complex else "spd"


def make_solver(crow, col, n, n_rhs, device="cuda", dtype=torch.float64,
                matrix_type=None, **kw):
    """Device-dispatching factory: the owned HIP solver on a GPU, SuperLU on cpu."""
    mt = _default_mtype(dtype, matrix_type)
    if torch.device(device).type == "cuda":
        return HipSolver(crow, col, n, n_rhs, device=device, dtype=dtype, matrix_type=mt, **kw)
    return SciPySolver(crow, col, n, n_rhs, dtype=dtype)


def make_batch_solver(crow, col, n, n_rhs, nbatch, device="cuda", dtype=torch.float64,
                      matrix_type=None, **kw):
    mt = _default_mtype(dtype, matrix_type)
    if torch.device(device).type == "cuda":
        return HipBatchSolver(crow, col, n, n_rhs, nbatch, device=device, dtype=dtype,
                              matrix_type=mt, **kw)
    return SciPyBatchSolver(crow, col, n, n_rhs, nbatch, dtype=dtype)


# reference class names, for drop-in imports
CuDSSSolver = HipSolver
CuDSSBatchSolver = HipBatchSolver
