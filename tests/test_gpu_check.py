"""The solve check and its fallback (GPU, through the C-ABI).

Every solve ends with a backward-error check (dcx_plan_set_check): an FP32
factor that lost accuracy is refined with FP64 residuals, and a system that
still fails, or met a non-positive pivot, is refactorised and solved in FP64.
The failures are forced from the host: `dcx_plan_inject_fault` makes the
plan factor (1 + rel) A_b for one system (a factor with backward error about
rel / 2), and a negated value set makes a system negative definite.  The
reference's counterpart is its guarded solve, a residual check followed by
a fallback (dctorch/optimization.py:753-761).
"""
import warnings

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from helpers import case, laplacian_fixture, noisy_data, rel

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(msg):
    print(msg, flush=True)


def _batch(nb=3, nx=30, n_rhs=9):
    A, b = laplacian_fixture(nx=nx, n_rhs=n_rhs)
    scales = [1.0, 2.5, 0.3, 1.7][:nb]
    vals = np.stack([A.data * s for s in scales])
    return A, b, vals


def _solver(A, n_rhs, nbatch, **kw):
    from simpeg_dc_torch_amd import make_batch_solver
    return make_batch_solver(torch.tensor(A.indptr, dtype=torch.int32),
                             torch.tensor(A.indices, dtype=torch.int32), A.shape[0], n_rhs,
                             nbatch, device=DEV, **kw)


def _exact(A, vals, b):
    return np.stack([spla.splu(sp.csr_matrix((v, A.indices, A.indptr), shape=A.shape).tocsc())
                     .solve(b) for v in vals])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_check_passes_on_ordinary_solves(precision):
    A, b, vals = _batch()
    _say(f"check on an ordinary solve, {precision}")
    slv = _solver(A, 9, 3, precision=precision, refine=0)
    slv.factorize(torch.tensor(vals, device=DEV))
    x = slv.solve(torch.tensor(b, device=DEV)).cpu().numpy()
    rep = slv.check_report()
    _say(f"rho {rep['rho']} eta {rep['eta']}")
    assert rep["status"] == "passed" and rep["passes"] == 0
    assert not rep["pivot_fail"].any()
    assert (rep["rho"] < (1e-12 if precision == "fp64" else 1e-5)).all()
    assert rel(x, _exact(A, vals, b)) < (1e-11 if precision == "fp64" else 1e-5)
    assert slv.fallbacks == []


def test_injected_fault_is_refined_away():
    """A factor of 1.001 A (relative residual 1e-3) fails a check at 1e-5;
    FP64 residual refinement (contraction 1e-3 per pass) brings every system
    back within the tolerance, without a fallback."""
    A, b, vals = _batch()
    slv = _solver(A, 9, 3, precision="mixed", refine=0)
    slv.set_check(tol=1e-5, max_passes=3)
    slv.inject_fault(1, 1e-3)
    slv.factorize(torch.tensor(vals, device=DEV))
    x = slv.solve(torch.tensor(b, device=DEV)).cpu().numpy()
    rep = slv.check_report()
    _say(f"refined: passes {rep['passes']} rho {rep['rho']}")
    assert rep["status"] == "passed" and 1 <= rep["passes"] <= 3
    assert slv.fallbacks == []
    assert rel(x[1], _exact(A, vals, b)[1]) < 1e-5


@pytest.mark.parametrize("f32_fields", [False, True])
def test_injected_fault_falls_back_to_fp64(f32_fields):
    """A factor of 2 A (relative residual 1/2, contraction 1/2 per pass) with no
    refinement allowed: the failing system is refactorised and solved in FP64
    on a second plan; its solution equals SciPy's to FP64 accuracy (FP32
    fields: to FP32 rounding), the others keep their FP32 factor."""
    from simpeg_dc_torch_amd import _lib
    A, b, vals = _batch()
    slv = _solver(A, 9, 3, precision="mixed", refine=0)
    slv.set_check(tol=1e-4, max_passes=0)
    slv.inject_fault(2, 1.0)
    v = torch.tensor(vals, device=DEV)
    ex = _exact(A, vals, b)
    if not f32_fields:
        slv.factorize(v)
        with pytest.warns(RuntimeWarning, match="systems \\[2\\]"):
            x = slv.solve(torch.tensor(b, device=DEV)).cpu().numpy()
        assert rel(x[2], ex[2]) < 1e-11
    else:
        perm = torch.tensor(slv.perm().astype(np.int64), device=DEV)
        bp = torch.tensor(b, device=DEV)[perm].contiguous()
        x32 = torch.empty(3, A.shape[0], 9, dtype=torch.float32, device=DEV)
        with torch.cuda.device(slv.device), pytest.warns(RuntimeWarning):
            slv.run_factor_solve(v, bp, x32, 9, _lib.DCX_PERMUTED | _lib.DCX_SHARED_RHS |
                                 _lib.DCX_F32_FIELDS)
        x = torch.empty_like(x32, dtype=torch.float64)
        x[:, perm] = x32.double()
        x = x.cpu().numpy()
        assert rel(x[2], ex[2]) < 1e-7
    assert slv.fallbacks and slv.fallbacks[-1][1] == [2]
    assert rel(x[0], ex[0]) < 1e-5 and rel(x[1], ex[1]) < 1e-5


def test_non_spd_system_raises():
    """A negated value set (negative definite system 1): the FP32 factor meets
    a non-positive pivot, the FP64 fallback fails too, and the solve raises
    NumericError naming the system; an FP64 plan raises at once."""
    from simpeg_dc_torch_amd._lib import NumericError
    A, b, vals = _batch()
    vals[1] = -vals[1]
    for prec in ("mixed", "fp64"):
        _say(f"non-SPD values, {prec}")
        slv = _solver(A, 9, 3, precision=prec, refine=0)
        slv.factorize(torch.tensor(vals, device=DEV))
        with pytest.raises(NumericError, match="system"):
            slv.solve(torch.tensor(b, device=DEV))
        rep = slv.check_report()
        assert rep["status"] == "failed" and rep["pivot_fail"][1]
        assert not rep["pivot_fail"][0] and not rep["pivot_fail"][2]


def test_engine_eval_grad_with_fallback_matches_fp64():
    """The 2.5-D engine in mixed precision with one wavenumber's FP32 factor
    forced to fail: that wavenumber runs in FP64 (forward and adjoint), and
    the eval+grad agrees with the FP64 engine to mixed-precision accuracy."""
    from simpeg_dc_torch_amd import TorchDC2D
    mesh, survey, m = case("suite")
    e64 = TorchDC2D(mesh, survey, nky=6, device=DEV, precision="fp64")
    emx = TorchDC2D(mesh, survey, nky=6, device=DEV, precision="mixed", refine=0)
    emx.solver.set_check(tol=1e-4, max_passes=0)
    emx.solver.inject_fault(3, 1.0)
    mt = torch.tensor(m, device=DEV)
    with torch.no_grad():
        d = e64.dpred(mt).cpu().numpy()
    dobs, w = noisy_data(d)
    out = {}
    for name, eng in (("fp64", e64), ("mixed", emx)):
        x = mt.clone().requires_grad_(True)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            phi = eng.misfit(x, torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV))
            phi.backward()
        out[name] = (float(phi), x.grad.cpu().numpy())
    assert [f[1] for f in emx.solver.fallbacks] == [[3], [3]]   # forward and adjoint
    assert abs(out["mixed"][0] / out["fp64"][0] - 1) < 1e-5
    assert rel(out["mixed"][1], out["fp64"][1]) < 1e-3
