"""GPU parity: libdcx.so kernels (through the C-ABI) against the oracle.

Tolerances follow the reference's tests: solver vs SciPy 1e-11
(test_solver.py:39), value/RHS gradients vs FD 1e-5 (:78), narrow RHS 1e-13
(:93), dpred vs the CPU path 1e-10 (test_simulation2d.py:80), gradient,
J^T v and J 1e-10 (the FP64 bar of the build's target); mixed precision
(FP32 factor + FP64 refinement) dpred <= 1e-6.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from helpers import case, laplacian_fixture, noisy_data, oracle_for, rel

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(msg):
    print(msg, flush=True)


def _solver(A, n_rhs, nbatch=1, **kw):
    from simpeg_dc_torch_amd import make_batch_solver
    return make_batch_solver(torch.tensor(A.indptr, dtype=torch.int32),
                             torch.tensor(A.indices, dtype=torch.int32), A.shape[0], n_rhs,
                             nbatch, device=DEV, **kw)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_mfma_lane_layout(dtype):
    """The 16x16x4 matrix instruction computes A(16x4) B(4x16) with the lane
    layout the sweep kernels assume (checked in bounds, one wavefront)."""
    from simpeg_dc_torch_amd import _lib
    lib = _lib.load()
    g = torch.Generator().manual_seed(0)
    A = torch.randn(16, 4, generator=g, dtype=dtype).to(DEV)
    B = torch.randn(4, 16, generator=g, dtype=dtype).to(DEV)
    C = torch.full((16, 16), float("nan"), dtype=dtype, device=DEV)
    _lib.check(lib.dcx_mma_probe(int(dtype == torch.float32), _lib.ptr(A), _lib.ptr(B),
                                 _lib.ptr(C), _lib.stream_of(torch.device(DEV))), "mma_probe")
    torch.cuda.synchronize()
    ref = A.double().cpu() @ B.double().cpu()
    tol = 1e-13 if dtype == torch.float64 else 1e-5
    assert float((C.double().cpu() - ref).abs().max()) < tol


def test_solver_matches_scipy():
    from simpeg_dc_torch_amd import CuDSSSolver, sparse_solve
    _say("solver fixture nx=40, 5 rhs")
    A, b = laplacian_fixture()
    n = A.shape[0]
    slv = CuDSSSolver(torch.tensor(A.indptr, dtype=torch.int32),
                      torch.tensor(A.indices, dtype=torch.int32), n, b.shape[1], device=DEV)
    x = sparse_solve(torch.tensor(A.data, device=DEV), torch.tensor(b, device=DEV), slv)
    assert rel(x.cpu().numpy(), spla.spsolve(A.tocsc(), b)) < 1e-11
    slv.free()


def test_solver_gradients_vs_fd():
    from simpeg_dc_torch_amd import CuDSSSolver, sparse_solve
    A, b = laplacian_fixture(nx=12, n_rhs=3)
    n = A.shape[0]
    slv = CuDSSSolver(torch.tensor(A.indptr, dtype=torch.int32),
                      torch.tensor(A.indices, dtype=torch.int32), n, 3, device=DEV)
    rng = np.random.default_rng(1)
    w = rng.standard_normal((n, 3))
    vals = torch.tensor(A.data, device=DEV, requires_grad=True)
    bt = torch.tensor(b, device=DEV, requires_grad=True)
    (torch.tensor(w, device=DEV) * sparse_solve(vals, bt, slv)).sum().backward()

    def loss(data, bb):
        Am = sp.csr_matrix((data, A.indices, A.indptr), shape=A.shape)
        return float((w * spla.spsolve(Am.tocsc(), bb)).sum())

    eps = 1e-6
    for p in rng.choice(A.nnz, 12, replace=False):
        d = A.data.copy()
        d[p] += eps
        fd = (loss(d, b) - loss(A.data, b)) / eps
        assert abs(fd - float(vals.grad[p])) / (abs(fd) + 1e-12) < 1e-5
    for _ in range(6):
        i, j = rng.integers(n), rng.integers(3)
        bb = b.copy()
        bb[i, j] += eps
        fd = (loss(A.data, bb) - loss(A.data, b)) / eps
        assert abs(fd - float(bt.grad[i, j])) / (abs(fd) + 1e-12) < 1e-5


def test_narrow_rhs_through_wide_plan():
    from simpeg_dc_torch_amd import CuDSSSolver, sparse_solve
    A, b = laplacian_fixture(nx=95, n_rhs=140)
    slv = CuDSSSolver(torch.tensor(A.indptr, dtype=torch.int32),
                      torch.tensor(A.indices, dtype=torch.int32), A.shape[0], 140, device=DEV)
    vals = torch.tensor(A.data, device=DEV)
    bt = torch.tensor(b, device=DEV)
    x = sparse_solve(vals, bt, slv)
    xn = sparse_solve(vals, bt[:, :7], slv)
    assert float(torch.linalg.norm(xn - x[:, :7]) / torch.linalg.norm(x[:, :7])) < 1e-13
    assert rel(x.cpu().numpy(), spla.splu(A.tocsc()).solve(b)) < 1e-11


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-9)])
def test_batch_solver_shared_and_batched(precision, tol):
    A, b = laplacian_fixture(nx=30, n_rhs=9)
    rng = np.random.default_rng(4)
    nb = 3
    scales = [1.0, 2.5, 0.3]
    vals = np.stack([A.data * s + (i * 0.05) * (A.indices == np.repeat(
        np.arange(A.shape[0]), np.diff(A.indptr))) for i, s in enumerate(scales)])
    slv = _solver(A, 9, nbatch=nb, precision=precision, refine=2)
    slv.factorize(torch.tensor(vals, device=DEV))
    xs = slv.solve(torch.tensor(b, device=DEV)).cpu().numpy()
    bb = rng.standard_normal((nb, A.shape[0], 9))
    xb = slv.solve(torch.tensor(bb, device=DEV)).cpu().numpy()
    for i in range(nb):
        Ai = sp.csr_matrix((vals[i], A.indices, A.indptr), shape=A.shape)
        lu = spla.splu(Ai.tocsc())
        assert rel(xs[i], lu.solve(b)) < tol
        assert rel(xb[i], lu.solve(bb[i])) < tol


@pytest.mark.parametrize("refine", [0, 1])
def test_f32_fields_match_f64_output(refine):
    """DCX_F32_FIELDS: the FP32 solution block of a mixed plan is the FP64
    output rounded to float (bit for bit), and the gather and projection read
    it exactly as they read its FP64 copy."""
    from simpeg_dc_torch_amd import _lib
    _say(f"f32 fields refine={refine}")
    A, _ = laplacian_fixture(nx=30, n_rhs=1)
    n, k, nb = A.shape[0], 70, 2
    rng = np.random.default_rng(11)
    slv = _solver(A, k, nbatch=nb, precision="mixed", refine=refine)
    slv.factorize(torch.tensor(np.stack([A.data, 1.7 * A.data]), device=DEV))
    b = torch.tensor(rng.standard_normal((n, k)), device=DEV)
    lib, st = slv._lib, slv._stream()
    base = _lib.DCX_PERMUTED | _lib.DCX_SHARED_RHS
    x64 = torch.empty(nb, n, k, dtype=torch.float64, device=DEV)
    x32 = torch.empty(nb, n, k, dtype=torch.float32, device=DEV)
    _lib.check(lib.dcx_solve(slv._h, _lib.ptr(b), _lib.ptr(x64), k, base, st), "solve")
    _lib.check(lib.dcx_solve(slv._h, _lib.ptr(b), _lib.ptr(x32), k,
                             base | _lib.DCX_F32_FIELDS, st), "solve f32")
    torch.cuda.synchronize()
    assert torch.equal(x32, x64.float())
    y32 = torch.tensor(rng.standard_normal((nb, n, k)), dtype=torch.float32, device=DEV)
    sc = torch.tensor([0.5, 2.0], dtype=torch.float64, device=DEV)
    g32 = slv.grad_values(y32, x32, scale=sc, permuted=True)
    g64 = slv.grad_values(y32.double(), x32.double(), scale=sc, permuted=True)
    assert torch.equal(g32, g64)
    # the FP32 result against SciPy at FP32-factor accuracy
    perm = slv.perm()  # elimination row -> original row
    xo, bo = np.empty((n, k)), np.empty((n, k))
    bo[perm] = b.cpu().numpy()
    for i, s in enumerate([1.0, 1.7]):
        Ai = sp.csr_matrix((s * A.data, A.indices, A.indptr), shape=A.shape)
        xo[perm] = x32[i].double().cpu().numpy()
        ref = spla.splu(Ai.tocsc()).solve(bo)
        assert rel(xo, ref) < (1e-5 if refine == 0 else 1e-6)


@pytest.mark.parametrize("k", [129, 301, 849, 1103])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gather_long_rows_against_reference(k, dtype):
    """The gather of rows longer than 128 columns (whole passes of each row's
    loads in flight, the last vector of a row read in place and masked):
    gv[b][e] = -scale[b] sum_c y[b][row_e][c] x[b][col_e][c] against the same
    sums in FP64 on the host, for widths that are not multiples of the vector
    width and wider than one pass."""
    _say(f"long-row gather k={k} {dtype}")
    A, _ = laplacian_fixture(nx=20, n_rhs=1)
    n, nb = A.shape[0], 2
    slv = _solver(A, k, nbatch=nb)
    rng = np.random.default_rng(k)
    y = torch.tensor(rng.standard_normal((nb, n, k)), dtype=dtype, device=DEV)
    x = torch.tensor(rng.standard_normal((nb, n, k)), dtype=dtype, device=DEV)
    sc = torch.tensor([0.5, 2.0], dtype=torch.float64, device=DEV)
    gv = slv.grad_values(y, x, scale=sc, permuted=True).cpu().numpy()
    iperm = np.argsort(slv.perm())  # original row -> elimination row
    Ac = A.tocoo()
    yy, xx = y.double().cpu().numpy(), x.double().cpu().numpy()
    ref = np.stack([-s * np.einsum("ec,ec->e", yy[b][iperm[Ac.row]], xx[b][iperm[Ac.col]])
                    for b, s in enumerate([0.5, 2.0])])
    assert rel(gv, ref) < 1e-14


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_factor_solve_matches_factorize_then_solve(precision):
    """dcx_factor_solve (factorisation overlapped with the forward sweep on a
    second stream) gives bit for bit the factorize + solve result, dense and
    sparse right-hand sides."""
    _say(f"factor_solve {precision}")
    A, _ = laplacian_fixture(nx=40, n_rhs=1)
    n = A.shape[0]
    rng = np.random.default_rng(5)
    vals = torch.tensor(np.stack([A.data, 1.3 * A.data, 0.7 * A.data]), device=DEV)
    b = torch.tensor(rng.standard_normal((n, 33)), device=DEV)
    rows = np.array([5, 77, 640, 1311])
    bs = torch.zeros(n, 12, dtype=torch.float64, device=DEV)
    bs[torch.tensor(rows, device=DEV)] = torch.tensor(rng.standard_normal((4, 12)), device=DEV)
    slv = _solver(A, 33, nbatch=3, precision=precision, refine=1)
    slv.set_rhs_support(rows)
    slv.factorize(vals)
    x_ref = slv.solve(b)
    from simpeg_dc_torch_amd import _lib
    xs_ref = slv.solve(bs, flags=_lib.DCX_SPARSE_RHS)
    x = slv.factor_solve(vals, b)
    xs = slv.factor_solve(vals, bs, flags=_lib.DCX_SPARSE_RHS)
    torch.cuda.synchronize()
    assert torch.equal(x, x_ref)
    assert torch.equal(xs, xs_ref)


def test_sparse_rhs_support_matches_dense():
    from simpeg_dc_torch_amd import _lib
    A, _ = laplacian_fixture(nx=40, n_rhs=1)
    n = A.shape[0]
    rows = np.array([3, 17, 400, 1200, 1599])
    b = np.zeros((n, 6))
    rng = np.random.default_rng(2)
    b[rows] = rng.standard_normal((rows.size, 6))
    slv = _solver(A, 6, nbatch=2)
    slv.factorize(torch.tensor(np.stack([A.data, 2 * A.data]), device=DEV))
    slv.set_rhs_support(rows)
    xd = slv.solve(torch.tensor(b, device=DEV)).cpu().numpy()
    xs = slv.solve(torch.tensor(b, device=DEV), flags=_lib.DCX_SPARSE_RHS).cpu().numpy()
    assert np.abs(xd - xs).max() <= 1e-14 * np.abs(xd).max()
    assert rel(xd[0], spla.spsolve(A.tocsc(), b)) < 1e-11


# ---------------------------------------------------------------------------
# the engine
# ---------------------------------------------------------------------------
def _engine(name, nky=6, **kw):
    from simpeg_dc_torch_amd import TorchDC2D
    mesh, survey, m = case(name)
    eng = TorchDC2D(mesh, survey, nky=nky, device=DEV, **kw)
    O = oracle_for(mesh, survey, eng.kys, eng.wq_np)
    return eng, O, m


@pytest.mark.parametrize("name", ["suite", "regression", "small_padded"])
def test_engine_fp64_dpred_and_gradient(name):
    _say(f"engine fp64 {name}")
    eng, O, m = _engine(name)
    d_o = O.dpred(m)
    mt = torch.tensor(m, device=DEV, requires_grad=True)
    d = eng.dpred(mt)
    assert rel(d.detach().cpu().numpy(), d_o) < 1e-10
    dobs, w = noisy_data(d_o)
    phi = eng.misfit(mt, torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV))
    phi.backward()
    phi_o, g_o, _ = O.misfit_and_grad(m, dobs, w)
    assert float(phi.detach()) == pytest.approx(phi_o, rel=1e-10)
    assert rel(mt.grad.cpu().numpy(), g_o) < 1e-10
    f, g = eng.misfit_and_grad(torch.tensor(m, device=DEV), torch.tensor(dobs, device=DEV),
                               torch.tensor(w, device=DEV), chunk=5)
    assert f == pytest.approx(phi_o, rel=1e-10)
    assert rel(g.cpu().numpy(), g_o) < 1e-10


@pytest.mark.parametrize("refine", [0, 2])
def test_engine_mixed_precision(refine):
    _say(f"engine mixed small_padded refine={refine}")
    eng, O, m = _engine("small_padded", precision="mixed", refine=refine)
    d_o = O.dpred(m)
    mt = torch.tensor(m, device=DEV, requires_grad=True)
    d = eng.dpred(mt)
    assert rel(d.detach().cpu().numpy(), d_o) < 1e-6
    dobs, w = noisy_data(d_o)
    eng.misfit(mt, torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV)).backward()
    _, g_o, _ = O.misfit_and_grad(m, dobs, w)
    # gradient: FP32-factor accuracy without refinement, 1e-5 with it
    assert rel(mt.grad.cpu().numpy(), g_o) < (1e-5 if refine else 1e-4)


def test_engine_jtvec_jmatrix_chunks():
    _say("engine Jtvec/Jmatrix suite")
    eng, O, m = _engine("suite", nky=4)
    mt = torch.tensor(m, device=DEV)
    J = eng.Jmatrix(mt, block=7).cpu().numpy()
    assert rel(J, O.Jmatrix(m)) < 1e-10
    rng = np.random.default_rng(2)
    v = rng.standard_normal(eng.nD)
    g1 = eng.Jtvec(mt, torch.tensor(v, device=DEV)).cpu().numpy()
    g2 = eng.Jtvec(mt, torch.tensor(v, device=DEV), chunk=4).cpu().numpy()
    assert rel(g1, J.T @ v) < 1e-10
    assert rel(g2, g1) < 1e-12
    u = rng.standard_normal(m.size)
    assert rel(eng.Jvec(mt, torch.tensor(u, device=DEV)).cpu().numpy(), J @ u) < 1e-10


@pytest.mark.parametrize("precision,tol,basis", [("fp64", 1e-10, "sources"),
                                                ("mixed", 1e-5, "sources"),
                                                ("fp64", 1e-10, "electrodes"),
                                                ("mixed", 1e-5, "electrodes")])
def test_linearization_products(precision, tol, basis):
    """Linearization2D on the GPU (dcx_spmm, dcx_jacobian_rows): J v, J^T v,
    diag(J^T D J) and the explicit J against the oracle's Jacobian; the
    adjoint identity u.(J^T v) = (J u).v.  basis "electrodes": every field
    from one solve with the electrode columns (no adjoint solves)."""
    from simpeg_dc_torch_amd import Linearization2D
    _say(f"linearization suite {precision} {basis}")
    eng, O, m = _engine("suite", nky=4, precision=precision, refine=0, rhs_basis=basis)
    mt = torch.tensor(m, device=DEV)
    lin = Linearization2D(eng, mt)
    J = O.Jmatrix(m)
    assert rel(lin.d.cpu().numpy(), O.dpred(m)) < tol
    rng = np.random.default_rng(3)
    u = rng.standard_normal(m.size)
    v = rng.standard_normal(eng.nD)
    Ju = lin.Jvec(torch.tensor(u, device=DEV)).cpu().numpy()
    Jtv = lin.Jtvec(torch.tensor(v, device=DEV)).cpu().numpy()
    assert rel(Ju, J @ u) < tol
    assert rel(Jtv, J.T @ v) < tol
    assert abs(Ju @ v - u @ Jtv) <= 10 * tol * abs(Ju @ v)
    dw = rng.random(eng.nD)
    dd = lin.diag_JtDJ(torch.tensor(dw, device=DEV), sub=33).cpu().numpy()
    assert rel(dd, (J ** 2 * dw[:, None]).sum(0)) < tol
    assert rel(lin.Jmatrix(sub=64).cpu().numpy(), J) < tol
    assert rel(eng.Jmatrix(mt, block=7).cpu().numpy(), J) < tol


@pytest.mark.parametrize("precision,tol_d,tol_g", [("fp64", 1e-10, 1e-10), ("mixed", 1e-6, 1e-4)])
def test_engine3d_against_oracle(precision, tol_d, tol_g):
    """TorchDC3D on the GPU (the 2.5-D kernels with one zero wavenumber)
    against the 3-D oracle: dpred, misfit and gradient."""
    from oracle.dc3d import Problem3D
    from simpeg_dc_torch_amd import TorchDC3D
    _say(f"engine 3-D small {precision}")
    mesh, survey, m = case("small_3d")
    eng = TorchDC3D(mesh, survey, device=DEV, precision=precision, refine=0)
    O = Problem3D(mesh.h[0], mesh.h[1], mesh.h[2], mesh.origin, survey)
    mt = torch.tensor(m, device=DEV)
    d_o = O.dpred(m)
    assert rel(eng.dpred(mt).cpu().numpy(), d_o) < tol_d
    dobs, w = noisy_data(d_o)
    phi, g = eng.misfit_and_grad(mt, torch.tensor(dobs, device=DEV),
                                 torch.tensor(w, device=DEV))
    phi_o, g_o, _ = O.misfit_and_grad(m, dobs, w)
    assert float(phi) == pytest.approx(phi_o, rel=max(tol_d, 1e-12) * 100)
    assert rel(g.cpu().numpy(), g_o) < tol_g


# mixed without refinement is the bare FP32 factor and sweeps (relative
# error a few 1e-6 on this grid); one FP64 refinement step brings dpred to 1e-6
@pytest.mark.parametrize("precision,refine,tol_d,tol_g", [("fp64", 0, 1e-10, 1e-10),
                                                          ("mixed", 0, 5e-6, 1e-3),
                                                          ("mixed", 1, 1e-6, 1e-3)])
def test_engine3d_large_fronts(precision, refine, tol_d, tol_g):
    """3-D fronts too tall for the in-LDS kernels (largest front 645 rows,
    870 with the inverse rows): the grid-wide factor launches, the row-group
    backward sweep and the two-step forward sweep, against the oracle."""
    from oracle.dc3d import Problem3D
    from simpeg_dc_torch_amd import TorchDC3D
    _say(f"engine 3-D medium {precision} refine {refine}")
    mesh, survey, m = case("medium_3d")
    eng = TorchDC3D(mesh, survey, device=DEV, precision=precision, refine=refine)
    O = Problem3D(mesh.h[0], mesh.h[1], mesh.h[2], mesh.origin, survey)
    mt = torch.tensor(m, device=DEV)
    d_o = O.dpred(m)
    assert rel(eng.dpred(mt).cpu().numpy(), d_o) < tol_d
    dobs, w = noisy_data(d_o)
    _, g = eng.misfit_and_grad(mt, torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV))
    _, g_o, _ = O.misfit_and_grad(m, dobs, w)
    assert rel(g.cpu().numpy(), g_o) < tol_g


def test_line25k_subset_against_oracle_and_fd():
    """Full 25k-cell line: dpred of a subset of sources against the oracle
    (same mesh and quadrature), and the gradient against central differences
    of the GPU misfit itself."""
    from simpeg_dc_torch_amd import TorchDC2D
    from simpeg_dc_torch_amd.survey import Survey
    _say("line 25k fp64")
    mesh, survey, m = case("line25k")
    eng = TorchDC2D(mesh, survey, nky=12, device=DEV)
    mt = torch.tensor(m, device=DEV)
    with torch.no_grad():
        d = eng.dpred(mt).cpu().numpy()
    pick = list(range(0, survey.nSrc, 53))
    sub = Survey([survey.source_list[i] for i in pick])
    O = oracle_for(mesh, sub, eng.kys, eng.wq_np)
    sel = np.concatenate([np.where(eng._maps.sidx == i)[0] for i in pick])
    assert rel(d[sel], O.dpred(m)) < 1e-10
    dobs, w = noisy_data(d, rel_noise=0.05)
    dobs_t, w_t = torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV)
    f, g = eng.misfit_and_grad(mt, dobs_t, w_t)
    rng = np.random.default_rng(9)
    eps = 1e-4
    for _ in range(2):
        v = rng.standard_normal(m.size)
        v /= np.linalg.norm(v)
        vt = torch.tensor(v, device=DEV)
        with torch.no_grad():
            fp = float(eng.misfit(mt + eps * vt, dobs_t, w_t))
            fm = float(eng.misfit(mt - eps * vt, dobs_t, w_t))
        fd = (fp - fm) / (2 * eps)
        assert abs(fd - float(g @ vt)) / abs(fd) < 1e-6
    _say("line 25k mixed")
    engm = TorchDC2D(mesh, survey, nky=12, device=DEV, precision="mixed", refine=2)
    with torch.no_grad():
        dm = engm.dpred(mt).cpu().numpy()
    assert rel(dm, d) < 1e-6


def test_grid3d_gradient_vs_fd():
    """The 56 x 56 x 38-cell 3-D workload (126,711 nodes, 200 pole sources,
    6,000 data, fronts up to 3,315 rows), FP64: the misfit gradient against
    central differences of the GPU misfit itself on the cells with the
    largest gradient and a few random ones."""
    from simpeg_dc_torch_amd import TorchDC3D, workloads
    _say("grid3d fp64 gradient vs central differences")
    mesh, survey, m = workloads.grid_3d_case()
    eng = TorchDC3D(mesh, survey, device=DEV)
    mt = torch.tensor(m, device=DEV)
    with torch.no_grad():
        d = eng.dpred(mt)
    rng = np.random.default_rng(3)
    dobs = d * (1 + 0.05 * torch.tensor(rng.standard_normal(eng.nD), device=DEV))
    w = 1.0 / (0.05 * dobs.abs() + 1e-4)
    _, g = eng.misfit_and_grad(mt, dobs, w)
    g = g.cpu().numpy()
    cells = np.r_[np.argsort(-np.abs(g))[:5], rng.choice(m.size, 3, replace=False)]
    def central(c, eps):
        e = torch.zeros_like(mt)
        e[c] = eps
        with torch.no_grad():
            fp = float(eng.misfit(mt + e, dobs, w))
            fm = float(eng.misfit(mt - e, dobs, w))
        return (fp - fm) / (2 * eps)

    # Richardson extrapolation of two central differences (truncation error
    # O(eps^4)): cells next to an electrode make the misfit strongly
    # nonlinear in their log-resistivity
    eps = 2e-3
    for c in cells:
        f1, f2 = central(c, eps), central(c, eps / 2)
        fd = (4 * f2 - f1) / 3
        assert abs(fd - g[c]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (c, f1, f2, fd, g[c])
    del eng
    torch.cuda.empty_cache()


def test_lbfgs_3d_gpu_decreases_objective():
    from simpeg_dc_torch_amd import TorchDC3D, lbfgs_minimize
    from simpeg_dc_torch_amd.gauss_newton import tensor_regularization
    _say("lbfgs 3-D small")
    mesh, survey, m = case("small_3d")
    eng = TorchDC3D(mesh, survey, device=DEV)
    with torch.no_grad():
        d = eng.dpred(torch.tensor(m, device=DEV)).cpu().numpy()
    dobs, w = noisy_data(d, seed=7)
    m0 = np.full(m.size, float(np.mean(m)))
    Hm = tensor_regularization(mesh, alpha_s=1e-3)
    m1, hist = lbfgs_minimize(eng, m0, dobs, w, Hm=Hm, beta=0.1, max_iter=6)
    assert hist[-1][0] + 0.1 * hist[-1][1] < hist[0][0] + 0.1 * hist[0][1]
    assert m1.device.type == "cuda"


@pytest.mark.parametrize("name,precision", [("suite", "fp64"), ("regression", "fp64"),
                                            ("suite", "mixed"), ("line25k", "fp64")])
def test_electrode_basis_matches_source_basis(name, precision):
    """rhs_basis="electrodes" (one solve with the electrode columns, adjoint
    by the electrode-space algebra) against the source basis (the
    reference's formulation): dpred, misfit gradient and J^T v."""
    from simpeg_dc_torch_amd import TorchDC2D
    _say(f"electrode basis {name} {precision}")
    mesh, survey, m = case(name)
    nky = 12 if name == "line25k" else 8
    e_s = TorchDC2D(mesh, survey, nky=nky, device=DEV, precision=precision, refine=0)
    e_e = TorchDC2D(mesh, survey, nky=nky, device=DEV, precision=precision, refine=0,
                    rhs_basis="electrodes")
    mt = torch.tensor(m, device=DEV)
    with torch.no_grad():
        d_s, d_e = e_s.dpred(mt), e_e.dpred(mt)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert float(torch.linalg.norm(d_e - d_s) / torch.linalg.norm(d_s)) < tol
    dobs, w = noisy_data(d_s.cpu().numpy())
    dobs, w = torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV)
    _, g_s = e_s.misfit_and_grad(mt, dobs, w)
    _, g_e = e_e.misfit_and_grad(mt, dobs, w)
    tolg = 1e-10 if precision == "fp64" else 1e-4
    assert float(torch.linalg.norm(g_e - g_s) / torch.linalg.norm(g_s)) < tolg
    v = torch.tensor(np.random.default_rng(4).standard_normal(e_s.nD), device=DEV)
    j_s, j_e = e_s.Jtvec(mt, v), e_e.Jtvec(mt, v)
    assert float(torch.linalg.norm(j_e - j_s) / torch.linalg.norm(j_s)) < tolg


@pytest.mark.parametrize("name", ["small_3d", "medium_3d"])
def test_electrode_basis_3d_matches_source_basis(name):
    from simpeg_dc_torch_amd import TorchDC3D
    _say(f"electrode basis 3-D {name}")
    mesh, survey, m = case(name)
    e_s = TorchDC3D(mesh, survey, device=DEV)
    e_e = TorchDC3D(mesh, survey, device=DEV, rhs_basis="electrodes")
    mt = torch.tensor(m, device=DEV)
    with torch.no_grad():
        d_s, d_e = e_s.dpred(mt), e_e.dpred(mt)
    assert float(torch.linalg.norm(d_e - d_s) / torch.linalg.norm(d_s)) < 1e-12
    dobs, w = noisy_data(d_s.cpu().numpy())
    dobs, w = torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV)
    _, g_s = e_s.misfit_and_grad(mt, dobs, w)
    _, g_e = e_e.misfit_and_grad(mt, dobs, w)
    assert float(torch.linalg.norm(g_e - g_s) / torch.linalg.norm(g_s)) < 1e-10


@pytest.mark.parametrize("precision,basis", [("fp64", "sources"), ("fp64", "electrodes"),
                                             ("mixed", "sources")])
def test_active_cells_restrict_the_full_model(precision, basis):
    """GPU: the active-cell engine (inactive cells held at val_inactive)
    against the full-mesh engine: dpred, gradient, J and the linearisation
    products on the active columns."""
    from simpeg_dc_torch_amd import Linearization2D, TorchDC2D
    _say(f"active cells {precision} {basis}")
    mesh, survey, m = case("suite")
    act = mesh.cell_centers[:, 1] < -1.0
    val = float(np.log(100.0))
    m_full = m.copy()
    m_full[~act] = val
    kw = dict(nky=6, device=DEV, precision=precision, refine=0, rhs_basis=basis)
    eng_f = TorchDC2D(mesh, survey, **kw)
    eng_a = TorchDC2D(mesh, survey, active_cells=act, val_inactive=val, **kw)
    mf = torch.tensor(m_full, device=DEV)
    ma = torch.tensor(m_full[act], device=DEV)
    with torch.no_grad():
        d_f, d_a = eng_f.dpred(mf), eng_a.dpred(ma)
    assert float(torch.linalg.norm(d_a - d_f) / torch.linalg.norm(d_f)) < 1e-13
    dobs, w = noisy_data(d_f.cpu().numpy())
    dobs, w = torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV)
    _, g_f = eng_f.misfit_and_grad(mf, dobs, w)
    _, g_a = eng_a.misfit_and_grad(ma, dobs, w)
    gfa = g_f[torch.tensor(act, device=DEV)]
    assert float(torch.linalg.norm(g_a - gfa) / torch.linalg.norm(gfa)) < 1e-12
    Jf = Linearization2D(eng_f, mf).Jmatrix()
    lin_a = Linearization2D(eng_a, ma)
    Ja = lin_a.Jmatrix()
    assert Ja.shape == (eng_a.nD, int(act.sum()))
    Jfa = Jf[:, torch.tensor(act, device=DEV)]
    assert float(torch.linalg.norm(Ja - Jfa) / torch.linalg.norm(Jfa)) < 1e-12
    u = torch.tensor(np.random.default_rng(2).standard_normal(int(act.sum())), device=DEV)
    tol_jv = 1e-10 if precision == "fp64" else 1e-5  # J v by FP32 solves vs the explicit J
    assert float(torch.linalg.norm(lin_a.Jvec(u) - Jfa @ u) / torch.linalg.norm(Jfa @ u)) < tol_jv


def test_active_cells_3d():
    from simpeg_dc_torch_amd import TorchDC3D
    _say("active cells 3-D")
    mesh, survey, m = case("small_3d")
    act = mesh.cell_centers[:, 2] < mesh.cell_centers[:, 2].max() - 1.5
    val = float(np.log(100.0))
    m_full = m.copy()
    m_full[~act] = val
    eng_f = TorchDC3D(mesh, survey, device=DEV)
    eng_a = TorchDC3D(mesh, survey, device=DEV, active_cells=act, val_inactive=val)
    mf = torch.tensor(m_full, device=DEV)
    ma = torch.tensor(m_full[act], device=DEV)
    with torch.no_grad():
        d_f, d_a = eng_f.dpred(mf), eng_a.dpred(ma)
    assert float(torch.linalg.norm(d_a - d_f) / torch.linalg.norm(d_f)) < 1e-13
    dobs, w = noisy_data(d_f.cpu().numpy())
    dobs, w = torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV)
    _, g_f = eng_f.misfit_and_grad(mf, dobs, w)
    _, g_a = eng_a.misfit_and_grad(ma, dobs, w)
    gfa = g_f[torch.tensor(act, device=DEV)]
    assert float(torch.linalg.norm(g_a - gfa) / torch.linalg.norm(gfa)) < 1e-12


def test_lbfgs_inversion_protocol_recovers_block():
    """The reference's inversion protocol (dctorch/tests/test_protocol2d.py,
    case A, L2) with this package's pieces in place of SimPEG's: the regression
    line (10 ohm-m block in 100 ohm-m, 21 electrodes, DD n = 1..6, nky 11
    SimPEG quadrature), 3% + 1e-6 noise, start at 100 ohm-m, beta0 by the
    Hessians' eigenvalue ratio (100), cooled by 2 per outer iteration to the
    chi^2 = 1 target, L-BFGS inner solves.  The same acceptance: chi^2 < 2, the
    conductor recovered (mean log-rho dip > 0.8 in the block) and the
    background kept (mean drift < 0.6)."""
    from simpeg_dc_torch_amd import Linearization2D, TorchDC2D, lbfgs_minimize
    from simpeg_dc_torch_amd.gauss_newton import beta_by_eig, tensor_regularization
    _say("inversion protocol (L-BFGS, regression line)")
    mesh, survey, m_true = case("regression")
    eng = TorchDC2D(mesh, survey, nky=11, device=DEV, quadrature="simpeg")
    with torch.no_grad():
        d_clean = eng.dpred(torch.tensor(m_true, device=DEV)).cpu().numpy()
    rng = np.random.default_rng(7)
    std = 0.03 * np.abs(d_clean) + 1e-6
    dobs = d_clean + std * rng.standard_normal(d_clean.size)
    w = 1.0 / std
    m0 = np.full(mesh.nC, np.log(100.0))
    Hm = tensor_regularization(mesh, alpha_s=0.16)
    lin = Linearization2D(eng, torch.tensor(m0, device=DEV))
    beta = beta_by_eig(lin.Jvec, lin.Jtvec, w, Hm, mesh.nC, ratio=100.0, device=DEV)
    m = m0.copy()
    nD = eng.nD
    for _ in range(15):
        mt, _ = lbfgs_minimize(eng, m, dobs, w, Hm=Hm, beta=beta, m_ref=m0, max_iter=25)
        m = mt.cpu().numpy()
        with torch.no_grad():
            d = eng.dpred(torch.tensor(m, device=DEV)).cpu().numpy()
        phi_d = float(((w * (d - dobs)) ** 2).sum())
        if phi_d <= nD:
            break
        beta /= 2.0
    cc = mesh.cell_centers
    blk = (np.abs(cc[:, 0] - 5) < 8) & (cc[:, 1] < -3) & (cc[:, 1] > -10)
    chi2 = phi_d / nD
    dip = float(np.log(100.0) - m[blk].mean())
    drift = float(np.abs(m[~blk] - np.log(100.0)).mean())
    print(f"chi2 {chi2:.3f} dip {dip:.2f} drift {drift:.2f}", flush=True)
    assert chi2 < 2.0 and dip > 0.8 and drift < 0.6, (chi2, dip, drift)


@pytest.mark.parametrize("basis", ["sources", "electrodes"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_eval_grad_device_equals_autograd(precision, basis):
    """TorchDC2D.eval_grad_device (no autograd graph, no host sync) gives the
    autograd misfit and gradient bit for bit, on the active-cell engine too;
    so does the two-phase form with both engines' passes queued before
    either is finished (eval_grad_device_begin, then finish)."""
    from simpeg_dc_torch_amd import TorchDC2D
    _say(f"eval_grad_device {precision} {basis}")
    mesh, survey, m = case("suite")
    act = mesh.cell_centers[:, 1] < -1.0
    runs = []
    for active in (None, act):
        eng = TorchDC2D(mesh, survey, nky=6, device=DEV, precision=precision, refine=0,
                        rhs_basis=basis, active_cells=active, val_inactive=float(np.log(100.0)))
        mm = m if active is None else m[act]
        with torch.no_grad():
            d = eng.dpred(torch.tensor(mm, device=DEV)).cpu().numpy()
        dobs, w = noisy_data(d, seed=3)
        dobs_t, w_t = torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV)
        x = torch.tensor(mm, device=DEV, requires_grad=True)
        phi = eng.misfit(x, dobs_t, w_t)
        phi.backward()
        phi2, g2 = eng.eval_grad_device(torch.tensor(mm, device=DEV), dobs_t, w_t)
        assert torch.equal(phi.detach(), phi2) and torch.equal(x.grad, g2)
        runs.append((eng, torch.tensor(mm, device=DEV), dobs_t, w_t, phi.detach(), x.grad))
    pending = [eng.eval_grad_device_begin(mm, dobs_t, w_t) for eng, mm, dobs_t, w_t, _, _ in runs]
    for p, (_, _, _, _, phi, g) in zip(pending, runs):
        phi3, g3 = p.finish()
        assert torch.equal(phi, phi3) and torch.equal(g, g3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_plan_without_overlap_matches(precision):
    """A plan with its extra streams off (dcx_plan_set_overlap: factorisation,
    split backward sweep and deferred checks all on the caller's stream)
    gives the eval+grad of the default plan to rounding (the unsplit backward
    sweep may pick other kernel variants), with its checks passed; switched
    back on, the default results again, bit for bit."""
    from simpeg_dc_torch_amd import TorchDC2D
    _say(f"plan without overlap {precision}")
    mesh, survey, m = case("suite")
    eng = TorchDC2D(mesh, survey, nky=6, device=DEV, precision=precision, refine=0)
    mt = torch.tensor(m, device=DEV)
    with torch.no_grad():
        d = eng.dpred(mt).cpu().numpy()
    dobs, w = (torch.tensor(a, device=DEV) for a in noisy_data(d, seed=5))
    sv = eng.solver
    out = []
    for ov in (True, False, True):
        sv.set_overlap(ov)
        phi, g = eng.eval_grad_device(mt, dobs, w)
        torch.cuda.synchronize()
        assert sv.check_report()["status"] == "passed" and sv.fallbacks == []
        out.append((float(phi), g.cpu().numpy()))
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert out[1][0] == pytest.approx(out[0][0], rel=tol) and rel(out[1][1], out[0][1]) < tol
    assert out[2][0] == out[0][0] and np.array_equal(out[2][1], out[0][1])


def test_eval_grad_device_equals_autograd_3d():
    from simpeg_dc_torch_amd import TorchDC3D
    _say("eval_grad_device 3-D")
    mesh, survey, m = case("small_3d")
    for precision in ("fp64", "mixed"):
        eng = TorchDC3D(mesh, survey, device=DEV, precision=precision, refine=0)
        with torch.no_grad():
            d = eng.dpred(torch.tensor(m, device=DEV)).cpu().numpy()
        dobs, w = noisy_data(d, seed=3)
        dobs_t, w_t = torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV)
        x = torch.tensor(m, device=DEV, requires_grad=True)
        phi = eng.misfit(x, dobs_t, w_t)
        phi.backward()
        phi2, g2 = eng.eval_grad_device(torch.tensor(m, device=DEV), dobs_t, w_t)
        assert torch.equal(phi.detach(), phi2) and torch.equal(x.grad, g2)


def test_plan_memory_through_torch_allocator():
    """The library's device memory comes from torch's caching allocator
    (dcx_set_allocator, registered by the ctypes layer): a plan's factor and
    workspaces show in torch.cuda.memory_allocated while it lives and go back
    when it is freed (the reference's cuDSS memory sat outside torch,
    dctorch/simulation3d.py:160-169)."""
    A, b = laplacian_fixture(nx=60, n_rhs=16)
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    slv = _solver(A, 16, nbatch=4, precision="mixed")
    slv.factorize(torch.tensor(np.stack([A.data] * 4), device=DEV))
    x = slv.solve(torch.tensor(b, device=DEV))
    torch.cuda.synchronize()
    held = torch.cuda.memory_allocated() - before - x.numel() * x.element_size()
    slv_bytes = slv.info()["device_bytes"]
    _say(f"plan memory through torch: {held} bytes (plan reports {slv_bytes})")
    assert held >= slv_bytes
    assert rel(x[0].cpu().numpy(), spla.splu(A.tocsc()).solve(b)) < 1e-5
    mid = torch.cuda.memory_allocated()
    slv.free()
    torch.cuda.synchronize()
    assert mid - torch.cuda.memory_allocated() >= slv_bytes


@pytest.mark.parametrize("dim", ["2d", "3d"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fields_budget_chunks_the_eval_grad(precision, dim):
    """A field-memory budget smaller than the fields of one eval+grad makes
    the fused misfit go through source chunks (fields recomputed per chunk
    instead of stored); the misfit and gradient equal the one-pass ones (to
    rounding), through autograd and through eval_grad_device."""
    from simpeg_dc_torch_amd import TorchDC2D, TorchDC3D
    from simpeg_dc_torch_amd.simulation2d import _field_chunk
    _say(f"fields budget {dim} {precision}")
    if dim == "2d":
        mesh, survey, m = case("suite")
        make = lambda **kw: TorchDC2D(mesh, survey, nky=6, device=DEV, precision=precision,
                                      refine=0, **kw)
    else:
        mesh, survey, m = case("small_3d")
        make = lambda **kw: TorchDC3D(mesh, survey, device=DEV, precision=precision,
                                      refine=0, **kw)
    full = make()
    per_col = 2 * full.nky * full.nN * (4 if precision == "mixed" else 8) + full.nN * 8
    want = max(1, full.nS // 3)
    small = make(fields_max_gib=(want + 0.5) * per_col / 2 ** 30)
    assert _field_chunk(full, True) == full.nS and _field_chunk(small, True) == want < full.nS
    mt = torch.tensor(m, device=DEV)
    with torch.no_grad():
        d = full.dpred(mt).cpu().numpy()
    dobs, w = (torch.tensor(a, device=DEV) for a in noisy_data(d))
    res = []
    for eng in (full, small):
        x = mt.clone().requires_grad_(True)
        phi = eng.misfit(x, dobs, w)
        phi.backward()
        p2, g2 = eng.eval_grad_device(mt, dobs, w)
        p4, g4 = eng.eval_grad_device_begin(mt, dobs, w).finish()   # chunked: no yield
        assert torch.equal(p4, p2) and torch.equal(g4, g2)
        with torch.no_grad():
            p3 = eng.misfit(mt, dobs, w)
        res.append((float(phi), x.grad.cpu().numpy(), float(p2), g2.cpu().numpy(), float(p3)))
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert res[1][0] == pytest.approx(res[0][0], rel=tol)
    assert rel(res[1][1], res[0][1]) < tol and rel(res[1][3], res[0][3]) < tol
    assert res[1][2] == pytest.approx(res[1][0], rel=1e-14)
    assert res[1][4] == pytest.approx(res[0][4], rel=tol)
