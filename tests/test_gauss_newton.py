"""Gauss-Newton step algebra (simpeg_dc_torch_amd/gauss_newton.py) against
dense FP64 linear algebra and the defining equations; the engine-level step on
the CPU path, and on the GPU through the C-ABI solver (marked gpu).

What pins each function: `pcg` and `direct_step` against numpy's dense solve
of (J^T D J + beta H_m) p = -g; `occam_step` against `direct_step` at the beta
it returns and against its own predicted-misfit definition, plus the search
rule (the largest beta reaching the target); `cg_step` against the dense
solve at a tight tolerance; `tensor_regularization` against its closed form
on a 3 x 2 mesh; the engine step by the misfit decrease of the full step and
the gradient against the engine's autograd gradient.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from helpers import case, noisy_data, rel


def _problem(nD=30, nM=120, seed=0):
    rng = np.random.default_rng(seed)
    J = rng.standard_normal((nD, nM))
    D = rng.uniform(0.5, 2.0, nD)
    # H_m: a 1-D second-difference chain plus a small smallness term (SPD)
    main = np.full(nM, 2.0)
    main[[0, -1]] = 1.0
    H = sp.diags([main + 1e-2, -np.ones(nM - 1), -np.ones(nM - 1)], [0, -1, 1], format="csr")
    g = rng.standard_normal(nM)
    return J, D, H, g


def _dense_step(J, D, H, g, beta):
    A = J.T @ np.diag(D) @ J + beta * H.toarray()
    return np.linalg.solve(A, -g)


def test_pcg_matches_dense_solve():
    from simpeg_dc_torch_amd.gauss_newton import pcg
    J, D, H, g = _problem()
    A = torch.tensor(J.T @ np.diag(D) @ J + 0.7 * H.toarray())
    b = torch.tensor(-g)
    x, it = pcg(lambda v: A @ v, b, 1.0 / torch.diagonal(A), 500, 1e-13)
    assert rel(x.numpy(), np.linalg.solve(A.numpy(), b.numpy())) < 1e-10
    assert it < 500


@pytest.mark.parametrize("beta", [0.05, 3.0])
def test_direct_step_matches_dense_solve(beta):
    from simpeg_dc_torch_amd.gauss_newton import HmSolver, direct_step
    J, D, H, g = _problem()
    p = direct_step(torch.tensor(J), torch.tensor(D), torch.tensor(g), H, beta, HmSolver("cpu"))
    assert rel(p.numpy(), _dense_step(J, D, H, g, beta)) < 1e-10


def test_direct_step_refuses_singular_hm():
    from simpeg_dc_torch_amd.gauss_newton import HmSolver, direct_step
    J, D, _, g = _problem()
    nM = J.shape[1]
    H = sp.csr_matrix((nM, nM))  # alpha_s = 0 and no smoothness: singular
    assert direct_step(torch.tensor(J), torch.tensor(D), torch.tensor(g), H, 1.0,
                       HmSolver("cpu")) is None


@pytest.mark.parametrize("scale,branch", [(1.0, "top"), (30.0, "bisect"), (1000.0, "lowest")])
def test_occam_step_consistency_and_rule(scale, branch):
    """scale sets the regularisation strength so that the target is reached
    at the top of the beta grid, between two grid points (bisection), or
    nowhere on the grid (the lowest prediction is taken)."""
    from simpeg_dc_torch_amd.gauss_newton import HmSolver, direct_step, occam_step
    rng = np.random.default_rng(2)
    J, D, H, _ = _problem(seed=2)
    H = H * scale
    nD, nM = J.shape
    w = np.sqrt(D / 2.0)                       # D = 2 W^2
    r = 3.0 * rng.standard_normal(nD)           # weighted residual, phi_d = 9 nD-ish
    g_d = J.T @ (w * 2.0 * r)
    g_m = H @ rng.standard_normal(nM)
    hs = HmSolver("cpu")
    phi = lambda rr: (rr ** 2).sum()
    T = lambda: 0.8 * 1.0 * nD
    res = occam_step(torch.tensor(J), torch.tensor(D), torch.tensor(g_d), torch.tensor(g_m),
                     torch.tensor(r), 1.0, H, hs, phi, torch.tensor(w))
    beta, p, pred = res
    # the step at the chosen beta is the exact step of that beta
    g = torch.tensor(g_d + beta * g_m)
    p_direct = direct_step(torch.tensor(J), torch.tensor(D), g, H, beta, hs)
    assert rel(p.numpy(), p_direct.numpy()) < 1e-8
    # its prediction is the linearised misfit of that step
    assert pred == pytest.approx(float(phi(torch.tensor(r) + torch.tensor(w) * (torch.tensor(J) @ p))),
                                 rel=1e-9)
    # the search rule: the chosen beta reaches the target and a slightly
    # larger one does not (bisection between grid points), or no grid point
    # reaches it
    def pred_of(b):
        gb = torch.tensor(g_d + b * g_m)
        pb = direct_step(torch.tensor(J), torch.tensor(D), gb, H, b, hs)
        return float(phi(torch.tensor(r) + torch.tensor(w) * (torch.tensor(J) @ pb)))
    grid = 1.0 * np.logspace(1, -1, 9)
    if branch == "top":
        assert beta == pytest.approx(10.0) and pred <= T()
    elif branch == "bisect":
        assert grid[-1] < beta < grid[0] and pred <= T() < pred_of(beta * 1.01)
    else:
        assert pred > T()
        assert min(pred_of(b) for b in grid) == pytest.approx(pred, rel=1e-8)


def test_cg_step_matches_dense_solve():
    from simpeg_dc_torch_amd.gauss_newton import cg_step
    J, D, H, g = _problem()
    Jt, Dt, gt = torch.tensor(J), torch.tensor(D), torch.tensor(g)
    Ht = torch.tensor(H.toarray())
    jtdj = torch.tensor(np.einsum("ij,i,ij->j", J, D, J))
    p, it = cg_step(lambda v: Jt @ v, lambda u: Jt.T @ u, Dt, gt, lambda v: Ht @ v,
                    torch.diagonal(Ht), 0.7, jtdj, maxiter=500, rtol=1e-12)
    assert rel(p.numpy(), _dense_step(J, D, H, g, 0.7)) < 1e-9


def test_tensor_regularization_closed_form():
    """3 x 2 cells, hx = 1, hz = 2, alpha_s = 0: across an x-face the weight
    is 2 * area * dist / dist^2 = 2 * 2 / 1 = 4, across a z-face
    2 * 1 * 2 / 4 = 1; rows sum to zero; with alpha_s the diagonal gains
    2 alpha_s vol."""
    from simpeg_dc_torch_amd.gauss_newton import tensor_regularization
    from simpeg_dc_torch_amd.mesh import TensorMesh
    mesh = TensorMesh([[(1.0, 3)], [(2.0, 2)]], origin="00")
    H = tensor_regularization(mesh, alpha_s=0.0).toarray()
    assert H[0, 1] == pytest.approx(-4.0) and H[1, 2] == pytest.approx(-4.0)
    assert H[0, 3] == pytest.approx(-1.0) and H[0, 2] == 0.0
    assert np.abs(H.sum(1)).max() < 1e-12 and np.abs(H - H.T).max() == 0.0
    H2 = tensor_regularization(mesh, alpha_s=0.5).toarray()
    assert np.allclose(np.diag(H2 - H), 2 * 0.5 * 2.0)


def _engine_problem(device):
    from simpeg_dc_torch_amd import TorchDC2D
    from simpeg_dc_torch_amd.gauss_newton import tensor_regularization
    mesh, survey, m = case("suite")
    eng = TorchDC2D(mesh, survey, nky=6, device=device)
    with torch.no_grad():
        d = eng.dpred(torch.tensor(m, device=device)).cpu().numpy()
    dobs, w = noisy_data(d, seed=5)
    m0 = np.full(m.size, np.log(100.0))
    Hm = tensor_regularization(mesh, alpha_s=1e-3)
    return eng, m0, dobs, w, Hm


@pytest.mark.parametrize("inner", ["direct", "occam", "cg"])
def test_engine_step_cpu(inner):
    """One step at a half-space start: the gradient is the engine's autograd
    gradient of phi_d + beta phi_m, and the full step lowers the objective
    (occam: the misfit)."""
    from simpeg_dc_torch_amd.gauss_newton import gauss_newton_step
    eng, m0, dobs, w, Hm = _engine_problem("cpu")
    beta = 1.0
    out = gauss_newton_step(eng, m0, dobs, w, Hm, beta, inner=inner, cg_maxiter=60,
                            cg_rtol=1e-3)
    p = out["p"].numpy()
    x = torch.tensor(m0, requires_grad=True)
    phi_d = eng.misfit(x, torch.tensor(dobs), torch.tensor(w))
    phi_d.backward()
    b = out["beta"]
    g_ref = x.grad.numpy() + b * (Hm @ m0)
    assert rel(out["g"].numpy(), g_ref) < 1e-9

    def f(mm):
        with torch.no_grad():
            dd = eng.dpred(torch.tensor(mm)).numpy()
        return float(((w * (dd - dobs)) ** 2).sum()), float(mm @ (Hm @ mm)) / 2.0
    fd0, fm0 = f(m0)
    fd1, fm1 = f(m0 + p)
    if inner == "occam":
        assert fd1 < fd0
    else:
        assert fd1 + beta * fm1 < fd0 + beta * fm0


@pytest.mark.gpu
@pytest.mark.parametrize("inner", ["direct", "occam"])
def test_engine_step_gpu_matches_cpu(inner):
    """The step on the GPU (J from the HIP linearisation, H_m through the HIP
    SPD solver, dense algebra in torch on the device) against the same step
    on the CPU path."""
    from simpeg_dc_torch_amd.gauss_newton import HmSolver, gauss_newton_step
    eng_c, m0, dobs, w, Hm = _engine_problem("cpu")
    eng_g, _, _, _, _ = _engine_problem("cuda")
    hs = HmSolver("cuda")
    out_g = gauss_newton_step(eng_g, m0, dobs, w, Hm, 1.0, inner=inner, hm_solve=hs)
    out_c = gauss_newton_step(eng_c, m0, dobs, w, Hm, 1.0, inner=inner)
    assert hs.last == "device"
    assert out_g["beta"] == pytest.approx(out_c["beta"], rel=1e-6)
    assert rel(out_g["p"].cpu().numpy(), out_c["p"].numpy()) < 1e-7


def test_lbfgs_3d_cpu_decreases_objective():
    """L-BFGS closures on the 3-D engine (CPU path): the objective of the
    accepted iterate is below the start and the history is one entry per
    closure."""
    from simpeg_dc_torch_amd import TorchDC3D, lbfgs_minimize
    from simpeg_dc_torch_amd.gauss_newton import tensor_regularization
    mesh, survey, m = case("small_3d")
    eng = TorchDC3D(mesh, survey, device="cpu")
    with torch.no_grad():
        d = eng.dpred(torch.tensor(m)).numpy()
    dobs, w = noisy_data(d, seed=7)
    m0 = np.full(m.size, float(np.mean(m)))
    Hm = tensor_regularization(mesh, alpha_s=1e-3)
    m1, hist = lbfgs_minimize(eng, m0, dobs, w, Hm=Hm, beta=0.1, max_iter=6)
    f = lambda pd, pm: pd + 0.1 * pm
    assert f(*hist[-1]) < f(*hist[0])
    with torch.no_grad():
        d1 = eng.dpred(m1).numpy()
    assert float(((w * (d1 - dobs)) ** 2).sum()) < hist[0][0]


def test_beta_by_eig_matches_dense_eigenvalues():
    """With enough power iterations the estimate is ratio * lam_max(J^T D J)
    / lam_max(H_m) from numpy's dense eigensolver (D = 2 W^2)."""
    from simpeg_dc_torch_amd.gauss_newton import beta_by_eig
    J, D, _, _ = _problem()
    # an H_m with a well separated largest eigenvalue, so the power iteration
    # converges to 1e-6 in a few hundred steps
    hd = np.random.default_rng(3).uniform(1.0, 2.0, J.shape[1])
    hd[7] = 4.0
    H = sp.diags(hd, format="csr")
    w = np.sqrt(D / 2.0)
    Jt = torch.tensor(J)
    beta = beta_by_eig(lambda v: Jt @ v, lambda u: Jt.T @ u, w, H, J.shape[1], ratio=100.0,
                       n_iter=400)
    lam_d = np.linalg.eigvalsh(J.T @ np.diag(D) @ J).max()
    lam_m = np.linalg.eigvalsh(H.toarray()).max()
    assert beta == pytest.approx(100.0 * lam_d / lam_m, rel=1e-6)


@pytest.mark.parametrize("inner", ["direct", "cg"])
def test_student_t_step_gradient_and_weights(inner):
    """nu > 0: the step's gradient is the autograd gradient of the Student-t
    misfit (nu + 1) sum log1p(r^2 / nu) plus beta H_m (m - m_ref), and the
    IRLS weight is its Gauss-Newton curvature; a huge nu gives the L2 step
    scaled as (nu + 1) / nu -> 1."""
    from simpeg_dc_torch_amd.gauss_newton import gauss_newton_step, student_t_weights
    eng, m0, dobs, w, Hm = _engine_problem("cpu")
    nu, beta = 4.0, 1.0
    out = gauss_newton_step(eng, m0, dobs, w, Hm, beta, inner=inner, nu=nu, cg_maxiter=60,
                            cg_rtol=1e-3)
    x = torch.tensor(m0, requires_grad=True)
    r = torch.tensor(w) * (eng.dpred(x) - torch.tensor(dobs))
    phi = (nu + 1.0) * torch.log1p(r ** 2 / nu).sum()
    phi.backward()
    assert out["phi_d"] == pytest.approx(float(phi), rel=1e-12)
    g_ref = x.grad.numpy() + beta * (Hm @ m0)
    assert rel(out["g"].numpy(), g_ref) < 1e-9
    rr = torch.tensor([0.0, 0.5, 3.0])
    _, wd, D = student_t_weights(rr, torch.ones(3), nu)
    # D is the derivative of the gradient weight at r = 0 and the secant
    # form 2 (nu + 1) / (nu + r^2) elsewhere
    assert float(D[0]) == pytest.approx(2.0 * (nu + 1.0) / nu)
    assert torch.allclose(wd, D * rr)


def test_lbfgs_student_t_history_is_the_robust_misfit():
    from simpeg_dc_torch_amd import lbfgs_minimize
    eng, m0, dobs, w, Hm = _engine_problem("cpu")
    nu = 3.0
    m1, hist = lbfgs_minimize(eng, m0, dobs, w, Hm=Hm, beta=1.0, max_iter=4, nu=nu)
    with torch.no_grad():
        r = torch.tensor(w) * (eng.dpred(torch.tensor(m0)) - torch.tensor(dobs))
    assert hist[0][0] == pytest.approx(float((nu + 1.0) * torch.log1p(r ** 2 / nu).sum()),
                                       rel=1e-12)
    assert hist[-1][0] + hist[-1][1] < hist[0][0] + hist[0][1]


def test_log_data_step_gradient():
    """log_data: the step's gradient is the autograd gradient of
    ||w |dobs| (ln d - ln dobs)||^2 + beta phi_m, and near the data the log
    residual is the relative-error L2 residual to first order."""
    from simpeg_dc_torch_amd.gauss_newton import data_residual, gauss_newton_step
    eng, m0, dobs, w, Hm = _engine_problem("cpu")
    out = gauss_newton_step(eng, m0, dobs, w, Hm, 1.0, inner="direct", log_data=True)
    x = torch.tensor(m0, requires_grad=True)
    r, _ = data_residual(eng.dpred(x), torch.tensor(dobs), torch.tensor(w), True)
    (r ** 2).sum().backward()
    assert out["phi_d"] == pytest.approx(float((r.detach() ** 2).sum()), rel=1e-12)
    assert rel(out["g"].numpy(), x.grad.numpy() + Hm @ m0) < 1e-9
    d0 = torch.tensor(dobs)
    r_small, lw = data_residual(d0 * (1 + 1e-6), d0, torch.tensor(w), True)
    assert torch.allclose(r_small, torch.tensor(w) * d0.abs() * 1e-6, rtol=1e-5)
    assert torch.allclose(lw, torch.tensor(w) * d0.abs() / (d0 * (1 + 1e-6)))
