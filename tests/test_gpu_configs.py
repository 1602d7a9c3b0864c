"""Parity at the benchmark configurations (GPU, through the C-ABI).

* The 25,088-cell line (849 dipole sources, 1,636 data, 10 electrode-fitted
  wavenumbers) -- the eval+grad benchmark's workload and data -- against the
  oracle on every source and wavenumber: FP64 dpred, misfit and gradient to
  1e-10 (the reference's parity bar, dctorch/tests/test_simulation2d.py:80),
  and the benchmark's own mixed setting (FP32 factor and fields, no fixed
  refinement pass, source basis) with dpred to 1e-6.
* The reference's CPU-runnable case (4,800 cells, nky = 11 SimPEG
  quadrature): FP64 against the oracle and the cpu engine to 1e-10, mixed
  dpred to 1e-6.
* The 62,016-cell Gauss-Newton line (270 sources, 2,430 data, 9
  wavenumbers): rows of the explicit J against the oracle's, J rows equal
  to J^T e_i through the engine's own adjoint (dctorch/tests/
  test_jmatrix.py:59-78), and the Gauss-Newton step (Woodbury direct step and
  Occam beta search) computed on the GPU equal to the same algebra on the
  host from the same J.

The oracle runs through the electrode positions here (exact by linearity,
pinned against its source-column path in tests/test_oracle.py): its
source-column restatement takes minutes at these sizes.
"""
import numpy as np
import pytest
import torch

from helpers import oracle_for, rel

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(msg):
    print(msg, flush=True)


def _eval_grad(eng, m, dobs, w):
    x = m.clone().requires_grad_(True)
    phi = eng.misfit(x, dobs, w)
    phi.backward()
    return float(phi), x.grad.cpu().numpy()


def test_line25k_full_parity():
    from oracle.dc25d import misfit_and_grad_electrodes
    from simpeg_dc_torch_amd import TorchDC2D, workloads
    mesh, survey, m, rng = workloads.line_25k_case()
    _say("line25k: FP64 engine")
    e64 = TorchDC2D(mesh, survey, nky=12, device=DEV, precision="fp64")
    assert (e64.nC, e64.nS, e64.nD, e64.nky) == (25088, 849, 1636, 10)
    mt = torch.tensor(m, device=DEV)
    with torch.no_grad():
        d64 = e64.dpred(mt)
    # the benchmark's data (bench.py build_engines)
    noise = torch.tensor(rng.standard_normal(e64.nD), device=DEV)
    dobs = d64 * (1 + 0.05 * noise)
    w = 1.0 / (0.05 * dobs.abs())
    _say("line25k: oracle (all 849 sources, 10 wavenumbers)")
    O = oracle_for(mesh, survey, e64.kys, e64.wq_np)
    phi_o, g_o, d_o = misfit_and_grad_electrodes(O, m, dobs.cpu().numpy(), w.cpu().numpy())
    assert rel(d64.cpu().numpy(), d_o) < 1e-10
    phi, g = _eval_grad(e64, mt, dobs, w)
    _say(f"FP64: dpred {rel(d64.cpu().numpy(), d_o):.1e} phi {abs(phi / phi_o - 1):.1e} "
         f"grad {rel(g, g_o):.1e}")
    assert abs(phi / phi_o - 1) < 1e-10
    assert rel(g, g_o) < 1e-10
    del e64
    _say("line25k: mixed engine, the benchmark's setting")
    emx = TorchDC2D(mesh, survey, nky=12, device=DEV, precision="mixed", refine=0)
    with torch.no_grad():
        dmx = emx.dpred(mt).cpu().numpy()
    phi_m, g_m = _eval_grad(emx, mt, dobs, w)
    rep = emx.solver.check_report()
    _say(f"mixed: dpred {rel(dmx, d_o):.1e} phi {abs(phi_m / phi_o - 1):.1e} "
         f"grad {rel(g_m, g_o):.1e}; check {rep['status']}, residual max {rep['rho'].max():.1e}, "
         f"refinement passes {rep['passes']}")
    assert rel(dmx, d_o) < 1e-6
    assert abs(phi_m / phi_o - 1) < 1e-5
    assert rel(g_m, g_o) < 1e-3
    assert rep["status"] == "passed" and emx.solver.fallbacks == []


def test_reference_case_parity():
    """The reference's CPU-runnable case as stated (80 x 40 core of 2.5 m
    cells plus padding, 4,800 cells; 21 electrodes dipole-dipole n = 1..6;
    two Gaussian blobs; nky = 11 SimPEG quadrature; dobs = dpred (1 + 0.05
    N(0, 1)) seed 0): FP64 dpred, misfit and gradient on the GPU against the
    oracle to 1e-10, the cpu engine (the reference chain, SciPy SuperLU) to
    the same bar, and the mixed engine's dpred to 1e-6."""
    from simpeg_dc_torch_amd import TorchDC2D, workloads
    mesh, survey, m = workloads.small_padded_case()
    _say("reference case: FP64 engine, nky 11 SimPEG quadrature")
    e64 = TorchDC2D(mesh, survey, nky=11, device=DEV, quadrature="simpeg")
    assert (e64.nC, e64.nky, e64.nD) == (4800, 11, 93)
    O = oracle_for(mesh, survey, e64.kys, e64.wq_np)
    d_o = O.dpred(m)
    rng = np.random.default_rng(0)
    dobs = d_o * (1 + 0.05 * rng.standard_normal(d_o.size))
    w = 1.0 / (0.05 * np.abs(dobs))
    phi_o, g_o, _ = O.misfit_and_grad(m, dobs, w)
    mt = torch.tensor(m, device=DEV)
    dt, wt = torch.tensor(dobs, device=DEV), torch.tensor(w, device=DEV)
    with torch.no_grad():
        d64 = e64.dpred(mt).cpu().numpy()
    phi, g = _eval_grad(e64, mt, dt, wt)
    _say(f"FP64: dpred {rel(d64, d_o):.1e} phi {abs(phi / phi_o - 1):.1e} grad {rel(g, g_o):.1e}")
    assert rel(d64, d_o) < 1e-10
    assert abs(phi / phi_o - 1) < 1e-10
    assert rel(g, g_o) < 1e-10
    ecpu = TorchDC2D(mesh, survey, nky=11, device="cpu", quadrature="simpeg")
    assert np.array_equal(ecpu.kys, e64.kys)
    phi_c, g_c = _eval_grad(ecpu, torch.tensor(m), torch.tensor(dobs), torch.tensor(w))
    assert abs(phi / phi_c - 1) < 1e-10
    assert rel(g, g_c) < 1e-10
    emx = TorchDC2D(mesh, survey, nky=11, device=DEV, quadrature="simpeg", precision="mixed",
                    refine=0)
    with torch.no_grad():
        dmx = emx.dpred(mt).cpu().numpy()
    phi_m, g_m = _eval_grad(emx, mt, dt, wt)
    _say(f"mixed: dpred {rel(dmx, d_o):.1e} grad {rel(g_m, g_o):.1e}")
    assert rel(dmx, d_o) < 1e-6
    assert rel(g_m, g_o) < 1e-3


def test_gn_line_jacobian_and_step():
    from oracle.dc25d import jacobian_rows_electrodes
    from simpeg_dc_torch_amd import TorchDC2D, workloads
    from simpeg_dc_torch_amd.gauss_newton import (HmSolver, direct_step, occam_step,
                                                  tensor_regularization)
    mesh, survey, m = workloads.gn_line_case()
    _say("GN line: FP64 engine, explicit J")
    eng = TorchDC2D(mesh, survey, nky=12, device=DEV, precision="fp64")
    assert (eng.nC, eng.nS, eng.nD) == (62016, 270, 2430)
    mt = torch.tensor(m, device=DEV)
    lin = eng.linearize(mt)
    J = lin.Jmatrix()
    rows = np.arange(0, eng.nD, 40)
    _say(f"GN line: oracle rows ({rows.size})")
    O = oracle_for(mesh, survey, eng.kys, eng.wq_np)
    J_o = jacobian_rows_electrodes(O, m, rows)
    Jr = J[torch.as_tensor(rows, device=DEV)].cpu().numpy()
    _say(f"J rows vs oracle {rel(Jr, J_o):.1e}")
    assert rel(Jr, J_o) < 1e-10
    for i in rows[[0, rows.size // 2, -1]]:
        e = torch.zeros(eng.nD, dtype=torch.float64, device=DEV)
        e[i] = 1.0
        jt = eng.Jtvec(mt, e).cpu().numpy()
        assert rel(jt, J[int(i)].cpu().numpy()) < 1e-10


def test_gn_line_step_device_vs_host():
    """One Gauss-Newton step from a 200 ohm-m half-space at the true model's
    data (5% noise), beta 10: the Woodbury direct step and the Occam search
    on the GPU equal the same algebra on the host from the same J."""
    from simpeg_dc_torch_amd import TorchDC2D, workloads
    from simpeg_dc_torch_amd.gauss_newton import (HmSolver, direct_step, occam_step,
                                                  tensor_regularization)
    mesh, survey, m = workloads.gn_line_case()
    _say("GN line: FP64 engine, J at the starting model")
    eng = TorchDC2D(mesh, survey, nky=12, device=DEV, precision="fp64")
    mt = torch.tensor(m, device=DEV)
    with torch.no_grad():
        d_true = eng.dpred(mt)
    rng = np.random.default_rng(3)
    dobs = d_true * (1 + 0.05 * torch.tensor(rng.standard_normal(eng.nD), device=DEV))
    w = 1.0 / (0.05 * dobs.abs() + 1e-4)
    m0 = torch.full((mesh.nC,), float(np.log(200.0)), dtype=torch.float64, device=DEV)
    J0 = eng.linearize(m0).Jmatrix()
    with torch.no_grad():
        r = w * (eng.dpred(m0) - dobs)
    D, wd = 2.0 * w ** 2, w * (2.0 * r)
    Hm = tensor_regularization(mesh, alpha_s=1e-4)
    g_m = torch.as_tensor(Hm @ m0.cpu().numpy(), device=DEV)
    g = J0.T @ wd + 10.0 * g_m
    _say("GN step on the GPU and on the host")
    p_dev = direct_step(J0, D, g, Hm, 10.0, HmSolver(DEV))
    p_host = direct_step(J0.cpu(), D.cpu(), g.cpu(), Hm, 10.0, HmSolver("cpu"))
    assert p_dev is not None and p_host is not None
    _say(f"direct step device vs host {rel(p_dev.cpu().numpy(), p_host.numpy()):.1e}")
    assert rel(p_dev.cpu().numpy(), p_host.numpy()) < 1e-8
    phi = lambda rr: (rr ** 2).sum()
    b_dev, q_dev, pred_dev = occam_step(J0, D, J0.T @ wd, g_m, r, 10.0, Hm, HmSolver(DEV), phi, w)
    b_host, q_host, pred_host = occam_step(J0.cpu(), D.cpu(), (J0.T @ wd).cpu(), g_m.cpu(),
                                           r.cpu(), 10.0, Hm, HmSolver("cpu"), phi, w.cpu())
    _say(f"occam beta {b_dev:.4g} / {b_host:.4g}, step {rel(q_dev.cpu().numpy(), q_host.numpy()):.1e}")
    assert b_dev == pytest.approx(b_host, rel=1e-6)
    assert rel(q_dev.cpu().numpy(), q_host.numpy()) < 1e-8
    assert pred_dev == pytest.approx(pred_host, rel=1e-8)
