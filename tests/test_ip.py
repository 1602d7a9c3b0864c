"""Induced polarisation on the DC engines (simpeg_dc_torch_amd/ip.py).

What pins each piece:
* TorchIP2D / TorchIP3D: the linearised IP datum is the derivative of the DC
  data along eta in m = log rho (d_IP = J_m eta), checked against central
  differences of the DC engine itself; for a uniform eta the systems scale
  exactly (A(sigma (1 - eta)) = (1 - eta) A(sigma), the Robin term included),
  so J_m 1 = V_dc and the apparent chargeability is eta exactly, for any
  background; the gradient is the adjoint (<J e, g> = <e, J^T g>); the
  active-cell mask is zero injection.
* TorchDCIP2D: for a uniform eta V1 = V0 / (1 - eta) exactly, so d_IP = eta
  and the "total" DC datum is V0 / (1 - eta); its gradient against central
  differences of its own misfit; to first order in eta it equals the
  linearised IP (the difference is O(eta)).
* TorchCR2D (cpu and GPU): m = 0 gives the DC data; a uniform m gives V* = V_dc /
  (1 - i m) exactly (phase ratio m, magnitude |V_dc| / sqrt(1 + m^2)); the
  complex-safe autograd solve against central differences; the GPU path
  (block Gauss-Seidel on the real factor) against the cpu path's complex
  SuperLU, data and gradient.
* GPU (marked gpu): every GPU class on the HIP path against the same class on
  the CPU path.
"""
import numpy as np
import pytest
import torch

from helpers import case, noisy_data


def _chargeability(survey):
    """The survey with every receiver re-labelled apparent_chargeability."""
    from simpeg_dc_torch_amd import survey as S
    srcs = []
    for s in survey.source_list:
        rxs = []
        for rx in s.receiver_list:
            if rx.locations_n is None:
                rxs.append(S.PoleReceiver(rx.locations_m, data_type="apparent_chargeability"))
            else:
                rxs.append(S.DipoleReceiver(rx.locations_m, rx.locations_n,
                                            data_type="apparent_chargeability"))
        if s.location_b is None:
            srcs.append(S.PoleSource(rxs, s.location_a))
        else:
            srcs.append(S.DipoleSource(rxs, s.location_a, s.location_b))
    return S.Survey(srcs)


def _fd_dc(dc, m, e, h=1e-4):
    with torch.no_grad():
        return (dc.dpred(torch.tensor(m) + h * e) - dc.dpred(torch.tensor(m) - h * e)) / (2 * h)


def test_ip2d_is_the_dc_derivative_along_eta():
    from simpeg_dc_torch_amd import TorchDC2D, TorchIP2D
    mesh, survey, m = case("suite")
    ip = TorchIP2D(mesh, survey, m, nky=6, device="cpu")
    e = torch.tensor(0.05 * np.random.default_rng(0).random(mesh.nC))
    d = ip.dpred(e)
    fd = _fd_dc(TorchDC2D(mesh, survey, nky=6, device="cpu"), m, e)
    assert float((d - fd).norm() / fd.norm()) < 1e-8


def test_ip3d_is_the_dc_derivative_along_eta():
    from simpeg_dc_torch_amd import TorchDC3D, TorchIP3D
    mesh, survey, m = case("small_3d")
    ip = TorchIP3D(mesh, survey, m, device="cpu")
    e = torch.tensor(0.05 * np.random.default_rng(1).random(mesh.nC))
    fd = _fd_dc(TorchDC3D(mesh, survey, device="cpu"), m, e)
    assert float((ip.dpred(e) - fd).norm() / fd.norm()) < 1e-8


@pytest.mark.parametrize("dim", [2, 3])
def test_ip_uniform_eta_is_the_apparent_chargeability(dim):
    from simpeg_dc_torch_amd import TorchIP2D, TorchIP3D
    mesh, survey, m = case("suite" if dim == 2 else "small_3d")
    make = (lambda sv: TorchIP2D(mesh, sv, m, nky=6, device="cpu")) if dim == 2 else \
        (lambda sv: TorchIP3D(mesh, sv, m, device="cpu"))
    eta = torch.full((mesh.nC,), 0.07, dtype=torch.float64)
    ipa = make(_chargeability(survey))
    assert float((ipa.dpred(eta) - 0.07).abs().max()) < 1e-10
    ipv = make(survey)
    assert float(((ipv.dpred(eta) - 0.07 * ipv.v_dc).norm()) / ipv.v_dc.norm()) < 1e-10


@pytest.mark.parametrize("dim", [2, 3])
def test_ip_gradient_is_the_adjoint(dim):
    from simpeg_dc_torch_amd import TorchIP2D, TorchIP3D
    mesh, survey, m = case("suite" if dim == 2 else "small_3d")
    sv = _chargeability(survey)
    ip = TorchIP2D(mesh, sv, m, nky=6, device="cpu") if dim == 2 else \
        TorchIP3D(mesh, sv, m, device="cpu")
    rng = np.random.default_rng(2)
    e = torch.tensor(rng.standard_normal(mesh.nC), requires_grad=True)
    g = torch.tensor(rng.standard_normal(ip.dc.nD))
    d = ip.dpred(e)
    (gr,) = torch.autograd.grad(d @ g, e)
    assert float(d.detach() @ g) == pytest.approx(float(e.detach() @ gr), rel=1e-10)


def test_ip2d_active_cells_inject_zero():
    from simpeg_dc_torch_amd import TorchIP2D
    mesh, survey, m = case("suite")
    act = mesh.cell_centers[:, 1] < -1.0
    e = torch.tensor(0.05 * np.random.default_rng(3).random(mesh.nC))
    full = TorchIP2D(mesh, survey, m, nky=6, device="cpu")
    part = TorchIP2D(mesh, survey, m, nky=6, device="cpu", active_cells=act)
    e0 = e.clone()
    e0[torch.tensor(~act)] = 0.0
    assert float((part.dpred(e[torch.tensor(act)]) - full.dpred(e0)).norm()) < \
        1e-12 * float(full.dpred(e0).norm())


def test_dcip_uniform_eta_closed_form():
    from simpeg_dc_torch_amd import TorchDCIP2D
    mesh, survey, m = case("suite")
    dcip = TorchDCIP2D(mesh, survey, nky=6, device="cpu")
    inst = TorchDCIP2D(mesh, survey, nky=6, device="cpu", dc_state="instantaneous")
    mr = torch.tensor(m)
    eta = torch.full((mesh.nC,), 0.3, dtype=torch.float64)
    d_dc, d_ip = dcip.dpred_from_fields(mr, eta)
    v0, _ = inst.dpred_from_fields(mr, eta)
    assert float((d_ip - 0.3).abs().max()) < 1e-10
    assert float((d_dc * 0.7 - v0).norm() / v0.norm()) < 1e-10


def test_dcip_gradient_vs_fd_and_linearised_ip():
    from simpeg_dc_torch_amd import TorchDCIP2D, TorchIP2D
    mesh, survey, m = case("suite")
    dcip = TorchDCIP2D(mesh, survey, nky=6, device="cpu", ip_scale=1000.0)
    rng = np.random.default_rng(4)
    eta = 0.01 + 0.02 * rng.random(mesh.nC)
    x = torch.tensor(np.r_[m, dcip.m_eta_of(eta)])
    d_dc, d_ip = dcip.dpred(x)
    dobs_dc, w_dc = noisy_data(d_dc.detach().numpy(), seed=5)
    dobs_ip, w_ip = noisy_data(d_ip.detach().numpy(), seed=6)
    args = [torch.tensor(a) for a in (dobs_dc, w_dc, dobs_ip, w_ip)]
    xg = x.clone().requires_grad_(True)
    dcip.misfit(xg, *args).backward()
    for k in range(3):
        v = torch.tensor(rng.standard_normal(x.numel()))
        h = 1e-5
        with torch.no_grad():
            fd = (dcip.misfit(x + h * v, *args) - dcip.misfit(x - h * v, *args)) / (2 * h)
        assert float(xg.grad @ v) == pytest.approx(float(fd), rel=1e-5)
    # first order: the linearised IP about sigma (apparent chargeability, mV/V)
    lin = TorchIP2D(mesh, _chargeability(survey), m, nky=6, device="cpu")
    d_lin = 1000.0 * lin.dpred(torch.tensor(eta))
    rel = float((d_ip - d_lin).norm() / d_lin.norm())
    assert rel < 3 * eta.max() and rel > 1e-4   # O(eta), and not identical


# --------------------------------------------------------------------- GPU
def _rel(a, b):
    a, b = torch.as_tensor(a).cpu(), torch.as_tensor(b).cpu()
    return float((a - b).norm() / b.norm())


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_ip_gpu_matches_cpu(dim):
    from simpeg_dc_torch_amd import TorchIP2D, TorchIP3D
    print(f"ip gpu {dim}-D", flush=True)
    mesh, survey, m = case("suite" if dim == 2 else "small_3d")
    sv = _chargeability(survey)
    act = mesh.cell_centers[:, -1] < -1.0
    for active in (None, act):
        if dim == 2:
            mk = lambda dev: TorchIP2D(mesh, sv, m, nky=6, device=dev, active_cells=active)
        else:
            mm = m if active is None else m[act]
            mk = lambda dev: TorchIP3D(mesh, sv, mm, device=dev, active_cells=active,
                                       val_inactive=None if active is None else float(np.log(100.0)))
        n = mesh.nC if active is None else int(act.sum())
        e = 0.05 * np.random.default_rng(5).random(n)
        outs = []
        for dev in ("cpu", "cuda"):
            ip = mk(dev)
            x = torch.tensor(e, device=dev, requires_grad=True)
            d = ip.dpred(x)
            (d ** 2).sum().backward()
            outs.append((d.detach(), x.grad))
        assert _rel(outs[1][0], outs[0][0]) < 1e-10
        assert _rel(outs[1][1], outs[0][1]) < 1e-9


@pytest.mark.gpu
def test_dcip_gpu_matches_cpu():
    from simpeg_dc_torch_amd import TorchDCIP2D
    print("dcip gpu", flush=True)
    mesh, survey, m = case("suite")
    rng = np.random.default_rng(6)
    eta = 0.05 + 0.2 * rng.random(mesh.nC)
    outs = []
    for dev in ("cpu", "cuda"):
        dcip = TorchDCIP2D(mesh, survey, nky=6, device=dev, ip_scale=1000.0)
        x = torch.tensor(np.r_[m, dcip.m_eta_of(eta)], device=dev, requires_grad=True)
        d_dc, d_ip = dcip.dpred(x)
        ((d_dc / d_dc.detach().abs()) ** 2 + (d_ip / 100.0) ** 2).sum().backward()
        outs.append((d_dc.detach(), d_ip.detach(), x.grad))
    for a, b in zip(outs[1], outs[0]):
        assert _rel(a, b) < 1e-9


def test_cr2d_zero_and_uniform_chargeability_closed_forms():
    """m = 0: the DC data (Re V* = V_dc, Im V* = 0).  Uniform m: sigma* =
    sigma (1 - i m) everywhere, so V* = V_dc / (1 - i m): d_ip = ip_scale m
    and |V*| = |V_dc| / sqrt(1 + m^2), exactly."""
    from simpeg_dc_torch_amd import TorchCR2D, TorchDC2D
    mesh, survey, m = case("suite")
    v_dc = TorchDC2D(mesh, survey, nky=6, device="cpu").dpred(torch.tensor(m)).detach()
    cr = TorchCR2D(mesh, survey, nky=6, device="cpu", ip_scale=1000.0, dc_datum="real")
    d_r, d_i = cr.dpred(torch.tensor(np.r_[m, np.zeros(mesh.nC)]))
    assert float((d_r - v_dc).norm() / v_dc.norm()) < 1e-12 and float(d_i.abs().max()) < 1e-12
    crm = TorchCR2D(mesh, survey, nky=6, device="cpu", ip_scale=1000.0)
    d_r, d_i = crm.dpred(torch.tensor(np.r_[m, np.full(mesh.nC, 0.05)]))
    assert float((d_i - 50.0).abs().max()) < 1e-9
    ref = v_dc.abs() / np.sqrt(1 + 0.05 ** 2)
    assert float((d_r - ref).norm() / ref.norm()) < 1e-12


def test_cr2d_gradient_vs_fd():
    """The complex-safe autograd solve (Wirtinger adjoint on the same
    factorisation) against central differences of the real misfit."""
    from simpeg_dc_torch_amd import TorchCR2D
    mesh, survey, m = case("suite")
    cr = TorchCR2D(mesh, survey, nky=6, device="cpu", ip_scale=1000.0)
    rng = np.random.default_rng(7)
    x = torch.tensor(np.r_[m, 0.02 + 0.05 * rng.random(mesh.nC)])
    d_r, d_i = cr.dpred(x)
    args = [torch.tensor(a) for a in (*noisy_data(d_r.detach().numpy(), seed=1),
                                      *noisy_data(d_i.detach().numpy(), seed=2))]
    xg = x.clone().requires_grad_(True)
    cr.misfit(xg, *args).backward()
    for _ in range(3):
        v = torch.tensor(rng.standard_normal(x.numel()))
        h = 1e-6
        with torch.no_grad():
            fd = (cr.misfit(x + h * v, *args) - cr.misfit(x - h * v, *args)) / (2 * h)
        assert float(xg.grad @ v) == pytest.approx(float(fd), rel=1e-5)


@pytest.mark.gpu
def test_cr2d_gpu_matches_cpu():
    """The GPU complex-resistivity forward (block Gauss-Seidel on the real
    FP64 factor, _ComplexResistivity) against the cpu path's complex SuperLU
    solves: both data and the misfit gradient with respect to both model
    halves (log rho and m), for chargeabilities up to 0.3."""
    from simpeg_dc_torch_amd import TorchCR2D
    mesh, survey, m = case("suite")
    rng = np.random.default_rng(11)
    x = np.r_[m, 0.3 * rng.random(mesh.nC)]
    outs = []
    for dev in ("cpu", "cuda"):
        cr = TorchCR2D(mesh, survey, nky=6, device=dev, ip_scale=1000.0)
        xt = torch.tensor(x, device=dev, requires_grad=True)
        d_r, d_i = cr.dpred(xt)
        if not outs:
            dobs = [torch.tensor(a) for a in (*noisy_data(d_r.detach().numpy(), seed=1),
                                              *noisy_data(d_i.detach().numpy(), seed=2))]
        args = [a.to(dev) for a in dobs]
        phi = ((args[1] * (d_r - args[0])) ** 2).sum() + ((args[3] * (d_i - args[2])) ** 2).sum()
        phi.backward()
        outs.append((d_r.detach().cpu(), d_i.detach().cpu(), xt.grad.cpu()))
    for a, b in zip(outs[1], outs[0]):
        assert _rel(a, b) < 1e-10


@pytest.mark.gpu
def test_cr2d_gpu_closed_forms():
    """On the GPU: m = 0 gives the DC data with Im V* = 0; a uniform m gives
    d_ip = ip_scale m exactly."""
    from simpeg_dc_torch_amd import TorchCR2D, TorchDC2D
    mesh, survey, m = case("suite")
    mt = torch.tensor(m, device="cuda")
    with torch.no_grad():
        v_dc = TorchDC2D(mesh, survey, nky=6, device="cuda").dpred(mt)
    cr = TorchCR2D(mesh, survey, nky=6, device="cuda", ip_scale=1000.0, dc_datum="real")
    with torch.no_grad():
        d_r, d_i = cr.dpred(torch.cat([mt, torch.zeros_like(mt)]))
    assert float((d_r - v_dc).norm() / v_dc.norm()) < 1e-12 and float(d_i.abs().max()) < 1e-12
    with torch.no_grad():
        d_r, d_i = cr.dpred(torch.cat([mt, torch.full_like(mt, 0.05)]))
    assert float((d_i - 50.0).abs().max()) < 1e-8
