"""The reference's linearisation and Jacobian contracts, restated.

dctorch/tests/test_linearization.py:34-68 and test_jmatrix.py:59-78 pin what
callers of `linearize` and `Jmatrix` rely on; they run here on the cpu (the
SuperLU chain) and on the GPU (the HIP path), FP64, on the suite case, with
source chunks None / 4 and with active cells:

(a) lin.d == dpred(m)                                      1e-12
(b) J v, J^T w against the explicit J                       1e-10
(c) adjoint identity (J v).w == v.(J^T w)                   1e-12
(d) diag(J^T D J) and the streamed J                        1e-10
(e) a linearisation without stored fields                   1e-10
(f) dpred at another model clobbers the shared factorisation; the
    linearisation detects it (stamp) and restores it        1e-10
and Jmatrix rows == J^T e_i one datum at a time, and the block width
(7 leaves a padded tail, nD) does not change J             1e-12.
"""
import numpy as np
import pytest
import torch

from simpeg_dc_torch_amd import TorchDC2D, workloads

DT = torch.float64
DEVICES = ["cpu", pytest.param("cuda", marks=pytest.mark.gpu)]


def _rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def _check(eng, m):
    dev = eng.device
    J = eng.Jmatrix(m)
    nM = J.shape[1]
    g = torch.Generator(device="cpu").manual_seed(3)
    v = torch.randn(nM, generator=g, dtype=DT).to(dev)
    w = torch.randn(eng.nD, generator=g, dtype=DT).to(dev)
    D = torch.rand(eng.nD, generator=g, dtype=DT).to(dev) + 0.1
    lin = eng.linearize(m)
    assert _rel(lin.d, eng.dpred(m)) < 1e-12                                   # (a)
    assert _rel(lin.Jvec(v), J @ v) < 1e-10                                    # (b)
    assert _rel(lin.Jtvec(w), J.T @ w) < 1e-10
    adj = abs(float(lin.Jvec(v) @ w - v @ lin.Jtvec(w))) \
        / float(torch.linalg.norm(J @ v) * torch.linalg.norm(w))
    assert adj < 1e-12                                                         # (c)
    assert _rel(lin.diag_JtDJ(D), (D[:, None] * J ** 2).sum(0)) < 1e-10       # (d)
    assert _rel(lin.Jmatrix(), J) < 1e-10
    lin_nf = eng.linearize(m, store_fields=False)                              # (e)
    assert _rel(lin_nf.Jvec(v), J @ v) < 1e-10
    assert _rel(lin_nf.Jtvec(w), J.T @ w) < 1e-10
    with torch.no_grad():
        eng.dpred(m + 0.3)                  # (f) clobbers the shared factorisation
    assert _rel(lin.Jvec(v), J @ v) < 1e-10
    assert _rel(lin.Jtvec(w), J.T @ w) < 1e-10


@pytest.mark.parametrize("chunk", [None, 4])
@pytest.mark.parametrize("device", DEVICES)
def test_linearization_contract(device, chunk):
    mesh, survey, m = workloads.suite_case()
    eng = TorchDC2D(mesh, survey, nky=11, device=device, chunk=chunk)
    _check(eng, torch.tensor(m, dtype=DT, device=eng.device))


@pytest.mark.parametrize("device", DEVICES)
def test_linearization_contract_active_cells(device):
    mesh, survey, m = workloads.suite_case()
    act = mesh.cell_centers[:, 1] < -1.0
    eng = TorchDC2D(mesh, survey, nky=11, device=device, active_cells=act,
                    val_inactive=float(np.log(100.0)), chunk=4)
    _check(eng, torch.tensor(m[act], dtype=DT, device=eng.device))


@pytest.mark.parametrize("device", DEVICES)
def test_jmatrix_rows_and_block_width(device):
    mesh, survey, m = workloads.suite_case()
    eng = TorchDC2D(mesh, survey, nky=11, device=device)
    mt = torch.tensor(m, dtype=DT, device=eng.device)
    J = eng.Jmatrix(mt)
    e = torch.zeros(eng.nD, dtype=DT, device=eng.device)
    rows = []
    for i in range(eng.nD):
        e.zero_()
        e[i] = 1.0
        rows.append(eng.Jtvec(mt, e))
    assert _rel(J, torch.stack(rows, 0)) < 1e-12
    for blk in (7, eng.nD):
        assert _rel(eng.Jmatrix(mt, block=blk), J) < 1e-12
