"""CPU checks of the product's host logic against the oracle and the fixtures."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

import emulate
from helpers import case, laplacian_fixture, noisy_data, oracle_for, quad, rel
from simpeg_dc_torch_amd import TorchDC2D, _lib
from simpeg_dc_torch_amd.operators2d import Problem2DMaps
from simpeg_dc_torch_amd.quadrature import k0_quadrature, transform_error

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "quadrature_golden.npz")


@pytest.mark.parametrize("name", ["suite", "small_padded", "regression"])
def test_assembly_map_matches_oracle(name):
    """Factored map == operator-form assembly at two models, every ky
    (the reference validates its map the same way, simulation2d.py:211-229)."""
    mesh, survey, m = case(name)
    kys, wq = quad(survey, 6)
    P = Problem2DMaps(mesh, survey, kys)
    O = oracle_for(mesh, survey, kys, wq)
    rng = np.random.default_rng(7)
    for mm in (m, m + 0.5 * rng.standard_normal(m.size)):
        sig = np.exp(-mm)
        V = P.values(sig)
        for k in range(len(kys)):
            A = O.A(sig, k)
            assert np.array_equal(A.indptr, P.crow) and np.array_equal(A.indices, P.col)
            assert np.abs(A.data - V[k]).max() <= 1e-12 * np.abs(A.data).max()
    assert np.abs(P.q.toarray() - O.q).max() == 0.0
    assert abs(P.Pd - O.Pd).max() == 0.0
    assert np.array_equal(P.sidx, O.sidx)


def test_quadrature_golden_vectors():
    """k0_quadrature against the reference's own outputs (tests/golden)."""
    g = np.load(GOLDEN)
    for i in range(int(g["ncases"][0])):
        lo, hi, n = g[f"case{i}_in"]
        p, w, info = k0_quadrature(lo, hi, int(n))
        assert p.shape == g[f"case{i}_points"].shape
        assert np.allclose(p, g[f"case{i}_points"], rtol=1e-10, atol=0)
        assert np.allclose(w, g[f"case{i}_weights"], rtol=1e-10, atol=0)


def test_quadrature_invariants():
    """Error never grows with nky; weights positive; err(31) < 1e-5
    (dctorch/tests/test_quadrature.py:27-40)."""
    p, w, info = k0_quadrature(5.0, 120.0, 31)
    eb = info["err_by_n"]
    ns = sorted(eb)
    assert max(eb[b] / eb[a] for a, b in zip(ns, ns[1:])) < 1.5
    assert np.all(w > 0) and info["amplification"] < 1 + 1e-12
    assert eb[31] < 1e-5
    assert transform_error(p, w, 5.0, 120.0) == pytest.approx(eb[31])


def test_library_exports_header():
    """libdcx.so loads (no GPU needed) and exports every declared entry point."""
    lib = _lib.load()
    hdr = open(os.path.join(ROOT, "include", "dcx.h")).read()
    names = set(re.findall(r"\b(dcx_[a-z_0-9]+)\s*\(", hdr))
    assert len(names) >= 20
    for nm in names:
        assert hasattr(lib, nm), nm
    assert set(_lib.PROTOTYPES) <= names


def test_rows_mix_rejects_bad_arguments():
    """dcx_rows_mix checks its sizes and segment table on the host and
    returns before any device work (the pointers below are never read); a
    table of empty segments is a no-op."""
    import ctypes as C
    lib = _lib.load()
    fake = C.c_void_p(256)

    def call(n, k, segs, nb=2):
        a, p = _lib.i32(np.asarray(segs))
        return lib.dcx_rows_mix(fake, fake, fake, nb, n, k, len(segs) - 1, p,
                                _lib.DCX_F32_FIELDS, None)
    assert call(100, 0, [0, 50, 100]) != 0 and b"k must" in lib.dcx_last_error()
    assert call(100, 129, [0, 50, 100]) != 0
    assert call(100, 8, [0]) != 0 and b"nseg" in lib.dcx_last_error()
    assert call(100, 8, list(range(258))) != 0
    assert call(100, 8, [0, 60, 40]) != 0 and b"ascend" in lib.dcx_last_error()
    assert call(100, 8, [0, 50, 101]) != 0
    assert call(100, 8, [-1, 50]) != 0
    assert call(100, 8, [0, 50], nb=0) != 0
    assert lib.dcx_rows_mix(None, fake, fake, 2, 100, 8, 1, _lib.i32(np.r_[0, 50])[1], 0,
                            None) != 0
    assert call(100, 8, [30, 30, 30]) == 0


@pytest.mark.parametrize("coords", [False, True])
def test_symbolic_structure_and_emulated_solve(coords):
    """Host analysis + a numpy run of the kernel algorithm == SciPy."""
    mesh, survey, m = case("small_padded")
    kys, wq = quad(survey, 3)
    P = Problem2DMaps(mesh, survey, kys)
    sym = _lib.symbolic(P.crow, P.col, P.node_coords() if coords else None, leaf_max=16)
    perm = sym["perm"]
    assert np.array_equal(np.sort(perm), np.arange(P.nN))
    # the extend-add condition: child rows inside the parent front
    for c in range(sym["nsup"]):
        p = sym["parent"][c]
        R = sym["ridx"][sym["rptr"][c]:sym["rptr"][c + 1]]
        if p < 0:
            assert R.size == 0
            continue
        F = np.r_[np.arange(sym["first"][p], sym["first"][p] + sym["ns"][p]),
                  sym["ridx"][sym["rptr"][p]:sym["rptr"][p + 1]]]
        rp = sym["relpos"][sym["rptr"][c]:sym["rptr"][c + 1]]
        assert np.array_equal(F[rp], R)
    V = P.values(np.exp(-m))
    rng = np.random.default_rng(0)
    B = rng.standard_normal((P.nN, 4))
    A = sp.csr_matrix((V[0], P.col, P.crow), shape=(P.nN, P.nN))
    Q = emulate.factor(sym, V[0])
    X = emulate.solve(sym, Q, B[perm])
    x = np.empty_like(X)
    x[perm] = X
    assert rel(x, spla.splu(A.tocsc()).solve(B)) < 1e-11


@pytest.mark.parametrize("fuse_rows", [120, 320, 600])
def test_fused_subtree_structure_and_emulated_solve(fuse_rows):
    """Fused bottom subtrees: structure invariants and a numpy run of the
    fused kernels' algorithm == SciPy."""
    mesh, survey, m = case("small_padded")
    kys, wq = quad(survey, 3)
    P = Problem2DMaps(mesh, survey, kys)
    sym = _lib.symbolic(P.crow, P.col, P.node_coords(), leaf_max=16, fuse_rows=fuse_rows)
    assert sym["nfused"] > 0
    assert sym["f_max_lds_rows"] <= fuse_rows
    nf = sym["ns"] + sym["rs"]
    covered = np.zeros(P.nN, bool)
    for t in range(sym["nfused"]):
        lo, hi = sym["f_lo"][t], sym["f_hi"][t]
        assert not covered[lo:hi].any()
        covered[lo:hi] = True
    upper = np.zeros(sym["nsup"], bool)
    upper[sym["ulvl_sup"]] = True
    mem = np.zeros(sym["nsup"], bool)
    mem[sym["f_mem"]] = True
    assert np.all(upper ^ mem)                   # every supernode exactly once
    V = P.values(np.exp(-m))
    rng = np.random.default_rng(0)
    B = rng.standard_normal((P.nN, 3))
    A = sp.csr_matrix((V[1], P.col, P.crow), shape=(P.nN, P.nN))
    Q = emulate.factor(sym, V[1])
    perm = sym["perm"]
    X = emulate.solve_fused(sym, Q, B[perm])
    x = np.empty_like(X)
    x[perm] = X
    assert rel(x, spla.splu(A.tocsc()).solve(B)) < 1e-11
    del nf


def test_symbolic_laplacian_fixture():
    A, b = laplacian_fixture()
    sym = _lib.symbolic(A.indptr, A.indices, None, leaf_max=16)
    Q = emulate.factor(sym, A.data)
    X = emulate.solve(sym, Q, b[sym["perm"]])
    x = np.empty_like(X)
    x[sym["perm"]] = X
    assert rel(x, np.linalg.solve(A.toarray(), b)) < 1e-11


def test_symbolic_rejects_bad_patterns():
    with pytest.raises(RuntimeError, match="symmetric"):
        _lib.symbolic(np.array([0, 2, 3]), np.array([0, 1, 1]))
    with pytest.raises(RuntimeError, match="diagonal"):
        _lib.symbolic(np.array([0, 1, 2]), np.array([1, 0]))


# ---------------------------------------------------------------------------
# the cpu engine (reference chain: torch sparse + SuperLU) against the oracle
# ---------------------------------------------------------------------------
def _engine_and_oracle(name, nky=5, **kw):
    mesh, survey, m = case(name)
    eng = TorchDC2D(mesh, survey, nky=nky, device="cpu", **kw)
    O = oracle_for(mesh, survey, eng.kys, eng.wq_np)
    return eng, O, m


def test_cpu_engine_dpred_and_gradient():
    eng, O, m = _engine_and_oracle("suite")
    mt = torch.tensor(m, requires_grad=True)
    d = eng.dpred(mt)
    d_o = O.dpred(m)
    assert rel(d.detach().numpy(), d_o) < 1e-10
    dobs, w = noisy_data(d_o)
    phi = eng.misfit(mt, torch.tensor(dobs), torch.tensor(w))
    phi.backward()
    phi_o, g_o, _ = O.misfit_and_grad(m, dobs, w)
    assert float(phi) == pytest.approx(phi_o, rel=1e-10)
    assert rel(mt.grad.numpy(), g_o) < 1e-9
    f, g = eng.misfit_and_grad(torch.tensor(m), torch.tensor(dobs), torch.tensor(w), chunk=5)
    assert rel(g.numpy(), mt.grad.numpy()) < 1e-12


def _config1_data(d):
    """dobs = dpred (1 + 0.05 N(0, 1)), seed 0; w = 1 / (0.05 |dobs|)."""
    rng = np.random.default_rng(0)
    dobs = d * (1 + 0.05 * rng.standard_normal(d.size))
    return dobs, 1.0 / (0.05 * np.abs(dobs))


def test_cpu_engine_reference_case():
    """The reference's CPU-runnable case as stated: the 80 x 40 core of 2.5 m
    cells with 8 (10 at depth) padding cells growing 1.3x (4,800 cells), 21
    electrodes dipole-dipole n = 1..6, two Gaussian blobs, nky = 11 SimPEG
    quadrature, dobs = dpred (1 + 0.05 N(0, 1)) seed 0: dpred, misfit and
    gradient of the cpu engine (the reference chain) against the oracle."""
    eng, O, m = _engine_and_oracle("small_padded", nky=11, quadrature="simpeg")
    assert (eng.nC, eng.nky, eng.nD) == (4800, 11, 93)
    d_o = O.dpred(m)
    dobs, w = _config1_data(d_o)
    mt = torch.tensor(m, requires_grad=True)
    assert rel(eng.dpred(mt).detach().numpy(), d_o) < 1e-10
    phi = eng.misfit(mt, torch.tensor(dobs), torch.tensor(w))
    phi.backward()
    phi_o, g_o, _ = O.misfit_and_grad(m, dobs, w)
    assert float(phi) == pytest.approx(phi_o, rel=1e-10)
    assert rel(mt.grad.numpy(), g_o) < 1e-10


def test_cpu_engine_jtvec_jmatrix_jvec():
    eng, O, m = _engine_and_oracle("suite", nky=4)
    mt = torch.tensor(m)
    J = eng.Jmatrix(mt, block=7).numpy()
    Jo = O.Jmatrix(m)
    assert rel(J, Jo) < 1e-9
    rng = np.random.default_rng(2)
    v = rng.standard_normal(eng.nD)
    assert rel(eng.Jtvec(mt, torch.tensor(v)).numpy(), J.T @ v) < 1e-10
    assert rel(eng.Jtvec(mt, torch.tensor(v), chunk=4).numpy(), J.T @ v) < 1e-10
    u = rng.standard_normal(m.size)
    assert rel(eng.Jvec(mt, torch.tensor(u)).numpy(), J @ u) < 1e-9


def test_cpu_linearization_products():
    """Linearization2D on the cpu engine: J v, J^T v, diag(J^T D J) and J
    against the oracle's operator-form Jacobian; the adjoint identity
    (dctorch/tests/test_linearization.py:48-64)."""
    from simpeg_dc_torch_amd import Linearization2D
    eng, O, m = _engine_and_oracle("suite", nky=4)
    mt = torch.tensor(m)
    lin = Linearization2D(eng, mt)
    J = O.Jmatrix(m)
    assert rel(lin.d.numpy(), O.dpred(m)) < 1e-10
    rng = np.random.default_rng(3)
    u = rng.standard_normal(m.size)
    v = rng.standard_normal(eng.nD)
    Ju = lin.Jvec(torch.tensor(u)).numpy()
    Jtv = lin.Jtvec(torch.tensor(v)).numpy()
    assert rel(Ju, J @ u) < 1e-9
    assert rel(Jtv, J.T @ v) < 1e-9
    assert abs(Ju @ v - u @ Jtv) <= 1e-11 * abs(Ju @ v)
    dw = rng.random(eng.nD)
    assert rel(lin.diag_JtDJ(torch.tensor(dw), sub=5).numpy(), (J ** 2 * dw[:, None]).sum(0)) < 1e-9
    assert rel(lin.Jmatrix(sub=7).numpy(), J) < 1e-9


# ---------------------------------------------------------------------------
# 3-D
# ---------------------------------------------------------------------------
def _oracle3d(mesh, survey):
    from oracle.dc3d import Problem3D
    return Problem3D(mesh.h[0], mesh.h[1], mesh.h[2], mesh.origin, survey)


def test_3d_assembly_map_matches_oracle():
    """Factored 3-D map == operator-form assembly at two models
    (dctorch/simulation3d.py:132-140 validates its map the same way)."""
    from simpeg_dc_torch_amd.operators3d import Problem3DMaps
    mesh, survey, m = case("small_3d")
    P = Problem3DMaps(mesh, survey)
    O = _oracle3d(mesh, survey)
    rng = np.random.default_rng(7)
    for mm in (m, m + 0.5 * rng.standard_normal(m.size)):
        sig = np.exp(-mm)
        A = O.A(sig)
        assert np.array_equal(A.indptr, P.crow) and np.array_equal(A.indices, P.col)
        assert np.abs(A.data - P.values(sig)[0]).max() <= 1e-12 * np.abs(A.data).max()
    assert abs(P.q - sp.csr_matrix(O.q)).max() == 0.0
    assert abs(P.Pd - O.Pd).max() == 0.0
    assert np.array_equal(P.sidx, O.sidx)


def test_3d_oracle_closed_form_stencil():
    """Uniform unit cells, sigma = 1: interior node 6 / -1; free-surface node
    3, -1/2 along the surface, -1 downwards; Grad-term rows sum to zero."""
    from simpeg_dc_torch_amd.mesh import TensorMesh
    from simpeg_dc_torch_amd import workloads
    mesh = TensorMesh([[(1.0, 4)], [(1.0, 4)], [(1.0, 4)]], origin="000")
    el = workloads.surface_grid(3, 1, 1.0, x0=1.0, y0=1.0, z=4.0)
    survey = workloads.pole_dipole_survey(el[:1], el, np.array([[1, 2]]), 1)
    O = _oracle3d(mesh, survey)
    K = (O.Grad.T @ sp.diags(O.Qe_ @ np.ones(mesh.nC)) @ O.Grad).tocsr()
    g = O.g
    c = g.node(2, 2, 2)
    assert K[c, c] == pytest.approx(6.0) and K[c, c + 1] == pytest.approx(-1.0)
    t = g.node(2, 2, 4)
    assert K[t, t] == pytest.approx(3.0)
    assert K[t, g.node(3, 2, 4)] == pytest.approx(-0.5)
    assert K[t, g.node(2, 2, 3)] == pytest.approx(-1.0)
    assert np.abs(np.asarray(K.sum(1))).max() < 1e-12
    assert abs(K - K.T).max() == 0.0


def test_3d_oracle_halfspace_pole_converges():
    """Pole source on a half space: the surface potential approaches
    rho/(2 pi r) as the cells halve (dctorch/tests/test_singularity.py:49-51
    pins the 2-D analogue)."""
    from simpeg_dc_torch_amd.mesh import TensorMesh
    from simpeg_dc_torch_amd.survey import PoleReceiver, PoleSource, Survey
    rs = np.array([4.0, 6.0])
    errs = []
    for h in (2.0, 1.0):
        ncore = (int(24 / h), int(24 / h), int(12 / h))
        hx = [(h, 10, -1.5), (h, ncore[0]), (h, 10, 1.5)]
        hz = [(h, 10, -1.5), (h, ncore[2])]
        mesh = TensorMesh([hx, hx, hz], origin="CCN")
        rx = PoleReceiver(np.c_[rs, np.zeros(2), np.zeros(2)])
        survey = Survey([PoleSource([rx], np.zeros(3))])
        O = _oracle3d(mesh, survey)
        d = O.dpred(np.full(mesh.nC, np.log(100.0)))
        errs.append(np.abs(d / (100.0 / (2 * np.pi * rs)) - 1))
    assert errs[1][0] < errs[0][0] / 3
    assert errs[1].max() < 0.02


def test_3d_cpu_engine_vs_oracle():
    from simpeg_dc_torch_amd import TorchDC3D
    mesh, survey, m = case("small_3d")
    eng = TorchDC3D(mesh, survey, device="cpu")
    O = _oracle3d(mesh, survey)
    mt = torch.tensor(m)
    d_o = O.dpred(m)
    assert rel(eng.dpred(mt).numpy(), d_o) < 1e-10
    dobs, w = noisy_data(d_o)
    phi, g = eng.misfit_and_grad(mt, torch.tensor(dobs), torch.tensor(w))
    phi_o, g_o, _ = O.misfit_and_grad(m, dobs, w)
    assert float(phi) == pytest.approx(phi_o, rel=1e-10)
    assert rel(g.numpy(), g_o) < 1e-9
    v = np.random.default_rng(1).standard_normal(eng.nD)
    assert rel(eng.Jtvec(mt, torch.tensor(v)).numpy(), O.Jtvec(m, v)) < 1e-9


def test_3d_oracle_gradient_and_jtvec_vs_fd():
    """The oracle's adjoint gradient and J^T v against central differences of
    its own misfit and dpred (independent of the adjoint formula)."""
    from simpeg_dc_torch_amd.mesh import TensorMesh
    from simpeg_dc_torch_amd import workloads
    mesh = TensorMesh([[(1.0, 6)], [(1.0, 5)], [(1.0, 4)]], origin="000")
    el = workloads.surface_grid(3, 2, 1.5, x0=1.2, y0=1.3, z=4.0)
    survey = workloads.pole_dipole_survey(el[:2], el, np.array([[2, 3], [3, 4], [4, 5]]), 2)
    O = _oracle3d(mesh, survey)
    rng = np.random.default_rng(5)
    m = np.log(50.0) + 0.3 * rng.standard_normal(mesh.nC)
    d0 = O.dpred(m)
    dobs, w = noisy_data(d0)
    _, g, _ = O.misfit_and_grad(m, dobs, w)
    v = rng.standard_normal(d0.size)
    jtv = O.Jtvec(m, v)
    eps = 1e-5
    for c in rng.choice(mesh.nC, 8, replace=False):
        e = np.zeros(mesh.nC)
        e[c] = eps
        phi = [float(np.sum((w * (O.dpred(m + s * e) - dobs)) ** 2)) for s in (1, -1)]
        fd = (phi[0] - phi[1]) / (2 * eps)
        assert abs(fd - g[c]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max())
        fdj = float(v @ (O.dpred(m + e) - O.dpred(m - e))) / (2 * eps)
        assert abs(fdj - jtv[c]) <= 1e-6 * max(abs(fdj), 1e-3 * np.abs(jtv).max())


def test_active_cells_restrict_the_full_model_cpu():
    """Inactive cells held at val_inactive: dpred, gradient and J of the
    active-cell engine equal the full-mesh engine's on the active columns
    (the reference's test_jmatrix.py:80-97 and test_linearization.py:83-88)."""
    from simpeg_dc_torch_amd import TorchDC2D
    mesh, survey, m = case("suite")
    act = mesh.cell_centers[:, 1] < -1.0
    val = float(np.log(100.0))
    m_full = m.copy()
    m_full[~act] = val
    eng_f = TorchDC2D(mesh, survey, nky=6, device="cpu")
    eng_a = TorchDC2D(mesh, survey, nky=6, device="cpu", active_cells=act, val_inactive=val)
    mf, ma = torch.tensor(m_full), torch.tensor(m_full[act])
    with torch.no_grad():
        d_f, d_a = eng_f.dpred(mf), eng_a.dpred(ma)
    assert rel(d_a.numpy(), d_f.numpy()) < 1e-13
    dobs, w = noisy_data(d_f.numpy())
    _, g_f = eng_f.misfit_and_grad(mf, torch.tensor(dobs), torch.tensor(w))
    _, g_a = eng_a.misfit_and_grad(ma, torch.tensor(dobs), torch.tensor(w))
    assert g_a.shape == (int(act.sum()),)
    assert rel(g_a.numpy(), g_f.numpy()[act]) < 1e-12
    Jf, Ja = eng_f.Jmatrix(mf).numpy(), eng_a.Jmatrix(ma).numpy()
    assert Ja.shape == (eng_a.nD, int(act.sum()))
    assert rel(Ja, Jf[:, act]) < 1e-13
