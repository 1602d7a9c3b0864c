"""Shared builders: the oracle and the product engine on the same inputs."""
import numpy as np

from oracle.dc25d import Problem2D
from simpeg_dc_torch_amd import workloads
from simpeg_dc_torch_amd.quadrature import electrode_distance_range, k0_quadrature


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def quad(survey, nky):
    lo, hi = electrode_distance_range(survey)
    p, w, _ = k0_quadrature(lo, hi, nky)
    return p, w


def oracle_for(mesh, survey, kys, wq):
    return Problem2D(mesh.h[0], mesh.h[1], mesh.origin, survey, kys, wq)


def case(name):
    if name == "line25k":
        mesh, survey, m, _ = workloads.line_25k_case()
        return mesh, survey, m
    return workloads.CASES[name]()


def noisy_data(d, seed=3, rel_noise=0.03):
    rng = np.random.default_rng(seed)
    dobs = d * (1 + rel_noise * rng.standard_normal(d.size))
    w = 1.0 / (rel_noise * np.abs(d) + 1e-12)
    return dobs, w


def laplacian_fixture(nx=40, n_rhs=5, seed=0):
    """The reference's solver fixture (dctorch/tests/test_solver.py:19-25)."""
    import scipy.sparse as sp
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(nx, nx))
    I = sp.eye(nx)
    A = (sp.kron(I, T) + sp.kron(T, I) + 0.1 * sp.eye(nx * nx)).tocsr()
    A.sort_indices()
    b = np.random.default_rng(seed).standard_normal((A.shape[0], n_rhs))
    return A, b
