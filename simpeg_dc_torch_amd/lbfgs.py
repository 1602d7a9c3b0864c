"""L-BFGS on an engine's misfit (TorchDC2D or TorchDC3D).

The reference's `TorchLBFGS` (dctorch/optimization.py:303-402) is a SimPEG
`Minimize` subclass whose search direction runs `torch.optim.LBFGS`
(strong-Wolfe line search) on the closure

    loss(m) = phi_d(m) + beta * phi_m(m),    phi_d = ||w (dpred(m) - dobs)||^2,

one forward and one reverse pass of the engine per closure.  SimPEG is not
a dependency of this package, so `lbfgs_minimize` provides that loop with the
regularisation given as its Hessian H_m (phi_m = (m - m_ref)^T H_m (m - m_ref)
/ 2, e.g. from `gauss_newton.tensor_regularization`) and optional bounds by
clamping, as the reference's `bind` does.
"""
import numpy as np
import scipy.sparse as sp
import torch

from .gauss_newton import data_residual


def lbfgs_minimize(eng, m0, dobs, w, Hm=None, beta=0.0, m_ref=None, max_iter=30,
                   history_size=12, lr=1.0, line_search_fn="strong_wolfe", lower=None,
                   upper=None, nu=0.0, log_data=False):
    """Minimise phi_d + beta phi_m from m0 with one torch L-BFGS run of at
    most `max_iter` iterations; phi_d = ||w (dpred - dobs)||^2, or the
    Student-t misfit (nu + 1) sum log1p(r^2 / nu) when nu > 0
    (dctorch/optimization.py:126-131); log_data: the residual of the
    log-amplitudes (`gauss_newton.data_residual`).  Returns (m, history) with
    history a list of (phi_d, phi_m) per closure."""
    dev = torch.device(eng.device)
    dt = torch.float64
    m = torch.tensor(np.asarray(m0, float), dtype=dt, device=dev, requires_grad=True)
    dobs = torch.as_tensor(dobs, dtype=dt, device=dev)
    w = torch.as_tensor(w, dtype=dt, device=dev)
    mref = torch.zeros_like(m) if m_ref is None else torch.as_tensor(m_ref, dtype=dt, device=dev)
    Ht = None
    if Hm is not None and beta != 0.0:
        H = sp.csr_matrix(Hm)
        Ht = torch.sparse_csr_tensor(torch.as_tensor(H.indptr, dtype=torch.int64, device=dev),
                                     torch.as_tensor(H.indices, dtype=torch.int64, device=dev),
                                     torch.as_tensor(H.data, dtype=dt, device=dev), H.shape)
    lo = -np.inf if lower is None else float(lower)
    hi = np.inf if upper is None else float(upper)

    def bind(x):
        return x if lower is None and upper is None else torch.clamp(x, lo, hi)

    opt = torch.optim.LBFGS([m], lr=lr, max_iter=max_iter, history_size=history_size,
                            line_search_fn=line_search_fn)
    history = []

    def closure():
        opt.zero_grad()
        mc = bind(m)
        if (nu and nu > 0) or log_data:
            r, _ = data_residual(eng.dpred(mc), dobs, w, log_data)
            phi_d = (nu + 1.0) * torch.log1p(r ** 2 / nu).sum() if nu and nu > 0 \
                else (r ** 2).sum()
        else:
            phi_d = eng.misfit(mc, dobs, w)
        loss = phi_d
        phi_m = torch.zeros((), dtype=dt, device=dev)
        if Ht is not None:
            dm = mc - mref
            phi_m = 0.5 * (dm @ (Ht @ dm))
            loss = loss + beta * phi_m
        loss.backward()
        history.append((float(phi_d.detach()), float(phi_m.detach())))
        return loss

    opt.step(closure)
    return bind(m).detach(), history
