"""Numpy emulation of the GPU factor and sweep kernels, step for step.

Test helper: it runs the multifrontal factorisation (extend-add, partial
Cholesky, panel P = [L_JJ^-1; L_RJ]) and the two sweeps exactly as the
kernels in csrc/dcx.hip do, level by level, on the structure that the host
symbolic analysis returns.  A mistake in the structure (fronts, relative
positions, entry maps) or in the sweep algebra shows up here, on the CPU.
"""
import numpy as np


def factor(sym, vals):
    ns_, rs_ = sym["ns"], sym["rs"]
    nsup = sym["nsup"]
    fronts = [None] * nsup
    Q = [None] * nsup
    for d in range(sym["ndepth"] - 1, -1, -1):
        for s in sym["lvl_sup"][sym["lvl_ptr"][d]:sym["lvl_ptr"][d + 1]]:
            ns, rs = ns_[s], rs_[s]
            nf = ns + rs
            F = np.zeros((nf, nf))
            a0, a1 = sym["aptr"][s], sym["aptr"][s + 1]
            F[sym["a_row"][a0:a1], sym["a_col"][a0:a1]] = vals[sym["a_e"][a0:a1]]
            for c in sym["cidx"][sym["cptr"][s]:sym["cptr"][s + 1]]:
                rp = sym["relpos"][sym["rptr"][c]:sym["rptr"][c + 1]]
                Uc = fronts[c][ns_[c]:, ns_[c]:]
                low = np.tril(Uc)
                F[np.ix_(rp, rp)] += low
            for j in range(ns):
                F[j, j] = np.sqrt(F[j, j])
                F[j + 1:, j] /= F[j, j]
                F[j + 1:, j + 1:] -= np.tril(np.outer(F[j + 1:, j], F[j + 1:, j]))
            L_JJ = np.tril(F[:ns, :ns])
            Linv = np.linalg.inv(L_JJ)
            Q[s] = np.vstack([np.tril(Linv), F[ns:, :ns]])
            fronts[s] = F
    return Q


def solve(sym, Q, B):
    """B: (n x k) in elimination order; returns X in elimination order."""
    W = np.array(B, dtype=float, copy=True)
    ns_, rs_, first = sym["ns"], sym["rs"], sym["first"]
    upd = [None] * sym["nsup"]
    for d in range(sym["ndepth"] - 1, -1, -1):
        for s in sym["lvl_sup"][sym["lvl_ptr"][d]:sym["lvl_ptr"][d + 1]]:
            ns, rs = ns_[s], rs_[s]
            f = np.zeros((ns + rs, W.shape[1]))
            f[:ns] = W[first[s]:first[s] + ns]
            for c in sym["cidx"][sym["cptr"][s]:sym["cptr"][s + 1]]:
                rp = sym["relpos"][sym["rptr"][c]:sym["rptr"][c + 1]]
                f[rp] += upd[c]
            y = Q[s][:ns] @ f[:ns]
            W[first[s]:first[s] + ns] = y
            upd[s] = f[ns:] - Q[s][ns:] @ y
    for d in range(sym["ndepth"]):
        for s in sym["lvl_sup"][sym["lvl_ptr"][d]:sym["lvl_ptr"][d + 1]]:
            ns, rs = ns_[s], rs_[s]
            R = sym["ridx"][sym["rptr"][s]:sym["rptr"][s + 1]]
            t = W[first[s]:first[s] + ns] - Q[s][ns:].T @ W[R]
            W[first[s]:first[s] + ns] = Q[s][:ns].T @ t
    return W
