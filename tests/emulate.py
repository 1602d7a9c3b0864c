"""Numpy emulation of the GPU factor and sweep kernels, step for step.

Test helper: it runs the multifrontal factorisation (extend-add, partial
Cholesky, panel P = [L_JJ^-1; Q_R], Q_R = L_RJ L_JJ^-1) and the two sweeps
exactly as the
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
            Q[s] = np.vstack([np.tril(Linv), F[ns:, :ns] @ np.tril(Linv)])
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
            out = Q[s] @ f[:ns]
            W[first[s]:first[s] + ns] = out[:ns]
            upd[s] = f[ns:] - out[ns:]
    for d in range(sym["ndepth"]):
        for s in sym["lvl_sup"][sym["lvl_ptr"][d]:sym["lvl_ptr"][d + 1]]:
            ns, rs = ns_[s], rs_[s]
            R = sym["ridx"][sym["rptr"][s]:sym["rptr"][s + 1]]
            W[first[s]:first[s] + ns] = Q[s].T @ np.vstack([W[first[s]:first[s] + ns], -W[R]])
    return W


def solve_fused(sym, Q, B):
    """The sweeps with fused bottom subtrees, as the fused kernels run them:
    each subtree works in a local block X (its rows [lo, hi) then the root's
    off-diagonal rows), a sublevel scratch Tb and colour groups; the other
    supernodes run level by level; fused roots pass their update through a
    separate buffer."""
    ns_, rs_, first = sym["ns"], sym["rs"], sym["first"]
    rptr, ridx, lridx, toff = sym["rptr"], sym["ridx"], sym["lridx"], sym["toff"]
    W = np.zeros_like(B, dtype=float)
    upd = {}

    def members(L):
        return sym["f_mem"][sym["f_lv"][L]:sym["f_lv"][L + 1]]

    # forward: fused subtrees first
    for t in range(sym["nfused"]):
        root, lo, hi = sym["f_root"][t], sym["f_lo"][t], sym["f_hi"][t]
        nloc, rsT = hi - lo, rs_[root]
        X = np.zeros((nloc + rsT, B.shape[1]))
        X[:nloc] = B[lo:hi]
        Tb = np.zeros((sym["f_tb_rows"][t], B.shape[1]))
        for L in range(sym["f_lv_end"][t] - 1, sym["f_lv_begin"][t] - 1, -1):
            for s in members(L):
                jl = first[s] - lo
                Tb[toff[s]:toff[s] + ns_[s]] = Q[s][:ns_[s]] @ X[jl:jl + ns_[s]]
            for g in range(sym["f_gptr"][L], sym["f_gptr"][L + 1]):
                grp = sym["f_mem"][sym["f_gb"][g]:sym["f_gb"][g + 1]]
                rows = [lridx[rptr[s]:rptr[s + 1]] for s in grp]
                flat = np.concatenate(rows) if rows else np.zeros(0, int)
                assert flat.size == np.unique(flat).size, "colour group rows overlap"
                for s in grp:
                    jl = first[s] - lo
                    lr = lridx[rptr[s]:rptr[s + 1]]
                    X[lr] -= Q[s][ns_[s]:] @ X[jl:jl + ns_[s]]
            for s in members(L):
                jl = first[s] - lo
                X[jl:jl + ns_[s]] = Tb[toff[s]:toff[s] + ns_[s]]
        W[lo:hi] = X[:nloc]
        upd[root] = X[nloc:].copy()
    # forward: the levels above
    up, usup = sym["ulvl_ptr"], sym["ulvl_sup"]
    for d in range(sym["ndepth"] - 1, -1, -1):
        for s in usup[up[d]:up[d + 1]]:
            ns, rs = ns_[s], rs_[s]
            f = np.zeros((ns + rs, B.shape[1]))
            f[:ns] = B[first[s]:first[s] + ns]
            for c in sym["cidx"][sym["cptr"][s]:sym["cptr"][s + 1]]:
                f[sym["relpos"][rptr[c]:rptr[c + 1]]] += upd[c]
            out = Q[s] @ f[:ns]
            W[first[s]:first[s] + ns] = out[:ns]
            upd[s] = f[ns:] - out[ns:]
    # backward: levels above, root first
    for d in range(sym["ndepth"]):
        for s in usup[up[d]:up[d + 1]]:
            ns = ns_[s]
            R = ridx[rptr[s]:rptr[s + 1]]
            W[first[s]:first[s] + ns] = Q[s].T @ np.vstack([W[first[s]:first[s] + ns], -W[R]])
    # backward: fused subtrees
    for t in range(sym["nfused"]):
        root, lo, hi = sym["f_root"][t], sym["f_lo"][t], sym["f_hi"][t]
        nloc, rsT = hi - lo, rs_[root]
        X = np.zeros((nloc + rsT, B.shape[1]))
        X[:nloc] = W[lo:hi]
        X[nloc:] = W[ridx[rptr[root]:rptr[root + 1]]]
        Tb = np.zeros((sym["f_tb_rows"][t], B.shape[1]))
        for L in range(sym["f_lv_begin"][t], sym["f_lv_end"][t]):
            for s in members(L):
                jl, ns = first[s] - lo, ns_[s]
                lr = lridx[rptr[s]:rptr[s + 1]]
                Tb[toff[s]:toff[s] + ns] = Q[s].T @ np.vstack([X[jl:jl + ns], -X[lr]])
            for s in members(L):
                jl, ns = first[s] - lo, ns_[s]
                X[jl:jl + ns] = Tb[toff[s]:toff[s] + ns]
        W[lo:hi] = X[:nloc]
    return W
