"""FP64 CPU restatement of the reference's 2.5-D DC forward-and-adjoint path.

TEST INFRASTRUCTURE ONLY.  Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import this module, and only as the
checker (or, for `cpu_baseline`, as the timed CPU restatement of the
reference).  Nothing in `simpeg_dc_torch_amd/` imports it.

What it restates
----------------
* The discretisation that stock SimPEG 0.25 `Simulation2DNodal` assembles
  (the reference extracts its map from `sim.getA(ky)`,
  dctorch/simulation2d.py:141-229):

      A(ky) = Grad^T diag(Qe sigma) Grad + ky^2 diag(Qn sigma)
              + diag(AvgBC[ky] sigma)                      (Robin, default)

  built here from the operators themselves (nodal gradient, edge and node
  inner products, boundary-face averages), not from a factored map.
* q = P_A^T - P_B^T (unit current), Pd = stacked P_M - P_N rows
  (dctorch/simulation2d.py:231-244).
* The CPU solver: one `scipy.sparse.linalg.splu(A.tocsc())` per wavenumber
  with SciPy defaults, then `.solve` on the dense RHS block
  (dctorch/solver.py:314-374).
* dpred = (Pd sum_k w_k U_k)[iD, sidx]  (dctorch/simulation2d.py:378-396);
  misfit = sum (w (d - dobs))^2, no 1/2  (:531-534).
* The gradient by the adjoint of the same chain (dctorch/solver.py:261-281
  plus the sparse-mm / exp backward of simulation2d.py:335-343), written
  here in operator form:

      dphi/dsigma_c = -sum_k w_k [ Qe^T sum_s (Grad Lam_k)(Grad U_k)
                                   + ky^2 Qn^T sum_s Lam_k U_k
                                   + AvgBC_k^T sum_s Lam_k U_k ]_c

  with Lam_k = A_k^{-1} G, G = Pd^T scatter(2 w^2 r) per source, and
  dphi/dm = -sigma * dphi/dsigma.
* Jtvec / Jmatrix (simulation2d.py:398-518) by the same formula with an
  arbitrary data-space vector in place of 2 w^2 r.

Pinning (what fixes each function independently of this file)
--------------------------------------------------------------
* operators: closed-form 5-point stencil on a uniform mesh
  (4 sigma / -sigma, + ky^2 sigma h^2 on the diagonal), symmetry, linearity in
  sigma, zero row sums of the Grad term;
* physics: the analytic half-space dipole-dipole potential
  V = rho/(2 pi) (1/AM - 1/BM - 1/AN + 1/BN) (dctorch/tests/test_singularity.py:49-51);
* solver: dense `numpy.linalg.solve` and the Laplacian fixture of
  dctorch/tests/test_solver.py:19-25;
* gradient / Jtvec / Jmatrix: central finite differences of `dpred`, the
  adjoint identity, Jmatrix rows == Jtvec(e_i).
* the Robin term (alpha(ky) and the boundary-face layout): per wavenumber,
  the solution for a surface pole at the Robin source point on an unpadded
  mesh converges at second order to the analytic (rho/pi) K0(ky r) on the
  side and bottom boundary nodes; alpha = 0, 2 alpha and alpha without
  r_hat . n_hat stall at 16-300% errors there (tests/test_oracle.py).
* Parity with stock SimPEG itself is unpinned (simpeg is not importable
  here): the formulas are recalled from SimPEG 0.25 and pinned by the
  physics above, not by SimPEG's own outputs.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.special import k0e, k1e


# --------------------------------------------------------------------------
# Robin coefficient (SimPEG 0.25 Simulation2DNodal, bc_type="Robin")
# --------------------------------------------------------------------------
def robin_alpha(ky, r, rdn):
    """alpha of one boundary face: ky K1(ky r)/K0(ky r) (r_hat . n_hat).

    The 2.5-D field of a surface pole at the source point decays as
    K0(ky r), so d u/d n = -alpha u holds exactly for it on every boundary
    face; tests/test_oracle.py pins this against the analytic K0 field
    (with mutants of this function that must fail)."""
    return ky * k1e(ky * r) / k0e(ky * r) * rdn


# --------------------------------------------------------------------------
# geometry (discretize conventions: x fastest; x-edges then z-edges)
# --------------------------------------------------------------------------
class Grid2D:
    """Node/cell/edge bookkeeping of a 2-D tensor mesh given by widths."""

    def __init__(self, hx, hz, origin):
        self.hx = np.asarray(hx, float)
        self.hz = np.asarray(hz, float)
        self.nx, self.nz = self.hx.size, self.hz.size
        self.xn = origin[0] + np.r_[0.0, np.cumsum(self.hx)]
        self.zn = origin[1] + np.r_[0.0, np.cumsum(self.hz)]
        self.xc = 0.5 * (self.xn[1:] + self.xn[:-1])
        self.zc = 0.5 * (self.zn[1:] + self.zn[:-1])
        self.nC = self.nx * self.nz
        self.nN = (self.nx + 1) * (self.nz + 1)
        self.nEx = self.nx * (self.nz + 1)
        self.nEz = (self.nx + 1) * self.nz
        self.nE = self.nEx + self.nEz
        self.vol = np.outer(self.hz, self.hx).ravel()      # c = ix + nx*iz

    def node(self, ix, iz):
        return ix + (self.nx + 1) * iz

    def cell(self, ix, iz):
        return ix + self.nx * iz

    @property
    def nodes(self):
        X, Z = np.meshgrid(self.xn, self.zn)
        return np.c_[X.ravel(), Z.ravel()]

    # nodal gradient: +-1/h on each edge (discretize `nodal_gradient`)
    def grad(self):
        r, c, v = [], [], []
        for iz in range(self.nz + 1):
            for ix in range(self.nx):
                e = ix + self.nx * iz
                r += [e, e]
                c += [self.node(ix, iz), self.node(ix + 1, iz)]
                v += [-1.0 / self.hx[ix], 1.0 / self.hx[ix]]
        for iz in range(self.nz):
            for ix in range(self.nx + 1):
                e = self.nEx + ix + (self.nx + 1) * iz
                r += [e, e]
                c += [self.node(ix, iz), self.node(ix, iz + 1)]
                v += [-1.0 / self.hz[iz], 1.0 / self.hz[iz]]
        return sp.csr_matrix((v, (r, c)), shape=(self.nE, self.nN))

    # diag of the edge inner product = Qe sigma: each edge gets vol/2 of every
    # cell it bounds (the isotropic tensor-mesh edge inner product)
    def Qe(self):
        r, c, v = [], [], []
        for iz in range(self.nz + 1):
            for ix in range(self.nx):
                e = ix + self.nx * iz
                for jz in (iz - 1, iz):
                    if 0 <= jz < self.nz:
                        cc = self.cell(ix, jz)
                        r.append(e); c.append(cc); v.append(self.vol[cc] / 2.0)
        for iz in range(self.nz):
            for ix in range(self.nx + 1):
                e = self.nEx + ix + (self.nx + 1) * iz
                for jx in (ix - 1, ix):
                    if 0 <= jx < self.nx:
                        cc = self.cell(jx, iz)
                        r.append(e); c.append(cc); v.append(self.vol[cc] / 2.0)
        return sp.csr_matrix((v, (r, c)), shape=(self.nE, self.nC))

    # node mass = aveN2CC^T diag(vol): vol/4 of each cell to each of its nodes
    def Qn(self):
        r, c, v = [], [], []
        for iz in range(self.nz):
            for ix in range(self.nx):
                cc = self.cell(ix, iz)
                for dx, dz in ((0, 0), (1, 0), (0, 1), (1, 1)):
                    r.append(self.node(ix + dx, iz + dz)); c.append(cc)
                    v.append(self.vol[cc] / 4.0)
        return sp.csr_matrix((v, (r, c)), shape=(self.nN, self.nC))

    def boundary_faces(self):
        """(nodes a, nodes b, adjacent cell, centre x, centre z, normal, length,
        is_top) of every boundary face.  Side faces first, then bottom/top."""
        out = []
        for iz in range(self.nz):
            for ix, cx, nrm in ((0, 0, (-1.0, 0.0)), (self.nx, self.nx - 1, (1.0, 0.0))):
                out.append((self.node(ix, iz), self.node(ix, iz + 1),
                            self.cell(cx, iz), self.xn[ix], self.zc[iz], nrm,
                            self.hz[iz], False))
        for ix in range(self.nx):
            for iz, cz, nrm, top in ((0, 0, (0.0, -1.0), False),
                                     (self.nz, self.nz - 1, (0.0, 1.0), True)):
                out.append((self.node(ix, iz), self.node(ix + 1, iz),
                            self.cell(ix, cz), self.xc[ix], self.zn[iz], nrm,
                            self.hx[ix], top))
        return out

    def AvgBC(self, ky):
        """AvgN2Fb^T diag(alpha * length) AvgCC2Fb: (nN x nC).  alpha = 0 on the
        top (free surface); elsewhere ky K1(ky r)/K0(ky r) (r_hat . n_hat), r
        from (median node x, top z) to the face centre."""
        sx = np.median(self.nodes[:, 0])
        sz = self.zn[-1]
        r, c, v = [], [], []
        for na, nb, cc, fx, fz, nrm, L, top in self.boundary_faces():
            if top:
                continue
            dx, dz = fx - sx, fz - sz
            rr = np.hypot(dx, dz)
            rdn = (dx * nrm[0] + dz * nrm[1]) / rr
            alpha = robin_alpha(ky, rr, rdn)
            for nd in (na, nb):
                r.append(nd); c.append(cc); v.append(0.5 * alpha * L)
        return sp.csr_matrix((v, (r, c)), shape=(self.nN, self.nC))

    def interp_row(self, x, z):
        """Bilinear nodal interpolation weights of one point: dict node->w."""
        ix = int(np.clip(np.searchsorted(self.xn, x, side="right") - 1, 0, self.nx - 1))
        iz = int(np.clip(np.searchsorted(self.zn, z, side="right") - 1, 0, self.nz - 1))
        tx = (x - self.xn[ix]) / self.hx[ix]
        tz = (z - self.zn[iz]) / self.hz[iz]
        w = {}
        for dx, wx in ((0, 1 - tx), (1, tx)):
            for dz, wz in ((0, 1 - tz), (1, tz)):
                if wx * wz != 0.0:
                    nd = self.node(ix + dx, iz + dz)
                    w[nd] = w.get(nd, 0.0) + wx * wz
        return w


# --------------------------------------------------------------------------
# the problem
# --------------------------------------------------------------------------
class Problem2D:
    """Operators, sources, receivers of one 2.5-D nodal problem (FP64).

    hx, hz : cell widths; origin : (x0, z0); survey : any object with
    SimPEG's source/receiver attribute names; kys, wq : quadrature.
    """

    def __init__(self, hx, hz, origin, survey, kys, wq):
        self.g = Grid2D(hx, hz, origin)
        self.kys = np.asarray(kys, float)
        self.wq = np.asarray(wq, float)
        self.nky = self.kys.size
        g = self.g
        self.Grad = g.grad()
        self.Qe_ = g.Qe()
        self.Qn_ = g.Qn()
        self.AvgBC_ = [g.AvgBC(k) for k in self.kys]
        # sources and receivers
        nS = len(survey.source_list)
        q = np.zeros((g.nN, nS))
        rows, sidx = [], []
        self._src_at, self._rx_at = [], []   # [(sign, (x, z))] per source / datum
        for s, src in enumerate(survey.source_list):
            for nd, w in g.interp_row(*src.location_a[:2]).items():
                q[nd, s] += w
            at = [(1.0, tuple(float(v) for v in src.location_a[:2]))]
            if getattr(src, "location_b", None) is not None:
                for nd, w in g.interp_row(*src.location_b[:2]).items():
                    q[nd, s] -= w
                at.append((-1.0, tuple(float(v) for v in src.location_b[:2])))
            self._src_at.append(at)
            for rx in src.receiver_list:
                M = np.atleast_2d(rx.locations_m)
                N = None if getattr(rx, "locations_n", None) is None \
                    else np.atleast_2d(rx.locations_n)
                for i in range(M.shape[0]):
                    row = dict(g.interp_row(*M[i, :2]))
                    at = [(1.0, tuple(float(v) for v in M[i, :2]))]
                    if N is not None:
                        for nd, w in g.interp_row(*N[i, :2]).items():
                            row[nd] = row.get(nd, 0.0) - w
                        at.append((-1.0, tuple(float(v) for v in N[i, :2])))
                    rows.append(row)
                    self._rx_at.append(at)
                    sidx.append(s)
        self.q = q
        self.nS = nS
        self.nD = len(rows)
        r, c, v = [], [], []
        for i, row in enumerate(rows):
            for nd, w in row.items():
                r.append(i); c.append(nd); v.append(w)
        self.Pd = sp.csr_matrix((v, (r, c)), shape=(self.nD, g.nN))
        self.sidx = np.asarray(sidx, dtype=np.int64)

    # -- assembly -------------------------------------------------------------
    def A(self, sigma, k):
        """System matrix of wavenumber index k, CSR with sorted indices."""
        ky = self.kys[k]
        A = (self.Grad.T @ sp.diags(self.Qe_ @ sigma) @ self.Grad
             + ky ** 2 * sp.diags(self.Qn_ @ sigma)
             + sp.diags(self.AvgBC_[k] @ sigma)).tocsr()
        A.sum_duplicates()
        A.sort_indices()
        return A

    # -- the CPU solver (SuperLU, SciPy defaults) ------------------------------
    def fields(self, sigma, rhs=None):
        """U_k = A_k^{-1} rhs for every k (rhs defaults to q): (nky, nN, k)."""
        rhs = self.q if rhs is None else rhs
        out = np.empty((self.nky,) + rhs.shape)
        lus = []
        for k in range(self.nky):
            lu = spla.splu(self.A(sigma, k).tocsc())
            out[k] = lu.solve(np.ascontiguousarray(rhs))
            lus.append(lu)
        return out, lus

    def dpred(self, m):
        sigma = np.exp(-m)
        U, _ = self.fields(sigma)
        Ubar = np.tensordot(self.wq, U, axes=1)
        return (self.Pd @ Ubar)[np.arange(self.nD), self.sidx]

    def misfit(self, m, dobs, w):
        r = w * (self.dpred(m) - dobs)
        return float(r @ r)

    def _adjoint_rhs(self, vd):
        """G (nN x nS): column s = sum over data of source s of Pd_i^T vd_i."""
        G = np.zeros((self.g.nN, self.nS))
        PdT = self.Pd.T.tocsc()
        for i in range(self.nD):
            lo, hi = PdT.indptr[i], PdT.indptr[i + 1]
            G[PdT.indices[lo:hi], self.sidx[i]] += PdT.data[lo:hi] * vd[i]
        return G

    def _sigma_grad(self, U, Lam):
        """sum_k w_k Lam_k^T (dA_k/dsigma) U_k summed over columns: (nC,)."""
        out = np.zeros(self.g.nC)
        for k in range(self.nky):
            gl = self.Grad @ Lam[k]
            gu = self.Grad @ U[k]
            nodal = (Lam[k] * U[k]).sum(1)
            out += self.wq[k] * (self.Qe_.T @ (gl * gu).sum(1)
                                 + self.kys[k] ** 2 * (self.Qn_.T @ nodal)
                                 + self.AvgBC_[k].T @ nodal)
        return out

    def misfit_and_grad(self, m, dobs, w):
        sigma = np.exp(-m)
        U, lus = self.fields(sigma)
        Ubar = np.tensordot(self.wq, U, axes=1)
        d = (self.Pd @ Ubar)[np.arange(self.nD), self.sidx]
        r = d - dobs
        phi = float(((w * r) ** 2).sum())
        G = self._adjoint_rhs(2.0 * w ** 2 * r)
        Lam = np.stack([lu.solve(G) for lu in lus])
        gsig = -self._sigma_grad(U, Lam)
        return phi, -sigma * gsig, d

    def Jtvec(self, m, v):
        sigma = np.exp(-m)
        U, lus = self.fields(sigma)
        G = self._adjoint_rhs(np.asarray(v, float))
        Lam = np.stack([lu.solve(G) for lu in lus])
        return -sigma * -self._sigma_grad(U, Lam)

    def Jmatrix(self, m):
        """(nD, nC): row i = J^T e_i, from one set of adjoint solves."""
        sigma = np.exp(-m)
        U, lus = self.fields(sigma)
        PdT = self.Pd.T.toarray()                  # (nN, nD)
        J = np.zeros((self.nD, self.g.nC))
        Y = [lu.solve(PdT) for lu in lus]
        for i in range(self.nD):
            s = self.sidx[i]
            acc = np.zeros(self.g.nC)
            for k in range(self.nky):
                lam = Y[k][:, i]
                u = U[k][:, s]
                acc += self.wq[k] * (
                    self.Qe_.T @ ((self.Grad @ lam) * (self.Grad @ u))
                    + self.kys[k] ** 2 * (self.Qn_.T @ (lam * u))
                    + self.AvgBC_[k].T @ (lam * u))
            J[i] = sigma * acc
        return J


# --------------------------------------------------------------------------
# the reference CPU chain wavenumber by wavenumber (the CPU baseline)
# --------------------------------------------------------------------------
def misfit_and_grad_by_ky(P, m, dobs, w, workers=1):
    """Problem2D.misfit_and_grad organised as the reference's CPU path runs it
    (dctorch/solver.py:314-374 with simulation2d.py:345-375): per wavenumber
    one SuperLU factorisation (SciPy defaults) and the forward solve of every
    source column; the projection and misfit; per wavenumber the adjoint
    solve on the kept factor and the gradient contribution.  workers > 1
    runs the independent wavenumbers of each phase on a thread pool (SuperLU
    releases the interpreter lock); workers = 1 is the reference's sequential
    loop.  Returns (phi, grad_m, dpred)."""
    from concurrent.futures import ThreadPoolExecutor
    sigma = np.exp(-m)
    q = np.ascontiguousarray(P.q)

    def run(fn):
        if workers <= 1:
            return [fn(k) for k in range(P.nky)]
        with ThreadPoolExecutor(min(workers, P.nky)) as ex:
            return list(ex.map(fn, range(P.nky)))

    def fwd(k):
        lu = spla.splu(P.A(sigma, k).tocsc())
        return lu, lu.solve(q)
    F = run(fwd)
    Ubar = sum(P.wq[k] * F[k][1] for k in range(P.nky))
    d = (P.Pd @ Ubar)[np.arange(P.nD), P.sidx]
    r = d - dobs
    phi = float(((w * r) ** 2).sum())
    G = np.ascontiguousarray(P._adjoint_rhs(2.0 * w ** 2 * r))

    def adj(k):
        lu, U = F[k]
        Lam = lu.solve(G)
        gl, gu = P.Grad @ Lam, P.Grad @ U
        nodal = (Lam * U).sum(1)
        return P.wq[k] * (P.Qe_.T @ (gl * gu).sum(1) + P.kys[k] ** 2 * (P.Qn_.T @ nodal)
                          + P.AvgBC_[k].T @ nodal)
    return phi, sigma * sum(run(adj)), d


# --------------------------------------------------------------------------
# the same quantities through the electrode positions (exact by linearity)
# --------------------------------------------------------------------------
def _electrode_basis(P):
    """(E, S, R): E (nN x ne) the interpolation columns of the ne distinct
    electrode positions, S (ne x nS) and R (ne x nD) their signed sums, so
    that q = E S and Pd = (E R)^T exactly (the same interpolation rows)."""
    pos = {}
    for at in P._src_at + P._rx_at:
        for _, loc in at:
            pos.setdefault(loc, len(pos))
    E = np.zeros((P.g.nN, len(pos)))
    for loc, j in pos.items():
        for nd, w in P.g.interp_row(*loc).items():
            E[nd, j] += w

    def combo(ats):
        M = np.zeros((len(pos), len(ats)))
        for c, at in enumerate(ats):
            for sign, loc in at:
                M[pos[loc], c] += sign
        return M
    return E, combo(P._src_at), combo(P._rx_at)


def misfit_and_grad_electrodes(P, m, dobs, w):
    """Problem2D.misfit_and_grad through the electrode positions: one SuperLU
    solve per wavenumber with the ne electrode columns, G_k = A_k^{-1} E,
    instead of the nS source and nS adjoint columns.  Exact by linearity:
    U_k = G_k S, the adjoint fields Lam_k = G_k T with T[:, s] the receiver
    combinations R weighted by the data residual of source s, and
    sum_s f(Lam_k[:, s]) g(U_k[:, s]) = rowsum((f(G_k) M) o g(G_k)) with
    M = T S^T.  Pinned against misfit_and_grad (tests/test_oracle.py); used
    where the source-column restatement would take minutes (the 25,088-cell
    line: 849 sources, 61 electrodes).  Returns (phi, grad_m, dpred)."""
    E, S, R = _electrode_basis(P)
    sigma = np.exp(-m)
    G = [spla.splu(P.A(sigma, k).tocsc()).solve(E) for k in range(P.nky)]
    K = sum(P.wq[k] * (E.T @ G[k]) for k in range(P.nky))          # ne x ne
    d = np.einsum("ei,ei->i", R, (K @ S)[:, P.sidx])
    r = d - dobs
    phi = float(((w * r) ** 2).sum())
    v = 2.0 * w ** 2 * r
    T = np.zeros((E.shape[1], P.nS))
    np.add.at(T.T, P.sidx, (R * v).T)
    M = T @ S.T
    out = np.zeros(P.g.nC)
    for k in range(P.nky):
        gG = P.Grad @ G[k]
        edge = ((gG @ M) * gG).sum(1)
        nodal = ((G[k] @ M) * G[k]).sum(1)
        out += P.wq[k] * (P.Qe_.T @ edge + P.kys[k] ** 2 * (P.Qn_.T @ nodal)
                          + P.AvgBC_[k].T @ nodal)
    return phi, sigma * out, d


def jacobian_rows_electrodes(P, m, rows):
    """Rows `rows` of J = d dpred / d m through the electrode positions (as
    misfit_and_grad_electrodes): J[i] = sigma * sum_k w_k [Qe^T((Grad G R_i)
    o (Grad G S_s)) + ky^2 Qn^T((G R_i) o (G S_s)) + AvgBC_k^T((G R_i) o
    (G S_s))], s the source of datum i.  Pinned against Jmatrix
    (tests/test_oracle.py)."""
    E, S, R = _electrode_basis(P)
    sigma = np.exp(-m)
    rows = np.asarray(rows)
    out = np.zeros((rows.size, P.g.nC))
    for k in range(P.nky):
        G = spla.splu(P.A(sigma, k).tocsc()).solve(E)
        gG = P.Grad @ G
        lam, u = G @ R[:, rows], G @ S[:, P.sidx[rows]]
        glam, gu = gG @ R[:, rows], gG @ S[:, P.sidx[rows]]
        out += P.wq[k] * ((P.Qe_.T @ (glam * gu)).T + P.kys[k] ** 2 * (P.Qn_.T @ (lam * u)).T
                          + (P.AvgBC_[k].T @ (lam * u)).T)
    return out * sigma[None, :]


def halfspace_dipole(rho, a, b, m, n):
    """Analytic surface potential difference of a dipole-dipole on a half
    space of resistivity rho, unit current: rho/(2 pi)(1/AM-1/BM-1/AN+1/BN)."""
    d = lambda p, q: np.linalg.norm(np.asarray(p, float) - np.asarray(q, float))
    return rho / (2 * np.pi) * (1 / d(a, m) - 1 / d(b, m) - 1 / d(a, n) + 1 / d(b, n))


def splu_solve(crow, col, vals, n, b):
    """The reference CPU solver on one CSR matrix: splu(A.tocsc()).solve(b)."""
    A = sp.csr_matrix((vals, col, crow), shape=(n, n))
    return spla.splu(A.tocsc()).solve(np.ascontiguousarray(b))
