"""Host build of the 2.5-D nodal problem in factored form (once per mesh+survey).

The reference extracts, per wavenumber, a linear map from conductivity to the
CSR values of SimPEG's system matrix (dctorch/simulation2d.py:141-229):

    A(ky).data = L_ky sigma,   L_ky = T1 Qe + ky^2 Td Qn + Td AvgBC[ky]

Here the same map is built directly from the tensor-mesh geometry, per cell,
and kept FACTORED: two wavenumber-independent coefficient sets on a CSR over
A's entries (``le``: the Grad^T Me Grad term, ``ln``: the node-mass term) plus
the small Robin boundary part ``bc[k]``.  The GPU assembly kernel computes all
nky value rows from one pass over sigma.

Discretisation (SimPEG 0.25 Simulation2DNodal, Robin boundary condition):

* x fastest; node (ix, iz) = ix + (nx+1) iz, cell (ix, iz) = ix + nx iz;
* each cell c with widths hx, hz, volume v = hx hz adds, for its four nodes,
  v/2 (1/hx^2 + 1/hz^2) sigma_c to each diagonal (the two cell edges through
  the node), -v/(2 hx^2) sigma_c to the entries along its two x-edges,
  -v/(2 hz^2) sigma_c along its two z-edges, and ky^2 v/4 sigma_c (node mass)
  to each diagonal;
* every boundary face except the top adds 0.5 alpha L sigma_c to the diagonal
  of both its nodes, with L the face length, c the adjacent cell and
  alpha = ky K1(ky r)/K0(ky r) (r_hat . n_hat), r measured from
  (median node x, top z) to the face centre.
"""
import numpy as np
from scipy.special import k0e, k1e


class Problem2DMaps:
    """Factored assembly map, pattern, sources and receivers of one problem."""

    def __init__(self, mesh, survey, kys):
        hx, hz = (np.asarray(h, float) for h in mesh.h)
        x0, z0 = (float(v) for v in mesh.origin)
        nx, nz = hx.size, hz.size
        self.nx, self.nz = nx, nz
        self.nC = nx * nz
        self.nN = (nx + 1) * (nz + 1)
        self.xn = x0 + np.r_[0.0, np.cumsum(hx)]
        self.zn = z0 + np.r_[0.0, np.cumsum(hz)]
        self.kys = np.asarray(kys, float)
        self.nky = self.kys.size
        self._pattern()
        self._map(hx, hz)
        self._robin(hx, hz)
        self._survey(survey)

    # -- A's CSR pattern: 5-point stencil, columns ascending ------------------
    def _pattern(self):
        nx1, nz1 = self.nx + 1, self.nz + 1
        ix, iz = np.meshgrid(np.arange(nx1), np.arange(nz1))
        ix, iz = ix.ravel(), iz.ravel()
        node = ix + nx1 * iz
        nbr = [(node - nx1, iz > 0), (node - 1, ix > 0), (node, np.ones_like(ix, bool)),
               (node + 1, ix < nx1 - 1), (node + nx1, iz < nz1 - 1)]
        rows = np.concatenate([node[m] for _, m in nbr])
        cols = np.concatenate([c[m] for c, m in nbr])
        order = np.lexsort((cols, rows))
        rows, cols = rows[order], cols[order]
        self.crow = np.r_[0, np.cumsum(np.bincount(rows, minlength=self.nN))].astype(np.int32)
        self.col = cols.astype(np.int32)
        self.nnz = cols.size
        self._keys = rows.astype(np.int64) * self.nN + cols

    def entry(self, r, c):
        k = np.asarray(r, np.int64) * self.nN + np.asarray(c, np.int64)
        p = np.searchsorted(self._keys, k)
        if np.any(p >= self._keys.size) or np.any(self._keys[np.minimum(p, self._keys.size - 1)] != k):
            raise AssertionError("entry missing from the A pattern")
        return p

    # -- conductivity -> values: le (Grad^T Me Grad), ln (node mass) ----------
    def _map(self, hx, hz):
        nx, nx1 = self.nx, self.nx + 1
        cx, cz = np.meshgrid(np.arange(nx), np.arange(self.nz))
        cx, cz = cx.ravel(), cz.ravel()
        cell = cx + nx * cz
        hxc, hzc = hx[cx], hz[cz]
        vol = hxc * hzc
        n00 = cx + nx1 * cz
        n10, n01, n11 = n00 + 1, n00 + nx1, n00 + nx1 + 1
        dg = 0.5 * vol * (1.0 / hxc ** 2 + 1.0 / hzc ** 2)
        ox = -0.5 * vol / hxc ** 2
        oz = -0.5 * vol / hzc ** 2
        zero = np.zeros_like(vol)
        R, Cc, CE, CN, CELL = [], [], [], [], []

        def add(r, c, ce, cn):
            R.append(r); Cc.append(c); CE.append(ce); CN.append(cn); CELL.append(cell)

        for nd in (n00, n10, n01, n11):
            add(nd, nd, dg, 0.25 * vol)
        for a, b in ((n00, n10), (n01, n11)):
            add(a, b, ox, zero); add(b, a, ox, zero)
        for a, b in ((n00, n01), (n10, n11)):
            add(a, b, oz, zero); add(b, a, oz, zero)
        e = self.entry(np.concatenate(R), np.concatenate(Cc))
        cells = np.concatenate(CELL)
        order = np.lexsort((cells, e))
        self.map_ptr = np.r_[0, np.cumsum(np.bincount(e, minlength=self.nnz))].astype(np.int32)
        self.map_cell = cells[order].astype(np.int32)
        self.map_le = np.concatenate(CE)[order]
        self.map_ln = np.concatenate(CN)[order]

    # -- Robin boundary term, per wavenumber ----------------------------------
    def _robin(self, hx, hz):
        nx, nz, nx1 = self.nx, self.nz, self.nx + 1
        xc = 0.5 * (self.xn[1:] + self.xn[:-1])
        zc = 0.5 * (self.zn[1:] + self.zn[:-1])
        sx = np.median(np.tile(self.xn, nz + 1))
        sz = self.zn[-1]
        iz = np.arange(nz)
        ix = np.arange(nx)
        # left, right, bottom faces (the top is a free surface: alpha = 0)
        fa = np.concatenate([0 + nx1 * iz, nx + nx1 * iz, ix])
        fb = np.concatenate([0 + nx1 * (iz + 1), nx + nx1 * (iz + 1), ix + 1])
        fc = np.concatenate([0 + nx * iz, (nx - 1) + nx * iz, ix])
        fx = np.concatenate([np.full(nz, self.xn[0]), np.full(nz, self.xn[-1]), xc])
        fz = np.concatenate([zc, zc, np.full(nx, self.zn[0])])
        nxv = np.concatenate([np.full(nz, -1.0), np.full(nz, 1.0), np.zeros(nx)])
        nzv = np.concatenate([np.zeros(nz), np.zeros(nz), np.full(nx, -1.0)])
        L = np.concatenate([hz, hz, hx])
        dx, dz = fx - sx, fz - sz
        r = np.hypot(dx, dz)
        rdn = (dx * nxv + dz * nzv) / r
        coef = np.empty((self.nky, 2 * fa.size))
        for k, ky in enumerate(self.kys):
            alpha = ky * k1e(ky * r) / k0e(ky * r) * rdn
            c = 0.5 * alpha * L
            coef[k] = np.concatenate([c, c])
        nodes = np.concatenate([fa, fb])
        self.bc_e = self.entry(nodes, nodes).astype(np.int32)
        self.bc_cell = np.concatenate([fc, fc]).astype(np.int32)
        self.bc_coef = coef

    # -- sources (q) and receivers (Pd) ----------------------------------------
    def interp(self, x, z):
        """Bilinear nodal interpolation of one point: (nodes, weights)."""
        nx, nz = self.nx, self.nz
        ix = int(np.clip(np.searchsorted(self.xn, x, side="right") - 1, 0, nx - 1))
        iz = int(np.clip(np.searchsorted(self.zn, z, side="right") - 1, 0, nz - 1))
        tx = (x - self.xn[ix]) / (self.xn[ix + 1] - self.xn[ix])
        tz = (z - self.zn[iz]) / (self.zn[iz + 1] - self.zn[iz])
        nodes, w = [], []
        for ddx, wx in ((0, 1.0 - tx), (1, tx)):
            for ddz, wz in ((0, 1.0 - tz), (1, tz)):
                if wx * wz != 0.0:
                    nodes.append(ix + ddx + (nx + 1) * (iz + ddz))
                    w.append(wx * wz)
        return nodes, w

    def _survey(self, survey):
        src = survey.source_list
        self.nS = len(src)
        q_rows, q_cols, q_vals = [], [], []
        P_rows, P_cols, P_vals, sidx = [], [], [], []
        self._src_at, self._rx_at = [], []  # (sign, (x, z)) per source / per datum
        nd = 0
        for s, so in enumerate(src):
            at = []
            for sign, loc in ((1.0, so.location_a), (-1.0, getattr(so, "location_b", None))):
                if loc is None:
                    continue
                loc = np.asarray(loc, float).reshape(-1)
                if not np.all(np.isfinite(loc)):
                    continue
                n_, w_ = self.interp(loc[0], loc[-1])
                q_rows += n_; q_cols += [s] * len(n_); q_vals += [sign * v for v in w_]
                at.append((sign, (float(loc[0]), float(loc[-1]))))
            self._src_at.append(at)
            for rx in so.receiver_list:
                M = np.atleast_2d(np.asarray(rx.locations_m, float))
                Nn = getattr(rx, "locations_n", None)
                Nn = None if Nn is None else np.atleast_2d(np.asarray(Nn, float))
                for i in range(M.shape[0]):
                    n_, w_ = self.interp(M[i, 0], M[i, -1])
                    P_rows += [nd] * len(n_); P_cols += n_; P_vals += w_
                    at = [(1.0, (float(M[i, 0]), float(M[i, -1])))]
                    if Nn is not None:
                        n_, w_ = self.interp(Nn[i, 0], Nn[i, -1])
                        P_rows += [nd] * len(n_); P_cols += n_; P_vals += [-v for v in w_]
                        at.append((-1.0, (float(Nn[i, 0]), float(Nn[i, -1]))))
                    self._rx_at.append(at)
                    sidx.append(s)
                    nd += 1
        self.nD = nd
        import scipy.sparse as sp
        self.q = sp.csr_matrix((q_vals, (q_rows, q_cols)), shape=(self.nN, self.nS))
        Pd = sp.csr_matrix((P_vals, (P_rows, P_cols)), shape=(nd, self.nN))
        Pd.sum_duplicates()
        Pd.sort_indices()
        self.Pd = Pd
        self.sidx = np.asarray(sidx, np.int64)

    def electrode_basis(self):
        """(E, S, R): E (nN x ne) the interpolation columns of the ne distinct
        electrode positions (sources and receivers), S (ne x nS) and R
        (ne x nD) their signed combinations, so that q = E S and
        Pd = (E R)^T exactly (the same interpolation weights)."""
        import scipy.sparse as sp
        pos = {}
        for at in self._src_at + self._rx_at:
            for _, loc in at:
                pos.setdefault(loc, len(pos))
        ne = len(pos)
        er, ec, ev = [], [], []
        for loc, e in pos.items():
            n_, w_ = self.interp(loc[0], loc[1])
            er += n_; ec += [e] * len(n_); ev += list(w_)
        E = sp.csr_matrix((ev, (er, ec)), shape=(self.nN, ne))

        def combo(ats, ncol):
            r, c, v = [], [], []
            for j, at in enumerate(ats):
                for sign, loc in at:
                    r.append(pos[loc]); c.append(j); v.append(sign)
            M = sp.csr_matrix((v, (r, c)), shape=(ne, ncol))
            M.sum_duplicates()
            return M
        return E, combo(self._src_at, self.nS), combo(self._rx_at, self.nD)

    # -- helpers ----------------------------------------------------------------
    def values(self, sigma):
        """Host evaluation of the map: (nky, nnz) values (numpy, FP64)."""
        ent = np.repeat(np.arange(self.nnz), np.diff(self.map_ptr))
        s = sigma[self.map_cell]
        ae = np.bincount(ent, self.map_le * s, minlength=self.nnz)
        an = np.bincount(ent, self.map_ln * s, minlength=self.nnz)
        out = ae[None, :] + (self.kys ** 2)[:, None] * an[None, :]
        for k in range(self.nky):
            out[k] += np.bincount(self.bc_e, self.bc_coef[k] * sigma[self.bc_cell],
                                  minlength=self.nnz)
        return out

    def node_coords(self):
        X, Z = np.meshgrid(self.xn, self.zn)
        return np.c_[X.ravel(), Z.ravel()]

    def electrode_rows(self):
        """Nodes carrying a source or receiver weight (RHS support)."""
        return np.unique(np.r_[self.q.tocoo().row, self.Pd.tocoo().col]).astype(np.int32)
