"""Host build of the 3-D nodal problem in factored form (once per mesh+survey).

The reference collapses stock SimPEG's `Simulation3DNodal` assembly into one
linear map from conductivity to the CSR values of the system matrix,

    A(sigma).data = L3 sigma,   L3 = T1 Qe + Td AvgBC

(dctorch/simulation3d.py:84-143).  Here the same map is built directly from
the tensor-mesh geometry, per cell, on a CSR over A's entries (``le``), plus
the Robin boundary part (``bc``), in the layout of the 2.5-D map with one
"wavenumber" of zero (no node-mass term), so the same GPU assembly, gradient
and transposed-map kernels serve both engines.

Discretisation (SimPEG 0.25 Simulation3DNodal, Robin boundary condition):

* x fastest; node (ix, iy, iz) = ix + (nx+1)(iy + (ny+1) iz), cell likewise;
* each cell c with widths hx, hy, hz and volume v adds, for each of its eight
  nodes, v/4 (1/hx^2 + 1/hy^2 + 1/hz^2) sigma_c to the diagonal (the three
  cell edges through the node), and -v/(4 h_a^2) sigma_c to the two entries
  of each of its four edges along axis a;
* every boundary face except the top adds alpha A sigma_c / 4 to the diagonal
  of each of its four nodes, A the face area, c the adjacent cell and
  alpha = (r_hat . n_hat) / r, r measured from (median node x, median node y,
  top z) to the face centre.
"""
import numpy as np
import scipy.sparse as sp


class Problem3DMaps:
    """Factored assembly map, pattern, sources and receivers of one problem;
    the attribute layout of `Problem2DMaps` with ``kys = [0]``."""

    def __init__(self, mesh, survey):
        if len(mesh.h) != 3:
            raise ValueError("Problem3DMaps needs a 3-D tensor mesh")
        self.h = [np.asarray(a, float) for a in mesh.h]
        self.n = [a.size for a in self.h]
        self.n1 = [k + 1 for k in self.n]
        self.nodes1 = [float(mesh.origin[i]) + np.r_[0.0, np.cumsum(self.h[i])] for i in range(3)]
        self.nC = int(np.prod(self.n))
        self.nN = int(np.prod(self.n1))
        self.kys = np.zeros(1)
        self.nky = 1
        self._pattern()
        self._map()
        self._robin()
        self._survey(survey)

    def _node(self, ix, iy, iz):
        return ix + self.n1[0] * (iy + self.n1[1] * iz)

    # -- A's CSR pattern: 7-point stencil, columns ascending ------------------
    def _pattern(self):
        nx1, ny1, nz1 = self.n1
        g = np.meshgrid(np.arange(nx1), np.arange(ny1), np.arange(nz1), indexing="ij")
        ix, iy, iz = (a.ravel(order="F") for a in g)
        node = self._node(ix, iy, iz)
        sxy = nx1 * ny1
        nbr = [(node - sxy, iz > 0), (node - nx1, iy > 0), (node - 1, ix > 0),
               (node, np.ones_like(ix, bool)), (node + 1, ix < nx1 - 1),
               (node + nx1, iy < ny1 - 1), (node + sxy, iz < nz1 - 1)]
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
        if np.any(p >= self._keys.size) or \
                np.any(self._keys[np.minimum(p, self._keys.size - 1)] != k):
            raise AssertionError("entry missing from the A pattern")
        return p

    # -- conductivity -> values ------------------------------------------------
    def _map(self):
        nx, ny, nz = self.n
        g = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        cx, cy, cz = (a.ravel(order="F") for a in g)
        cell = cx + nx * (cy + ny * cz)
        hh = [self.h[0][cx], self.h[1][cy], self.h[2][cz]]
        vol = hh[0] * hh[1] * hh[2]
        w = [vol / (4.0 * a ** 2) for a in hh]          # per edge of the cell, per axis
        corner = {}
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    corner[(dx, dy, dz)] = self._node(cx + dx, cy + dy, cz + dz)
        R, Cc, V, CELL = [], [], [], []
        dg = w[0] + w[1] + w[2]
        for nd in corner.values():
            R.append(nd); Cc.append(nd); V.append(dg); CELL.append(cell)
        for axis in range(3):
            for key, a in corner.items():
                if key[axis] == 1:
                    continue
                k2 = list(key)
                k2[axis] = 1
                b = corner[tuple(k2)]
                R += [a, b]; Cc += [b, a]; V += [-w[axis], -w[axis]]; CELL += [cell, cell]
        e = self.entry(np.concatenate(R), np.concatenate(Cc))
        cells = np.concatenate(CELL)
        order = np.lexsort((cells, e))
        self.map_ptr = np.r_[0, np.cumsum(np.bincount(e, minlength=self.nnz))].astype(np.int32)
        self.map_cell = cells[order].astype(np.int32)
        self.map_le = np.concatenate(V)[order]
        self.map_ln = np.zeros_like(self.map_le)

    # -- Robin boundary term ------------------------------------------------------
    def _robin(self):
        nodes = self.node_coords()
        src = np.r_[np.median(nodes[:, 0]), np.median(nodes[:, 1]), self.nodes1[2][-1]]
        cent1 = [0.5 * (a[1:] + a[:-1]) for a in self.nodes1]
        E, CELL, COEF = [], [], []
        for axis in range(3):
            others = [i for i in range(3) if i != axis]
            for side in (0, 1):
                if axis == 2 and side == 1:
                    continue                         # the top: free surface
                g = np.meshgrid(np.arange(self.n[others[0]]), np.arange(self.n[others[1]]),
                                indexing="ij")
                u, v = (a.ravel(order="F") for a in g)
                fixed = 0 if side == 0 else self.n[axis]
                cfix = 0 if side == 0 else self.n[axis] - 1
                idx_c = [None, None, None]
                idx_c[axis] = np.full(u.size, cfix)
                idx_c[others[0]], idx_c[others[1]] = u, v
                cell = idx_c[0] + self.n[0] * (idx_c[1] + self.n[1] * idx_c[2])
                centre = np.zeros((u.size, 3))
                centre[:, axis] = self.nodes1[axis][fixed]
                centre[:, others[0]] = cent1[others[0]][u]
                centre[:, others[1]] = cent1[others[1]][v]
                normal = np.zeros(3)
                normal[axis] = -1.0 if side == 0 else 1.0
                area = self.h[others[0]][u] * self.h[others[1]][v]
                rv = centre - src
                r = np.linalg.norm(rv, axis=1)
                alpha = (rv @ normal) / (r * r)
                for d0 in (0, 1):
                    for d1 in (0, 1):
                        idx_n = [None, None, None]
                        idx_n[axis] = np.full(u.size, fixed)
                        idx_n[others[0]], idx_n[others[1]] = u + d0, v + d1
                        nd = self._node(*idx_n)
                        E.append(nd); CELL.append(cell); COEF.append(0.25 * alpha * area)
        nodes_ = np.concatenate(E)
        self.bc_e = self.entry(nodes_, nodes_).astype(np.int32)
        self.bc_cell = np.concatenate(CELL).astype(np.int32)
        self.bc_coef = np.concatenate(COEF)[None, :]

    # -- sources (q) and receivers (Pd) ----------------------------------------
    def interp(self, p):
        """Trilinear nodal interpolation of one point: (nodes, weights)."""
        ii, tt = [], []
        for a in range(3):
            n1 = self.nodes1[a]
            i = int(np.clip(np.searchsorted(n1, p[a], side="right") - 1, 0, self.n[a] - 1))
            ii.append(i)
            tt.append((p[a] - n1[i]) / (n1[i + 1] - n1[i]))
        nodes, w = [], []
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    wt = ((tt[0] if dx else 1 - tt[0]) * (tt[1] if dy else 1 - tt[1]) *
                          (tt[2] if dz else 1 - tt[2]))
                    if wt != 0.0:
                        nodes.append(self._node(ii[0] + dx, ii[1] + dy, ii[2] + dz))
                        w.append(wt)
        return nodes, w

    def _survey(self, survey):
        src = survey.source_list
        self.nS = len(src)
        q_rows, q_cols, q_vals = [], [], []
        P_rows, P_cols, P_vals, sidx = [], [], [], []
        self._src_at, self._rx_at = [], []  # (sign, position) per source / per datum
        nd = 0
        for s, so in enumerate(src):
            at = []
            for sign, loc in ((1.0, so.location_a), (-1.0, getattr(so, "location_b", None))):
                if loc is None:
                    continue
                loc = np.asarray(loc, float).reshape(-1)
                if loc.size != 3 or not np.all(np.isfinite(loc)):
                    continue
                n_, w_ = self.interp(loc)
                q_rows += n_; q_cols += [s] * len(n_); q_vals += [sign * v for v in w_]
                at.append((sign, tuple(float(x) for x in loc)))
            self._src_at.append(at)
            for rx in so.receiver_list:
                M = np.atleast_2d(np.asarray(rx.locations_m, float))
                Nn = getattr(rx, "locations_n", None)
                Nn = None if Nn is None else np.atleast_2d(np.asarray(Nn, float))
                for i in range(M.shape[0]):
                    n_, w_ = self.interp(M[i])
                    P_rows += [nd] * len(n_); P_cols += n_; P_vals += w_
                    at = [(1.0, tuple(float(x) for x in M[i]))]
                    if Nn is not None:
                        n_, w_ = self.interp(Nn[i])
                        P_rows += [nd] * len(n_); P_cols += n_; P_vals += [-v for v in w_]
                        at.append((-1.0, tuple(float(x) for x in Nn[i])))
                    self._rx_at.append(at)
                    sidx.append(s)
                    nd += 1
        self.nD = nd
        self.q = sp.csr_matrix((q_vals, (q_rows, q_cols)), shape=(self.nN, self.nS))
        Pd = sp.csr_matrix((P_vals, (P_rows, P_cols)), shape=(nd, self.nN))
        Pd.sum_duplicates()
        Pd.sort_indices()
        self.Pd = Pd
        self.sidx = np.asarray(sidx, np.int64)

    def electrode_basis(self):
        """(E, S, R) with q = E S and Pd = (E R)^T: E (nN x ne) the
        interpolation columns of the distinct electrode positions, S and R
        their signed combinations per source and per datum."""
        pos = {}
        for at in self._src_at + self._rx_at:
            for _, loc in at:
                pos.setdefault(loc, len(pos))
        ne = len(pos)
        er, ec, ev = [], [], []
        for loc, e in pos.items():
            n_, w_ = self.interp(np.asarray(loc))
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
        """Host evaluation of the map: (1, nnz) values (numpy, FP64)."""
        ent = np.repeat(np.arange(self.nnz), np.diff(self.map_ptr))
        out = np.bincount(ent, self.map_le * sigma[self.map_cell], minlength=self.nnz)
        out = out + np.bincount(self.bc_e, self.bc_coef[0] * sigma[self.bc_cell],
                                minlength=self.nnz)
        return out[None, :]

    def node_coords(self):
        g = np.meshgrid(*self.nodes1, indexing="ij")
        return np.column_stack([a.ravel(order="F") for a in g])

    def electrode_rows(self):
        """Nodes carrying a source or receiver weight (RHS support)."""
        return np.unique(np.r_[self.q.tocoo().row, self.Pd.tocoo().col]).astype(np.int32)
