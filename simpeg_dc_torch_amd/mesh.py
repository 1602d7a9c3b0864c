"""Tensor meshes with discretize's conventions, for hosts without discretize.

`TorchDC2D` reads only a handful of mesh attributes (`h`, `origin`, `nC`,
`n_nodes`, `nodes`, `cell_centers`, `cell_volumes`), so a `discretize.TensorMesh`
works unchanged where discretize is installed.  This class provides the same
attributes with the same conventions where it is not:

* cell widths are given per axis as scalars, ``(h, n)`` or ``(h, n, factor)``
  tuples (a negative factor grows the padding towards the low end);
* the origin is a vector or one character per axis: ``"0"`` (origin at 0),
  ``"C"`` (axis centred on 0) or ``"N"`` (axis ends at 0);
* every flattened array runs x fastest (Fortran order), nodes and cells alike.
"""
import numpy as np


def unpack_widths(spec):
    """Cell widths of one axis from discretize's width specification."""
    out = []
    for v in spec:
        if np.isscalar(v):
            out.append(float(v))
        elif len(v) == 2:
            out += [float(v[0])] * int(v[1])
        elif len(v) == 3:
            h, n, f = float(v[0]), int(v[1]), float(v[2])
            pad = h * np.abs(f) ** np.arange(1, n + 1)
            out += list(pad[::-1] if f < 0 else pad)
        else:
            raise ValueError(f"cannot read width specification {v!r}")
    return np.asarray(out, dtype=float)


class TensorMesh:
    """A rectilinear 2-D or 3-D mesh (x fastest), discretize-compatible."""

    _meshType = "TENSOR"

    def __init__(self, h, origin=None):
        hs = []
        for ax in h:
            if np.isscalar(ax):
                raise ValueError("each axis needs a list of widths")
            hs.append(unpack_widths(ax))
        self.h = hs
        self.dim = len(hs)
        if self.dim not in (2, 3):
            raise ValueError("only 2-D and 3-D tensor meshes are supported")
        if origin is None:
            origin = "0" * self.dim
        if isinstance(origin, str):
            if len(origin) != self.dim:
                raise ValueError(f"origin {origin!r} does not match dim {self.dim}")
            origin = list(origin)
        o = []
        for ax, v in zip(hs, origin):
            if isinstance(v, str):
                if v == "0":
                    o.append(0.0)
                elif v == "C":
                    o.append(-ax.sum() / 2.0)
                elif v == "N":
                    o.append(-ax.sum())
                else:
                    raise ValueError(f"unknown origin code {v!r}")
            else:
                o.append(float(v))
        self.origin = np.asarray(o, dtype=float)

    # -- sizes ---------------------------------------------------------------
    @property
    def shape_cells(self):
        return tuple(len(a) for a in self.h)

    @property
    def shape_nodes(self):
        return tuple(len(a) + 1 for a in self.h)

    @property
    def nC(self):
        return int(np.prod(self.shape_cells))

    n_cells = nC

    @property
    def n_nodes(self):
        return int(np.prod(self.shape_nodes))

    nN = n_nodes

    # -- coordinates -----------------------------------------------------------
    def _axis_nodes(self, i):
        return self.origin[i] + np.r_[0.0, np.cumsum(self.h[i])]

    def _axis_centers(self, i):
        n = self._axis_nodes(i)
        return 0.5 * (n[1:] + n[:-1])

    @property
    def nodes_x(self):
        return self._axis_nodes(0)

    @property
    def nodes_y(self):
        return self._axis_nodes(1)

    @property
    def nodes_z(self):
        return self._axis_nodes(self.dim - 1)

    @property
    def cell_centers_x(self):
        return self._axis_centers(0)

    @property
    def cell_centers_y(self):
        return self._axis_centers(1)

    @property
    def cell_centers_z(self):
        return self._axis_centers(self.dim - 1)

    def _grid(self, axes):
        g = np.meshgrid(*axes, indexing="ij")
        return np.column_stack([a.ravel(order="F") for a in g])

    @property
    def nodes(self):
        return self._grid([self._axis_nodes(i) for i in range(self.dim)])

    gridN = nodes

    @property
    def cell_centers(self):
        return self._grid([self._axis_centers(i) for i in range(self.dim)])

    gridCC = cell_centers

    @property
    def cell_volumes(self):
        g = np.meshgrid(*self.h, indexing="ij")
        v = np.ones_like(g[0])
        for a in g:
            v = v * a
        return v.ravel(order="F")
