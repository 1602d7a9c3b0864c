"""Synthetic 2.5-D workloads: meshes, electrode arrays and models.

* `regression_case`  -- the reference's 2.5-D regression problem
  (dctorch/tests/test_simulation2d.py:24-56): 100 x 50 cells of 1 m, 21
  electrodes at -30..30, dipole-dipole n = 1..6, a 10 ohm-m block in 100.
* `suite_case`       -- the reference's 2d-tensor coverage case
  (dctorch/tests/test_suite.py:62-87): 64 x 32 cells, 15 electrodes, DD
  n = 1..4, rough model (seed 11, sigma_log 0.3).
* `small_padded_case` -- 4.8k cells, 2.5 m core with 1.3x padding, 21
  electrodes at 5 m, DD n = 1..6, 100 ohm-m with 10 and 1000 ohm-m Gaussian
  blobs (seed 0).
* `line_25k_case`    -- 25,088 cells (200 x 100 core of 1 m, 12 padding cells
  growing 1.3x), 61 electrodes at 2 m, dipole-dipole (a = 1, 2, 3 electrode
  spacings, n = 1..6) plus Wenner-Schlumberger (a = 1, 2, n = 1..8), sources
  merged by electrode pair: 849 sources, 1,636 data; m = log(rho) a Gaussian
  random field (mean log 100, std 1, correlation length 10 m, seed 1).
* `gn_line_case`     -- the Gauss-Newton line: 62,016 cells, 96 electrodes at
  5 m, dipole-dipole a = 1..3, n = 1..10 (270 sources, 2,430 data), blocks of
  5, 50 and 2000 ohm-m in 200 ohm-m.
* `uneven_lines_case` -- 16 independent lines of 3k-40k cells (log-uniform),
  21-96 electrodes each, dipole-dipole, independent random fields.
"""
import numpy as np

from .mesh import TensorMesh
from .survey import Survey, receivers, sources


def dipole_dipole(xe, nmax, spacings=(1,), z=0.0):
    """{(A, B): [(M, N), ...]} electrode-index pairs of a DD array."""
    out = {}
    ne = len(xe)
    for a in spacings:
        for i in range(ne):
            for n in range(1, nmax + 1):
                jm, jn = i + a + n * a, i + 2 * a + n * a
                if jn >= ne:
                    break
                out.setdefault((i, i + a), []).append((jm, jn))
    return out


def wenner_schlumberger(xe, nmax, spacings=(1,)):
    out = {}
    ne = len(xe)
    for a in spacings:
        for n in range(1, nmax + 1):
            for i in range(ne):
                ib = i + (2 * n + 1) * a
                if ib >= ne:
                    break
                im = i + n * a
                out.setdefault((i, ib), []).append((im, im + a))
    return out


def build_survey(xe, pairs, z=0.0):
    """Survey of dipole sources with merged receiver lists (sorted by A, B)."""
    srcs = []
    for (ia, ib) in sorted(pairs):
        rx = sorted(set(pairs[(ia, ib)]))
        M = np.array([[xe[m], z] for m, _ in rx])
        N = np.array([[xe[n], z] for _, n in rx])
        srcs.append(sources.Dipole([receivers.Dipole(M, N)], [xe[ia], z], [xe[ib], z]))
    return Survey(srcs)


def merge(*dicts):
    out = {}
    for d in dicts:
        for k, v in d.items():
            out.setdefault(k, []).extend(v)
    return out


def regression_case():
    mesh = TensorMesh([[(1.0, 100)], [(1.0, 50)]], origin="CN")
    xe = np.linspace(-30, 30, 21)
    srcs = []
    for i in range(len(xe) - 1):
        lm, ln = [], []
        for n in range(1, 7):
            jm, jn = i + 1 + n, i + 2 + n
            if jn >= len(xe):
                break
            lm.append([xe[jm], 0.0])
            ln.append([xe[jn], 0.0])
        if lm:
            srcs.append(sources.Dipole([receivers.Dipole(np.array(lm), np.array(ln))],
                                       np.r_[xe[i], 0.0], np.r_[xe[i + 1], 0.0]))
    survey = Survey(srcs)
    m = np.log(100.0) * np.ones(mesh.nC)
    cc = mesh.cell_centers
    blk = (np.abs(cc[:, 0] - 5) < 8) & (cc[:, 1] < -3) & (cc[:, 1] > -10)
    m[blk] = np.log(10.0)
    return mesh, survey, m


def suite_case(seed=11):
    rng = np.random.default_rng(seed)
    mesh = TensorMesh([[(1.0, 64)], [(1.0, 32)]], origin="CN")
    xe = np.linspace(-20, 20, 15)
    srcs = []
    for i in range(len(xe) - 1):
        lm, ln = [], []
        for n in range(1, 5):
            jm, jn = i + 1 + n, i + 2 + n
            if jn >= len(xe):
                break
            lm.append([xe[jm], 0.0])
            ln.append([xe[jn], 0.0])
        if lm:
            srcs.append(sources.Dipole([receivers.Dipole(np.array(lm), np.array(ln))],
                                       np.r_[xe[i], 0.0], np.r_[xe[i + 1], 0.0]))
    m = np.log(100.0) * np.ones(mesh.nC) + 0.3 * rng.standard_normal(mesh.nC)
    return mesh, Survey(srcs), m


def small_padded_case(seed=0):
    mesh = TensorMesh([[(2.5, 8, -1.3), (2.5, 80), (2.5, 8, 1.3)],
                       [(2.5, 10, -1.3), (2.5, 40)]], origin="CN")
    xe = np.linspace(-50, 50, 21)
    survey = build_survey(xe, dipole_dipole(xe, 6))
    rng = np.random.default_rng(seed)
    cc = mesh.cell_centers
    c1 = np.array([rng.uniform(-60, 60), rng.uniform(-60, -15)])
    c2 = np.array([rng.uniform(-60, 60), rng.uniform(-60, -15)])
    g1 = np.exp(-((cc - c1) ** 2).sum(1) / (2 * 8.0 ** 2))
    g2 = np.exp(-((cc - c2) ** 2).sum(1) / (2 * 8.0 ** 2))
    m = np.log(100.0) + np.log(10.0 / 100.0) * g1 + np.log(1000.0 / 100.0) * g2
    return mesh, survey, m


def gaussian_random_field(mesh, rng, corr_len, std=1.0, mean=np.log(100.0), dx=0.5):
    """Smooth field at the cell centres: white noise on a dx grid over the mesh,
    Gaussian filter of std corr_len/2 (covariance exp(-r^2/corr_len^2)),
    standardised, scaled, sampled at the nearest grid point."""
    from scipy.ndimage import gaussian_filter
    lo = mesh.origin
    ext = np.array([np.sum(h) for h in mesh.h])
    shape = tuple(int(np.ceil(e / dx)) + 1 for e in ext)
    noise = rng.standard_normal(shape)
    f = gaussian_filter(noise, sigma=(corr_len / 2.0) / dx, mode="reflect")
    f = (f - f.mean()) / f.std()
    cc = mesh.cell_centers
    idx = tuple(np.clip(np.rint((cc[:, i] - lo[i]) / dx).astype(int), 0, shape[i] - 1)
                for i in range(cc.shape[1]))
    return mean + std * f[idx]


def line_25k_case(seed=1):
    mesh = TensorMesh([[(1.0, 12, -1.3), (1.0, 200), (1.0, 12, 1.3)],
                       [(1.0, 12, -1.3), (1.0, 100)]], origin="CN")
    xe = np.arange(-60.0, 60.0 + 1e-9, 2.0)
    pairs = merge(dipole_dipole(xe, 6, spacings=(1, 2, 3)),
                  wenner_schlumberger(xe, 8, spacings=(1, 2)))
    survey = build_survey(xe, pairs)
    rng = np.random.default_rng(seed)
    m = gaussian_random_field(mesh, rng, corr_len=10.0, std=1.0)
    return mesh, survey, m, rng


def gn_line_case():
    """The Gauss-Newton line: 62,016 cells (384 x 140 core of 1.25 m cells plus
    12 padding cells growing 1.3x), 96 electrodes at 5 m, dipole-dipole with
    a = 1, 2, 3 electrode spacings and n = 1..10 (a = 1 alone: 93 sources and
    885 data; all three: 270 sources, 2,430 data); true model 200 ohm-m with
    blocks of 5, 50 and 2000 ohm-m.  Returns (mesh, survey, m_true)."""
    mesh = TensorMesh([[(1.25, 12, -1.3), (1.25, 384), (1.25, 12, 1.3)],
                       [(1.25, 12, -1.3), (1.25, 140)]], origin="CN")
    xe = np.arange(96) * 5.0 - 237.5
    survey = build_survey(xe, dipole_dipole(xe, 10, spacings=(1, 2, 3)))
    cc = mesh.cell_centers
    m = np.full(mesh.nC, np.log(200.0))
    for (x0, x1, z0, z1), rho in (((-150.0, -90.0, -40.0, -15.0), 5.0),
                                  ((-30.0, 30.0, -60.0, -20.0), 50.0),
                                  ((80.0, 150.0, -35.0, -10.0), 2000.0)):
        inside = (cc[:, 0] > x0) & (cc[:, 0] < x1) & (cc[:, 1] > z0) & (cc[:, 1] < z1)
        m[inside] = np.log(rho)
    return mesh, survey, m


def uneven_lines_case(seed=5, nlines=16, cells=(3000, 40000)):
    """An uneven batch of independent 2.5-D lines: cell counts log-uniform in
    `cells`; each line a 2:1 core of 1 m cells with 12 padding cells growing
    1.3x on the sides and below; 21-96 electrodes on nodes across the core, a
    dipole-dipole array (n = 1..6); m = log(rho) an independent Gaussian random
    field (mean log 100, std 1, correlation length a tenth of the core width).
    Returns a list of (mesh, survey, m)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nlines):
        N = float(np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1]))))
        nx = max(20, int(round(np.sqrt(2.0 * N) - 24.0)))   # (nx + 24)(nx / 2 + 12) = N
        nz = max(10, nx // 2)
        mesh = TensorMesh([[(1.0, 12, -1.3), (1.0, nx), (1.0, 12, 1.3)],
                           [(1.0, 12, -1.3), (1.0, nz)]], origin="CN")
        ne = int(np.clip(round(nx / 2.7), 21, 96))
        step = max(1, (nx - 4) // (ne - 1))
        start = 12 + (nx - (ne - 1) * step) // 2          # core nodes, centred array
        xe = mesh.nodes_x[start + step * np.arange(ne)]
        survey = build_survey(xe, dipole_dipole(xe, 6))
        m = gaussian_random_field(mesh, rng, corr_len=max(5.0, nx / 10.0), std=1.0)
        out.append((mesh, survey, m))
    return out


# ---------------------------------------------------------------------------
# 3-D surveys (pole sources, dipole receivers on a surface electrode grid)
# ---------------------------------------------------------------------------
def surface_grid(nx_e, ny_e, spacing, x0=None, y0=None, z=0.0):
    """(n, 3) electrode positions on a surface grid, x fastest, centred."""
    xs = (np.arange(nx_e) - (nx_e - 1) / 2.0) * spacing if x0 is None else x0 + spacing * np.arange(nx_e)
    ys = (np.arange(ny_e) - (ny_e - 1) / 2.0) * spacing if y0 is None else y0 + spacing * np.arange(ny_e)
    X, Y = np.meshgrid(xs, ys)
    return np.c_[X.ravel(), Y.ravel(), np.full(X.size, z)]


def pole_dipole_survey(sources, electrodes, dipoles, n_rx):
    """One pole source per row of `sources`; its receivers are the `n_rx`
    dipoles (index pairs into `electrodes`) whose midpoints lie nearest to
    it, skipping dipoles that touch the source position."""
    from .survey import DipoleReceiver, PoleSource, Survey
    e = np.asarray(electrodes, float)
    dp = np.asarray(dipoles, int)
    mid = 0.5 * (e[dp[:, 0]] + e[dp[:, 1]])
    out = []
    for a in np.asarray(sources, float):
        ok = (np.linalg.norm(e[dp[:, 0]] - a, axis=1) > 1e-9) & \
             (np.linalg.norm(e[dp[:, 1]] - a, axis=1) > 1e-9)
        cand = np.where(ok)[0]
        cand = cand[np.argsort(np.linalg.norm(mid[cand] - a, axis=1), kind="stable")][:n_rx]
        rx = DipoleReceiver(e[dp[cand, 0]], e[dp[cand, 1]])
        out.append(PoleSource([rx], a))
    return Survey(out)


def x_dipoles(nx_e, ny_e):
    """Adjacent electrode pairs along x of a surface_grid."""
    return np.array([(j * nx_e + i, j * nx_e + i + 1)
                     for j in range(ny_e) for i in range(nx_e - 1)])


def small_3d_case(seed=0):
    """Small 3-D problem for parity tests: 10 x 10 x 5 core of 1 m cells, 3
    padding cells growing 1.3x; a 5 x 3 electrode grid at 2 m, every
    electrode a pole source read by its 6 nearest x-dipoles; m = log 100 plus
    a seeded perturbation."""
    mesh = TensorMesh([[(1.0, 3, -1.3), (1.0, 10), (1.0, 3, 1.3)],
                       [(1.0, 3, -1.3), (1.0, 10), (1.0, 3, 1.3)],
                       [(1.0, 3, -1.3), (1.0, 5)]], origin="CCN")
    el = surface_grid(5, 3, 2.0)
    survey = pole_dipole_survey(el, el, x_dipoles(5, 3), 6)
    rng = np.random.default_rng(seed)
    m = np.log(100.0) + 0.3 * rng.standard_normal(mesh.nC)
    return mesh, survey, m


def medium_3d_case(seed=1):
    """3-D problem whose upper fronts exceed the in-LDS kernels (FP64 factor
    and forward sweep take the large-front path): 16 x 16 x 8 core of 1 m
    cells, 6 padding cells growing 1.3x (28 x 28 x 14 cells); a 5 x 5
    electrode grid at 2 m, every electrode a pole source read by its 8
    nearest x-dipoles."""
    mesh = TensorMesh([[(1.0, 6, -1.3), (1.0, 16), (1.0, 6, 1.3)],
                       [(1.0, 6, -1.3), (1.0, 16), (1.0, 6, 1.3)],
                       [(1.0, 6, -1.3), (1.0, 8)]], origin="CCN")
    el = surface_grid(5, 5, 2.0)
    survey = pole_dipole_survey(el, el, x_dipoles(5, 5), 8)
    rng = np.random.default_rng(seed)
    m = gaussian_random_field(mesh, rng, corr_len=4.0, std=0.5, dx=0.5)
    return mesh, survey, m


def grid_3d_case(seed=3):
    """The 3-D L-BFGS workload: 40 x 40 x 30 core of 5 m cells plus 8
    padding cells per side growing 1.3x (56 x 56 x 38 = 119,168 cells); 64
    surface electrodes on an 8 x 8 grid at 10 m; pole-dipole with 200 source
    positions (a 20 x 10 grid at 5 m over the core) each read by its 30
    nearest x-dipoles of the electrode grid: 6,000 data; m = log(rho) a 3-D
    Gaussian random field (mean log 100, std 0.8, correlation length 25 m)."""
    pad = (5.0, 8, 1.3)
    mesh = TensorMesh([[(pad[0], pad[1], -pad[2]), (5.0, 40), pad],
                       [(pad[0], pad[1], -pad[2]), (5.0, 40), pad],
                       [(pad[0], pad[1], -pad[2]), (5.0, 30)]], origin="CCN")
    el = surface_grid(8, 8, 10.0)
    src = surface_grid(20, 10, 5.0)
    survey = pole_dipole_survey(src, el, x_dipoles(8, 8), 30)
    rng = np.random.default_rng(seed)
    m = gaussian_random_field(mesh, rng, corr_len=25.0, std=0.8, dx=2.5)
    return mesh, survey, m


CASES = {"regression": regression_case, "suite": suite_case,
         "small_padded": small_padded_case, "small_3d": small_3d_case,
         "medium_3d": medium_3d_case}
