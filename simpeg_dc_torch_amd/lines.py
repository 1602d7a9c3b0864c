"""An uneven batch of independent 2.5-D lines on one GPU.

A set of survey lines of different sizes (e.g. 16 lines of 3k-40k cells)
cannot share one batched plan: their patterns differ.  `LineBatch` keeps one
`TorchDC2D` (one plan) per line and gets throughput from concurrency and
launch economy:

* streams: every line's eval+grad is issued on one of `streams`
  default-priority streams, so the small lines' level launches overlap and
  fill the chip together instead of running one line after another;
* host threads: with `threads=True` one host thread per stream issues its
  lines (`TorchDC2D.eval_grad_device`: no autograd graph, no host
  synchronisation), so the launch work of several lines proceeds in
  parallel on the host as well: the library's entry points are C calls,
  which run without the interpreter lock;
* graphs: `capture()` records each line's whole eval+grad as a HIP graph
  through `torch.cuda.CUDAGraph` and `eval_grad` then replays them, for FP64
  engines only: replays of FP64 graphs match the eager results bit for bit,
  while a replay of mixed-precision graphs ended the process with a host
  segmentation fault inside hipGraphLaunch whose cause is not found, so
  capture refuses mixed-precision engines (`LineForest.capture` records the
  one-plan batch as a single graph, under the same restriction);
* autotuning: `autotune_leaf` times the nested-dissection leaf size per size
  class (cells rounded to a power of two) on one line of that class and
  `LineBatch(leaf="auto")` builds every line with its class's choice;
  `autotune_tiles` then times the kernel-variant thresholds of a plan (sweep
  column tile, split-K threshold of the backward sweep, wide-factor
  threshold, factor panel width; `dcx_plan_set_tuning`) by coordinate search, per size class in
  `LineBatch(tiles="auto")` and for the one-plan batch
  (`LineForest.autotune_tiles`).  Tuning is not part of a timed step.

The nearest reference analogue is cuDSS's batched sequence (dctorch/
solver.py:181-185), which needs one pattern; there is no uneven-batch code in
the reference.
"""
import ctypes as C
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.sparse as sp
import torch

from .simulation2d import TorchDC2D


def size_class(n_cells):
    """Power-of-two bucket of a line's cell count."""
    return int(np.round(np.log2(max(n_cells, 1))))


def _eval_grad(eng, x, dobs, w):
    x.grad = None
    phi = eng.misfit(x, dobs, w)
    phi.backward()
    return phi, x.grad


#: candidate values of each kernel-variant threshold (dcx_plan_set_tuning)
TILE_GRID = {"sweep_tc": (16, 32, 64), "splitk_below": (1024, 4096, 16384),
             "rowsplit_below": (0, 4096, 16384, 65536),
             "wide_below": (0, 321, 4096), "panel_width": (16, 32)}


def tile_key(setting):
    """Canonical text of a tuning setting (the keys of the tuners' timings)."""
    return ",".join(f"{k}={v}" for k, v in sorted(setting.items()))


def tune_solver_tiles(solver, step, reps=3, grid=None, device=None):
    """Coordinate search over the plan's kernel-variant thresholds: each knob
    in turn takes every value of its grid with the others at the best so
    far; `step()` issues one eval+grad, timed (median of `reps` after one
    warm-up) with device synchronisation.  Leaves the best setting on the
    solver.  Returns (best setting, {setting: ms})."""
    grid = TILE_GRID if grid is None else grid
    best = solver.tuning()
    times = {}

    def timed():
        step()
        torch.cuda.synchronize(device)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            step()
            torch.cuda.synchronize(device)
            ts.append(1e3 * (time.perf_counter() - t0))
        return sorted(ts)[len(ts) // 2]

    key = tile_key
    t_best = times[key(best)] = timed()
    for name, values in grid.items():
        for v in values:
            trial = dict(best, **{name: v})
            if key(trial) in times:
                continue
            solver.set_tuning(**trial)
            times[key(trial)] = t = timed()
            if t < t_best:
                best, t_best = trial, t
        solver.set_tuning(**best)
    return best, times


def autotune_tiles(eng, m, dobs, w, reps=3, grid=None):
    """`tune_solver_tiles` on one GPU engine's eval+grad (TorchDC2D or
    TorchDC3D); the engine keeps the best setting."""
    dev = eng.device
    x = torch.as_tensor(m, dtype=torch.float64, device=dev)
    d_t = torch.as_tensor(dobs, dtype=torch.float64, device=dev)
    w_t = torch.as_tensor(w, dtype=torch.float64, device=dev)
    return tune_solver_tiles(eng._gpu.solver, lambda: eng.eval_grad_device(x, d_t, w_t),
                             reps=reps, grid=grid, device=dev)


def autotune_leaf(mesh, survey, m, dobs, w, candidates=(32, 48, 64), reps=3, nky=12,
                  precision="mixed", refine=0, device="cuda"):
    """Leaf size with the fastest eval+grad on this line (median of `reps`
    after one warm-up).  Returns (best leaf, {leaf: ms})."""
    dev = torch.device(device)
    times = {}
    for leaf in candidates:
        eng = TorchDC2D(mesh, survey, nky=nky, device=dev, precision=precision,
                        refine=refine, leaf=leaf)
        x = torch.tensor(m, dtype=torch.float64, device=dev, requires_grad=True)
        d_t = torch.as_tensor(dobs, dtype=torch.float64, device=dev)
        w_t = torch.as_tensor(w, dtype=torch.float64, device=dev)
        _eval_grad(eng, x, d_t, w_t)
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize(dev)
            t0 = time.perf_counter()
            _eval_grad(eng, x, d_t, w_t)
            torch.cuda.synchronize(dev)
            ts.append(1e3 * (time.perf_counter() - t0))
        times[leaf] = sorted(ts)[len(ts) // 2]
        del eng
    return min(times, key=times.get), times


class LineBatch:
    """eval+grad of many independent lines per call.

    lines: list of (mesh, survey); leaf: None (the solver default), an int, a
    {size_class: leaf} dict, or "auto" together with `tune_data` =
    [(m, dobs, w) per line] to run `autotune_leaf` on one line per class;
    tiles: None, a {size_class: setting} dict, or "auto" (with `tune_data`:
    `autotune_tiles` on one line per class, its setting given to every line
    of the class); overlap: whether each line's plan uses its own extra
    streams (factor / forward-sweep overlap, split backward sweep, check
    stream; `set_overlap`), None = only when the batch runs on one stream
    (several batch streams already overlap the lines, and each plan's two
    extra streams would share the device's few hardware queues with them);
    further keywords (e.g. rhs_basis="electrodes") go to every TorchDC2D."""

    def __init__(self, lines, device="cuda", nky=12, precision="mixed", refine=0, streams=4,
                 leaf=None, tune_data=None, threads=False, tiles=None, overlap=None,
                 **engine_kw):
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.tuning = {}
        if leaf == "auto":
            if tune_data is None:
                raise ValueError("leaf='auto' needs tune_data")
            leaf = {}
            for i, (mesh, survey) in enumerate(lines):
                c = size_class(mesh.nC)
                if c not in leaf:
                    m, d, w = tune_data[i]
                    leaf[c], self.tuning[c] = autotune_leaf(mesh, survey, m, d, w, nky=nky,
                                                            precision=precision, refine=refine,
                                                            device=self.device)
        self.leaf = leaf
        self.engines = []
        for mesh, survey in lines:
            lf = leaf.get(size_class(mesh.nC)) if isinstance(leaf, dict) else leaf
            self.engines.append(TorchDC2D(mesh, survey, nky=nky, device=self.device,
                                          precision=precision, refine=refine, leaf=lf,
                                          **engine_kw))
        self.tiles = {}
        if tiles == "auto":
            if tune_data is None:
                raise ValueError("tiles='auto' needs tune_data")
            for i, (mesh, _) in enumerate(lines):
                c = size_class(mesh.nC)
                if c not in self.tiles:
                    m, d, w = tune_data[i]
                    self.tiles[c], _ = autotune_tiles(self.engines[i], m, d, w)
        elif isinstance(tiles, dict):
            self.tiles = dict(tiles)
        for (mesh, _), eng in zip(lines, self.engines):
            t = self.tiles.get(size_class(mesh.nC))
            if t:
                eng._gpu.solver.set_tuning(**t)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(max(1, streams))]
        self._pool = ThreadPoolExecutor(max_workers=len(self.streams)) if threads else None
        self._data = None
        self._graphs = None
        self.overlap = overlap
        self._overlap_set = None
        self._apply_overlap()

    def _apply_overlap(self):
        want = self.overlap if self.overlap is not None else len(self.streams) == 1
        if want != self._overlap_set:
            for e in self.engines:
                e._gpu.solver.set_overlap(want)
            self._overlap_set = want

    def set_data(self, dobs, w):
        """Observed data and weights per line (kept on the device)."""
        self._data = [(torch.as_tensor(d, dtype=torch.float64, device=self.device),
                       torch.as_tensor(x, dtype=torch.float64, device=self.device))
                      for d, x in zip(dobs, w)]
        self._graphs = None

    def _stream(self, i):
        return self.streams[i % len(self.streams)]

    def capture(self, ms, warmup=2):
        """Record each line's eval+grad at models `ms` as a HIP graph (FP64
        engines only); later `eval_grad` calls replay them.  Mixed-precision
        engines are refused: a replay of their graphs ended the process with
        a host segmentation fault inside hipGraphLaunch, cause not found."""
        if any(e.precision != "fp64" for e in self.engines):
            raise NotImplementedError(
                "graph capture is limited to FP64 engines: a replay of mixed-precision "
                "graphs ended in a host segmentation fault inside hipGraphLaunch "
                "(cause not found)")
        self._capture_graphs(ms, warmup)

    def _capture_graphs(self, ms, warmup=2):
        """Record each line's eval+grad at models `ms` as a graph (after
        `warmup` eager runs on its stream, which also sizes every buffer the
        library keeps, so nothing is allocated during capture)."""
        if self._data is None:
            raise RuntimeError("set_data first")
        self._apply_overlap()
        graphs = []
        for i, eng in enumerate(self.engines):
            s = self._stream(i)
            s.wait_stream(torch.cuda.current_stream(self.device))
            x = torch.tensor(np.asarray(ms[i], float), dtype=torch.float64, device=self.device,
                             requires_grad=True)
            dobs, w = self._data[i]
            with torch.cuda.stream(s):
                for _ in range(warmup):
                    _eval_grad(eng, x, dobs, w)
            s.synchronize()
            g = torch.cuda.CUDAGraph()
            x.grad = None
            with torch.cuda.graph(g, stream=s):
                phi = eng.misfit(x, dobs, w)
                phi.backward()
            graphs.append((g, x, phi, x.grad))
        torch.cuda.synchronize(self.device)
        self._graphs = graphs
        self._graph_gens = self._buffer_generations()

    def _buffer_generations(self):
        # a plan reallocates a workspace buffer when a later call needs more
        # (a wider solve, a Jacobian block): graphs recorded before hold the
        # old addresses, so the generations are compared before each replay
        return [e._gpu.solver.info()["buffer_generation"] for e in self.engines]

    def _issue(self, t, ms, out):
        """The lines of stream t on that stream: every line's kernels queued
        first, then their solve checks waited for."""
        s = self.streams[t]
        with torch.cuda.device(self.device), torch.cuda.stream(s):
            idx = range(t, len(self.engines), len(self.streams))
            pending = [self._begin(i, ms) for i in idx]
            for i, p in zip(idx, pending):
                out[i] = p.finish()

    def _begin(self, i, ms):
        x = torch.as_tensor(ms[i], dtype=torch.float64, device=self.device)
        dobs, w = self._data[i]
        return self.engines[i].eval_grad_device_begin(x, dobs, w)

    def eval_grad(self, ms):
        """[(phi, grad)] for the models `ms` (one per line), line i on stream
        i mod streams, issued round-robin from this thread, or from one host
        thread per stream with `threads=True`; graph replays when captured.
        The results are ordered after the caller's stream: phi and grad of a
        captured line are its graph's output buffers, overwritten by the next
        call."""
        if self._data is None:
            raise RuntimeError("set_data first")
        if self._graphs is None:
            self._apply_overlap()
        cur = torch.cuda.current_stream(self.device)
        for s in self.streams:
            s.wait_stream(cur)
        if self._graphs is not None and self._buffer_generations() != self._graph_gens:
            self._graphs = None
            raise RuntimeError("a plan reallocated its workspace after capture(): the recorded "
                               "graphs are dropped; call capture() again")
        if self._graphs is not None:
            out = []
            for i, (g, x, phi, grad) in enumerate(self._graphs):
                with torch.cuda.stream(self._stream(i)):
                    with torch.no_grad():
                        x.copy_(torch.as_tensor(ms[i], dtype=torch.float64, device=self.device))
                    g.replay()
                out.append((phi, grad))
        else:
            out = [None] * len(self.engines)
            if self._pool is not None:
                for f in [self._pool.submit(self._issue, t, ms, out)
                          for t in range(len(self.streams))]:
                    f.result()
            else:
                # every line queued before the host waits for any line's checks
                pending = []
                for i in range(len(self.engines)):
                    with torch.cuda.stream(self._stream(i)):
                        pending.append(self._begin(i, ms))
                for i, p in enumerate(pending):
                    with torch.cuda.stream(self._stream(i)):
                        out[i] = p.finish()
            for phi, grad in out:
                phi.record_stream(cur)
                grad.record_stream(cur)
        for s in self.streams:
            cur.wait_stream(s)
        return out


# ---------------------------------------------------------------------------
# every line in one plan
# ---------------------------------------------------------------------------
def merge_line_maps(maps_list, kys_list, wq_list):
    """Stack the operator maps of independent 2.5-D lines into one block-
    diagonal problem (pure numpy, host).

    maps_list: Problem2DMaps per line; kys_list / wq_list: each line's
    wavenumbers and quadrature weights.  Lines with fewer wavenumbers than the
    longest list are padded with their last wavenumber at weight 0 (a valid
    SPD system whose fields are multiplied by 0).  Returns a dict with the
    merged CSR pattern, assembly map (entry group = line), per-(system, line)
    ky^2 and weights, stacked survey (Pd, sidx: line-local source column),
    sources (q, columns 0..nS_i-1 of line i), node coordinates, electrode rows,
    and the offsets of nodes, cells, entries and data per line."""
    L = len(maps_list)
    nb = max(len(k) for k in kys_list)
    node_off = np.r_[0, np.cumsum([m.nN for m in maps_list])]
    cell_off = np.r_[0, np.cumsum([m.nC for m in maps_list])]
    ent_off = np.r_[0, np.cumsum([m.nnz for m in maps_list])]
    term_off = np.r_[0, np.cumsum([m.map_ptr[-1] for m in maps_list])]
    data_off = np.r_[0, np.cumsum([m.nD for m in maps_list])]
    K = max(m.nS for m in maps_list)
    ky2 = np.empty((nb, L))
    wq = np.zeros((nb, L))
    bc_coef = []
    for i, (m, ky, w) in enumerate(zip(maps_list, kys_list, wq_list)):
        ky = np.asarray(ky, float)
        n = ky.size
        ky2[:n, i] = ky ** 2
        ky2[n:, i] = ky[-1] ** 2
        wq[:n, i] = np.asarray(w, float)
        c = np.asarray(m.bc_coef, float)
        bc_coef.append(np.vstack([c, np.repeat(c[-1:], nb - n, axis=0)]))
    cat = np.concatenate
    out = dict(
        nlines=L, nbatch=nb, K=K, n=int(node_off[-1]), nC=int(cell_off[-1]),
        nnz=int(ent_off[-1]), nD=int(data_off[-1]),
        node_off=node_off, cell_off=cell_off, ent_off=ent_off, data_off=data_off,
        crow=cat([m.crow[:-1] + ent_off[i] for i, m in enumerate(maps_list)]
                 + [[ent_off[-1]]]).astype(np.int32),
        col=cat([m.col + node_off[i] for i, m in enumerate(maps_list)]).astype(np.int32),
        map_ptr=cat([m.map_ptr[:-1] + term_off[i] for i, m in enumerate(maps_list)]
                    + [[term_off[-1]]]).astype(np.int32),
        map_cell=cat([m.map_cell + cell_off[i] for i, m in enumerate(maps_list)]).astype(np.int32),
        map_le=cat([m.map_le for m in maps_list]),
        map_ln=cat([m.map_ln for m in maps_list]),
        bc_e=cat([m.bc_e + ent_off[i] for i, m in enumerate(maps_list)]).astype(np.int32),
        bc_cell=cat([m.bc_cell + cell_off[i] for i, m in enumerate(maps_list)]).astype(np.int32),
        bc_coef=np.hstack(bc_coef),
        group=cat([np.full(m.nnz, i) for i, m in enumerate(maps_list)]).astype(np.int32),
        ky2=ky2, wq=wq,
        Pd=sp.block_diag([m.Pd for m in maps_list], format="csr"),
        sidx=cat([m.sidx for m in maps_list]).astype(np.int64),
        data_line=cat([np.full(m.nD, i) for i, m in enumerate(maps_list)]).astype(np.int64),
        node_line=cat([np.full(m.nN, i) for i, m in enumerate(maps_list)]).astype(np.int64),
        coords=np.vstack([m.node_coords() for m in maps_list]),
        rows=cat([m.electrode_rows() + node_off[i] for i, m in enumerate(maps_list)]).astype(np.int32),
    )
    blocks = []
    for i, m in enumerate(maps_list):
        qc = sp.coo_matrix(m.q)
        blocks.append((qc.row + node_off[i], qc.col, qc.data))
    r = cat([b[0] for b in blocks])
    c = cat([b[1] for b in blocks])
    v = cat([b[2] for b in blocks])
    out["q"] = sp.csr_matrix((v, (r, c)), shape=(out["n"], K))
    return out


def merged_values_host(M, sigma):
    """Host evaluation of the merged assembly map (nbatch, nnz): entry e of
    system k is sum_p (le + ky2[k, group e] ln) sigma + boundary terms."""
    ent = np.repeat(np.arange(M["nnz"]), np.diff(M["map_ptr"]))
    s = sigma[M["map_cell"]]
    ae = np.bincount(ent, M["map_le"] * s, minlength=M["nnz"])
    an = np.bincount(ent, M["map_ln"] * s, minlength=M["nnz"])
    out = ae[None, :] + M["ky2"][:, M["group"]] * an[None, :]
    for k in range(M["nbatch"]):
        out[k] += np.bincount(M["bc_e"], M["bc_coef"][k] * sigma[M["bc_cell"]],
                              minlength=M["nnz"])
    return out


class LineForest:
    """eval+grad of many independent 2.5-D lines through ONE plan.

    The lines' systems are stacked block-diagonally: the nested dissection
    splits the disconnected pattern into its components first, so the plan's
    elimination tree is a forest of the lines' own separator trees and every
    level launch (factor and sweeps) serves all lines at once: one launch
    sequence per step instead of one per line.  Each line keeps its own
    wavenumbers (a grouped assembly map, `dcx_lmap_set_groups`) and
    quadrature weights, which enter through per-system forward right-hand
    sides B_k = w_(k, line) q, so the projected data and the gradient gathers
    take unit weights:
        U'_k = A_k^-1 B_k,  d = sum_k Pd U'_k,
        dphi/dA_k = -Lambda_k U'_k^T with Lambda_k = A_k^-1 Pd^T v.
    Every line is solved with K = max nS source columns (a line with fewer
    sources carries zero columns).
    The weighted right-hand sides are kept on the device: nbatch x n x K FP64
    (1.7 GB for 16 lines of 251k cells in all, 10 systems, K = 81).

    rhs_basis="electrodes" solves once, with the lines' electrode columns as
    one right-hand side shared by every system (column j of the block holds
    electrode j of every line that has one; K = max ne), and no adjoint
    solve: with G_k = A_k^-1 E, per line i
        K_i = sum_k w_(k,i) E_i^T G_k,  d_i = (R_i o (K_i S_i)[:, sidx_i]) summed
        over electrodes,
        dphi/dA_k = -G_k w_(k,i) M_i G_k^T on line i's rows,
        M_i = T_i S_i^T,  T_i[:, s] = sum over the data of source s of R_i v,
    so the adjoint field is H_k = G_k (w_(k,i) M_i)^T, one small product per
    line over its rows of the solution (the lines' rows are contiguous in
    the elimination order when the dissection orders component by component;
    otherwise they are gathered).  The same operators as TorchDC2D's
    electrode basis, per line.

    lines: list of (mesh, survey); nky: an int or one per line; precision:
    "mixed" (FP32 factor, sweeps and fields) or "fp64"; refine, leaf and
    quadrature as TorchDC2D; rhs_basis "sources" or "electrodes";
    active_cells: None or one entry per line (None or a boolean mask of the
    line's cells), with val_inactive (a number or one per line) the fixed
    log-resistivity of the inactive cells, as TorchDC2D: each line's model
    and gradient then cover its active cells only."""

    def __init__(self, lines, device="cuda", nky=12, precision="mixed", refine=0, leaf=None,
                 quadrature=None, rhs_basis="sources", active_cells=None, val_inactive=None):
        from .solver import HipBatchSolver
        from . import _lib as lib_mod
        self._libm = lib_mod
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        dev = self.device
        if rhs_basis not in ("sources", "electrodes"):
            raise ValueError(f"rhs_basis must be sources|electrodes, got {rhs_basis!r}")
        self.rhs_basis = rhs_basis
        nkys = list(nky) if np.iterable(nky) else [int(nky)] * len(lines)
        maps, kys, wqs = [], [], []
        for (mesh, survey), nk in zip(lines, nkys):
            e = TorchDC2D(mesh, survey, nky=nk, device="cpu", quadrature=quadrature)
            maps.append(e._maps)
            kys.append(e.kys)
            wqs.append(e.wq_np)
        M = self.M = merge_line_maps(maps, kys, wqs)
        self.nlines, self.K, self.nbatch = M["nlines"], M["K"], M["nbatch"]
        self.nS = [m.nS for m in maps]
        self.nD = [m.nD for m in maps]
        self.nC = [m.nC for m in maps]
        self.precision = precision
        self._graph = None
        self._active_map(active_cells, val_inactive)
        self.solver = HipBatchSolver(
            torch.tensor(M["crow"]), torch.tensor(M["col"]), M["n"], M["K"], M["nbatch"],
            device=dev, dtype=torch.float64, precision=precision, refine=refine,
            coords=M["coords"], leaf=leaf)
        lib = self.lib = self.solver._lib
        self.solver.set_rhs_support(M["rows"])
        perm = self.solver.perm()
        L = lib_mod
        ky2_0, ky2_0p = L.f64(M["ky2"][:, 0])
        ptr, ptrp = L.i32(M["map_ptr"])
        cell, cellp = L.i32(M["map_cell"])
        le, lep = L.f64(M["map_le"])
        ln, lnp = L.f64(M["map_ln"])
        be, bep = L.i32(M["bc_e"])
        bcl, bclp = L.i32(M["bc_cell"])
        bco, bcop = L.f64(M["bc_coef"])
        grp, grpp = L.i32(M["group"])
        ky2, ky2p = L.f64(M["ky2"])
        h = C.c_void_p()
        with torch.cuda.device(dev):
            L.check(lib.dcx_lmap_create(M["nnz"], M["nC"], M["nbatch"], ky2_0p, ptrp, cellp, lep,
                                        lnp, int(be.size), bep, bclp, bcop, C.byref(h)),
                    "lmap_create")
            self.lmap = h
            L.check(lib.dcx_lmap_set_groups(h, M["nlines"], grpp, ky2p), "lmap_set_groups")
        Pd = M["Pd"]
        pp, ppp = L.i32(Pd.indptr)
        pc, pcp = L.i32(Pd.indices)
        pv, pvp = L.f64(Pd.data)
        si, sip = L.i32(M["sidx"])
        h2 = C.c_void_p()
        with torch.cuda.device(dev):
            L.check(lib.dcx_survey_create(self.solver._h, M["nD"], M["K"], ppp, pcp, pvp, sip,
                                          C.byref(h2)), "survey_create")
        self.survey = h2
        self.ones = torch.ones(M["nbatch"], dtype=torch.float64, device=dev)
        self.data_line = torch.tensor(M["data_line"], device=dev)
        self._data = None
        if rhs_basis == "electrodes":
            self._electrode_basis(maps, perm)
            return
        # weighted forward right-hand sides, elimination order: (nbatch, n, K)
        q = torch.tensor(M["q"][perm].toarray(), dtype=torch.float64, device=dev)
        wrow = torch.tensor(M["wq"][:, M["node_line"][perm]], dtype=torch.float64, device=dev)
        self.B = (wrow.unsqueeze(2) * q.unsqueeze(0)).contiguous()
        del q, wrow

    def _electrode_basis(self, maps, perm):
        """Electrode-basis operators of every line, padded to common sizes:
        the shared right-hand side E (elimination order, n x Ke), per line the
        support rows of its electrodes and E_i^T on them, S_i, R_i and the
        line's source index per datum, and the line's rows of the solution."""
        M, dev, L = self.M, self.device, self.nlines
        off = M["node_off"]
        basis = [m.electrode_basis() for m in maps]
        ne = [E.shape[1] for E, _, _ in basis]
        Ke = self.K = max(ne)
        pos = np.empty(M["n"], np.int64)
        pos[perm] = np.arange(M["n"])
        er, ec, ev = [], [], []
        sups, esel = [], []
        for i, (E, _, _) in enumerate(basis):
            c = sp.coo_matrix(E)
            prow = pos[c.row + off[i]]
            er.append(prow); ec.append(c.col); ev.append(c.data)
            sup = np.unique(prow)
            sups.append(sup)
            Es = sp.csr_matrix((c.data, (np.searchsorted(sup, prow), c.col)),
                               shape=(sup.size, Ke)).toarray().T
            esel.append(Es)
        Ep = sp.csr_matrix((np.concatenate(ev), (np.concatenate(er), np.concatenate(ec))),
                           shape=(M["n"], Ke))
        self.E = torch.tensor(Ep.toarray(), dtype=torch.float64, device=dev)
        smax = max(s.size for s in sups)
        sup_pad = np.zeros((L, smax), np.int64)
        esel_pad = np.zeros((L, Ke, smax))
        for i, (sup, Es) in enumerate(zip(sups, esel)):
            sup_pad[i, :sup.size] = sup
            sup_pad[i, sup.size:] = sup[0]
            esel_pad[i, :, :sup.size] = Es
        self.e_sup = torch.tensor(sup_pad.reshape(-1), device=dev)
        self.e_sel = torch.tensor(esel_pad, dtype=torch.float64, device=dev)
        nS_max, nD_max = max(self.nS), max(self.nD)
        S_pad = np.zeros((L, Ke, nS_max))
        R_pad = np.zeros((L, Ke, nD_max))
        sidx_pad = np.zeros((L, nD_max), np.int64)
        dmask = np.zeros((L, nD_max), bool)
        for i, ((_, S, R), m) in enumerate(zip(basis, maps)):
            S_pad[i, :S.shape[0], :S.shape[1]] = S.toarray()
            R_pad[i, :R.shape[0], :R.shape[1]] = R.toarray()
            sidx_pad[i, :m.nD] = m.sidx
            dmask[i, :m.nD] = True
        self.S_pad = torch.tensor(S_pad, dtype=torch.float64, device=dev)
        self.R_pad = torch.tensor(R_pad, dtype=torch.float64, device=dev)
        self.sidx_pad = torch.tensor(sidx_pad, device=dev).unsqueeze(1).expand(L, Ke, nD_max)
        # flat positions of the real data in the padded (line, datum) block:
        # integer indexing, where a boolean mask would synchronise the host
        self.dflat = torch.tensor(np.flatnonzero(dmask), dtype=torch.int64, device=dev)
        self.wq_lines = torch.tensor(M["wq"], dtype=torch.float64, device=dev)  # (nbatch, L)
        # each line's rows in elimination order: one row segment per line
        # (dcx_rows_mix) when every line's rows are contiguous, else gathers
        self.line_rows = []
        for i in range(L):
            r = np.sort(pos[off[i]:off[i + 1]])
            if r[-1] - r[0] + 1 == r.size:
                self.line_rows.append(slice(int(r[0]), int(r[-1]) + 1))
            else:
                self.line_rows.append(torch.tensor(r, device=dev))
        self._mix_segs = None
        if all(isinstance(r, slice) for r in self.line_rows):
            order = np.argsort([r.start for r in self.line_rows])
            starts = [self.line_rows[i].start for i in order]
            stops = [self.line_rows[i].stop for i in order]
            if L <= 256 and Ke <= 128 and starts[0] == 0 and stops[-1] == M["n"] and all(
                    a == b for a, b in zip(stops[:-1], starts[1:])):
                self._mix_order = torch.tensor(order, device=dev)
                self._mix_segs = np.ascontiguousarray(np.r_[starts, stops[-1]], np.int32)

    def _active_map(self, active_cells, val_inactive):
        """Indices of the active cells in the stacked full model and the fill
        of the inactive ones (None when every line is full-mesh)."""
        L, co = self.nlines, self.M["cell_off"]
        self._act = None
        self.n_model = list(self.nC)
        if active_cells is None or all(a is None for a in active_cells):
            return
        if len(active_cells) != L:
            raise ValueError("active_cells needs one entry (None or a mask) per line")
        vals = list(val_inactive) if np.iterable(val_inactive) else [val_inactive] * L
        fill = np.zeros(int(co[-1]))
        idx = []
        for i, a in enumerate(active_cells):
            if a is None:
                idx.append(np.arange(co[i], co[i + 1]))
                continue
            a = np.asarray(a, dtype=bool).reshape(-1)
            if a.size != self.nC[i]:
                raise ValueError(f"active_cells[{i}] has {a.size} entries, line has {self.nC[i]} cells")
            if vals[i] is None:
                raise ValueError(f"val_inactive is needed for line {i}")
            fill[co[i]:co[i + 1]][~a] = float(vals[i])
            idx.append(co[i] + np.where(a)[0])
            self.n_model[i] = int(a.sum())
        self._act = torch.tensor(np.concatenate(idx), dtype=torch.int64, device=self.device)
        self._fill = torch.tensor(fill, dtype=torch.float64, device=self.device)
        self._moff = np.r_[0, np.cumsum(self.n_model)]

    def _full_model(self, m):
        """Stacked full-mesh model (log-resistivity) from the stacked models
        of the lines' active cells (`_model_vector`)."""
        if self._act is None:
            return m
        if m.numel() != self._act.numel():
            raise ValueError(f"models hold {m.numel()} values, the active cells {self._act.numel()}")
        full = self._fill.clone()
        full[self._act] = m
        return full

    def _per_line(self, phi, grad):
        """[(phi_i, grad_i)], the gradient restricted to each line's model."""
        if self._act is None:
            co = self.M["cell_off"]
        else:
            grad, co = grad[self._act], self._moff
        return [(phi[i], grad[co[i]:co[i + 1]]) for i in range(self.nlines)]

    def _stream(self):
        return self._libm.stream_of(self.device)

    def autotune_tiles(self, ms, reps=3, grid=None):
        """`tune_solver_tiles` on this plan's eval+grad at the models `ms`
        (set_data first); the plan keeps the best setting."""
        if self._data is None:
            raise RuntimeError("set_data first")
        return tune_solver_tiles(self.solver, lambda: self.eval_grad(ms), reps=reps, grid=grid,
                                 device=self.device)

    def set_data(self, dobs, w):
        """Observed data and weights per line (stacked on the device)."""
        cat = lambda xs: torch.cat([torch.as_tensor(x, dtype=torch.float64).reshape(-1).to(
            self.device) for x in xs])
        self._data = (cat(dobs), cat(w))
        self._graph = None  # a captured graph reads the data it was recorded with

    def capture(self, ms, warmup=2):
        """Record the eval+grad of every line at the models `ms` as ONE HIP
        graph (through `torch.cuda.CUDAGraph`, after `warmup` eager runs on
        its stream that size every workspace the plan keeps); later
        `eval_grad` calls copy their models into the graph's input tensor and
        replay it.  FP64 plans only, as `LineBatch.capture` (a replay of
        mixed-precision graphs ended a process inside hipGraphLaunch, cause
        not found).  The graph takes the stacked model from a device tensor:
        the host-side stacking of `ms` stays outside it (a captured copy from
        pageable host memory would re-read that host buffer at every
        replay).  The solve checks do not run inside a captured graph."""
        if self.precision != "fp64":
            raise NotImplementedError(
                "graph capture is limited to FP64 plans: a replay of mixed-precision graphs "
                "ended in a host segmentation fault inside hipGraphLaunch (cause not found)")
        if self._data is None:
            raise RuntimeError("set_data first")
        dev = self.device
        self._graph = None
        m_in = self._model_vector(ms).clone()
        s = torch.cuda.Stream(device=dev)  # default priority
        s.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self._eval_grad_vec(m_in)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            out = self._eval_grad_vec(m_in)
        torch.cuda.synchronize(dev)
        self._graph = (g, m_in, out, self.solver.info()["buffer_generation"])

    def _model_vector(self, ms):
        """The lines' models (active cells) stacked on the device."""
        return torch.cat([torch.as_tensor(x, dtype=torch.float64).reshape(-1).to(self.device)
                          for x in ms])

    def eval_grad(self, ms):
        """[(phi_i, grad_i)] for the models `ms` (one per line, m = log rho on
        the line's active cells):
        phi_i = ||w_i (dpred_i - dobs_i)||^2 and its gradient, all lines in
        one launch sequence on the current stream, no host synchronisation.
        After `capture` it replays the graph: phi_i and grad_i are then views
        of the graph's output buffers, overwritten by the next call."""
        if self._data is None:
            raise RuntimeError("set_data first")
        m = self._model_vector(ms)
        graph = getattr(self, "_graph", None)
        if graph is not None:
            g, m_in, out, gen = graph
            if self.solver.info()["buffer_generation"] != gen:
                self._graph = None
                raise RuntimeError("the plan reallocated its workspace after capture(): the "
                                   "recorded graph is dropped; call capture() again")
            if m.numel() != m_in.numel():
                raise ValueError(f"models hold {m.numel()} values, the graph {m_in.numel()}")
            with torch.no_grad():
                m_in.copy_(m)
            g.replay()  # on torch's current stream, after the copy
            return self._per_line(*out)
        return self._per_line(*self._eval_grad_vec(m))

    def _eval_grad_vec(self, m):
        """(phi per line, full-mesh gradient) for the stacked active model m."""
        if self.rhs_basis == "electrodes":
            return self._eval_grad_electrodes(m)
        dobs, w = self._data
        L, lib, sv, dev = self._libm, self.lib, self.solver, self.device
        f32 = self.precision == "mixed"
        fl32 = L.DCX_F32_FIELDS if f32 else 0
        n, K, nb = sv.n, self.K, self.nbatch
        st = self._stream()
        with torch.no_grad(), torch.cuda.device(dev):
            sig = torch.exp(-self._full_model(m))
            vals = torch.empty(nb, self.M["nnz"], dtype=torch.float64, device=dev)
            L.check(lib.dcx_assemble(self.lmap, L.ptr(sig), L.ptr(vals), st), "assemble")
            U = torch.empty(nb, n, K, dtype=torch.float32 if f32 else torch.float64, device=dev)
            # the solves' checks are deferred and waited for where their results
            # are used; systems refactorised in FP64 have their results redone
            with sv._lock:
                sv.run_factor_solve(vals, self.B, U, K, L.DCX_PERMUTED | L.DCX_SPARSE_RHS | fl32,
                                    defer=True)
            d = torch.empty(self.M["nD"], dtype=torch.float64, device=dev)
            G = torch.empty(n, K, dtype=torch.float64, device=dev)

            def project():
                L.check(lib.dcx_survey_dpred(self.survey, L.ptr(U), int(K), nb, L.ptr(self.ones),
                                             0, int(K), L.ptr(d), L.DCX_PERMUTED | fl32, st),
                        "survey_dpred")
                r = w * (d - dobs)
                gd = ((2.0 * r) * w).contiguous()
                L.check(lib.dcx_survey_adjoint_rhs(self.survey, L.ptr(gd), L.ptr(G), int(K), 0,
                                                   int(K), L.DCX_PERMUTED, st), "adjoint_rhs")
                return r
            r = project()
            if sv.check_wait():
                r = project()
            phi = torch.zeros(self.nlines, dtype=torch.float64, device=dev)
            phi.index_add_(0, self.data_line, r ** 2)
            Lam = torch.empty_like(U)
            with sv._lock:
                sv.run_solve(G, Lam, K,
                             L.DCX_PERMUTED | L.DCX_SHARED_RHS | L.DCX_SPARSE_RHS | fl32,
                             defer=True)
            gv = sv.grad_values(Lam, U, scale=self.ones, permuted=True)
            if sv.check_wait():
                gv = sv.grad_values(Lam, U, scale=self.ones, permuted=True)
            grad = torch.empty_like(sig)
            L.check(lib.dcx_lmap_transpose(self.lmap, L.ptr(gv), L.ptr(sig), L.ptr(grad), st),
                    "lmap_transpose")
        return phi, grad

    def _eval_grad_electrodes(self, m):
        """eval_grad in the electrode basis: one solve, no adjoint solve."""
        dobs, w = self._data
        L, lib, sv, dev = self._libm, self.lib, self.solver, self.device
        f32 = self.precision == "mixed"
        fl32 = L.DCX_F32_FIELDS if f32 else 0
        n, K, nb, nl = sv.n, self.K, self.nbatch, self.nlines
        st = self._stream()
        with torch.no_grad(), torch.cuda.device(dev):
            sig = torch.exp(-self._full_model(m))
            vals = torch.empty(nb, self.M["nnz"], dtype=torch.float64, device=dev)
            L.check(lib.dcx_assemble(self.lmap, L.ptr(sig), L.ptr(vals), st), "assemble")
            G = torch.empty(nb, n, K, dtype=torch.float32 if f32 else torch.float64, device=dev)
            with sv._lock:
                sv.run_factor_solve(vals, self.E, G, K, L.DCX_PERMUTED | L.DCX_SHARED_RHS |
                                    L.DCX_SPARSE_RHS | fl32, defer=True)
            phi, gv = self._electrode_rest(G, dobs, w)
            if sv.check_wait():     # systems of G were solved again in FP64
                phi, gv = self._electrode_rest(G, dobs, w)
            grad = torch.empty_like(sig)
            L.check(lib.dcx_lmap_transpose(self.lmap, L.ptr(gv), L.ptr(sig), L.ptr(grad), st),
                    "lmap_transpose")
        return phi, grad

    def _electrode_rest(self, G, dobs, w):
        """Everything of the electrode-basis eval+grad after the solve G:
        per-line phi and the value gradient gv."""
        L, lib, sv, dev = self._libm, self.lib, self.solver, self.device
        fl32 = L.DCX_F32_FIELDS if G.dtype == torch.float32 else 0
        n, K, nb, nl = sv.n, self.K, self.nbatch, self.nlines
        st = self._stream()
        with torch.no_grad(), torch.cuda.device(dev):
            # K_i = sum_k w_(k,i) E_i^T G_k on each line's electrode rows
            Gs = G[:, self.e_sup, :].to(torch.float64).view(nb, nl, -1, K)
            Ke = torch.matmul(self.e_sel, Gs)                      # (nb, nl, K, K)
            Kl = (self.wq_lines.view(nb, nl, 1, 1) * Ke).sum(0)    # (nl, K, K)
            KS = torch.matmul(Kl, self.S_pad)                      # (nl, K, nS_max)
            dpad = (self.R_pad * torch.gather(KS, 2, self.sidx_pad)).sum(1)
            d = dpad.reshape(-1)[self.dflat]
            r = w * (d - dobs)
            phi = torch.zeros(nl, dtype=torch.float64, device=dev)
            phi.index_add_(0, self.data_line, r ** 2)
            vpad = torch.zeros_like(dpad)
            vpad.view(-1)[self.dflat] = (2.0 * r) * w
            T = torch.zeros(nl, K, self.S_pad.shape[2], dtype=torch.float64, device=dev)
            T.scatter_add_(2, self.sidx_pad, self.R_pad * vpad.unsqueeze(1))
            Mt = torch.matmul(self.S_pad, T.transpose(1, 2))       # M_i^T = S_i T_i^T
            Mw = (self.wq_lines.view(nb, nl, 1, 1) * Mt.unsqueeze(0)).to(G.dtype)
            H = torch.empty_like(G)
            if self._mix_segs is not None:
                # one launch: each line's rows times its own matrix
                Wseg = Mw[:, self._mix_order].contiguous()
                L.check(lib.dcx_rows_mix(L.ptr(G), L.ptr(Wseg), L.ptr(H), nb, n, int(K), nl,
                                         self._mix_segs.ctypes.data_as(L._i32p), fl32, st),
                        "rows_mix")
            else:
                for i, rows in enumerate(self.line_rows):
                    H[:, rows, :] = torch.matmul(G[:, rows, :], Mw[:, i])
            gv = sv.grad_values(G, H, scale=self.ones, permuted=True)
        return phi, gv

    def free(self):
        if getattr(self, "lmap", None):
            self.lib.dcx_lmap_destroy(self.lmap)
            self.lmap = None
        if getattr(self, "survey", None):
            self.lib.dcx_survey_destroy(self.survey)
            self.survey = None
        self.solver.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
