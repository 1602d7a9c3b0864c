"""Uneven batch of independent lines (simpeg_dc_torch_amd/lines.py).

CPU: the workload generator's sizes and electrode placement.  GPU: the
stream-concurrent batch (one issuing host thread, or one per stream) against
each line's own eval+grad on the default stream, the captured graphs (one per
line, and one for every line of the one-plan batch) against eager runs at
several models with eager work and synchronisations between the replays
(every replay checked), and the leaf autotuner's contract.
"""
import numpy as np
import pytest
import torch

from helpers import case, noisy_data


def _lines(n=4, cells=(3000, 8000), seed=5):
    from simpeg_dc_torch_amd import workloads
    return workloads.uneven_lines_case(seed=seed, nlines=n, cells=cells)


def test_uneven_lines_workload():
    L = _lines(16, (3000, 40000))
    sizes = np.array([mesh.nC for mesh, _, _ in L])
    assert sizes.min() >= 2500 and sizes.max() <= 45000 and len(set(sizes)) > 8
    for mesh, survey, m in L:
        xa = np.array([s.location_a[0] for s in survey.source_list])
        assert np.all(np.isin(np.round(xa, 9), np.round(mesh.nodes_x, 9)))
        pos = set()
        for s in survey.source_list:
            pos |= {round(float(s.location_a[0]), 9), round(float(s.location_b[0]), 9)}
            for r in s.receiver_list:
                pos |= {round(float(v), 9) for v in np.r_[r.locations_m[:, 0], r.locations_n[:, 0]]}
        ne = len(pos)
        assert 21 <= ne <= 96 and m.size == mesh.nC


def _setup(n=4, nky=6, precision="fp64", threads=False):
    from simpeg_dc_torch_amd import TorchDC2D
    from simpeg_dc_torch_amd.lines import LineBatch
    L = _lines(n)
    data = []
    for mesh, survey, m in L:
        eng = TorchDC2D(mesh, survey, nky=nky, device="cuda")
        with torch.no_grad():
            d = eng.dpred(torch.tensor(m, device="cuda")).cpu().numpy()
        data.append(noisy_data(d, seed=9))
    lb = LineBatch([(mesh, survey) for mesh, survey, _ in L], nky=nky, precision=precision,
                   streams=4, threads=threads)
    lb.set_data([d for d, _ in data], [w for _, w in data])
    return L, data, lb


def _single(eng, m, dobs, w):
    x = torch.tensor(m, device="cuda", requires_grad=True)
    phi = eng.misfit(x, torch.tensor(dobs, device="cuda"), torch.tensor(w, device="cuda"))
    phi.backward()
    return float(phi.detach()), x.grad.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("threads", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_line_batch_streams_match_single(precision, threads):
    """Every line of the batch (4 streams, issued from one host thread or
    from one per stream) equals its own autograd eval+grad on the default
    stream; twice in a row, the second at other models."""
    print(f"line batch streams {precision} threads={threads}", flush=True)
    L, data, lb = _setup(precision=precision, threads=threads)
    rng = np.random.default_rng(4)
    for rep in range(2):
        ms = [m + 0.1 * rep * rng.standard_normal(m.size) for _, _, m in L]
        out = lb.eval_grad(ms)
        torch.cuda.synchronize()
        for i in range(len(L)):
            phi, g = _single(lb.engines[i], ms[i], *data[i])
            assert float(out[i][0]) == pytest.approx(phi, rel=1e-13), (rep, i)
            assert np.abs(out[i][1].cpu().numpy() - g).max() <= 1e-13 * np.abs(g).max(), (rep, i)


@pytest.mark.gpu
def test_line_batch_electrode_basis_matches_sources():
    """Engine keywords reach every line: a batch built with
    rhs_basis="electrodes" gives each line's misfit and gradient of the
    source-basis batch (FP64; the two bases solve different right-hand sides,
    so agreement is to round-off, not bit-exact)."""
    from simpeg_dc_torch_amd.lines import LineBatch
    print("line batch electrode basis fp64", flush=True)
    L, data, lb = _setup(precision="fp64")
    lbe = LineBatch([(mesh, survey) for mesh, survey, _ in L], nky=6, precision="fp64",
                    streams=4, rhs_basis="electrodes")
    lbe.set_data([d for d, _ in data], [w for _, w in data])
    assert all(e.rhs_basis == "electrodes" for e in lbe.engines)
    ms = [m for _, _, m in L]
    a, b = lbe.eval_grad(ms), lb.eval_grad(ms)
    torch.cuda.synchronize()
    for i in range(len(L)):
        assert float(a[i][0]) == pytest.approx(float(b[i][0]), rel=1e-10), i
        g = b[i][1].cpu().numpy()
        assert np.abs(a[i][1].cpu().numpy() - g).max() <= 1e-9 * np.abs(g).max(), i


_MIXED_GRAPHS = pytest.mark.skip(
    reason="capture refuses mixed-precision engines: a replay of their graphs ended the "
           "process with a host segmentation fault inside hipGraphLaunch, cause not found")


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", pytest.param("mixed", marks=_MIXED_GRAPHS)])
def test_line_batch_graph_replays(precision):
    """Capture at one model, then replay at three models, each replay
    preceded by eager eval+grad work and a synchronisation; every replay
    equals the eager result at its model."""
    print(f"line batch graphs {precision}", flush=True)
    L, data, lb = _setup(precision=precision)
    rng = np.random.default_rng(1)
    ms0 = [m for _, _, m in L]
    lb.capture(ms0)
    for rep in range(3):
        ms = [m + 0.1 * rep * rng.standard_normal(m.size) for m in ms0]
        ref = [_single(lb.engines[i], ms[i], *data[i]) for i in range(len(L))]
        torch.cuda.synchronize()
        out = lb.eval_grad(ms)
        torch.cuda.synchronize()
        for i in range(len(L)):
            assert float(out[i][0]) == pytest.approx(ref[i][0], rel=1e-13), (rep, i)
            g = out[i][1].cpu().numpy()
            assert np.abs(g - ref[i][1]).max() <= 1e-13 * np.abs(ref[i][1]).max(), (rep, i)


@pytest.mark.gpu
def test_graphs_dropped_when_a_plan_reallocates():
    """A wider solve after capture (an explicit Jacobian in blocks wider
    than the sources) reallocates plan workspace: the next eval_grad refuses
    to replay (the graphs hold the old addresses) instead of replaying, and a
    new capture replays correctly again."""
    print("graphs dropped on reallocation", flush=True)
    L, data, lb = _setup(precision="fp64")
    ms0 = [m for _, _, m in L]
    lb.capture(ms0)
    eng = lb.engines[0]
    gen0 = eng._gpu.solver.info()["buffer_generation"]
    eng.linearize(torch.tensor(ms0[0], device="cuda")).Jmatrix(sub=eng.nD)
    torch.cuda.synchronize()
    assert eng._gpu.solver.info()["buffer_generation"] > gen0
    with pytest.raises(RuntimeError, match="capture"):
        lb.eval_grad(ms0)
    assert lb._graphs is None
    lb.capture(ms0)
    ref = [_single(lb.engines[i], ms0[i], *data[i]) for i in range(len(L))]
    torch.cuda.synchronize()
    out = lb.eval_grad(ms0)
    torch.cuda.synchronize()
    for i in range(len(L)):
        assert float(out[i][0]) == pytest.approx(ref[i][0], rel=1e-13), i


@pytest.mark.gpu
def test_capture_refuses_mixed_engines():
    from simpeg_dc_torch_amd.lines import LineBatch
    print("capture refuses mixed", flush=True)
    L = _lines(1)
    lb = LineBatch([(mesh, survey) for mesh, survey, _ in L], nky=6, precision="mixed")
    with pytest.raises(NotImplementedError):
        lb.capture([m for _, _, m in L])


@pytest.mark.gpu
def test_autotune_leaf_contract():
    from simpeg_dc_torch_amd.lines import LineBatch, autotune_leaf, size_class
    print("autotune leaf", flush=True)
    L = _lines(2)
    mesh, survey, m = L[0]
    from simpeg_dc_torch_amd import TorchDC2D
    eng = TorchDC2D(mesh, survey, nky=6, device="cuda")
    with torch.no_grad():
        d = eng.dpred(torch.tensor(m, device="cuda")).cpu().numpy()
    dobs, w = noisy_data(d, seed=2)
    best, times = autotune_leaf(mesh, survey, m, dobs, w, candidates=(32, 48), reps=2, nky=6)
    assert best in (32, 48) and set(times) == {32, 48} and times[best] == min(times.values())
    lb = LineBatch([(mesh, survey)], nky=6, leaf="auto", tune_data=[(m, dobs, w)])
    assert size_class(mesh.nC) in lb.leaf and lb.leaf[size_class(mesh.nC)] in (32, 48, 64)


def _small_lines():
    """Three small lines of 2.3k-5.2k cells with 18, 26 and 23 sources and
    6, 4 and 5 wavenumbers, so the merge pads both source columns and
    systems."""
    L = _lines(3, (2000, 6000), seed=11)
    return L, [6, 4, 5]


def test_merge_line_maps_host():
    """The merged problem, solved on the host (scipy, per system), gives each
    line's data from its own CPU engine: the merged pattern, grouped
    wavenumbers, boundary terms, padded systems, weighted right-hand sides,
    survey rows and source columns all land where they belong.  The merged
    values equal the lines' own values block by block."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from simpeg_dc_torch_amd import TorchDC2D
    from simpeg_dc_torch_amd.lines import merge_line_maps, merged_values_host
    L, nkys = _small_lines()
    engs = [TorchDC2D(mesh, survey, nky=nk, device="cpu") for (mesh, survey, _), nk in zip(L, nkys)]
    M = merge_line_maps([e._maps for e in engs], [e.kys for e in engs], [e.wq_np for e in engs])
    assert M["nbatch"] == max(e.nky for e in engs) and M["K"] == max(e.nS for e in engs)
    m = np.concatenate([mm for _, _, mm in L])
    V = merged_values_host(M, np.exp(-m))
    for i, e in enumerate(engs):
        own = e._maps.values(np.exp(-L[i][2]))
        blk = V[:, M["ent_off"][i]:M["ent_off"][i + 1]]
        assert np.array_equal(blk[:own.shape[0]], own)
        assert np.array_equal(blk[own.shape[0]:], np.repeat(own[-1:], M["nbatch"] - own.shape[0], 0))
    rows = np.repeat(np.arange(M["n"]), np.diff(M["crow"]))
    q = M["q"].toarray()
    d = np.zeros(M["nD"])
    for k in range(M["nbatch"]):
        A = sp.csc_matrix((V[k], (rows, M["col"])), shape=(M["n"], M["n"]))
        Bk = M["wq"][k, M["node_line"]][:, None] * q
        Uk = spla.spsolve(A, Bk).reshape(M["n"], M["K"])
        d += (M["Pd"] @ Uk)[np.arange(M["nD"]), M["sidx"]]
    for i, e in enumerate(engs):
        with torch.no_grad():
            di = e.dpred(torch.tensor(L[i][2])).numpy()
        got = d[M["data_off"][i]:M["data_off"][i + 1]]
        assert np.abs(got - di).max() <= 1e-10 * np.abs(di).max(), i


def test_merged_lines_order_component_by_component():
    """The host symbolic analysis of a merged batch (the ordering the plan
    builds) keeps each line's rows contiguous in the elimination order, keeps
    every supernode and its parent inside one line (the elimination tree is a
    forest of the lines' own trees), and has one root per line: the layout
    the one-plan batch relies on for one row-mix launch per step."""
    from simpeg_dc_torch_amd import TorchDC2D, _lib
    from simpeg_dc_torch_amd.lines import merge_line_maps
    from simpeg_dc_torch_amd.solver import DEFAULT_LEAF
    for L in (_small_lines()[0], _lines(6, (2000, 9000), seed=2)):
        engs = [TorchDC2D(mesh, survey, nky=4, device="cpu") for mesh, survey, _ in L]
        M = merge_line_maps([e._maps for e in engs], [e.kys for e in engs],
                            [e.wq_np for e in engs])
        sym = _lib.symbolic(M["crow"], M["col"], coords=M["coords"], leaf_max=DEFAULT_LEAF)
        pos = np.empty(M["n"], np.int64)
        pos[sym["perm"]] = np.arange(M["n"])
        line_of_pos = M["node_line"][sym["perm"]]
        for i in range(len(L)):
            r = np.sort(pos[M["node_off"][i]:M["node_off"][i + 1]])
            assert r[-1] - r[0] + 1 == r.size, i
        sup_line = line_of_pos[sym["first"]]
        for s_, (f, ns) in enumerate(zip(sym["first"], sym["ns"])):
            assert np.all(line_of_pos[f:f + ns] == sup_line[s_]), s_
            p_ = sym["parent"][s_]
            if p_ >= 0:
                assert sup_line[p_] == sup_line[s_], s_
        roots = sup_line[sym["parent"] < 0]
        assert sorted(roots.tolist()) == list(range(len(L)))


# mixed: with 3% noise the residual is ~0.03 |d|, so a relative data error e
# moves phi by ~2 e / 0.03; the 1e-6 data bar allows phi to move by ~7e-5
@pytest.mark.gpu
@pytest.mark.parametrize("basis", ["sources", "electrodes"])
@pytest.mark.parametrize("precision,tol_phi,tol_g", [("fp64", 1e-11, 1e-10),
                                                     ("mixed", 1e-4, 1e-3)])
def test_line_forest_matches_lines(precision, tol_phi, tol_g, basis):
    """All lines in one plan (source or electrode basis) against each line's
    own FP64 engine in the source basis (autograd eval+grad): three small
    lines with 6, 4 and 5 wavenumbers (two padded), then the four-line batch
    of 3k-8k cells; twice in a row, the second at other models."""
    from simpeg_dc_torch_amd import TorchDC2D
    from simpeg_dc_torch_amd.lines import LineForest
    for name, (L, nkys) in (("small", _small_lines()), ("batch4", (_lines(4), [6] * 4))):
        print(f"line forest {precision} {basis} {name}", flush=True)
        engs = [TorchDC2D(mesh, survey, nky=nk, device="cuda")
                for (mesh, survey, _), nk in zip(L, nkys)]
        data = []
        for e, (_, _, m) in zip(engs, L):
            with torch.no_grad():
                d = e.dpred(torch.tensor(m, device="cuda")).cpu().numpy()
            data.append(noisy_data(d, seed=9))
        lf = LineForest([(mesh, survey) for mesh, survey, _ in L], nky=nkys, precision=precision,
                        rhs_basis=basis)
        lf.set_data([d for d, _ in data], [w for _, w in data])
        rng = np.random.default_rng(6)
        for rep in range(2):
            ms = [m + 0.1 * rep * rng.standard_normal(m.size) for _, _, m in L]
            out = lf.eval_grad(ms)
            torch.cuda.synchronize()
            for i, e in enumerate(engs):
                phi, g = _single(e, ms[i], *data[i])
                assert float(out[i][0]) == pytest.approx(phi, rel=tol_phi), (name, rep, i)
                gi = out[i][1].cpu().numpy()
                assert np.linalg.norm(gi - g) <= tol_g * np.linalg.norm(g), (name, rep, i)
        lf.free()


@pytest.mark.gpu
@pytest.mark.parametrize("basis", ["sources", "electrodes"])
def test_line_forest_graph_replays(basis):
    """Every line's eval+grad as one captured graph (FP64 plan): captured at
    one model, replayed at three, each replay preceded by an eager eval+grad
    of the same plan at another model and a synchronisation; every replay
    equals the eager result at its model.  A mixed-precision plan's capture
    is refused."""
    from simpeg_dc_torch_amd import TorchDC2D
    from simpeg_dc_torch_amd.lines import LineForest
    print(f"line forest graphs {basis}", flush=True)
    L, nkys = _lines(4), [6] * 4
    data = []
    for mesh, survey, m in L:
        e = TorchDC2D(mesh, survey, nky=6, device="cuda")
        with torch.no_grad():
            d = e.dpred(torch.tensor(m, device="cuda")).cpu().numpy()
        data.append(noisy_data(d, seed=9))
    lf = LineForest([(mesh, survey) for mesh, survey, _ in L], nky=nkys, precision="fp64",
                    rhs_basis=basis)
    lf.set_data([d for d, _ in data], [w for _, w in data])
    rng = np.random.default_rng(2)
    ms0 = [m for _, _, m in L]
    lf.capture(ms0)

    def eager(ms):
        phi, grad = lf._eval_grad_vec(lf._model_vector(ms))
        return [(float(p), g.cpu().numpy()) for p, g in lf._per_line(phi, grad)]
    for rep in range(3):
        ms = [m + 0.1 * (rep + 1) * rng.standard_normal(m.size) for m in ms0]
        other = [m - 0.05 * rng.standard_normal(m.size) for m in ms0]
        eager(other)                      # eager work on the same plan between replays
        ref = eager(ms)
        torch.cuda.synchronize()
        out = lf.eval_grad(ms)
        torch.cuda.synchronize()
        for i in range(len(L)):
            assert float(out[i][0]) == pytest.approx(ref[i][0], rel=1e-13), (rep, i)
            g = out[i][1].cpu().numpy()
            assert np.abs(g - ref[i][1]).max() <= 1e-13 * np.abs(ref[i][1]).max(), (rep, i)
    lf.free()
    lm = LineForest([(mesh, survey) for mesh, survey, _ in L[:2]], nky=nkys[:2],
                    precision="mixed", rhs_basis=basis)
    lm.set_data([d for d, _ in data[:2]], [w for _, w in data[:2]])
    with pytest.raises(NotImplementedError):
        lm.capture(ms0[:2])
    lm.free()


@pytest.mark.gpu
@pytest.mark.parametrize("basis", ["sources", "electrodes"])
def test_line_forest_active_cells(basis):
    """Lines with inactive cells (two of three lines masked, one full-mesh,
    per-line fill values) in one FP64 plan against each line's own engine
    with the same mask: misfit and active-cell gradient; wrong model sizes
    and a mask without its fill are refused."""
    from simpeg_dc_torch_amd import TorchDC2D
    from simpeg_dc_torch_amd.lines import LineForest
    print(f"line forest active cells {basis}", flush=True)
    L, nkys = _small_lines()
    rng = np.random.default_rng(3)
    masks = [rng.random(L[0][0].nC) > 0.15, None, rng.random(L[2][0].nC) > 0.3]
    fills = [np.log(300.0), None, np.log(30.0)]
    engs, data, ms = [], [], []
    for (mesh, survey, m), nk, a, v in zip(L, nkys, masks, fills):
        e = TorchDC2D(mesh, survey, nky=nk, device="cuda", active_cells=a, val_inactive=v)
        mi = m if a is None else m[a]
        with torch.no_grad():
            d = e.dpred(torch.tensor(mi, device="cuda")).cpu().numpy()
        engs.append(e)
        data.append(noisy_data(d, seed=4))
        ms.append(mi + 0.05 * rng.standard_normal(mi.size))
    lf = LineForest([(mesh, survey) for mesh, survey, _ in L], nky=nkys, precision="fp64",
                    rhs_basis=basis, active_cells=masks, val_inactive=fills)
    assert lf.n_model == [m.size for m in ms]
    lf.set_data([d for d, _ in data], [w for _, w in data])
    out = lf.eval_grad(ms)
    torch.cuda.synchronize()
    for i, e in enumerate(engs):
        phi, g = _single(e, ms[i], *data[i])
        assert float(out[i][0]) == pytest.approx(phi, rel=1e-11), i
        gi = out[i][1].cpu().numpy()
        assert gi.shape == g.shape and np.linalg.norm(gi - g) <= 1e-10 * np.linalg.norm(g), i
    with pytest.raises(ValueError):
        lf.eval_grad([L[0][2], ms[1], ms[2]])
    lf.free()
    with pytest.raises(ValueError):
        LineForest([(mesh, survey) for mesh, survey, _ in L], nky=nkys, precision="fp64",
                   active_cells=masks, val_inactive=[None, None, 1.0])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("k", [1, 17, 40, 84, 128])
def test_rows_mix_matches_per_segment_products(dtype, k):
    """dcx_rows_mix against torch per segment: one and two column groups
    (k <= 64, k > 64), k not a multiple of 4, segments longer and shorter
    than a workgroup's 64 rows, an empty segment, and rows outside every
    segment left as they were."""
    from simpeg_dc_torch_amd import _lib
    print(f"rows mix {dtype} k={k}", flush=True)
    lib = _lib.load()
    nb, n = 3, 1000
    segs = np.array([5, 6, 200, 200, 333, 900, 997], np.int32)
    nseg = segs.size - 1
    gen = torch.Generator(device="cpu").manual_seed(k)
    G = torch.randn(nb, n, k, generator=gen, dtype=torch.float64).to("cuda", dtype)
    W = torch.randn(nb, nseg, k, k, generator=gen, dtype=torch.float64).to("cuda", dtype)
    H = torch.full_like(G, 7.0)
    flags = _lib.DCX_F32_FIELDS if dtype == torch.float32 else 0
    _lib.check(lib.dcx_rows_mix(_lib.ptr(G), _lib.ptr(W), _lib.ptr(H), nb, n, k, nseg,
                                segs.ctypes.data_as(_lib._i32p), flags,
                                _lib.stream_of(G.device)), "rows_mix")
    torch.cuda.synchronize()
    ref = torch.full_like(G, 7.0, dtype=torch.float64)
    G64, W64 = G.double(), W.double()
    for l in range(nseg):
        a, b = int(segs[l]), int(segs[l + 1])
        ref[:, a:b] = torch.matmul(G64[:, a:b], W64[:, l])
    tol = 1e-5 if dtype == torch.float32 else 1e-13
    scale = max(1.0, float(np.sqrt(k)))
    assert torch.equal(H[:, :5], ref[:, :5].to(dtype)) and torch.equal(H[:, 997:], ref[:, 997:].to(dtype))
    assert float((H.double() - ref).abs().max()) <= tol * scale * 4


@pytest.mark.gpu
def test_tile_tuning_keeps_results_and_contract():
    """Every kernel-variant setting of the tuning grid gives the default
    setting's FP64 eval+grad (same kernels up to summation order: 1e-12);
    the tuner returns a setting of its grid, leaves it on the plan, and
    timed every setting it reports; bad values are refused."""
    from simpeg_dc_torch_amd import TorchDC2D
    from simpeg_dc_torch_amd.lines import TILE_GRID, autotune_tiles, tile_key
    print("tile tuning", flush=True)
    mesh, survey, m = _lines(1, (6000, 9000), seed=3)[0]
    eng = TorchDC2D(mesh, survey, nky=6, device="cuda")
    with torch.no_grad():
        d = eng.dpred(torch.tensor(m, device="cuda")).cpu().numpy()
    dobs, w = noisy_data(d, seed=4)
    args = [torch.tensor(a, device="cuda") for a in (m, dobs, w)]
    phi0, g0 = eng.eval_grad_device(*args)
    sv = eng._gpu.solver
    default = sv.tuning()
    for name, values in TILE_GRID.items():
        for v in values:
            sv.set_tuning(**dict(default, **{name: v}))
            phi, g = eng.eval_grad_device(*args)
            assert float(phi) == pytest.approx(float(phi0), rel=1e-12), (name, v)
            assert float((g - g0).norm() / g0.norm()) < 1e-12, (name, v)
    sv.set_tuning(**default)
    best, times = autotune_tiles(eng, m, dobs, w, reps=1)
    assert sv.tuning() == best and tile_key(best) in times
    assert all(best[k] in TILE_GRID[k] or best[k] == default[k] for k in best)
    with pytest.raises(RuntimeError):
        sv.set_tuning(sweep_tc=48)


@pytest.mark.gpu
def test_line_forest_full_batch_fp64():
    """The full uneven workload (16 lines of 3k-31k cells, 251k cells, 18-81
    sources) in one FP64 plan against each line's own FP64 engine (the
    sync-free eval+grad), at 5% noise."""
    from simpeg_dc_torch_amd import TorchDC2D, workloads
    from simpeg_dc_torch_amd.lines import LineForest
    print("line forest full batch fp64", flush=True)
    L = workloads.uneven_lines_case()
    lines = [(mesh, survey) for mesh, survey, _ in L]
    lf = LineForest(lines, precision="fp64")
    data, refs = [], []
    for mesh, survey, m in L:
        eng = TorchDC2D(mesh, survey, nky=12, device="cuda")
        x = torch.tensor(m, device="cuda")
        with torch.no_grad():
            d = eng.dpred(x).cpu().numpy()
        dobs, w = noisy_data(d, seed=11, rel_noise=0.05)
        data.append((dobs, w))
        refs.append(eng.eval_grad_device(x, torch.tensor(dobs, device="cuda"),
                                         torch.tensor(w, device="cuda")))
        del eng
    lf.set_data([d for d, _ in data], [w for _, w in data])
    out = lf.eval_grad([m for _, _, m in L])
    torch.cuda.synchronize()
    for i, ((phi, g), (phi_r, g_r)) in enumerate(zip(out, refs)):
        assert float(phi) == pytest.approx(float(phi_r), rel=1e-11), i
        assert float((g - g_r).norm() / g_r.norm()) < 1e-10, i
    lf.free()


@pytest.mark.gpu
def test_grouped_map_arguments_and_refusals():
    """dcx_lmap_set_groups rejects out-of-range groups and keeps the map
    usable; a two-group table whose columns both hold the map's own
    wavenumbers leaves the assembled values unchanged; dcx_jacobian_rows
    refuses a map with more than one group (its kernel reads one wavenumber
    per system) before launching anything."""
    import ctypes as C
    from simpeg_dc_torch_amd import TorchDC2D, _lib
    print("grouped map arguments", flush=True)
    mesh, survey, m = case("suite")
    eng = TorchDC2D(mesh, survey, nky=6, device="cuda")
    g = eng._gpu
    sig = torch.exp(-torch.tensor(m, device="cuda"))
    v0 = g.assemble(sig)
    nnz = eng.nnzA
    bad, badp = _lib.i32(np.full(nnz, 3))
    ky2, ky2p = _lib.f64((eng.kys ** 2).reshape(-1, 1))
    assert g.lib.dcx_lmap_set_groups(g.lmap, 1, badp, ky2p) != 0
    ky22, ky22p = _lib.f64(np.repeat((eng.kys ** 2).reshape(-1, 1), 2, axis=1))
    grp, grpp = _lib.i32(np.arange(nnz) % 2)
    _lib.check(g.lib.dcx_lmap_set_groups(g.lmap, 2, grpp, ky22p), "set_groups")
    assert torch.equal(g.assemble(sig), v0)
    Y = torch.zeros(eng.nky, eng.nN, 1, dtype=torch.float64, device="cuda")
    U = torch.zeros(eng.nky, eng.nN, eng.nS, dtype=torch.float64, device="cuda")
    src = torch.zeros(1, dtype=torch.int32, device="cuda")
    J = torch.zeros(eng.nC, 1, dtype=torch.float64, device="cuda")
    rc = g.lib.dcx_jacobian_rows(g.lmap, g.solver._h, _lib.ptr(Y), 1, _lib.ptr(U), eng.nS,
                                 _lib.ptr(src), _lib.ptr(g.wq), _lib.ptr(sig), _lib.ptr(J),
                                 _lib.DCX_PERMUTED, g.stream())
    assert rc != 0 and b"group" in g.lib.dcx_last_error()
