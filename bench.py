"""eval+grad benchmark of the 2.5-D forward-and-adjoint path on one GPU.

Workload: the 25,088-cell line (200 x 100 core of 1 m cells plus 12 padding
cells growing 1.3x; 61 electrodes at 2 m; dipole-dipole a = 1..3, n = 1..6
plus Wenner-Schlumberger a = 1..2, n = 1..8: 849 sources, 1,636 data;
12 electrode-fitted wavenumbers requested, 10 carry weight; m = log(rho) a
Gaussian random field, seed 1).  One step = misfit(m).backward(): assembly,
factorisation of every wavenumber, forward solve of all sources, projection,
misfit, adjoint solve, gradient gather and transposed map -- all resident on
the GPU, inputs in HBM before the timed region.

Prints one JSON line.  `value` is the mixed-precision eval+grad time in ms
(FP32 factorisation, sweeps and fields; FP64 assembly, projection, sums and
the residual check of every solve, which refines or refactorises in FP64 a
wavenumber whose check fails; `--refine N` adds N fixed FP64-residual
refinement passes per solve).  Beside it: the FP64 time, the errors of the
mixed result against the FP64 one and the check's report, the roofline of
the dominant kernel (the backward sweep: its algorithmic bytes per launch
over its average launch time, from HIP events on the solve stream; HBM
traffic per launch from the committed rocprofv3 counter summary), and the
CPU baseline (the oracle's restatement of the reference CPU path, the whole
eval+grad at every wavenumber, median of 3 runs, sequential and with the
wavenumbers on threads), with the Gauss-Newton building blocks
(`linearization`), the 3-D engine on a 56 x 56 x 38-cell grid (`grid3d`) and
an uneven batch of 16 independent lines (`uneven_lines`).
`--summarize-trace` and `--summarize-pmc` turn rocprofv3 output of this
script into the tables kept under profiles/.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import torch  # noqa: E402

METRIC = "eval+grad ms (2.5-D 25k cells, 12 ky): mixed vs FP64 GPU at 1e-6 rel-err"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK = {"fp64": 78.6, "fp32": 157.3}  # dense TFLOP/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def tune_tiles(engines, mt, dobs, w):
    """Kernel-variant thresholds of each engine's plan by the tile autotuner
    (outside every timed region); {label: chosen setting} and the seconds
    spent."""
    from simpeg_dc_torch_amd.lines import autotune_tiles
    t0 = time.perf_counter()
    out = {label: autotune_tiles(eng, mt, dobs, w, reps=5)[0] for label, eng in engines.items()}
    return out, round(time.perf_counter() - t0, 2)


def build_engines(refine):
    from simpeg_dc_torch_amd import TorchDC2D, workloads
    mesh, survey, m, rng = workloads.line_25k_case()
    t0 = time.perf_counter()
    e64 = TorchDC2D(mesh, survey, nky=12, device="cuda:0", precision="fp64")
    t_build = time.perf_counter() - t0
    emx = TorchDC2D(mesh, survey, nky=12, device="cuda:0", precision="mixed", refine=refine)
    mt = torch.tensor(m, device="cuda:0")
    with torch.no_grad():
        d64 = e64.dpred(mt)
    noise = torch.tensor(rng.standard_normal(e64.nD), device="cuda:0")
    dobs = d64 * (1 + 0.05 * noise)
    w = 1.0 / (0.05 * dobs.abs())
    return mesh, survey, m, e64, emx, mt, dobs, w, t_build


def eval_grad(eng, mt, dobs, w):
    x = mt.detach().clone().requires_grad_(True)
    phi = eng.misfit(x, dobs, w)
    phi.backward()
    return phi.detach(), x.grad


def time_steps(eng, mt, dobs, w, steps, warmup):
    for _ in range(warmup):
        eval_grad(eng, mt, dobs, w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eval_grad(eng, mt, dobs, w)
    torch.cuda.synchronize()
    return 1e3 * (time.perf_counter() - t0) / steps


def sweep_roofline(eng, mt, dobs, w, refine, reps=3):
    """Backward-sweep kernel (k_bwd_wave, one launch per tree level) of the
    engine's solves: average launch duration from HIP events recorded on the
    solve stream, and its algorithmic bytes and flops per launch."""
    sv = eng.solver
    info = sv.info()
    # a mixed plan writes an FP64 copy of x_J unless the fields stay FP32
    x64_copy = info["precision"] == 1 and (not eng._gpu.fields32 or refine > 0)
    sv.set_timing(True)
    for _ in range(reps):
        eval_grad(eng, mt, dobs, w)
    torch.cuda.synchronize()
    f_ms, b_ms, nsolve = sv.sweep_times()
    sv.set_timing(False)
    launches = nsolve * info["level_launches"]
    ts = 4 if info["precision"] == 1 else 8
    fs = 4 if (info["precision"] == 1 and eng._gpu.fields32) else 8
    nb, k, n = info["nbatch"], eng.nS, info["n"]
    # algorithmic bytes of one backward sweep: the factor once (nnz_L entries
    # of s_f bytes) and the field block in and out (n x k of s bytes each way)
    # per wavenumber, plus the FP64 copy of x when the plan writes one
    per_sweep = nb * (ts * info["nnz_L"] + 2 * fs * n * k + (8 * n * k if x64_copy else 0))
    # the bytes the kernels address: every supernode's front rows (x_R
    # re-read by each supernode that borders them, mostly from L2) and the
    # stored sweep panels
    access = nb * k * (ts * (info["sweep_rows_nf"] + info["sweep_rows_ns"])
                       + (8 * info["sweep_rows_ns"] if x64_copy else 0)) \
        + nb * info["panel_size"] * ts
    flops = 2.0 * nb * k * info["solve_flops_per_rhs"] / 4.0
    avg_ms = b_ms / launches
    achieved = (per_sweep / info["level_launches"]) / (avg_ms * 1e-3) / 1e9
    return {"kernel": "k_bwd_wave", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "avg_ms": round(avg_ms, 4),
            "bytes_per_launch": int(per_sweep / info["level_launches"]),
            "bytes_per_sweep": int(per_sweep),
            "access_bytes_per_launch": int(access / info["level_launches"]),
            "tflops": round(flops / (b_ms / nsolve * 1e-3) / 1e12, 2),
            "tflops_peak": MFMA_PEAK["fp32" if ts == 4 else "fp64"],
            "launches_per_sweep": info["level_launches"],
            "fwd_sweep_ms": round(f_ms / nsolve, 3), "bwd_sweep_ms": round(b_ms / nsolve, 3)}


def factor_roofline(eng, mt, reps=10):
    """The numeric factorisation (K2: `k_factor_level`, one launch per tree
    level, all wavenumbers at once) alone: its time per factorisation from
    HIP events on the solve stream (torch's current stream, where
    dcx_factorize runs every level), its flops (partial Cholesky of every
    front plus the inverse of its diagonal block, the symbolic phase's
    count) against the factor precision's MFMA peak, and its algorithmic
    bytes nky (8 nnzA + s_f nnz(L)) against HBM peak."""
    sv, g = eng.solver, eng._gpu
    info = sv.info()
    with torch.no_grad():
        vals = g.assemble(torch.exp(-mt))
    sv.factorize(vals)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        sv.factorize(vals)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    fp32 = info["precision"] == 1
    nb = info["nbatch"]
    flops = nb * info["factor_flops"]
    bytes_ = nb * (8 * sv.col.numel() + (4 if fp32 else 8) * info["nnz_L"])
    peak = MFMA_PEAK["fp32" if fp32 else "fp64"]
    tflops = flops / (ms * 1e-3) / 1e12
    return {"kernel": "k_factor_level", "bound": "mfma", "ms": round(ms, 3),
            "gflop": round(flops / 1e9, 3), "tflops": round(tflops, 2), "peak_tflops": peak,
            "frac": round(tflops / peak, 4), "bytes": int(bytes_),
            "achieved_gbs": round(bytes_ / (ms * 1e-3) / 1e9, 1), "launches": info["ndepth"]}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_line25k_eval_grad.json")


def measured_traffic(kernel):
    """HBM bytes per dispatch of the kernel family `kernel` (name up to the
    first template argument) from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, calibrated on a 1 GiB copy), or None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            return int(json.load(fh)["families"][kernel]["traffic_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def linearization_times(eng, mt, reps=3):
    """Gauss-Newton building blocks on the FP64 engine at m: the linearisation
    (factorisation + forward fields), J v, J^T v, the explicit J and
    diag(J^T D J), each the median of `reps` runs."""
    from simpeg_dc_torch_amd import Linearization2D

    def med(fn):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ts.append(1e3 * (time.perf_counter() - t0))
        return sorted(ts)[len(ts) // 2], out

    t_lin, lin = med(lambda: Linearization2D(eng, mt))
    rng = np.random.default_rng(5)
    u = torch.tensor(rng.standard_normal(eng.nC), device="cuda:0")
    v = torch.tensor(rng.standard_normal(eng.nD), device="cuda:0")
    t_jv, _ = med(lambda: lin.Jvec(u))
    t_jtv, _ = med(lambda: lin.Jtvec(v))
    t_j, J = med(lambda: lin.Jmatrix())
    t_dg, _ = med(lambda: lin.diag_JtDJ(torch.ones_like(v)))
    return {"linearize_ms": round(t_lin, 2), "jvec_ms": round(t_jv, 2),
            "jtvec_ms": round(t_jtv, 2), "jmatrix_ms": round(t_j, 2),
            "diag_jtdj_ms": round(t_dg, 2), "J_shape": list(J.shape),
            "precision": "fp64"}


def grid3d_times(steps, warmup, tune=True):
    """The 3-D nodal engine (TorchDC3D) on a 56 x 56 x 38-cell grid with 200
    pole sources and 6,000 pole-dipole data: FP64 and mixed eval+grad ms and
    the mixed result's error against FP64 (one system, fronts up to 3,315
    rows: the grid-wide large-front kernels).  Data uncertainty 5% of |dobs|
    plus a floor of 1e-4 (0.1% of the median datum; the data span 2.5e-5 to
    18), the rel * |d| + floor form of the reference's own scripts: without a
    floor the smallest data get weights near 1e6 and the misfit gradient
    measures FP32's absolute dpred error on them rather than the solver.  The
    J^T v error (random v) isolates the adjoint path."""
    from simpeg_dc_torch_amd import TorchDC3D, workloads
    mesh, survey, m = workloads.grid_3d_case()
    mt = torch.tensor(m, device="cuda:0")
    out, ref = {}, None
    for prec, refine in (("fp64", 0), ("mixed", 0), ("mixed", 1)):
        eng = TorchDC3D(mesh, survey, device="cuda:0", precision=prec, refine=refine)
        if ref is None:
            with torch.no_grad():
                d = eng.dpred(mt)
            rng = np.random.default_rng(3)
            dobs = d * (1 + 0.05 * torch.tensor(rng.standard_normal(eng.nD), device="cuda:0"))
            w = 1.0 / (0.05 * dobs.abs() + 1e-4)
            v = torch.tensor(rng.standard_normal(eng.nD), device="cuda:0")
            info = eng.solver.info()
            out.update(cells=eng.nC, nodes=eng.nN, sources=eng.nS, data=eng.nD,
                       max_front=info["max_front"], nnz_L=info["nnz_L"])
        key = prec if prec == "fp64" else f"mixed_refine{refine}"
        if tune:
            out.setdefault("tiles", {})[key] = tune_tiles({key: eng}, mt, dobs, w)[0][key]
        ms = time_steps(eng, mt, dobs, w, steps, warmup)
        with torch.no_grad():
            d = eng.dpred(mt)
        _, g = eval_grad(eng, mt, dobs, w)
        jt = eng.Jtvec(mt, v)
        out[f"{key}_ms"] = round(ms, 2)
        if ref is None:
            ref = (d, g, jt)
        else:
            for nm, a, b in (("dpred", d, ref[0]), ("grad", g, ref[1]), ("jtvec", jt, ref[2])):
                out[f"{key}_{nm}_rel_err"] = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        del eng
    for prec in ("fp64", "mixed"):
        eng = TorchDC3D(mesh, survey, device="cuda:0", precision=prec, refine=0,
                        rhs_basis="electrodes")
        out[f"electrode_basis_{prec}_ms"] = round(time_steps(eng, mt, dobs, w, steps, warmup), 2)
        out["electrodes"] = eng._gpu.ne
        del eng
    return out


def electrode_basis_times(mesh, survey, mt, dobs, w, d64, g64, steps, warmup):
    """The same 25k-cell eval+grad with rhs_basis="electrodes": one solve per
    wavenumber with the 61 electrode columns instead of the 849 source
    columns, the adjoint by electrode-space algebra (no second solve).  Not
    the headline (which keeps the source formulation of the reference);
    reported beside it, with its results' distance from the FP64 source-basis
    ones."""
    from simpeg_dc_torch_amd import TorchDC2D
    out = {}
    for prec in ("fp64", "mixed"):
        eng = TorchDC2D(mesh, survey, nky=12, device="cuda:0", precision=prec, refine=0,
                        rhs_basis="electrodes")
        out["electrodes"] = eng._gpu.ne
        out[f"{prec}_ms"] = round(time_steps(eng, mt, dobs, w, steps, warmup), 3)
        with torch.no_grad():
            d = eng.dpred(mt)
        _, g = eval_grad(eng, mt, dobs, w)
        out[f"{prec}_dpred_rel_err"] = float(torch.linalg.norm(d - d64) / torch.linalg.norm(d64))
        out[f"{prec}_grad_rel_err"] = float(torch.linalg.norm(g - g64) / torch.linalg.norm(g64))
        del eng
    return out


def gn_step_times(reps=3):
    """One Gauss-Newton step on the 62,016-cell, 2,430-datum line (FP64): data
    from the true model with 5% + 1e-4 V noise (seed 3), start at a 200 ohm-m
    half-space, SimPEG-style tensor regularisation (alpha_s 1e-4), beta 10.
    Times (median of `reps`) of the linearisation, the explicit J (batched
    adjoint), the Woodbury direct step (H_m solve of J^T through the HIP
    solver, S = J K, dense Cholesky) and the Occam beta search (one eigh), and
    the misfit before and after the Occam step (one forward)."""
    from simpeg_dc_torch_amd import TorchDC2D, workloads
    from simpeg_dc_torch_amd.gauss_newton import (HmSolver, direct_step, occam_step,
                                                  tensor_regularization)
    mesh, survey, m_true = workloads.gn_line_case()
    dev = torch.device("cuda:0")
    eng = TorchDC2D(mesh, survey, nky=12, device=dev, precision="fp64")
    with torch.no_grad():
        d_true = eng.dpred(torch.tensor(m_true, device=dev))
    rng = np.random.default_rng(3)
    std = 0.05 * d_true.abs() + 1e-4
    dobs = d_true + std * torch.tensor(rng.standard_normal(eng.nD), device=dev)
    w = 1.0 / (0.05 * dobs.abs() + 1e-4)
    m0 = torch.full((mesh.nC,), float(np.log(200.0)), dtype=torch.float64, device=dev)
    Hm = tensor_regularization(mesh, alpha_s=1e-4)
    hs = HmSolver(dev)
    beta = 10.0

    def med(fn):
        ts, out = [], None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ts.append(1e3 * (time.perf_counter() - t0))
        return sorted(ts)[len(ts) // 2], out

    t_lin, lin = med(lambda: eng.linearize(m0))
    t_j, J = med(lambda: lin.Jmatrix())
    with torch.no_grad():
        d0 = eng.dpred(m0)
    r = w * (d0 - dobs)
    D, wd = 2.0 * w ** 2, w * (2.0 * r)
    g_d = J.T @ wd
    import scipy.sparse as sp_
    H = sp_.csr_matrix(Hm)
    Ht = torch.sparse_csr_tensor(torch.as_tensor(H.indptr, dtype=torch.int64, device=dev),
                                 torch.as_tensor(H.indices, dtype=torch.int64, device=dev),
                                 torch.as_tensor(H.data, dtype=torch.float64, device=dev),
                                 H.shape)
    g_m = Ht @ m0
    t_dir, p_dir = med(lambda: direct_step(J, D, g_d + beta * g_m, H, beta, hs))
    t_occ, occ = med(lambda: occam_step(J, D, g_d, g_m, r, beta, H, hs,
                                        lambda rr: (rr ** 2).sum(), w))
    b_occ, p_occ, pred = occ
    with torch.no_grad():
        d1 = eng.dpred(m0 + p_occ)
    del lin
    engE = TorchDC2D(mesh, survey, nky=12, device=dev, precision="fp64",
                     rhs_basis="electrodes")
    t_linE, linE = med(lambda: engE.linearize(m0))
    t_jE, JE = med(lambda: linE.Jmatrix())
    j_err = float(torch.linalg.norm(JE - J) / torch.linalg.norm(J))
    del linE, JE, engE
    return {"cells": mesh.nC, "data": eng.nD, "sources": eng.nS, "nky": eng.nky,
            "precision": "fp64", "linearize_ms": round(t_lin, 2),
            "jmatrix_ms": round(t_j, 2), "direct_step_ms": round(t_dir, 2),
            "occam_step_ms": round(t_occ, 2), "hm_solve": hs.last,
            "electrode_basis_linearize_ms": round(t_linE, 2),
            "electrode_basis_jmatrix_ms": round(t_jE, 2), "electrode_basis_J_rel_err": j_err,
            "phi_d_start": float((r ** 2).sum()),
            "occam_beta": b_occ, "occam_predicted_phi_d": pred,
            "phi_d_after_occam_step": float(((w * (d1 - dobs)) ** 2).sum()),
            "direct_step_finite": bool(p_dir is not None and torch.isfinite(p_dir).all())}


def readme_case_times(steps, warmup):
    """The reference's published eval+grad case (README.md:33-39: 18.5 ms,
    FP64, RTX 4070 Laptop): its 2.5-D regression problem, 100 x 50 cells of
    1 m, 21 electrodes, dipole-dipole n = 1..6 (93 data), nky = 11 with the
    SimPEG quadrature, a 10 ohm-m block in 100 ohm-m."""
    from simpeg_dc_torch_amd import TorchDC2D, workloads
    mesh, survey, m = workloads.regression_case()
    mt = torch.tensor(m, device="cuda:0")
    out = {"cells": mesh.nC, "reference_published_ms": 18.5,
           "reference_hardware": "NVIDIA RTX 4070 Laptop (README.md:39)"}
    for prec in ("fp64", "mixed"):
        eng = TorchDC2D(mesh, survey, nky=11, device="cuda:0", precision=prec,
                        quadrature="simpeg")
        with torch.no_grad():
            d = eng.dpred(mt)
        rng = np.random.default_rng(0)
        dobs = d * (1 + 0.05 * torch.tensor(rng.standard_normal(eng.nD), device="cuda:0"))
        w = 1.0 / (0.05 * dobs.abs())
        out[f"{prec}_ms"] = round(time_steps(eng, mt, dobs, w, steps, warmup), 3)
        out["data"] = eng.nD
    return out


def ip_times(steps, warmup):
    """Induced polarisation, FP64, eval+grad = misfit and its gradient:
    linearised IP on the 25k-cell line (TorchIP2D: one J v and one J^T v of
    the DC engine frozen at the background, apparent-chargeability data;
    also with the electrode basis, where neither product solves again) and
    on the 56 x 56 x 38 grid (TorchIP3D), and the joint non-linear DC-IP
    forward on the 25k-cell line (TorchDCIP2D: two DC eval+grads, sigma and
    sigma (1 - eta))."""
    from simpeg_dc_torch_amd import TorchDCIP2D, TorchIP2D, TorchIP3D, workloads
    from simpeg_dc_torch_amd import survey as S
    dev = "cuda:0"

    def charge(sv):
        srcs = []
        for src in sv.source_list:
            rxs = [S.DipoleReceiver(r.locations_m, r.locations_n,
                                    data_type="apparent_chargeability")
                   if r.locations_n is not None else
                   S.PoleReceiver(r.locations_m, data_type="apparent_chargeability")
                   for r in src.receiver_list]
            srcs.append(S.DipoleSource(rxs, src.location_a, src.location_b)
                        if src.location_b is not None else S.PoleSource(rxs, src.location_a))
        return S.Survey(srcs)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return round(1e3 * (time.perf_counter() - t0) / steps, 3)

    out = {}
    rng = np.random.default_rng(8)
    line = workloads.line_25k_case()[:3]
    for name, (mesh, survey, m), kw in (("line25k", line, {}),
                                        ("line25k_electrodes", line, {"rhs_basis": "electrodes"}),
                                        ("grid3d", workloads.grid_3d_case(), {})):
        t0 = time.perf_counter()
        ip = (TorchIP3D(mesh, charge(survey), m, device=dev) if name == "grid3d"
              else TorchIP2D(mesh, charge(survey), m, nky=12, device=dev, **kw))
        t_build = time.perf_counter() - t0
        eta = torch.tensor(0.01 + 0.05 * rng.random(mesh.nC), device=dev)
        with torch.no_grad():
            d = ip.dpred(eta)
        dobs = d * (1 + 0.05 * torch.tensor(rng.standard_normal(d.numel()), device=dev))
        w = 1.0 / (0.05 * dobs.abs() + 1e-6)

        def step():
            x = eta.clone().requires_grad_(True)
            ip.misfit(x, dobs, w).backward()
        out[f"ip_{name}_ms"] = timed(step)
        out[f"ip_{name}_build_s"] = round(t_build, 1)
        del ip
    mesh, survey, m = workloads.line_25k_case()[:3]
    dcip = TorchDCIP2D(mesh, survey, nky=12, device=dev, ip_scale=1000.0)
    x0 = torch.tensor(np.r_[m, dcip.m_eta_of(0.02 + 0.1 * rng.random(mesh.nC))], device=dev)
    with torch.no_grad():
        d_dc, d_ip = dcip.dpred(x0)
    w_dc, w_ip = 1.0 / (0.05 * d_dc.abs()), torch.full_like(d_ip, 1.0)

    def step_dcip():
        x = x0.clone().requires_grad_(True)
        dcip.misfit(x, d_dc * 1.01, w_dc, d_ip + 1.0, w_ip).backward()
    out["dcip_line25k_ms"] = timed(step_dcip)
    out["precision"] = "fp64"
    return out


def lines_times(steps, warmup, autotune=True, per_line=True):
    """The uneven batch: 16 independent 2.5-D lines of 3k-40k cells
    (log-uniform, seed 5; 21-96 electrodes, dipole-dipole, 12 wavenumbers
    requested), mixed precision.  One step = eval+grad of every line.  Timed
    four times: the lines one after another on one stream; concurrently on 4
    default-priority streams, issued from one host thread; on the 4 streams
    issued from 4 host threads (leaf size autotuned per size class, tuning
    outside the timed region); and all lines in one plan (`LineForest`: one
    launch sequence per step).  The one-plan results are compared with the
    per-line engines' (both mixed precision).  Beside them, FP64: the
    per-line batch eager and as one captured graph per line, and the one-plan
    batch eager and as one captured graph for every line.  per_line=False:
    the mixed one-plan leg alone (for profiling it)."""
    from simpeg_dc_torch_amd import workloads
    from simpeg_dc_torch_amd.lines import LineBatch
    L = workloads.uneven_lines_case()
    lines = [(mesh, survey) for mesh, survey, _ in L]
    ms = [torch.tensor(m, device="cuda:0") for _, _, m in L]
    probe = LineBatch(lines, precision="fp64", streams=1)
    data = []
    for eng, m in zip(probe.engines, ms):
        with torch.no_grad():
            d = eng.dpred(m)
        rng = np.random.default_rng(7)
        dobs = d * (1 + 0.05 * torch.tensor(rng.standard_normal(d.numel()), device="cuda:0"))
        data.append((dobs, 1.0 / (0.05 * dobs.abs())))
    del probe
    if not per_line:
        one = one_plan_times(lines, ms, data, steps, warmup)
        one_e = one_plan_times(lines, ms, data, steps, warmup, rhs_basis="electrodes")
        return {**one, "one_plan_electrode_basis_ms": one_e["one_plan_ms"],
                "one_plan_electrode_basis": one_e["one_plan"]}
    tune = [(m.cpu().numpy(), d.cpu().numpy(), w.cpu().numpy()) for m, (d, w) in zip(ms, data)]
    t0 = time.perf_counter()
    lb = LineBatch(lines, precision="mixed", streams=4, leaf="auto" if autotune else None,
                   tune_data=tune if autotune else None, threads=True,
                   tiles="auto" if autotune else None)
    pool = lb._pool
    t_tune = time.perf_counter() - t0
    lb.set_data([d for d, _ in data], [w for _, w in data])

    def timed(nstreams, threads=False):
        saved = lb.streams
        lb.streams = saved[:nstreams]
        lb._pool = pool if threads else None
        for _ in range(warmup):
            lb.eval_grad(ms)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            lb.eval_grad(ms)
        torch.cuda.synchronize()
        lb.streams = saved
        return 1e3 * (time.perf_counter() - t1) / steps

    ms_seq, ms_par, ms_thr = timed(1), timed(4), timed(4, threads=True)
    one = one_plan_times(lines, ms, data, steps, warmup, compare=lb)
    one_e = one_plan_times(lines, ms, data, steps, warmup, compare=lb, rhs_basis="electrodes")

    # the same per-line batch with every engine in the electrode basis
    lbe = LineBatch(lines, precision="mixed", streams=4, leaf=lb.leaf, rhs_basis="electrodes")
    lbe.set_data([d for d, _ in data], [w for _, w in data])
    for _ in range(warmup):
        lbe.eval_grad(ms)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        lbe.eval_grad(ms)
    torch.cuda.synchronize()
    ms_elec = 1e3 * (time.perf_counter() - t1) / steps
    oe, ob = lbe.eval_grad(ms), lb.eval_grad(ms)
    torch.cuda.synchronize()
    elec_diff = max(float(torch.linalg.norm(a[1] - b[1]) / torch.linalg.norm(b[1]))
                    for a, b in zip(oe, ob))
    del lbe

    # FP64 engines: eager on 4 streams, then each line's eval+grad replayed
    # as a captured HIP graph (capture is limited to FP64 engines)
    lb64 = LineBatch(lines, precision="fp64", streams=4)
    lb64.set_data([d for d, _ in data], [w for _, w in data])

    def timed64():
        for _ in range(warmup):
            lb64.eval_grad(ms)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            lb64.eval_grad(ms)
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t1) / steps

    ms64_eager = timed64()
    ref64 = [(float(p), g.clone()) for p, g in lb64.eval_grad(ms)]
    lb64.capture([m.cpu().numpy() for m in ms])
    ms64_graph = timed64()
    out64 = lb64.eval_grad(ms)
    torch.cuda.synchronize()
    graph_diff = max(abs(float(p) / r[0] - 1.0) for (p, _), r in zip(out64, ref64))
    del lb64
    forest64 = forest_graph_times(lines, ms, data, steps, warmup)
    ms_forest = one["one_plan_ms"]
    cells = sum(mesh.nC for mesh, _ in lines)
    return {"lines": len(lines), "cells_total": cells,
            "cells_range": [min(mesh.nC for mesh, _ in lines), max(mesh.nC for mesh, _ in lines)],
            "precision": "mixed", "sequential_ms": round(ms_seq, 2),
            "streams4_ms": round(ms_par, 2), "streams4_threads4_ms": round(ms_thr, 2),
            "one_plan_ms": ms_forest,
            "streams4_electrode_basis_ms": round(ms_elec, 2),
            "electrode_basis_grad_max_rel_diff": float(f"{elec_diff:.1e}"),
            "fp64_streams4_ms": round(ms64_eager, 2), "fp64_graphs_ms": round(ms64_graph, 2),
            "fp64_graph_phi_max_rel_diff": float(f"{graph_diff:.1e}"),
            **forest64,
            "one_plan_electrode_basis_ms": one_e["one_plan_ms"],
            "lines_per_s": round(1e3 * len(lines) / min(ms_par, ms_thr, ms_forest, ms_elec,
                                                        one_e["one_plan_ms"]), 1),
            "one_plan": one["one_plan"],
            "one_plan_electrode_basis": one_e["one_plan"],
            "leaf_by_size_class": {str(k): v for k, v in (lb.leaf or {}).items()},
            "tiles_by_size_class": {str(k): v for k, v in lb.tiles.items()},
            "build_and_tune_s": round(t_tune, 1),
            "graph_capture": "FP64 engines and FP64 one-plan batches (mixed-precision "
                             "capture refused)"}


def forest_graph_times(lines, ms, data, steps, warmup):
    """Every line in one FP64 plan (`LineForest`), eager and as one captured
    HIP graph replayed per step (tiles at their defaults)."""
    from simpeg_dc_torch_amd.lines import LineForest
    lf = LineForest(lines, precision="fp64")
    lf.set_data([d for d, _ in data], [w for _, w in data])

    def timed():
        for _ in range(warmup):
            lf.eval_grad(ms)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            lf.eval_grad(ms)
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t1) / steps

    eager = timed()
    ref = [float(p) for p, _ in lf.eval_grad(ms)]
    lf.capture([m.cpu().numpy() for m in ms])
    graph = timed()
    out = lf.eval_grad(ms)
    torch.cuda.synchronize()
    diff = max(abs(float(p) / r - 1.0) for (p, _), r in zip(out, ref))
    lf.free()
    return {"one_plan_fp64_ms": round(eager, 2), "one_plan_fp64_graph_ms": round(graph, 2),
            "one_plan_fp64_graph_phi_max_rel_diff": float(f"{diff:.1e}")}


def one_plan_times(lines, ms, data, steps, warmup, compare=None, rhs_basis="sources"):
    """Every line in one plan (a forest of the lines' separator trees), mixed
    precision, source or electrode basis; with `compare` (a LineBatch) the
    largest relative differences of phi and of the gradient against the
    per-line engines."""
    from simpeg_dc_torch_amd.lines import LineForest
    t0 = time.perf_counter()
    lf = LineForest(lines, precision="mixed", rhs_basis=rhs_basis)
    t_forest = time.perf_counter() - t0
    lf.set_data([d for d, _ in data], [w for _, w in data])
    t0 = time.perf_counter()
    tiles, _ = lf.autotune_tiles(ms)
    t_tiles = time.perf_counter() - t0
    for _ in range(warmup):
        lf.eval_grad(ms)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        lf.eval_grad(ms)
    torch.cuda.synchronize()
    ms_forest = 1e3 * (time.perf_counter() - t1) / steps
    one = {"rhs_columns": lf.K, "systems": lf.nbatch,
           "level_launches": lf.solver.info().get("level_launches"),
           "build_s": round(t_forest, 1), "tiles": tiles, "tile_tuning_s": round(t_tiles, 1)}
    if compare is not None:
        out_f, out_b = lf.eval_grad(ms), compare.eval_grad(ms)
        torch.cuda.synchronize()
        one["phi_max_rel_diff_vs_per_line"] = float(
            f"{max(abs(float(a[0]) / float(b[0]) - 1.0) for a, b in zip(out_f, out_b)):.2e}")
        one["grad_max_rel_diff_vs_per_line"] = float(f"""{max(
            float(torch.linalg.norm(a[1] - b[1]) / torch.linalg.norm(b[1]))
            for a, b in zip(out_f, out_b)):.2e}""")
    lf.free()
    return {"one_plan_ms": round(ms_forest, 2), "one_plan": one}


def host_cores():
    """(cores this process may use, CPU model name)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def cpu_baseline(mesh, survey, m, kys, wq, dobs, w, reps=3):
    """The reference CPU path restated by the oracle (SuperLU with SciPy
    defaults per wavenumber, forward and adjoint solves of all 849 source
    columns, projection, misfit and the gradient) on the whole eval+grad at
    every wavenumber: the median of `reps` runs one wavenumber after another
    (the reference's loop, one core for the factorisations and solves), and
    the median of `reps` runs with the independent wavenumbers on a thread
    pool (min(nky, 16, cores) threads)."""
    from oracle.dc25d import Problem2D, misfit_and_grad_by_ky
    O = Problem2D(mesh.h[0], mesh.h[1], mesh.origin, survey, kys, wq)
    ncores, model = host_cores()
    workers = max(1, min(len(kys), 16, ncores))
    out = {}
    for label, nw in (("sequential", 1), ("threaded", workers)):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            misfit_and_grad_by_ky(O, m, dobs, w, workers=nw)
            ts.append(1e3 * (time.perf_counter() - t0))
        out[label] = (sorted(ts)[len(ts) // 2], ts, nw)
    return out, ncores, model


def summarize_trace(path, steps):
    """Per-kernel and per-family tables of the mixed-precision eval+grad steps
    that bench.py timed, from a rocprofv3 --kernel-trace CSV of a bench.py run
    (markdown on stdout).  A step starts at k_assemble; steps are classed by
    their kernels (FP64 or FP32 factor, electrode basis or source basis); the
    timed mixed steps are the last `steps` of the first run of source-basis
    FP32 steps that follows a run of source-basis FP64 steps after the tile
    tuning (the tuner's trial runs, which come first, are excluded)."""
    import csv
    import re
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    # a step: k_assemble up to k_lmap_t (the transposed map closes the
    # gradient) and the kernels that follow it back to back (autograd's tail)
    steps_all, cur, closed, last_end = [], None, False, 0
    for r in rows:
        name = r["Kernel_Name"]
        if "k_assemble" in name:
            cur, closed = [], False
            steps_all.append(cur)
        elif cur is not None and closed and int(r["Start_Timestamp"]) - last_end > 30_000:
            cur = None
        if cur is not None:
            cur.append(r)
            last_end = max(last_end, int(r["End_Timestamp"])) if closed else int(r["End_Timestamp"])
            closed = closed or "k_lmap_t" in name

    def kind(st):
        names = " ".join(r["Kernel_Name"] for r in st)
        if "k_rows_mix" in names or "k_jpair" in names or "k_spmm" in names:
            return "other"
        if "k_factor_level<float" in names:
            return "mixed"
        if "k_factor_level<double" in names:
            return "fp64"
        return "other"
    kinds = [kind(st) for st in steps_all]
    # runs of equal kind
    runs, i = [], 0
    while i < len(kinds):
        j = i
        while j < len(kinds) and kinds[j] == kinds[i]:
            j += 1
        runs.append((kinds[i], i, j))
        i = j
    # the timed pair is the last fp64 run followed by a mixed run of at least
    # `steps` steps (tuning: fp64 trials, mixed trials; timing: fp64, mixed)
    sel = None
    for a, b in zip(runs, runs[1:]):
        if a[0] == "fp64" and b[0] == "mixed" and b[2] - b[1] >= steps:
            sel = b
    if sel is None:
        raise SystemExit("no timed mixed steps found in the trace")
    chosen = steps_all[sel[2] - steps:sel[2]]

    def short(n):
        n = re.sub(r"^void |\(anonymous namespace\)::", "", n)
        return n.split("(")[0]
    per, fam = {}, {}
    for st in chosen:
        for r in st:
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            k = short(r["Kernel_Name"])
            f = k.split("<")[0] + ("<float>" if "<float" in k else "<double>" if "<double" in k else "")
            for tab, key in ((per, k), (fam, f)):
                c, t = tab.get(key, (0, 0.0))
                tab[key] = (c + 1, t + d)
    span = [(int(st[-1]["End_Timestamp"]) - int(st[0]["Start_Timestamp"])) / 1e3 for st in chosen]
    print(f"Timed mixed steps: {steps} (trace steps {sel[2] - steps}..{sel[2] - 1}); "
          f"span per step {np.mean(span):.1f} us (first kernel start to last kernel end)\n")
    for title, tab in (("Per kernel family (all template instances)", fam),
                       ("Per kernel instance", per)):
        print(f"### {title}\n")
        print("| kernel | dispatches per step | us per step | average us per dispatch |")
        print("|---|---|---|---|")
        for k, (c, t) in sorted(tab.items(), key=lambda kv: -kv[1][1]):
            print(f"| `{k}` | {c / steps:g} | {t / steps:.1f} | {t / c:.2f} |")
        print()
    # the sweep-roofline reps (sweep timing on: one backward launch per tree
    # level, no split sweep): the last 3 steps of the first mixed run after
    # the timed one that an FP64 run follows (the electrode basis comes
    # between), and the first 3 steps of that FP64 run
    reps = 3
    after = [r for r in runs if r[1] >= sel[2]]
    pair = next(((a, b) for a, b in zip(after, after[1:])
                 if a[0] == "mixed" and b[0] == "fp64" and a[2] - a[1] >= reps
                 and b[2] - b[1] >= reps), None)
    if pair is None:
        return
    print("### Sweep-roofline reps (sweep timing on, one backward launch per level)\n")
    print("| kernel | steps | dispatches per step | average us per dispatch |")
    print("|---|---|---|---|")
    for fam_name, sl in (("k_bwd_wave<float>", steps_all[pair[0][2] - reps:pair[0][2]]),
                         ("k_bwd_wave<double>", steps_all[pair[1][1]:pair[1][1] + reps])):
        ds = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
              for st in sl for r in st if short(r["Kernel_Name"]).startswith(fam_name[:-1])]
        if ds:
            print(f"| `{fam_name}` | {reps} | {len(ds) / reps:g} | {np.mean(ds):.2f} |")
    print()


def summarize_pmc(fetch_csv, write_csv, out_json):
    """HBM bytes per dispatch from two rocprofv3 --pmc runs of `bench.py
    --pmc-pass` (FETCH_SIZE, then WRITE_SIZE; counter_collection CSVs): each
    counter (KiB) scaled by what it reads for the pass's 1 GiB calibration
    copy (the largest copy dispatch moves 2^30 bytes each way); per kernel and
    per family (name up to its template arguments, all instances).  Writes
    `out_json` (read by the benchmark's roofline) and prints markdown."""
    import csv
    import re

    def load(path, counter):
        vals, names = {}, {}
        for r in csv.DictReader(open(path)):
            if r.get("Counter_Name") != counter:
                continue
            d = r["Dispatch_Id"]
            vals[d] = vals.get(d, 0.0) + float(r["Counter_Value"])
            names[d] = r["Kernel_Name"]
        return vals, names
    fv, fn = load(fetch_csv, "FETCH_SIZE")
    wv, wn = load(write_csv, "WRITE_SIZE")

    def calib(vals, names):
        copies = [vals[d] for d in vals if "copyBuffer" in names[d] or "copy" in names[d].lower()]
        return (1 << 30) / (1024.0 * max(copies))
    fs, ws = calib(fv, fn), calib(wv, wn)

    def short(n):
        n = re.sub(r"^void |\(anonymous namespace\)::", "", n)
        return n.split("(")[0]

    def per_kernel(vals, names, scale):
        out = {}
        for d, v in vals.items():
            k = short(names[d])
            c, t = out.get(k, (0, 0.0))
            out[k] = (c + 1, t + v * 1024.0 * scale)
        return out
    pf, pw = per_kernel(fv, fn, fs), per_kernel(wv, wn, ws)
    kernels, fams = {}, {}
    for k in set(pf) | set(pw):
        cf, tf = pf.get(k, (0, 0.0))
        cw, tw = pw.get(k, (0, 0.0))
        c = max(cf, cw)
        f = tf / cf if cf else 0.0
        w_ = tw / cw if cw else 0.0
        kernels[k] = {"fetch_bytes": f, "write_bytes": w_, "traffic_bytes": f + w_,
                      "dispatches": c}
        fam = k.split("<")[0] + ("<float>" if "<float" in k else "<double>" if "<double" in k else "")
        a = fams.setdefault(fam, {"fetch": 0.0, "write": 0.0, "dispatches": 0})
        a["fetch"] += tf
        a["write"] += tw
        a["dispatches"] += c
    families = {k: {"fetch_bytes": v["fetch"] / v["dispatches"],
                    "write_bytes": v["write"] / v["dispatches"],
                    "traffic_bytes": (v["fetch"] + v["write"]) / v["dispatches"],
                    "dispatches": v["dispatches"]} for k, v in fams.items()}
    with open(out_json, "w") as fh:
        json.dump({"calibration": {"fetch_scale": fs, "write_scale": ws},
                   "kernels": kernels, "families": families}, fh, indent=1)
    print(f"Calibration (1 GiB copy): FETCH_SIZE x {fs:.3f}, WRITE_SIZE x {ws:.3f}\n")
    for title, tab in (("Per kernel family (all template instances)", families),
                       ("Per kernel instance", kernels)):
        print(f"### {title}\n")
        print("| kernel | dispatches | fetch MB/dispatch | write MB/dispatch | traffic MB/dispatch |")
        print("|---|---|---|---|---|")
        for k, v in sorted(tab.items(), key=lambda kv: -kv[1]["traffic_bytes"] * kv[1]["dispatches"]):
            print(f"| `{k}` | {v['dispatches']} | {v['fetch_bytes'] / 1e6:.1f} | "
                  f"{v['write_bytes'] / 1e6:.1f} | {v['traffic_bytes'] / 1e6:.1f} |")
        print()


def pmc_pass(args):
    """Short run for hardware-counter passes: a 1 GiB device copy (the
    counters' calibration) and two eval+grad steps of each engine (mixed,
    then FP64), with sweep timing on as in the roofline's reps."""
    log("pmc pass: calibration copy of 1 GiB")
    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    mesh, survey, m, e64, emx, mt, dobs, w, _ = build_engines(args.refine)
    # default kernel-variant thresholds: a tuning search here would add its
    # trial runs to the per-dispatch counter averages.  Sweep timing on, as in
    # the roofline's timed reps: one backward launch per tree level (no split
    # sweep), so a dispatch's bytes are a level's bytes
    for eng in (emx, e64):
        eng.solver.set_timing(True)
        for _ in range(2):
            eval_grad(eng, mt, dobs, w)
        torch.cuda.synchronize()
        eng.solver.set_timing(False)
    info = emx.solver.info()
    log(f"pmc pass done: nbatch={info['nbatch']} k={emx.nS} rows_nf={info['sweep_rows_nf']} "
        f"rows_ns={info['sweep_rows_ns']} panel={info['panel_size']}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--refine", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-pass", action="store_true")
    ap.add_argument("--summarize-pmc", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT_JSON"),
                    help="HBM traffic per dispatch from the two --pmc-pass counter CSVs "
                         "(FETCH_SIZE, WRITE_SIZE); writes OUT_JSON, prints markdown, exits")
    ap.add_argument("--summarize-trace", metavar="CSV",
                    help="print per-kernel tables of the timed mixed steps from a rocprofv3 "
                         "kernel-trace CSV of a bench.py run (--steps as in that run) and exit")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default kernel-variant thresholds (no tile autotuning)")
    ap.add_argument("--workload", choices=("all", "line25k", "grid3d", "lines",
                                           "lines_one_plan", "gn", "ip"),
                    default="all",
                    help="all (default): the 2.5-D line with the 3-D grid and the uneven "
                         "batch of lines beside it; line25k, grid3d, lines or gn alone; "
                         "lines_one_plan: the uneven batch in one plan only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the path runs on one GPU; there is no multi-GPU code path)", file=sys.stderr)
        sys.exit(2)
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.steps)
        return
    if args.pmc_pass:
        pmc_pass(args)
        return
    if args.workload == "grid3d":
        log("3-D grid only (TorchDC3D, 56 x 56 x 38 cells, 200 sources)")
        print(json.dumps({"workload": "grid3d",
                          **grid3d_times(args.steps, args.warmup, tune=not args.no_tune)}),
              flush=True)
        return
    if args.workload == "gn":
        log("Gauss-Newton step only (62k cells, 2,430 data)")
        print(json.dumps({"workload": "gn", **gn_step_times()}), flush=True)
        return
    if args.workload == "ip":
        log("induced polarisation only (IP2D, IP3D, DCIP2D)")
        print(json.dumps({"workload": "ip", **ip_times(args.steps, args.warmup)}), flush=True)
        return
    if args.workload == "lines_one_plan":
        log("uneven batch in one plan only (16 lines, 3k-40k cells)")
        print(json.dumps({"workload": "lines_one_plan",
                          **lines_times(args.steps, args.warmup, per_line=False)}), flush=True)
        return
    if args.workload == "lines":
        log("uneven batch only (16 lines, 3k-40k cells)")
        print(json.dumps({"workload": "lines", **lines_times(args.steps, args.warmup)}),
              flush=True)
        return
    pr = torch.cuda.get_device_properties(0)
    log(f"device: {pr.name} CUs={pr.multi_processor_count} "
        f"L2={getattr(pr, 'L2_cache_size', 0) >> 20} MiB")
    log("building engines (config: 25k-cell line, 12 ky requested)")
    mesh, survey, m, e64, emx, mt, dobs, w, t_build = build_engines(args.refine)
    info = e64.solver.info()
    log(f"plan: n={info['n']} nsup={info['nsup']} depth={info['ndepth']} "
        f"max_front={info['max_front']} nnz_L={info['nnz_L']} nky={e64.nky} nS={e64.nS}")
    phi64, g64 = eval_grad(e64, mt, dobs, w)
    phimx, gmx = eval_grad(emx, mt, dobs, w)
    with torch.no_grad():
        d64 = e64.dpred(mt)
        dmx = emx.dpred(mt)
    err_d = float(torch.linalg.norm(dmx - d64) / torch.linalg.norm(d64))
    err_g = float(torch.linalg.norm(gmx - g64) / torch.linalg.norm(g64))
    rep = emx.solver.check_report()
    check = {"status": rep["status"], "refinement_passes": rep["passes"],
             "relative_residual_max": float(rep["rho"].max()),
             "tolerance": emx.solver.check_tol, "fp64_fallbacks": len(emx.solver.fallbacks)}
    log(f"mixed vs fp64: dpred {err_d:.2e} grad {err_g:.2e}; solve check {check}")
    tiles = None
    if not args.no_tune:
        tiles, t_tune = tune_tiles({"fp64": e64, "mixed": emx}, mt, dobs, w)
        tiles["tune_s"] = t_tune
        log(f"tile tuning {tiles}")
    log("timing fp64")
    ms64 = time_steps(e64, mt, dobs, w, args.steps, args.warmup)
    log(f"fp64 eval+grad {ms64:.2f} ms")
    log("timing mixed")
    msmx = time_steps(emx, mt, dobs, w, args.steps, args.warmup)
    log(f"mixed eval+grad {msmx:.2f} ms")
    log("electrode basis (61 electrode columns instead of 849 sources)")
    eb = electrode_basis_times(mesh, survey, mt, dobs, w, d64, g64, args.steps, args.warmup)
    log(f"electrode basis {eb}")
    roof = sweep_roofline(emx, mt, dobs, w, args.refine)
    roof64 = sweep_roofline(e64, mt, dobs, w, 0)
    roof["traffic"] = measured_traffic("k_bwd_wave<float>")
    roof64["traffic"] = measured_traffic("k_bwd_wave<double>")
    froof, froof64 = factor_roofline(emx, mt), factor_roofline(e64, mt)
    log(f"factorisation: mixed {froof}, fp64 {froof64}")
    log("linearisation (FP64): J v, J^T v, J, diag(J^T D J)")
    lin = linearization_times(e64, mt)
    log(f"linearisation {lin}")
    g3 = lb = rc = gn = ipt = None
    if args.workload == "all":
        log("3-D grid (TorchDC3D, 56 x 56 x 38 cells, 200 sources)")
        g3 = grid3d_times(args.steps, args.warmup, tune=not args.no_tune)
        log(f"3-D {g3}")
        log("uneven batch (16 lines, 3k-40k cells, leaf autotuned per size class)")
        lb = lines_times(args.steps, args.warmup)
        log(f"lines {lb}")
        log("Gauss-Newton step (62k cells, 2,430 data, FP64)")
        gn = gn_step_times()
        log(f"gn step {gn}")
        log("induced polarisation (IP2D, IP3D, DCIP2D)")
        ipt = ip_times(args.steps, args.warmup)
        log(f"ip {ipt}")
        log("the reference's published eval+grad case (5k cells, nky 11)")
        rc = readme_case_times(args.steps, args.warmup)
        log(f"readme case {rc}")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (the oracle's restatement of the reference CPU path, all "
            "wavenumbers, median of 3, sequential and threaded)")
        res, ncores, model = cpu_baseline(mesh, survey, m, e64.kys, e64.wq_np,
                                          dobs.cpu().numpy(), w.cpu().numpy())
        seq, thr = res["sequential"], res["threaded"]
        cpu = {"value": round(seq[0], 1), "unit": "ms", "cores": 1, "kind": "port",
               "sample": f"the whole eval+grad: SuperLU (SciPy defaults) factorisation, forward "
                         f"and adjoint solves of all {e64.nS} source columns and the gradient "
                         f"at all {e64.nky} wavenumbers, one after another; median of "
                         f"{len(seq[1])} runs {[round(t) for t in seq[1]]} ms",
               "threaded": {"value": round(thr[0], 1), "unit": "ms", "cores": thr[2],
                            "runs_ms": [round(t) for t in thr[1]],
                            "sample": f"the same with the wavenumbers on {thr[2]} threads"},
               "host": {"cores_available": ncores, "cpu_model": model}}
        log(f"cpu baseline {cpu}")
    out = {
        "metric": METRIC, "value": round(msmx, 3), "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(msmx, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 factor, sweeps and fields; f64 assembly, projection, ky sum, gather sums",
        "data": "synthetic",
        "config": {"workload": "2.5-D line: 25,088 cells, 61 electrodes DD+WS, "
                               f"{e64.nS} sources, {e64.nD} data, nky={e64.nky} "
                               "(12 requested, electrode-fitted)",
                   "cells": e64.nC, "nodes": e64.nN, "sources": e64.nS, "data": e64.nD,
                   "nky": e64.nky, "refine_passes": args.refine},
        "fp64_ms": round(ms64, 3), "mixed_speedup_vs_fp64": round(ms64 / msmx, 3),
        "mixed_dpred_rel_err": err_d, "mixed_grad_rel_err": err_g,
        "mixed_dpred_within_1e-6": bool(err_d <= 1e-6),
        "mixed_solve_check": check,
        "build_s": round(t_build, 2),
        "tile_tuning": tiles,
        "roofline": roof, "roofline_fp64": roof64,
        "factor_roofline": froof, "factor_roofline_fp64": froof64,
        "electrode_basis": eb,
        "linearization": lin,
        "grid3d": g3,
        "uneven_lines": lb,
        "reference_readme_case": rc,
        "gn_step": gn,
        "induced_polarisation": ipt,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(cpu["value"] / msmx, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
