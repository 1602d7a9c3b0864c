"""eval+grad benchmark of the 2.5-D forward-and-adjoint path on one GPU.

Workload: the 25,088-cell line (200 x 100 core of 1 m cells plus 12 padding
cells growing 1.3x; 61 electrodes at 2 m; dipole-dipole a = 1..3, n = 1..6
plus Wenner-Schlumberger a = 1..2, n = 1..8: 849 sources, 1,636 data;
12 electrode-fitted wavenumbers requested; m = log(rho) a Gaussian random
field, seed 1).  One step = misfit(m).backward(): assembly, factorisation of
every wavenumber, forward solve of all sources, projection, misfit, adjoint
solve, gradient gather and transposed map -- everything resident on the GPU.

Prints one JSON line.  `value` is the mixed-precision (FP32 factor + FP64
refinement) eval+grad time in ms; the FP64 time, the errors of the mixed
result against FP64, the roofline of the dominant kernel and the CPU
baseline (the oracle's SuperLU restatement of the reference CPU path on a
bounded sample) are reported beside it.
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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


def kernel_roofline(eng, mt, dobs, w, reps=5):
    """Average duration of the gradient gather-dot (one launch per eval+grad,
    HBM-bound) on the stream it runs on, and its algorithmic bytes."""
    g = eng._gpu
    with torch.no_grad():
        sig = torch.exp(-mt)
        vals = g.assemble(sig)
        g.solver.factorize(vals)
        U = g.solve_perm(g.q_perm, eng.nS)
        Lam = g.solve_perm(g.q_perm, eng.nS)
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        g.solver.grad_values(Lam, U, scale=g.wq, permuted=True)
        ev[0].record(st)
        for _ in range(reps):
            g.solver.grad_values(Lam, U, scale=g.wq, permuted=True)
        ev[1].record(st)
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / reps
    nky, n, k, nnz = eng.nky, eng.nN, eng.nS, eng.nnzA
    algo = nky * (2 * 8 * n * k + 8 * nnz)
    return ms, algo


def cpu_baseline(mesh, survey, m, kys, wq, sample_ky=1):
    """The oracle's SuperLU path (reference CPU solver semantics) on a sample
    of `sample_ky` wavenumbers, scaled to all of them."""
    from oracle.dc25d import Problem2D
    import scipy.sparse.linalg as spla
    O = Problem2D(mesh.h[0], mesh.h[1], mesh.origin, survey, kys, wq)
    sigma = np.exp(-m)
    rng = np.random.default_rng(0)
    G = O._adjoint_rhs(rng.standard_normal(O.nD))
    t0 = time.perf_counter()
    for k in range(sample_ky):
        lu = spla.splu(O.A(sigma, k).tocsc())
        U = lu.solve(np.ascontiguousarray(O.q))
        Lam = lu.solve(np.ascontiguousarray(G))
        gl, gu = O.Grad @ Lam, O.Grad @ U
        _ = O.Qe_.T @ (gl * gu).sum(1) + O.Qn_.T @ (Lam * U).sum(1)
    t = time.perf_counter() - t0
    return 1e3 * t * len(kys) / sample_ky, t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--refine", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)
    pr = torch.cuda.get_device_properties(0)
    log(f"device: {pr.name} CUs={pr.multi_processor_count} "
        f"L2={getattr(pr, 'L2_cache_size', 0) >> 20} MiB "
        f"lds/block={getattr(pr, 'shared_memory_per_block', 0) >> 10} KiB "
        f"lds/CU={getattr(pr, 'shared_memory_per_multiprocessor', 0) >> 10} KiB")
    log("building engines (config: 25k-cell line, 12 ky)")
    mesh, survey, m, e64, emx, mt, dobs, w, t_build = build_engines(args.refine)
    info = e64.solver.info()
    log(f"plan: n={info['n']} nsup={info['nsup']} depth={info['ndepth']} "
        f"max_front={info['max_front']} nnz_L={info['nnz_L']} nky={e64.nky} nS={e64.nS}")
    # accuracy of mixed against fp64
    phi64, g64 = eval_grad(e64, mt, dobs, w)
    phimx, gmx = eval_grad(emx, mt, dobs, w)
    with torch.no_grad():
        d64 = e64.dpred(mt)
        dmx = emx.dpred(mt)
    err_d = float(torch.linalg.norm(dmx - d64) / torch.linalg.norm(d64))
    err_g = float(torch.linalg.norm(gmx - g64) / torch.linalg.norm(g64))
    log(f"mixed vs fp64: dpred {err_d:.2e} grad {err_g:.2e}")
    log("timing fp64")
    ms64 = time_steps(e64, mt, dobs, w, args.steps, args.warmup)
    log(f"fp64 eval+grad {ms64:.2f} ms")
    log("timing mixed")
    msmx = time_steps(emx, mt, dobs, w, args.steps, args.warmup)
    log(f"mixed eval+grad {msmx:.2f} ms")
    kms, kbytes = kernel_roofline(e64, mt, dobs, w)
    ach = kbytes / (kms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle SuperLU, 1 wavenumber sample)")
        cms, ct = cpu_baseline(mesh, survey, m, e64.kys, e64.wq_np)
        cpu = {"value": round(cms, 1), "unit": "ms", "cores": 1, "kind": "port",
               "sample": f"1 of {e64.nky} wavenumbers (factor + forward and adjoint solves "
                         f"of all {e64.nS} sources + gradient gather), {ct:.1f} s, "
                         f"scaled x{e64.nky}"}
    out = {
        "metric": METRIC, "value": round(msmx, 3), "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(msmx, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 factor+sweeps, f64 refinement/residual/gather",
        "data": "synthetic",
        "config": {"workload": "2.5-D line, 25,088 cells, 61 electrodes DD+WS, "
                               f"{e64.nS} sources, {e64.nD} data, nky={e64.nky} "
                               "(12 requested, electrode-fitted)",
                   "cells": e64.nC, "nodes": e64.nN, "sources": e64.nS, "data": e64.nD,
                   "nky": e64.nky, "refine_passes": args.refine},
        "fp64_ms": round(ms64, 3), "mixed_speedup_vs_fp64": round(ms64 / msmx, 3),
        "mixed_dpred_rel_err": err_d, "mixed_grad_rel_err": err_g,
        "build_s": round(t_build, 2),
        "roofline": {"kernel": "k_grad_values", "bound": "hbm", "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_ms": round(kms, 4)},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(cpu["value"] / msmx, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
