"""dcx_jacobian_rows_noalloc (the allocation-free variant of
dcx_jacobian_rows) against dcx_jacobian_rows, through the C-ABI: bit-identical
rows of J on two calls in a row, the second with fewer data, and a clear
refusal of a block larger than its fixed workspace (checked on the host
before any launch)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _inputs(lin, a, b):
    eng, g = lin.eng, lin.g
    U = lin._fields_gpu()
    r = b - a
    sel = (lin._pd_datum >= a) & (lin._pd_datum < b)
    E = torch.zeros(eng.nN, r, dtype=torch.float64, device=eng.device)
    E.index_put_((lin._pd_row[sel], lin._pd_datum[sel] - a), lin._pd_val[sel], accumulate=True)
    lin._ensure()
    Y = g.solve_perm(E, r, sparse=True, shared=True)
    return U, Y, lin._src32[a:b].contiguous()


def _rows(fn, lin, U, Y, src, r):
    from simpeg_dc_torch_amd import _lib
    eng, g = lin.eng, lin.g
    Jt = torch.full((eng.nC, r), float("nan"), dtype=torch.float64, device=eng.device)
    _lib.check(fn(g.lmap, g.solver._h, _lib.ptr(Y), r, _lib.ptr(U), eng.nS, _lib.ptr(src),
                  _lib.ptr(g.wq), _lib.ptr(lin.sig), _lib.ptr(Jt), _lib.DCX_PERMUTED,
                  g.stream()), "jacobian_rows")
    return Jt


def test_jacobian_rows_noalloc_matches():
    from simpeg_dc_torch_amd import Linearization2D, TorchDC2D, _lib, workloads
    mesh, survey, m = workloads.suite_case()
    eng = TorchDC2D(mesh, survey, nky=6, device=DEV)
    lin = Linearization2D(eng, torch.tensor(m, device=DEV))
    lib = _lib.load()
    with torch.no_grad(), torch.cuda.device(eng.device):
        for a, b in ((0, min(40, eng.nD)), (3, 10)):       # the second call smaller
            print(f"noalloc rows {a}..{b}", flush=True)
            U, Y, src = _inputs(lin, a, b)
            ref = _rows(lib.dcx_jacobian_rows, lin, U, Y, src, b - a)
            out = _rows(lib.dcx_jacobian_rows_noalloc, lin, U, Y, src, b - a)
            torch.cuda.synchronize()
            assert torch.isfinite(out).all()
            assert torch.equal(out, ref)
        # a block beyond the 64 MiB workspace is refused before any launch
        per_row = (2 * eng.nnzA + 2 * (mesh.shape_cells[0] + 2 * mesh.shape_cells[1])) * 8
        huge = int((64 << 20) // per_row) + 1000
        with pytest.raises(RuntimeError, match="allocation-free variant holds 64 MiB"):
            _rows(lib.dcx_jacobian_rows_noalloc, lin, U, Y, src, huge)
