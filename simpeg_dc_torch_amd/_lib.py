"""ctypes binding of libdcx.so (the C-ABI declared in include/dcx.h).

The library is loaded from this package directory only.  A missing or
unloadable library raises ImportError with the reason the first time a GPU
object is built; the CPU path of the package never needs it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; DCX_LIBRARY names another build of the same library
LIB_PATH = os.environ.get("DCX_LIBRARY") or os.path.join(_HERE, "libdcx.so")

DCX_FP64, DCX_FP32 = 0, 1
DCX_SHARED_RHS, DCX_PERMUTED, DCX_SPARSE_RHS, DCX_F32_FIELDS = 1, 2, 4, 8
DCX_CHECK_DEFER, DCX_SPLIT_BWD = 16, 32
DCX_TUNE_SWEEP_TC, DCX_TUNE_SPLITK_BELOW, DCX_TUNE_WIDE_BELOW, DCX_TUNE_PANEL_WIDTH = 1, 2, 3, 4
DCX_TUNE_ROWSPLIT_BELOW = 5

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p


class PlanInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("nbatch", C.c_int), ("nsup", C.c_int),
                ("ndepth", C.c_int), ("max_front", C.c_int), ("max_ns", C.c_int),
                ("precision", C.c_int), ("nnz_L", C.c_int64), ("factor_flops", C.c_double),
                ("solve_flops_per_rhs", C.c_double), ("device_bytes", C.c_int64),
                ("sweep_rows_nf", C.c_int64), ("sweep_rows_ns", C.c_int64),
                ("panel_size", C.c_int64), ("level_launches", C.c_int), ("nfused", C.c_int),
                ("buffer_generation", C.c_int64)]


class SymbolicView(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("nsup", C.c_int), ("ndepth", C.c_int),
                ("max_front", C.c_int), ("max_ns", C.c_int)] + \
        [(f, _i32p) for f in ("perm", "first", "ns", "rs", "parent", "depth", "rptr", "ridx",
                              "cptr", "cidx", "relpos", "aptr", "a_e", "a_row", "a_col",
                              "lvl_ptr", "lvl_sup")] + \
        [("nnz_L", C.c_int64), ("factor_flops", C.c_double),
         ("solve_flops_per_rhs", C.c_double)] + \
        [(f, C.c_int) for f in ("nfused", "nfused_lv", "nfused_mem", "nfused_grp",
                                "f_max_lds_rows")] + \
        [(f, _i32p) for f in ("f_root", "f_lo", "f_hi", "f_lv_begin", "f_lv_end", "f_tb_rows",
                              "f_lv", "f_mem", "f_gptr", "f_gb", "toff", "lridx", "ulvl_ptr",
                              "ulvl_sup")]


# name -> (restype, argtypes); every entry point returns an int status
PROTOTYPES = {
    "dcx_last_error": (C.c_char_p, []),
    "dcx_set_allocator": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "dcx_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "dcx_symbolic_analyse": (C.c_int, [C.c_int, _i32p, _i32p, _f64p, C.c_int, C.c_int,
                                       C.c_int, C.POINTER(_vp)]),
    "dcx_symbolic_view_get": (C.c_int, [_vp, C.POINTER(SymbolicView)]),
    "dcx_symbolic_destroy": (C.c_int, [_vp]),
    "dcx_plan_create": (C.c_int, [C.c_int, _i32p, _i32p, C.c_int, _f64p, C.c_int, C.c_int,
                                  C.c_int, C.POINTER(_vp)]),
    "dcx_plan_destroy": (C.c_int, [_vp]),
    "dcx_plan_get_info": (C.c_int, [_vp, C.POINTER(PlanInfo)]),
    "dcx_plan_get_perm": (C.c_int, [_vp, _i32p]),
    "dcx_plan_set_refinement": (C.c_int, [_vp, C.c_int]),
    "dcx_plan_set_fusion": (C.c_int, [_vp, C.c_int]),
    "dcx_plan_set_check": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
    "dcx_plan_get_check": (C.c_int, [_vp, _f64p, _f64p, _i32p, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
    "dcx_plan_inject_fault": (C.c_int, [_vp, C.c_int, C.c_double]),
    "dcx_plan_check_wait": (C.c_int, [_vp]),
    "dcx_plan_join": (C.c_int, [_vp, _vp]),
    "dcx_plan_set_timing": (C.c_int, [_vp, C.c_int]),
    "dcx_plan_set_overlap": (C.c_int, [_vp, C.c_int]),
    "dcx_plan_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int]),
    "dcx_plan_get_tuning": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "dcx_plan_sweep_times": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int)]),
    "dcx_factorize": (C.c_int, [_vp, _vp, _vp]),
    "dcx_solve": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "dcx_factor_solve": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "dcx_spmm": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_int, _vp]),
    "dcx_jacobian_rows": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp,
                                    C.c_int, _vp]),
    "dcx_jacobian_rows_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.c_int, _vp, _vp,
                                            _vp, _vp, C.c_int, _vp]),
    "dcx_set_rhs_support": (C.c_int, [_vp, _i32p, C.c_int]),
    "dcx_rows_mix": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _i32p,
                               C.c_int, _vp]),
    "dcx_grad_values": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    "dcx_mma_probe": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp]),
    "dcx_permute_rows": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "dcx_lmap_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _f64p, _i32p, _i32p, _f64p,
                                  _f64p, C.c_int, _i32p, _i32p, _f64p, C.POINTER(_vp)]),
    "dcx_lmap_destroy": (C.c_int, [_vp]),
    "dcx_lmap_set_groups": (C.c_int, [_vp, C.c_int, _i32p, _f64p]),
    "dcx_assemble": (C.c_int, [_vp, _vp, _vp, _vp]),
    "dcx_lmap_transpose": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "dcx_survey_create": (C.c_int, [_vp, C.c_int, C.c_int, _i32p, _i32p, _f64p, _i32p,
                                    C.POINTER(_vp)]),
    "dcx_survey_destroy": (C.c_int, [_vp]),
    "dcx_survey_dpred": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp,
                                   C.c_int, _vp]),
    "dcx_survey_adjoint_rhs": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                         _vp]),
}

_lib = None
_load_error = None

# the library's device memory through torch's caching allocator
# (dcx_set_allocator); DCX_TORCH_ALLOCATOR=0 keeps plain hipMalloc
_ALLOC_T = C.CFUNCTYPE(C.c_void_p, C.c_size_t, C.c_int, C.c_void_p)
_FREE_T = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p)


def _torch_alloc(nbytes, device, ctx):
    try:
        import torch
        return int(torch.cuda.caching_allocator_alloc(int(nbytes), int(device)))
    except Exception:
        return None


def _torch_free(ptr, nbytes, device, ctx):
    try:
        import torch
        torch.cuda.caching_allocator_delete(int(ptr))
    except Exception:
        pass


_ALLOC_CB = _ALLOC_T(_torch_alloc)   # kept alive for the library's lifetime
_FREE_CB = _FREE_T(_torch_free)


def load():
    """The loaded library; ImportError (with the reason) when it is absent."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if _load_error is not None:
        raise ImportError(_load_error)
    if not os.path.exists(LIB_PATH):
        _load_error = (f"{LIB_PATH} is not built: run __graft_entry__.build() "
                       "(or python -m simpeg_dc_torch_amd._build)")
        raise ImportError(_load_error)
    try:
        lib = C.CDLL(LIB_PATH)
    except OSError as exc:
        _load_error = f"cannot load {LIB_PATH}: {exc}"
        raise ImportError(_load_error) from exc
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if os.environ.get("DCX_TORCH_ALLOCATOR", "1") != "0":
        lib.dcx_set_allocator(C.cast(_ALLOC_CB, C.c_void_p), C.cast(_FREE_CB, C.c_void_p), None)
    _lib = lib
    return lib


DCX_ERR_NUMERIC = -4


class NumericError(RuntimeError):
    """A solve failed its backward-error check or met a non-positive pivot
    (status DCX_ERR_NUMERIC); the plan's check report names the systems."""


def check(status, what=""):
    if status != 0:
        msg = _lib.dcx_last_error().decode(errors="replace") if _lib else "unknown"
        cls = NumericError if status == DCX_ERR_NUMERIC else RuntimeError
        raise cls(f"dcx {what} failed ({status}): {msg}")


def i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_f64p)


def ptr(t):
    """Device address of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_of(device):
    import torch
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def symbolic(crow, col, coords=None, leaf_max=16, fuse_rows=0):
    """Host-only symbolic analysis: a dict of numpy arrays (copies)."""
    lib = load()
    crow_a, crow_p = i32(crow)
    col_a, col_p = i32(col)
    n = crow_a.size - 1
    if coords is not None:
        co, co_p = f64(coords)
        dim = co.shape[1] if co.ndim == 2 else 1
    else:
        co_p, dim = None, 0
    h = _vp()
    check(lib.dcx_symbolic_analyse(n, crow_p, col_p, co_p, dim, int(leaf_max),
                                   int(fuse_rows), C.byref(h)), "symbolic analysis")
    try:
        v = SymbolicView()
        check(lib.dcx_symbolic_view_get(h, C.byref(v)), "symbolic view")
        ns = v.nsup

        def arr(p, m):
            return np.ctypeslib.as_array(p, shape=(m,)).copy() if m else np.zeros(0, np.int32)

        rptr = arr(v.rptr, ns + 1)
        cptr = arr(v.cptr, ns + 1)
        aptr = arr(v.aptr, ns + 1)
        lvl_ptr = arr(v.lvl_ptr, v.ndepth + 1)
        out = dict(n=v.n, nnz=v.nnz, nsup=ns, ndepth=v.ndepth, max_front=v.max_front,
                   max_ns=v.max_ns, nnz_L=v.nnz_L, factor_flops=v.factor_flops,
                   solve_flops_per_rhs=v.solve_flops_per_rhs,
                   perm=arr(v.perm, v.n), first=arr(v.first, ns), ns=arr(v.ns, ns),
                   rs=arr(v.rs, ns), parent=arr(v.parent, ns), depth=arr(v.depth, ns),
                   rptr=rptr, ridx=arr(v.ridx, rptr[-1]), cptr=cptr, cidx=arr(v.cidx, cptr[-1]),
                   relpos=arr(v.relpos, rptr[-1]), aptr=aptr, a_e=arr(v.a_e, aptr[-1]),
                   a_row=arr(v.a_row, aptr[-1]), a_col=arr(v.a_col, aptr[-1]),
                   lvl_ptr=lvl_ptr, lvl_sup=arr(v.lvl_sup, lvl_ptr[-1]))
        nf, nlv = v.nfused, v.nfused_lv
        out.update(nfused=nf, f_max_lds_rows=v.f_max_lds_rows,
                   f_root=arr(v.f_root, nf), f_lo=arr(v.f_lo, nf), f_hi=arr(v.f_hi, nf),
                   f_lv_begin=arr(v.f_lv_begin, nf), f_lv_end=arr(v.f_lv_end, nf),
                   f_tb_rows=arr(v.f_tb_rows, nf), f_lv=arr(v.f_lv, nlv + 1),
                   f_mem=arr(v.f_mem, v.nfused_mem), f_gptr=arr(v.f_gptr, nlv + 1),
                   f_gb=arr(v.f_gb, v.nfused_grp + 1), toff=arr(v.toff, ns),
                   lridx=arr(v.lridx, rptr[-1]), ulvl_ptr=arr(v.ulvl_ptr, v.ndepth + 1),
                   ulvl_sup=arr(v.ulvl_sup, int(arr(v.ulvl_ptr, v.ndepth + 1)[-1])))
        return out
    finally:
        lib.dcx_symbolic_destroy(h)
