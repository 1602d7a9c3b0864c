/* dcx — C-ABI of the MI355X-native forward-and-adjoint path for 2.5-D / 3-D
 * DC-resistivity modelling.
 *
 * Every entry point returns 0 on success and a negative status on failure;
 * dcx_last_error() then gives the reason (thread-local).  No C++ exception
 * crosses this boundary.  Device pointers are plain device addresses (for
 * example torch's tensor.data_ptr()); host pointers are marked "host".  Every
 * call that launches work takes the HIP stream to order it on (void* so that
 * the header needs no HIP include; pass torch.cuda.current_stream().cuda_stream).
 * One plan must not be used from two threads or streams at once.
 *
 * Layouts: dense right-hand sides and solutions are row-major (n x k) blocks,
 * rows = matrix rows, k columns contiguous; a batch of them is nbatch such
 * blocks back to back.  Values of the nbatch matrices are (nbatch x nnz), in
 * the order of the CSR pattern given to dcx_plan_create.
 */
#ifndef DCX_H_
#define DCX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dcx_plan dcx_plan;
typedef struct dcx_lmap dcx_lmap;
typedef struct dcx_survey dcx_survey;

enum {
  DCX_OK = 0,
  DCX_ERR_ARG = -1,
  DCX_ERR_HIP = -2,
  DCX_ERR_STATE = -3,
  DCX_ERR_NUMERIC = -4,
  DCX_ERR_INTERNAL = -5
};

/* factor precision */
enum { DCX_FP64 = 0, DCX_FP32 = 1 };
/* keys of dcx_plan_set_tuning / dcx_plan_get_tuning */
enum { DCX_TUNE_SWEEP_TC = 1, DCX_TUNE_SPLITK_BELOW = 2, DCX_TUNE_WIDE_BELOW = 3,
       DCX_TUNE_PANEL_WIDTH = 4, DCX_TUNE_ROWSPLIT_BELOW = 5 };

/* dcx_solve flags */
enum {
  DCX_SHARED_RHS = 1,  /* b is one (n x k) block shared by every batch member */
  DCX_PERMUTED = 2,    /* b and x are in the plan's elimination order */
  DCX_SPARSE_RHS = 4,  /* b is zero outside the rows set by dcx_set_rhs_support */
  DCX_F32_FIELDS = 8,  /* the field blocks (dcx_solve's x, dcx_grad_values' y and
                          x, dcx_survey_dpred's U) are float, not double */
  DCX_CHECK_DEFER = 16, /* the solve's check is launched but not waited for: its
                          verdict (and any refinement) comes from
                          dcx_plan_check_wait; b and x must stay valid until then */
  DCX_SPLIT_BWD = 32   /* with DCX_SPARSE_RHS: the backward sweep computes the
                          supernodes whose subtree holds a support row on the call's
                          stream, the others on the plan's second stream after them;
                          x is final on the support rows (and their supernodes) when
                          the call's stream gets there, on every row only after
                          dcx_plan_join.  Ignored (one stream) unless x is the
                          sweeps' own block: FP64 with DCX_PERMUTED, or DCX_F32_FIELDS
                          without fixed refinement passes */
};

typedef struct {
  int n, nnz, nbatch, nsup, ndepth, max_front, max_ns, precision;
  int64_t nnz_L;              /* stored factor entries (dense supernode blocks) */
  double factor_flops;        /* per matrix */
  double solve_flops_per_rhs; /* forward + backward sweep, per column */
  int64_t device_bytes;       /* factor + front storage held by the plan */
  int64_t sweep_rows_nf;      /* sum of front rows over the level-scheduled supernodes */
  int64_t sweep_rows_ns;      /* sum of their own rows */
  int64_t panel_size;         /* factor panel elements per batch member */
  int level_launches;         /* level launches per sweep */
  int nfused;                 /* fused bottom subtrees */
  int64_t buffer_generation;  /* bumped whenever a workspace buffer is reallocated: work
                                 recorded in a HIP graph before a change holds stale
                                 addresses */
} dcx_plan_info;

typedef struct dcx_symbolic dcx_symbolic;

/* Read-only view of a host symbolic analysis (arrays owned by the object).
 * Supernodes are in postorder; columns use the elimination numbering. */
typedef struct {
  int n, nnz, nsup, ndepth, max_front, max_ns;
  const int32_t *perm, *first, *ns, *rs, *parent, *depth;
  const int32_t *rptr, *ridx, *cptr, *cidx, *relpos;
  const int32_t *aptr, *a_e, *a_row, *a_col, *lvl_ptr, *lvl_sup;
  int64_t nnz_L;
  double factor_flops, solve_flops_per_rhs;
  /* fused bottom subtrees of the sweeps */
  int nfused, nfused_lv, nfused_mem, nfused_grp, f_max_lds_rows;
  const int32_t *f_root, *f_lo, *f_hi, *f_lv_begin, *f_lv_end, *f_tb_rows;
  const int32_t *f_lv, *f_mem, *f_gptr, *f_gb, *toff, *lridx, *ulvl_ptr, *ulvl_sup;
} dcx_symbolic_view;

const char* dcx_last_error(void);

/* Device memory of plans, maps and surveys (factors, fronts, workspaces,
 * device tables) comes from hipMalloc unless an allocator is registered:
 * alloc(bytes, device, ctx) returns device memory on `device` (the current
 * device) or NULL, and free_fn(ptr, bytes, device, ctx) takes it back after
 * the library has waited for the device to be idle.  Registering (or
 * clearing, with two NULLs) affects later allocations; each allocation is
 * freed the way it was made.  The ctypes layer registers torch's caching
 * allocator, so plan memory shows in torch.cuda.memory_allocated and
 * torch.cuda.empty_cache can reclaim it: the ownership the reference could not
 * have with cuDSS, whose memory sat outside torch (dctorch/simulation3d.py:
 * 160-169).  The fixed array of dcx_jacobian_rows_noalloc is not allocated. */
typedef void* (*dcx_alloc_fn)(size_t bytes, int device, void* ctx);
typedef void (*dcx_free_fn)(void* ptr, size_t bytes, int device, void* ctx);
int dcx_set_allocator(dcx_alloc_fn alloc, dcx_free_fn free_fn, void* ctx);

/* Host-only symbolic analysis (no device work): the same ordering and
 * supernode structure dcx_plan_create builds, for inspection and tests. */
int dcx_symbolic_analyse(int n, const int32_t* crow, const int32_t* col,
                         const double* coords, int dim, int leaf_max, int fuse_rows,
                         dcx_symbolic** out);
int dcx_symbolic_view_get(const dcx_symbolic* sym, dcx_symbolic_view* out);
int dcx_symbolic_destroy(dcx_symbolic* sym);
int dcx_device_count(int* out);

/* Symbolic phase, once per sparsity pattern: nested-dissection ordering,
 * supernodal front structure, device maps.  Replaces cuDSS plan() in
 * CuDSSBatchSolver.__init__ (reference dctorch/solver.py:187-213).
 * crow/col: host, the FULL symmetric pattern (n+1, nnz) with a stored
 * diagonal.  coords: host (n x dim) node coordinates or NULL (graph-only
 * ordering).  precision: DCX_FP64 or DCX_FP32 (FP32 factor; dcx_solve then
 * refines in FP64).  Runs on the current HIP device. */
int dcx_plan_create(int n, const int32_t* crow, const int32_t* col, int nbatch,
                    const double* coords, int dim, int leaf_max, int precision,
                    dcx_plan** out);
int dcx_plan_destroy(dcx_plan* plan);
int dcx_plan_get_info(const dcx_plan* plan, dcx_plan_info* out);
/* host (n): perm[new] = old */
int dcx_plan_get_perm(const dcx_plan* plan, int32_t* perm);
/* Record HIP events around the forward and backward sweeps of every solve
 * (on the solve's stream); dcx_plan_sweep_times sums them (synchronising on
 * the last event) over the solves since timing was enabled. */
int dcx_plan_set_timing(dcx_plan* plan, int enable);
int dcx_plan_sweep_times(dcx_plan* plan, double* fwd_ms, double* bwd_ms, int* nsweeps);
/* The plan's own streams beside the caller's (default on): the factorisation
 * of dcx_factor_solve overlapped with the first forward sweep on a second
 * stream, the split backward sweep (DCX_SPLIT_BWD) and deferred checks on a
 * check stream.  Off (enable = 0), every kernel of a call goes on the caller's
 * stream in order (DCX_SPLIT_BWD is ignored, a deferred check runs after the
 * solve on that stream); for callers that run many plans on streams of their
 * own, where each plan's two extra streams would share the device's few
 * hardware queues with them.  It may be switched between any two calls.  It
 * has no counterpart in the reference. */
int dcx_plan_set_overlap(dcx_plan* plan, int enable);
/* Kernel-variant thresholds of one plan (defaults from the environment at
 * dcx_plan_create): DCX_TUNE_SWEEP_TC, the widest column tile of the in-LDS
 * sweeps (16, 32, 64); DCX_TUNE_SPLITK_BELOW, backward-sweep levels with fewer
 * wavefronts take the split-K variant; DCX_TUNE_WIDE_BELOW, factor levels with
 * fewer workgroups take 1024-thread fronts; DCX_TUNE_PANEL_WIDTH (16 or 32),
 * columns per panel of the in-LDS factor kernel; DCX_TUNE_ROWSPLIT_BELOW,
 * backward-sweep levels of supernodes of at most 64 rows with fewer wavefronts
 * (of the column-split kernel) take the row-split variant, four wavefronts per
 * column block (checked before the split-K threshold).  Every value selects among
 * kernels that are correct for any level; only speed and summation order
 * change.  No counterpart in the reference (cuDSS chooses internally behind
 * dctorch/solver.py:187-213); a tile autotuner times them per size class. */
int dcx_plan_set_tuning(dcx_plan* plan, int key, int value);
int dcx_plan_get_tuning(const dcx_plan* plan, int key, int* value);

/* LDS row budget of the fused bottom subtrees of the sweeps (0: none);
 * rebuilds the fused structure (host) and its device tables. */
int dcx_plan_set_fusion(dcx_plan* plan, int fuse_rows);
/* number of FP64 refinement passes after the first FP32 solve (mixed mode) */
int dcx_plan_set_refinement(dcx_plan* plan, int passes);

/* Residual check of every solve (on by default; tol <= 0 turns it off;
 * skipped while the stream is being captured).  After each dcx_solve /
 * dcx_factor_solve the FP64 residual R = B - A_b X of the first sample_cols
 * <= 64 columns is formed and rho_b = |R| / |B| (Frobenius norms over the
 * sample) and the pivot flags of the factorisation are read back (one stream
 * synchronisation, or none with DCX_CHECK_DEFER).  An FP32 plan whose rho_b
 * exceeds tol for some b adds FP64-residual refinement passes, up to
 * max_passes; then, or at once for a non-positive pivot or an FP64 plan, the
 * call returns DCX_ERR_NUMERIC with the solution of the last pass in x, and
 * dcx_plan_get_check names the failing systems (the caller refactorises
 * those in FP64).  The residual bounds the backward error, not the forward
 * error: the check catches a factorisation that broke down or lost accuracy
 * well beyond rounding, not a 1e-6 change in the data.  Defaults: tol 1e-3
 * (FP32) / 1e-8 (FP64), or DCX_CHECK_TOL at plan creation; max_passes 2;
 * sample_cols 8.  Counterpart of the reference's guarded solve, a residual
 * check followed by a fallback (dctorch/optimization.py:753-761); cuDSS
 * reports pivot failures itself. */
int dcx_plan_set_check(dcx_plan* plan, double tol, int max_passes, int sample_cols);
/* Verdict of the oldest pending deferred check (DCX_CHECK_DEFER; at most two
 * are pending, a third solve completes the oldest first): waits for it, runs
 * the refinement passes a failing FP32 solve gets, and returns DCX_OK or
 * DCX_ERR_NUMERIC as the solve itself would have; DCX_OK when nothing is
 * pending.  A factorisation of the plan (and a solve of an FP32 plan with
 * fixed refinement passes) completes the pending checks first.  No reference
 * counterpart (see dcx_plan_set_check). */
int dcx_plan_check_wait(dcx_plan* plan);
/* Orders `stream` after the second half of the last split backward sweep
 * (DCX_SPLIT_BWD) of the plan, so every row of its x may be read there; no
 * host synchronisation.  No reference counterpart: the work split is this
 * library's (dpred needs the fields on the receivers only, which the first
 * half gives; the rest overlaps the misfit and the adjoint solve). */
int dcx_plan_join(dcx_plan* plan, void* stream);
/* Result of the last check (host arrays of nbatch, each may be NULL): rho
 * (the tested relative residual), eta (the normwise backward error
 * |R| / (|A| |X| + |B|), for information), pivot_fail; refinement passes run;
 * status (0 none yet, 1 passed, 2 failed, 3 skipped under stream capture or
 * with the check off). */
int dcx_plan_get_check(const dcx_plan* plan, double* rho, double* eta, int* pivot_fail,
                       int* passes, int* status);
/* Fault injection for tests: later factorisations of batch member `batch`
 * factor (1 + rel) A_b instead of A_b (the solve's FP64 residual keeps A_b),
 * a factor with backward error about rel / 2; rel = 0 disarms.  No reference
 * counterpart (the reference's failure test clobbers a shared factor,
 * dctorch/tests/test_linearization.py:56-57). */
int dcx_plan_inject_fault(dcx_plan* plan, int batch, double rel);

/* Numeric factorisation of all nbatch matrices (values: device, FP64,
 * nbatch x nnz).  Replaces CuDSSBatchSolver.factorize
 * (dctorch/solver.py:215-221).  The values are copied (the FP64 copy feeds the
 * refinement residual). */
int dcx_factorize(dcx_plan* plan, const double* values, void* stream);

/* Multi-RHS solve of every batch member, FP64 in and out.  Replaces
 * CuDSSBatchSolver.solve (dctorch/solver.py:223-230).  b: device, (n x k)
 * with DCX_SHARED_RHS, else (nbatch x n x k); x: device (nbatch x n x k),
 * double, or float with DCX_F32_FIELDS (FP32 plans and DCX_PERMUTED only:
 * x is then the sweeps' own working block, the FP64 refinement passes, if
 * any, accumulate in a plan buffer and are rounded into x at the end).
 * b and x may alias when not shared (double x only). */
int dcx_solve(dcx_plan* plan, const double* b, void* x, int k, int flags,
              void* stream);

/* dcx_factorize followed by dcx_solve, with the factorisation and the
 * solve's forward sweep overlapped: the forward sweep of each tree level runs
 * on a second stream of the plan (default priority) as soon as that level is
 * factored; the backward sweep and everything after it are ordered on
 * `stream`.  Same arguments and results as the two calls.  Replaces the
 * factorize + solve pair of _SparseSolveBatchFn.forward
 * (dctorch/solver.py:251-259). */
int dcx_factor_solve(dcx_plan* plan, const double* values, const double* b, void* x, int k,
                     int flags, void* stream);

/* Rows (host, original numbering) outside of which right-hand sides passed
 * with DCX_SPARSE_RHS are zero; lets the forward sweep skip the supernodes
 * whose subtree holds none of them. */
int dcx_set_rhs_support(dcx_plan* plan, const int32_t* rows, int nrows);

/* Backward gather of the solve: gv[b][e] = -scale[b] * sum_c y[b][row_e][c] *
 * x[b][col_e][c] over the pattern's entries e.  Replaces the chunked gather of
 * _SparseSolveBatchFn.backward (dctorch/solver.py:268-276).  y, x: device
 * (nbatch x n x k), double, or float with DCX_F32_FIELDS (the sums are FP64
 * either way); scale: device (nbatch) or NULL (= 1); gv: device
 * (nbatch x nnz).  With DCX_PERMUTED the blocks are in elimination order. */
int dcx_grad_values(dcx_plan* plan, const void* y, const void* x, int k,
                    const double* scale, double* gv, int flags, void* stream);

/* y[b] = alpha * A_b x[b] for every batch member, A_b the plan's pattern with
 * the values vals[b] (device, nbatch x nnz, the caller's CSR order); x, y:
 * device (nbatch x n x k), rows in elimination order with DCX_PERMUTED.
 * Replaces the per-wavenumber sparse products -A_k(dsigma) U_k of
 * Linearization2D.Jvec (dctorch/simulation2d.py:689-710). */
int dcx_spmm(dcx_plan* plan, const double* vals, const double* x, double* y, int k,
             double alpha, int flags, void* stream);

/* Lane-layout probe of the 16x16x4 matrix instruction (fp32 = 0: FP64):
 * C (16x16, row-major) = A (16x4, row-major) * B (4x16, row-major), one
 * wavefront.  Device pointers.  A test of the instruction layout the sweep
 * kernels assume; no counterpart in the reference. */
int dcx_mma_probe(int fp32, const void* A, const void* B, void* C, void* stream);

/* Row permutations between the caller's order and the elimination order:
 * out[b][new][c] = in[b][perm[new]][c] (to_elim=1) or the inverse. */
int dcx_permute_rows(dcx_plan* plan, const double* in, double* out, int k,
                     int nblocks, int to_elim, void* stream);

/* ---- the sigma -> A-values assembly map (factored form) -------------------
 * A_k(sigma).values[e] = sum_p (le[p] + ky_k^2 ln[p]) sigma[cell[p]]
 *                        + sum_q bc[k][q] sigma[bc_cell[q]]  (bc_e[q] == e)
 * Replaces the stacked COO map Lstack and torch.sparse.mm of
 * TorchDC2D._extract / dpred (dctorch/simulation2d.py:249-260, :392). */
int dcx_lmap_create(int nnz, int ncell, int nky, const double* ky2,
                    const int32_t* ptr, const int32_t* cell, const double* le,
                    const double* ln, int nbc, const int32_t* bc_e,
                    const int32_t* bc_cell, const double* bc_coef,
                    dcx_lmap** out);
int dcx_lmap_destroy(dcx_lmap* map);
/* Wavenumber groups: entry e of the values takes ky^2 = ky2[k * ngroup +
 * group[e]] in system k (host arrays: group has nnz entries, ky2 nky x
 * ngroup), for a map that stacks independent problems with their own
 * wavenumbers (the uneven batch of lines in one plan).  No counterpart in the
 * reference, whose engines hold one problem each (dctorch/simulation2d.py:
 * 249-260 builds one ky list per engine).  Maps with groups are refused by
 * dcx_jacobian_rows. */
int dcx_lmap_set_groups(dcx_lmap* map, int ngroup, const int32_t* group,
                        const double* ky2);
/* values (device, nky x nnz) from sigma (device, ncell) */
int dcx_assemble(dcx_lmap* map, const double* sigma, double* values, void* stream);
/* grad_m[c] = -sigma[c] * (L^T gv)[c]: the transposed map applied to the value
 * gradient (device, nky x nnz), chained through sigma = exp(-m).  Replaces the
 * sparse-mm backward + exp backward (dctorch/simulation2d.py:392, autograd). */
int dcx_lmap_transpose(dcx_lmap* map, const double* gv, const double* sigma,
                       double* grad_m, void* stream);

/* Row mix per segment: h[b][r][:] = g[b][r][:] w[b][l] for the rows r of
 * segment l, rows [seg_rows[l], seg_rows[l+1]) (host array of nseg + 1
 * ascending row offsets within 0..n; rows outside every segment are left
 * untouched).  g, h: device (nbatch x n x k); w: device (nbatch x nseg x k x
 * k), row-major; double, or float with DCX_F32_FIELDS.  k <= 128, nseg <=
 * 256.  The adjoint field of the electrode basis when the systems hold
 * several independent lines, each with its own k x k data matrix: replaces
 * one small batched GEMM per line.  No reference counterpart (the reference
 * simulates one line per call: dpred, dctorch/simulation2d.py:391-396). */
int dcx_rows_mix(const void* g, const void* w, void* h, int nbatch, int n, int k, int nseg,
                 const int32_t* seg_rows, int flags, void* stream);

/* Rows of the Jacobian of dpred for r data at once, transposed:
 * J[c][j] = d dpred_j / d m_c (m = log rho on all cells).  Y: device
 * (nbatch x n x r), A_b^{-1} applied to the data's projection rows (Pd^T
 * columns); U: device (nbatch x n x nS), the forward fields; src: device (r)
 * source of each datum; wq: device (nbatch) quadrature weights; sigma: device
 * (ncell); J: device (ncell x r).  With DCX_PERMUTED, Y and U rows are in the
 * plan's elimination order.  The map and the plan must describe the same
 * systems (nky = nbatch).  Replaces _rows_of_J's pairwise gather and L^T
 * (dctorch/simulation2d.py:734-776, :808-818) and TorchDC2D.Jmatrix's block
 * loop (:439-518). */
int dcx_jacobian_rows(dcx_lmap* map, dcx_plan* plan, const double* Y, int r, const double* U,
                      int nS, const int32_t* src, const double* wq, const double* sigma,
                      double* J, int flags, void* stream);

/* dcx_jacobian_rows without allocating: the same arguments, kernels and
 * results (bit for bit), with the pairwise-sum workspace in a fixed 64 MiB
 * device array of the library (one per GPU, held until the process ends)
 * instead of a buffer grown with hipMalloc.  Refuses (DCX_ERR_ARG, naming the
 * largest r that fits) a block needing more than 64 MiB:
 * (2 nnz + max(nbc, 1)) r 8 bytes, e.g. 2 MB per datum on the 25,088-cell
 * line, at most 33 data per call.  Calls that overlap on one GPU share the
 * array and race: order them on one stream.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). */
int dcx_jacobian_rows_noalloc(dcx_lmap* map, dcx_plan* plan, const double* Y, int r,
                              const double* U, int nS, const int32_t* src, const double* wq,
                              const double* sigma, double* J, int flags, void* stream);

/* ---- receivers and adjoint sources ----------------------------------------
 * Pd (nD x nN, CSR, host) and the source index of each datum.  Replaces the
 * projection torch.sparse.mm(Pd, Ubar)[iD, sidx] and its backward
 * (dctorch/simulation2d.py:394-395). */
int dcx_survey_create(dcx_plan* plan, int nD, int nS, const int32_t* pd_ptr,
                      const int32_t* pd_col, const double* pd_val,
                      const int32_t* sidx, dcx_survey** out);
int dcx_survey_destroy(dcx_survey* sv);
/* d[i] = sum_k w[k] sum_j Pd[i,j] U[k][j][sidx_i - s0] for the data whose
 * source lies in [s0, s1) (the others are left untouched); U device
 * (nky x n x k), double or float (DCX_F32_FIELDS), rows in elimination order
 * (DCX_PERMUTED) or original order. */
int dcx_survey_dpred(dcx_survey* sv, const void* U, int k, int nky,
                     const double* w, int s0, int s1, double* d, int flags,
                     void* stream);
/* G[j][s - s0] = sum over data i of source s in [s0, s1) of Pd[i,j] v[i]:
 * device (n x k), zero elsewhere; rows in elimination order (DCX_PERMUTED) or
 * original order.  With DCX_SPARSE_RHS the block is not cleared first: only
 * the survey's target entries (the receiver rows of each source's data, the
 * same set on every call with the same k, s0, s1 and order) are written, for
 * a G that is already zero elsewhere -- a block cleared once and reused. */
int dcx_survey_adjoint_rhs(dcx_survey* sv, const double* v, double* G, int k,
                           int s0, int s1, int flags, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DCX_H_ */
