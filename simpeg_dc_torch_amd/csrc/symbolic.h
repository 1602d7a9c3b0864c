// Host-side symbolic analysis of a sparse SPD pattern for the supernodal
// multifrontal solver: nested-dissection ordering, the separator tree (whose
// nodes are the supernodes), each supernode's front structure, and the maps
// the GPU kernels index with.  Pure C++; no HIP types.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dcx {

struct Symbolic {
  int n = 0;
  int nnz = 0;                      // entries of the full (symmetric) pattern
  std::vector<int> perm;            // new -> old
  std::vector<int> iperm;           // old -> new

  // supernodes in postorder (children before parents)
  int nsup = 0;
  std::vector<int> first;           // first column (new numbering)
  std::vector<int> ns;              // columns
  std::vector<int> rs;              // off-diagonal rows (ancestor nodes)
  std::vector<int> parent;          // -1 for a root
  std::vector<int> depth;           // root = 0
  std::vector<int> rptr, ridx;      // R_s (new numbering, ascending), CSR
  std::vector<int> cptr, cidx;      // children of s, CSR
  std::vector<int> relpos;          // per child: rows of R_c inside F_parent,
                                    // stored at rptr[c] .. rptr[c+1]
  std::vector<int> snode_of;        // new column -> supernode

  // lower-triangle entries of A assembled into fronts: for supernode s,
  // entries aptr[s]..aptr[s+1] give (source entry e, local row, local col)
  std::vector<int> aptr, a_e, a_row, a_col;

  // level schedule by depth: levels[d] lists the supernodes at depth d
  int ndepth = 0;
  std::vector<int> lvl_ptr, lvl_sup;

  // storage: panels (nf x ns, col-major) are disjoint per supernode; fronts
  // and solve-update rows are double-buffered by depth parity
  std::vector<int64_t> panel_off;   // per supernode
  int64_t panel_size = 0;
  std::vector<int64_t> front_off;   // within buffer depth % 2
  int64_t front_size[2] = {0, 0};
  std::vector<int64_t> upd_off;     // rows within buffer depth % 2
  int64_t upd_rows[2] = {0, 0};

  // permuted full pattern (for residual SpMV): prow_ptr/pcol, pval_idx maps to
  // the caller's value array; e_prow/e_pcol give each source entry's permuted
  // row/column (for the gradient gather)
  std::vector<int> pcrow, pcol, pval_idx;
  std::vector<int> e_prow, e_pcol;

  // fused bottom subtrees (sweeps): maximal subtrees whose LDS block
  // (subtree rows + root off-diagonal rows + widest sublevel) fits fuse_rows.
  // Sublevels are root first; each sublevel's members are ordered by colour
  // so that members of one colour have disjoint off-diagonal row sets.
  int nfused = 0;
  std::vector<int> f_root, f_lo, f_hi, f_lv_begin, f_lv_end, f_tb_rows;
  std::vector<int64_t> f_uoff;      // rows in the fused update buffer
  int64_t f_upd_rows = 0;
  int f_max_lds_rows = 0;
  std::vector<int> f_lv;            // sublevel -> first member (+ end sentinel per subtree)
  std::vector<int> f_mem;           // members
  std::vector<int> f_gptr, f_gb;    // colour groups: sublevel -> [f_gptr, f_gptr+1) into f_gb
  std::vector<int> in_fused;        // supernode -> subtree index or -1
  std::vector<int> fused_of_root;   // supernode -> subtree index if it is a fused root, else -1
  std::vector<int> toff;            // member's offset in the sublevel scratch
  std::vector<int> lridx;           // like ridx: row inside the subtree's LDS block
  std::vector<int> ulvl_ptr, ulvl_sup;  // levels of the supernodes above the subtrees

  // statistics
  int64_t nnz_L = 0;                // dense supernode blocks, lower incl. diag
  double factor_flops = 0.0;        // per matrix
  double solve_flops_per_rhs = 0.0; // forward + backward sweeps, per column
  int max_front = 0;
  int max_ns = 0;
};

// Build the analysis.  coords (n x dim, row-major) may be null: the ordering
// then falls back to level-structure bisection.  Returns "" or an error.
// fuse_rows: LDS row budget of a fused subtree (0: no fusion).
std::string analyse(int n, const int* crow, const int* col, const double* coords,
                    int dim, int leaf_max, int fuse_rows, Symbolic& out);
void build_fused(Symbolic& S, int fuse_rows);

}  // namespace dcx
