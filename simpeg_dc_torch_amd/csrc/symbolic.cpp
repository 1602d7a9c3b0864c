// Symbolic phase of the supernodal multifrontal Cholesky (host, once per
// sparsity pattern).  Replaces the reordering + symbolic factorisation that the
// reference delegates to cuDSS `plan()` (dctorch/solver.py:109, :207).
//
// Ordering: nested dissection (George 1973).  With node coordinates each piece
// is split at the median of its widest axis and the separator is the set of
// high-side nodes adjacent to the low side: on a tensor grid that is exactly
// one grid line (plane in 3-D).  Without coordinates the separator is the
// middle level of a breadth-first level structure rooted at a
// pseudo-peripheral node (George & Liu 1978), thinned to the nodes that touch
// the far side.  Pieces of at most `leaf_max` nodes become dense leaves.
//
// Supernodes: every separator and every leaf is one supernode with a dense
// diagonal block, so the separator tree is the assembly tree.  The
// off-diagonal row set of supernode s is
//     R_s = { i in adj(J_s) U (U_children R_c) : i > max J_s },
// which is always a subset of the ancestors' columns, so R_c is contained in
// J_p U R_p for every child c of p (the extend-add condition).
#include "symbolic.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace dcx {
namespace {

struct Graph {
  int n = 0;
  std::vector<int> ptr, adj;  // neighbours, no self loops
};

class NestedDissection {
 public:
  NestedDissection(const Graph& g, const double* coords, int dim, int leaf_max)
      : g_(g), xyz_(coords), dim_(dim), leaf_max_(std::max(1, leaf_max)),
        stamp_(g.n, 0), lvl_(g.n, -1) {}

  std::vector<int> order;                    // new -> old
  std::vector<int> t_first, t_ns, t_parent;  // tree, creation = postorder
  std::vector<std::vector<int>> t_children;

  void run() {
    std::vector<int> all(g_.n);
    std::iota(all.begin(), all.end(), 0);
    dissect(all);
  }

 private:
  const Graph& g_;
  const double* xyz_;
  int dim_, leaf_max_;
  std::vector<int> stamp_;
  int cur_ = 0;
  std::vector<int> lvl_;

  int make_node(std::vector<int> cols, const std::vector<int>& kids) {
    std::sort(cols.begin(), cols.end());
    int id = static_cast<int>(t_first.size());
    t_first.push_back(static_cast<int>(order.size()));
    t_ns.push_back(static_cast<int>(cols.size()));
    order.insert(order.end(), cols.begin(), cols.end());
    t_parent.push_back(-1);
    t_children.push_back(kids);
    for (int k : kids) t_parent[k] = id;
    return id;
  }

  void mark(const std::vector<int>& v) {
    ++cur_;
    for (int x : v) stamp_[x] = cur_;
  }

  // connected components of the subgraph induced by v
  std::vector<std::vector<int>> components(const std::vector<int>& v) {
    mark(v);
    const int in = cur_;
    ++cur_;
    const int seen = cur_;
    std::vector<std::vector<int>> comps;
    for (int s : v) {
      if (stamp_[s] != in) continue;
      std::vector<int> c{s};
      stamp_[s] = seen;
      for (size_t h = 0; h < c.size(); ++h) {
        int u = c[h];
        for (int p = g_.ptr[u]; p < g_.ptr[u + 1]; ++p) {
          int w = g_.adj[p];
          if (stamp_[w] == in) {
            stamp_[w] = seen;
            c.push_back(w);
          }
        }
      }
      comps.push_back(std::move(c));
    }
    return comps;
  }

  // breadth-first levels inside the marked set (stamp == in); returns depth
  int bfs(int root, int in, std::vector<int>& seq) {
    seq.clear();
    seq.push_back(root);
    lvl_[root] = 0;
    int h = 0;
    ++cur_;
    const int seen = cur_;
    stamp_[root] = seen;
    for (size_t i = 0; i < seq.size(); ++i) {
      int u = seq[i];
      for (int p = g_.ptr[u]; p < g_.ptr[u + 1]; ++p) {
        int w = g_.adj[p];
        if (stamp_[w] == in) {
          stamp_[w] = seen;
          lvl_[w] = lvl_[u] + 1;
          h = std::max(h, lvl_[w]);
          seq.push_back(w);
        }
      }
    }
    for (int x : seq) stamp_[x] = in;  // restore membership
    return h;
  }

  bool split_geometric(const std::vector<int>& v, std::vector<int>& sep,
                       std::vector<int>& a, std::vector<int>& b) {
    std::vector<std::pair<double, int>> ext;
    for (int d = 0; d < dim_; ++d) {
      double lo = 1e300, hi = -1e300;
      for (int x : v) {
        double c = xyz_[static_cast<size_t>(x) * dim_ + d];
        lo = std::min(lo, c);
        hi = std::max(hi, c);
      }
      ext.push_back({hi - lo, d});
    }
    std::sort(ext.rbegin(), ext.rend());
    for (auto [span, d] : ext) {
      if (span <= 0) break;
      std::vector<double> c(v.size());
      for (size_t i = 0; i < v.size(); ++i) c[i] = xyz_[static_cast<size_t>(v[i]) * dim_ + d];
      std::vector<double> s = c;
      std::nth_element(s.begin(), s.begin() + s.size() / 2, s.end());
      double med = s[s.size() / 2];
      double lo = *std::min_element(c.begin(), c.end());
      if (!(med > lo)) {  // pick the next value above the minimum
        double nxt = 1e300;
        for (double x : c)
          if (x > lo) nxt = std::min(nxt, x);
        if (nxt == 1e300) continue;
        med = nxt;
      }
      ++cur_;
      const int left = cur_;
      std::vector<int> l, r;
      for (size_t i = 0; i < v.size(); ++i) {
        if (c[i] < med) {
          stamp_[v[i]] = left;
          l.push_back(v[i]);
        } else {
          r.push_back(v[i]);
        }
      }
      if (l.empty() || r.empty()) continue;
      sep.clear();
      b.clear();
      for (int x : r) {
        bool touch = false;
        for (int p = g_.ptr[x]; p < g_.ptr[x + 1] && !touch; ++p) touch = stamp_[g_.adj[p]] == left;
        (touch ? sep : b).push_back(x);
      }
      a = std::move(l);
      return true;
    }
    return false;
  }

  bool split_levels(const std::vector<int>& v, std::vector<int>& sep,
                    std::vector<int>& a, std::vector<int>& b) {
    mark(v);
    const int in = cur_;
    std::vector<int> seq;
    int root = v[0];
    int h = bfs(root, in, seq);
    for (int it = 0; it < 8; ++it) {  // pseudo-peripheral node
      int best = -1, bdeg = 1 << 30;
      for (int x : seq)
        if (lvl_[x] == h) {
          int deg = g_.ptr[x + 1] - g_.ptr[x];
          if (deg < bdeg) bdeg = deg, best = x;
        }
      std::vector<int> seq2;
      int h2 = bfs(best, in, seq2);
      if (h2 <= h) break;
      h = h2;
      root = best;
      seq.swap(seq2);
    }
    h = bfs(root, in, seq);
    if (h < 2) return false;
    std::vector<int> cnt(h + 1, 0);
    for (int x : seq) cnt[lvl_[x]]++;
    int m = 1, acc = cnt[0];
    while (m < h - 1 && acc + cnt[m] < static_cast<int>(v.size()) / 2) acc += cnt[m++];
    sep.clear();
    a.clear();
    b.clear();
    for (int x : seq) {
      int l = lvl_[x];
      if (l < m) {
        a.push_back(x);
      } else if (l > m) {
        b.push_back(x);
      } else {
        bool touch = false;
        for (int p = g_.ptr[x]; p < g_.ptr[x + 1] && !touch; ++p) {
          int w = g_.adj[p];
          touch = stamp_[w] == in && lvl_[w] == m + 1;
        }
        (touch ? sep : a).push_back(x);
      }
    }
    return !sep.empty();
  }

  std::vector<int> dissect(const std::vector<int>& v) {
    if (static_cast<int>(v.size()) <= leaf_max_) return {make_node(v, {})};
    auto comps = components(v);
    if (comps.size() > 1) {
      std::vector<int> roots;
      for (auto& c : comps) {
        auto r = dissect(c);
        roots.insert(roots.end(), r.begin(), r.end());
      }
      return roots;
    }
    std::vector<int> sep, a, b;
    bool ok = xyz_ && split_geometric(v, sep, a, b);
    if (!ok) ok = split_levels(v, sep, a, b);
    if (!ok || sep.size() == v.size()) return {make_node(v, {})};  // dense piece
    std::vector<int> kids;
    if (!a.empty()) {
      auto r = dissect(a);
      kids.insert(kids.end(), r.begin(), r.end());
    }
    if (!b.empty()) {
      auto r = dissect(b);
      kids.insert(kids.end(), r.begin(), r.end());
    }
    return {make_node(sep, kids)};
  }
};

}  // namespace

// Fused bottom subtrees of the sweeps (see symbolic.h).
void build_fused(Symbolic& S, int fuse_rows) {
  const int ns_ = S.nsup;
  S.in_fused.assign(ns_, -1);
  S.fused_of_root.assign(ns_, -1);
  S.toff.assign(ns_, 0);
  S.lridx.assign(S.ridx.size(), 0);
  S.f_root.clear(); S.f_lo.clear(); S.f_hi.clear(); S.f_lv_begin.clear();
  S.f_lv_end.clear(); S.f_tb_rows.clear(); S.f_uoff.clear();
  S.f_lv.clear(); S.f_mem.clear(); S.f_gptr.clear(); S.f_gb.clear();
  S.f_max_lds_rows = 0;
  std::vector<int> rows(ns_, 0), lo(ns_, 0);
  for (int s = 0; s < ns_; ++s) {  // postorder: children first
    rows[s] = S.ns[s];
    lo[s] = S.first[s];
    for (int q = S.cptr[s]; q < S.cptr[s + 1]; ++q) {
      rows[s] += rows[S.cidx[q]];
      lo[s] = std::min(lo[s], lo[S.cidx[q]]);
    }
  }
  auto sublevels = [&](int root) {
    std::vector<std::vector<int>> lv{{root}};
    for (size_t d = 0; d < lv.size(); ++d) {
      std::vector<int> nxt;
      for (int s : lv[d])
        for (int q = S.cptr[s]; q < S.cptr[s + 1]; ++q) nxt.push_back(S.cidx[q]);
      if (!nxt.empty()) lv.push_back(nxt);
    }
    return lv;
  };
  auto tb_rows = [&](const std::vector<std::vector<int>>& lv) {
    int tb = 0;
    for (auto& l : lv) {
      int t = 0;
      for (int s : l) t += S.ns[s];
      tb = std::max(tb, t);
    }
    return tb;
  };
  std::vector<char> fits(ns_, 0);
  for (int s = 0; s < ns_; ++s)
    fits[s] = rows[s] + S.rs[s] + tb_rows(sublevels(s)) <= fuse_rows;
  std::vector<int> roots;
  for (int s = 0; s < ns_; ++s)
    if (fits[s] && (S.parent[s] < 0 || !fits[S.parent[s]]) && S.cptr[s + 1] > S.cptr[s])
      roots.push_back(s);
  S.nfused = static_cast<int>(roots.size());
  std::vector<std::vector<int>> marks;  // per colour: sublevel stamp of each row
  int stamp = 0;
  int64_t uoff = 0;
  for (int t = 0; t < S.nfused; ++t) {
    const int root = roots[t];
    const auto lv = sublevels(root);
    const int flo = lo[root], fhi = S.first[root] + S.ns[root];
    const int rsT = S.rs[root];
    const int tb = tb_rows(lv);
    S.f_root.push_back(root);
    S.f_lo.push_back(flo);
    S.f_hi.push_back(fhi);
    S.f_tb_rows.push_back(tb);
    S.f_uoff.push_back(uoff);
    uoff += rsT;
    S.f_max_lds_rows = std::max(S.f_max_lds_rows, (fhi - flo) + rsT + tb);
    S.fused_of_root[root] = t;
    S.f_lv_begin.push_back(static_cast<int>(S.f_lv.size()));
    const int* Rt = S.ridx.data() + S.rptr[root];
    for (const auto& l : lv) {
      ++stamp;
      std::vector<std::vector<int>> groups;
      for (int s : l) {
        S.in_fused[s] = t;
        size_t c = 0;
        for (; c < groups.size(); ++c) {
          bool clash = false;
          for (int q = S.rptr[s]; q < S.rptr[s + 1] && !clash; ++q) clash = marks[c][S.ridx[q]] == stamp;
          if (!clash) break;
        }
        if (c == groups.size()) {
          groups.emplace_back();
          if (marks.size() <= c) marks.emplace_back(S.n, 0);
        }
        groups[c].push_back(s);
        for (int q = S.rptr[s]; q < S.rptr[s + 1]; ++q) marks[c][S.ridx[q]] = stamp;
      }
      S.f_lv.push_back(static_cast<int>(S.f_mem.size()));
      S.f_gptr.push_back(static_cast<int>(S.f_gb.size()));
      int off = 0;
      for (const auto& g : groups) {
        S.f_gb.push_back(static_cast<int>(S.f_mem.size()));
        for (int s : g) {
          S.f_mem.push_back(s);
          S.toff[s] = off;
          off += S.ns[s];
          for (int q = S.rptr[s]; q < S.rptr[s + 1]; ++q) {
            const int r = S.ridx[q];
            if (r >= flo && r < fhi) {
              S.lridx[q] = r - flo;
            } else {
              const int pos = static_cast<int>(std::lower_bound(Rt, Rt + rsT, r) - Rt);
              S.lridx[q] = (fhi - flo) + pos;  // rows outside lie in R_root
            }
          }
        }
      }
    }
    S.f_lv_end.push_back(static_cast<int>(S.f_lv.size()));
  }
  S.f_lv.push_back(static_cast<int>(S.f_mem.size()));
  S.f_gptr.push_back(static_cast<int>(S.f_gb.size()));
  S.f_gb.push_back(static_cast<int>(S.f_mem.size()));
  S.f_upd_rows = uoff;
  // levels of the supernodes above the subtrees
  S.ulvl_ptr.assign(S.ndepth + 1, 0);
  for (int s = 0; s < ns_; ++s)
    if (S.in_fused[s] < 0) S.ulvl_ptr[S.depth[s] + 1]++;
  for (int d = 0; d < S.ndepth; ++d) S.ulvl_ptr[d + 1] += S.ulvl_ptr[d];
  S.ulvl_sup.assign(S.ulvl_ptr[S.ndepth], 0);
  std::vector<int> fill(S.ulvl_ptr.begin(), S.ulvl_ptr.end() - 1);
  for (int s = 0; s < ns_; ++s)
    if (S.in_fused[s] < 0) S.ulvl_sup[fill[S.depth[s]]++] = s;
}

std::string analyse(int n, const int* crow, const int* col, const double* coords,
                    int dim, int leaf_max, int fuse_rows, Symbolic& S) {
  if (n <= 0) return "matrix order must be positive";
  if (crow[0] != 0) return "crow[0] must be 0";
  const int nnz = crow[n];
  for (int i = 0; i < n; ++i)
    if (crow[i + 1] < crow[i]) return "crow is not non-decreasing";
  Graph g;
  g.n = n;
  g.ptr.assign(n + 1, 0);
  std::vector<char> diag(n, 0);
  for (int i = 0; i < n; ++i)
    for (int p = crow[i]; p < crow[i + 1]; ++p) {
      int j = col[p];
      if (j < 0 || j >= n) return "column index out of range";
      if (j == i) diag[i] = 1;
      else g.ptr[i + 1]++;
    }
  for (int i = 0; i < n; ++i)
    if (!diag[i]) return "pattern has an empty diagonal entry (row " + std::to_string(i) + ")";
  for (int i = 0; i < n; ++i) g.ptr[i + 1] += g.ptr[i];
  g.adj.resize(g.ptr[n]);
  {
    std::vector<int> fill(g.ptr.begin(), g.ptr.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int p = crow[i]; p < crow[i + 1]; ++p)
        if (col[p] != i) g.adj[fill[i]++] = col[p];
  }
  // symmetric pattern check
  for (int i = 0; i < n; ++i) std::sort(g.adj.begin() + g.ptr[i], g.adj.begin() + g.ptr[i + 1]);
  for (int i = 0; i < n; ++i)
    for (int p = g.ptr[i]; p < g.ptr[i + 1]; ++p) {
      int j = g.adj[p];
      if (!std::binary_search(g.adj.begin() + g.ptr[j], g.adj.begin() + g.ptr[j + 1], i))
        return "pattern is not symmetric";
    }

  NestedDissection nd(g, coords, dim, leaf_max);
  nd.run();
  S = Symbolic();
  S.n = n;
  S.nnz = nnz;
  S.perm = nd.order;
  S.iperm.assign(n, -1);
  for (int k = 0; k < n; ++k) S.iperm[S.perm[k]] = k;
  S.nsup = static_cast<int>(nd.t_first.size());
  S.first = nd.t_first;
  S.ns = nd.t_ns;
  S.parent = nd.t_parent;
  const int ns_ = S.nsup;

  S.depth.assign(ns_, 0);
  for (int s = ns_ - 1; s >= 0; --s)
    S.depth[s] = S.parent[s] < 0 ? 0 : S.depth[S.parent[s]] + 1;
  S.cptr.assign(ns_ + 1, 0);
  for (int s = 0; s < ns_; ++s) S.cptr[s + 1] = S.cptr[s] + static_cast<int>(nd.t_children[s].size());
  for (int s = 0; s < ns_; ++s) S.cidx.insert(S.cidx.end(), nd.t_children[s].begin(), nd.t_children[s].end());
  S.snode_of.assign(n, -1);
  for (int s = 0; s < ns_; ++s)
    for (int j = 0; j < S.ns[s]; ++j) S.snode_of[S.first[s] + j] = s;

  // row structures in postorder
  S.rptr.assign(ns_ + 1, 0);
  std::vector<int> mark(n, -1);
  std::vector<int> buf;
  for (int s = 0; s < ns_; ++s) {
    const int last = S.first[s] + S.ns[s] - 1;
    buf.clear();
    for (int j = S.first[s]; j <= last; ++j) {
      int o = S.perm[j];
      for (int p = g.ptr[o]; p < g.ptr[o + 1]; ++p) {
        int r = S.iperm[g.adj[p]];
        if (r > last && mark[r] != s) mark[r] = s, buf.push_back(r);
      }
    }
    for (int q = S.cptr[s]; q < S.cptr[s + 1]; ++q) {
      int c = S.cidx[q];
      for (int p = S.rptr[c]; p < S.rptr[c + 1]; ++p) {
        int r = S.ridx[p];
        if (r > last && mark[r] != s) mark[r] = s, buf.push_back(r);
      }
    }
    std::sort(buf.begin(), buf.end());
    S.ridx.insert(S.ridx.end(), buf.begin(), buf.end());
    S.rptr[s + 1] = static_cast<int>(S.ridx.size());
  }
  S.rs.resize(ns_);
  for (int s = 0; s < ns_; ++s) S.rs[s] = S.rptr[s + 1] - S.rptr[s];

  auto front_pos = [&](int s, int r) -> int {  // row r (new) inside F_s, or -1
    if (r >= S.first[s] && r < S.first[s] + S.ns[s]) return r - S.first[s];
    auto b = S.ridx.begin() + S.rptr[s], e = S.ridx.begin() + S.rptr[s + 1];
    auto it = std::lower_bound(b, e, r);
    if (it == e || *it != r) return -1;
    return S.ns[s] + static_cast<int>(it - b);
  };

  S.relpos.assign(S.ridx.size(), 0);
  for (int c = 0; c < ns_; ++c) {
    int p = S.parent[c];
    if (p < 0) {
      if (S.rs[c] != 0) return "internal: a root supernode has off-diagonal rows";
      continue;
    }
    for (int q = S.rptr[c]; q < S.rptr[c + 1]; ++q) {
      int pos = front_pos(p, S.ridx[q]);
      if (pos < 0) return "internal: child rows not contained in the parent front";
      S.relpos[q] = pos;
    }
  }

  // lower entries of A -> fronts; permuted CSR for the residual
  S.e_prow.resize(nnz);
  S.e_pcol.resize(nnz);
  std::vector<int> cnt(ns_ + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int p = crow[i]; p < crow[i + 1]; ++p) {
      int pi = S.iperm[i], pj = S.iperm[col[p]];
      S.e_prow[p] = pi;
      S.e_pcol[p] = pj;
      if (pi >= pj) cnt[S.snode_of[pj] + 1]++;
    }
  S.aptr.assign(ns_ + 1, 0);
  for (int s = 0; s < ns_; ++s) S.aptr[s + 1] = S.aptr[s] + cnt[s + 1];
  S.a_e.resize(S.aptr[ns_]);
  S.a_row.resize(S.aptr[ns_]);
  S.a_col.resize(S.aptr[ns_]);
  {
    std::vector<int> fill(S.aptr.begin(), S.aptr.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int p = crow[i]; p < crow[i + 1]; ++p) {
        int pi = S.e_prow[p], pj = S.e_pcol[p];
        if (pi < pj) continue;
        int s = S.snode_of[pj];
        int lr = front_pos(s, pi);
        if (lr < 0) return "internal: entry of A outside its front";
        int k = fill[s]++;
        S.a_e[k] = p;
        S.a_row[k] = lr;
        S.a_col[k] = pj - S.first[s];
      }
    // column order within a supernode: the large-front assembly finds the
    // entries of a column range by binary search
    std::vector<int> ord;
    for (int s = 0; s < ns_; ++s) {
      const int a0 = S.aptr[s], a1 = S.aptr[s + 1];
      ord.resize(a1 - a0);
      for (int k = a0; k < a1; ++k) ord[k - a0] = k;
      std::sort(ord.begin(), ord.end(), [&](int x, int y) {
        return S.a_col[x] != S.a_col[y] ? S.a_col[x] < S.a_col[y] : S.a_row[x] < S.a_row[y];
      });
      std::vector<int> e(a1 - a0), r(a1 - a0), c(a1 - a0);
      for (int k = 0; k < a1 - a0; ++k) {
        e[k] = S.a_e[ord[k]];
        r[k] = S.a_row[ord[k]];
        c[k] = S.a_col[ord[k]];
      }
      std::copy(e.begin(), e.end(), S.a_e.begin() + a0);
      std::copy(r.begin(), r.end(), S.a_row.begin() + a0);
      std::copy(c.begin(), c.end(), S.a_col.begin() + a0);
    }
  }
  S.pcrow.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) S.pcrow[S.iperm[i] + 1] = crow[i + 1] - crow[i];
  for (int i = 0; i < n; ++i) S.pcrow[i + 1] += S.pcrow[i];
  S.pcol.resize(nnz);
  S.pval_idx.resize(nnz);
  for (int i = 0; i < n; ++i) {
    int pi = S.iperm[i];
    std::vector<std::pair<int, int>> row;
    for (int p = crow[i]; p < crow[i + 1]; ++p) row.push_back({S.iperm[col[p]], p});
    std::sort(row.begin(), row.end());
    for (size_t k = 0; k < row.size(); ++k) {
      S.pcol[S.pcrow[pi] + k] = row[k].first;
      S.pval_idx[S.pcrow[pi] + k] = row[k].second;
    }
  }

  // level schedule, storage
  S.ndepth = 0;
  for (int s = 0; s < ns_; ++s) S.ndepth = std::max(S.ndepth, S.depth[s] + 1);
  S.lvl_ptr.assign(S.ndepth + 1, 0);
  for (int s = 0; s < ns_; ++s) S.lvl_ptr[S.depth[s] + 1]++;
  for (int d = 0; d < S.ndepth; ++d) S.lvl_ptr[d + 1] += S.lvl_ptr[d];
  S.lvl_sup.resize(ns_);
  {
    std::vector<int> fill(S.lvl_ptr.begin(), S.lvl_ptr.end() - 1);
    for (int s = 0; s < ns_; ++s) S.lvl_sup[fill[S.depth[s]]++] = s;
  }
  S.panel_off.resize(ns_);
  S.front_off.resize(ns_);
  S.upd_off.resize(ns_);
  int64_t po = 0;
  for (int s = 0; s < ns_; ++s) {
    int64_t nf = S.ns[s] + S.rs[s];
    S.panel_off[s] = po;
    po += nf * S.ns[s];
    S.max_front = std::max<int>(S.max_front, static_cast<int>(nf));
    S.max_ns = std::max(S.max_ns, S.ns[s]);
    double a = S.ns[s], r = S.rs[s];
    S.nnz_L += static_cast<int64_t>(a * (a + 1) / 2 + r * a);
    S.factor_flops += a * a * a / 3.0 + r * a * a + r * r * a  // partial Cholesky
                      + a * a * a / 3.0;                       // inverse of L_JJ
    S.solve_flops_per_rhs += 4.0 * (a + r) * a;
  }
  S.panel_size = po;
  build_fused(S, fuse_rows);
  for (int d = 0; d < S.ndepth; ++d) {
    int64_t fo = 0, uo = 0;
    for (int q = S.lvl_ptr[d]; q < S.lvl_ptr[d + 1]; ++q) {
      int s = S.lvl_sup[q];
      int64_t nf = S.ns[s] + S.rs[s];
      S.front_off[s] = fo;
      fo += (nf + S.ns[s]) * nf;  // front with the identity rows below it
      S.upd_off[s] = uo;
      uo += S.rs[s];
    }
    S.front_size[d & 1] = std::max(S.front_size[d & 1], fo);
    S.upd_rows[d & 1] = std::max(S.upd_rows[d & 1], uo);
  }
  return "";
}

}  // namespace dcx
