// dcx: supernodal multifrontal Cholesky, multi-RHS sweeps, gather-dot
// backward, sigma->A assembly map and receiver projection for gfx950.
//
// Factor storage per supernode s (front F_s = J_s U R_s, nf = ns + rs):
//   Q_s = [ L_JJ^{-1} ; L_RJ L_JJ^{-1} ]   (nf x ns, column-major)
// so that one sweep step is one small GEMM:
//   forward  (L y = b):   [y_J ; w] = Q_s f_J,  u_s = f_R - w  (passed up)
//   backward (L^T x = y): x_J = Q_s^T [ y_J ; -x_R ]
// Fronts and update rows are double-buffered by tree depth parity; each
// depth level is one launch over (supernodes x batch [x column tiles]).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/dcx.h"
#include "symbolic.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(DCX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct SupDev {
  int first, ns, rs, rptr;
  int c_begin, c_end, a_begin, a_end;
  int64_t panel_off, front_off, upd_off;
  int64_t fz_uoff;  // row of its update in the fused-subtree buffer, -1 if not a fused root
};

struct FusedDev {
  int root, lo, hi, rsT, rptrT, lv_begin, lv_end, pad;
  int64_t uoff;
};

// ---------------------------------------------------------------------------
// device memory of plans, maps and surveys: hipMalloc, or the caller's
// allocator once one is registered (dcx_set_allocator); each allocation
// remembers how it was made, so it is freed the same way
// ---------------------------------------------------------------------------
struct AllocRec {
  size_t bytes;
  int device;
  dcx_free_fn free_fn;  // null: hipMalloc
  void* ctx;
};
std::mutex g_alloc_mu;
dcx_alloc_fn g_alloc_fn = nullptr;
dcx_free_fn g_free_fn = nullptr;
void* g_alloc_ctx = nullptr;
std::unordered_map<void*, AllocRec> g_allocs;

hipError_t dev_malloc_bytes(void** p, size_t bytes) {
  *p = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  dcx_alloc_fn af;
  dcx_free_fn ff;
  void* ctx;
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    af = g_alloc_fn;
    ff = g_free_fn;
    ctx = g_alloc_ctx;
  }
  if (af) {
    void* q = af(bytes, dev, ctx);  // outside the lock: the callback may take others
    if (!q) return hipErrorOutOfMemory;
    *p = q;
  } else {
    e = hipMalloc(p, bytes);
    if (e != hipSuccess) return e;
    ff = nullptr;
  }
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_allocs[*p] = AllocRec{bytes, dev, ff, ctx};
  return hipSuccess;
}

template <typename T>
hipError_t dev_malloc(T** p, size_t bytes) {
  void* q = nullptr;
  const hipError_t e = dev_malloc_bytes(&q, bytes);
  *p = static_cast<T*>(q);
  return e;
}

// memory from the caller's allocator goes back to it after the device is
// idle (the allocator may hand it to other work at once)
hipError_t dev_free(void* p) {
  if (!p) return hipSuccess;
  AllocRec r{0, 0, nullptr, nullptr};
  bool known = false;
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    auto it = g_allocs.find(p);
    if (it != g_allocs.end()) {
      r = it->second;
      known = true;
      g_allocs.erase(it);
    }
  }
  if (known && r.free_fn) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != r.device) (void)hipSetDevice(r.device);
    const hipError_t e = hipDeviceSynchronize();
    r.free_fn(p, r.bytes, r.device, r.ctx);
    if (cur != r.device) (void)hipSetDevice(cur);
    return e;
  }
  return hipFree(p);
}

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& v) {
  *dst = nullptr;
  if (v.empty()) return hipSuccess;
  hipError_t e = dev_malloc(dst, v.size() * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// 16x16x4 matrix instruction of the element type, one wavefront:
//   A operand: lane l holds A[r0 + l%16][k0 + l/16]
//   B operand: lane l holds B[k0 + l/16][c0 + l%16]
//   result   : lane l, register i holds C[r0 + Mma<T>::row(l, i)][c0 + l%16]
//              (FP32: row 4*(l/16) + i; FP64: row l/16 + 4*i, as measured by
//              the dcx_mma_probe layout test)
// ---------------------------------------------------------------------------
template <typename T>
struct Mma;
template <>
struct Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static inline v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int i) { return (lane >> 4) + 4 * i; }
};
template <>
struct Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static inline v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int i) { return 4 * (lane >> 4) + i; }
};

// ---------------------------------------------------------------------------
// numeric factorisation, one launch per depth level (deepest first)
// grid (supernodes at level, batch), NT threads.
// The front F (column-major, leading dimension nf + ns) holds the lower
// triangle of [J; R] in its first nf rows and, below them, ns extra rows E
// that start as the identity in the J columns.  Eliminating the ns J columns
// leaves L_JJ, L_RJ, the Schur complement in the (R,R) block -- this front's
// update block -- and L_EJ = L_JJ^{-T} in the E rows, i.e. the inverse the
// sweeps need, without a separate inversion.
// Per panel of kNB columns: every thread holds RPT panel rows in registers;
// wavefront 0 factors the kNB x kNB diagonal block (pivot columns read across
// lanes, no barrier), then every other row solves against it on its own
// registers (one barrier); the panel goes to LDS and the trailing update --
// lower triangle of the J/R rows and the E rows times the J/R columns --
// runs as 16x16 tiles on the matrix unit, two tiles per wavefront at a time.
// Global gathers (assembly, extend-add, outputs) take 16 entries per thread
// at a time, loads first.
// The panel written for the sweeps is P_s = [ L_JJ^{-1} ; L_RJ L_JJ^{-1} ]
// (nf x ns) and its transpose.
// ---------------------------------------------------------------------------
constexpr int kNB = 32;        // panel width
constexpr int kLDP = kNB + 4;  // LDS row stride of the panel
constexpr int kLDG = kNB + 4;  // LDS row stride of the diagonal block

// value of v in lane l (l uniform) for every lane of the wavefront
__device__ inline float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double lane_value(double v, int l) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(x), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(x >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// zero the front, seed the E rows with the identity, scatter the A entries
// and extend-add the children's update blocks (lower triangles; relative
// positions staged in the LDS array rpl).  Global gathers take kGather
// entries per thread at a time, loads first, so that their latencies overlap
// (8, not 16: the FP64 narrow-panel factor then needs 85 instead of 107
// VGPRs; FP64 eval+grad 5.41 -> 5.39 ms, the 16-line batch 9.96 -> 9.90 ms).
constexpr int kGather = 8;
template <typename T, int NT>
__device__ void front_init(const SupDev& S, int b, const SupDev* __restrict__ sup,
                           const int* __restrict__ relpos, const int* __restrict__ cidx,
                           const int* __restrict__ a_e, const int* __restrict__ a_row,
                           const int* __restrict__ a_col, const double* __restrict__ vals,
                           int64_t nnz, T* __restrict__ F, const T* __restrict__ fprev,
                           int64_t fprev_bs, int* rpl) {
  constexpr int kU = kGather;
  const int tid = threadIdx.x;
  const int ns = S.ns, nf = S.ns + S.rs, ld = nf + S.ns;
  const int64_t fsz = static_cast<int64_t>(ld) * nf;
  {
    // the strict upper triangle of the J/R rows is never read (panel loads
    // use the diagonal block's lower triangle, the trailing update and the
    // extend-add write i >= l only), so it is not written either: the front's
    // dirty bytes, written back to HBM, shrink by up to a third
    const int sq = NT / ld, sr = NT % ld;
    int zr = tid % ld, zc = tid / ld;
    for (int64_t i = tid; i < fsz; i += NT) {
      if (zr >= zc || zr >= nf) F[i] = T(0);
      zr += sr;
      zc += sq;
      if (zr >= ld) {
        zr -= ld;
        ++zc;
      }
    }
  }
  __syncthreads();
  for (int j = tid; j < ns; j += NT) F[(nf + j) + static_cast<int64_t>(j) * ld] = T(1);
  const double* v = vals + b * nnz;
  for (int q0 = S.a_begin + tid; q0 < S.a_end; q0 += NT * kU) {
    T av[kU];
    int ad[kU];  // front offsets: in-LDS fronts hold < 2^31 entries
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int q = q0 + u * NT;
      ad[u] = -1;
      if (q < S.a_end) {
        ad[u] = a_row[q] + a_col[q] * ld;
        av[u] = static_cast<T>(v[a_e[q]]);
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (ad[u] >= 0) F[ad[u]] = av[u];
  }
  for (int cc = S.c_begin; cc < S.c_end; ++cc) {
    const SupDev C = sup[cidx[cc]];
    const int rc = C.rs, ldc = C.ns + C.rs + C.ns;
    const T* Fc = fprev + b * fprev_bs + C.front_off;
    __syncthreads();  // F writes before, rpl reads of the previous child
    for (int i = tid; i < rc; i += NT) rpl[i] = relpos[C.rptr + i];
    __syncthreads();
    if (rc == 0) continue;  // uniform
    // entry t = i + j rc of the child's square update block; (i, j) is
    // carried along from one entry of the thread to the next (steps of NT
    // and NT kU entries) instead of divided out per entry
    const int rc2 = rc * rc;  // in-LDS fronts: rc <= 1024
    const int sq = NT / rc, sr = NT % rc;
    const int bq = (NT * kU) / rc, br = (NT * kU) % rc;
    int ib = tid % rc, jbk = tid / rc;
    for (int t0 = tid; t0 < rc2; t0 += NT * kU) {
      T cv[kU], fv[kU];
      int fa[kU];  // front offsets: in-LDS fronts hold < 2^31 entries
      int i = ib, j = jbk;
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        fa[u] = -1;
        if (t0 + u * NT < rc2 && i >= j) {
          fa[u] = rpl[i] + rpl[j] * ld;
          cv[u] = Fc[(C.ns + i) + static_cast<int64_t>(C.ns + j) * ldc];
          fv[u] = F[fa[u]];
        }
        i += sr;
        j += sq;
        if (i >= rc) {
          i -= rc;
          ++j;
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (fa[u] >= 0) F[fa[u]] = fv[u] + cv[u];
      ib += br;
      jbk += bq;
      if (ib >= rc) {
        ib -= rc;
        ++jbk;
      }
    }
  }
  __syncthreads();
}

// pivot d -> (sqrt d, 1 / sqrt d) from the hardware reciprocal square root
// (about 23 bits in FP64) refined by two Newton steps
// y <- y + y (1 - d y^2) / 2, each doubling the correct bits; an FP32 pivot
// goes through the FP64 steps and is rounded once
__device__ inline void pivot(double d, double& sq, double& dinv) {
  double y = __builtin_amdgcn_rsq(d);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-(d * y), y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  dinv = y;
  sq = d * y;
}
__device__ inline void pivot(float d, float& sq, float& dinv) {
  double s, y;
  pivot(static_cast<double>(d), s, y);
  sq = static_cast<float>(s);
  dinv = static_cast<float>(y);
}

// the jb x jb diagonal block of a panel, by one wavefront: lanes < jb hold its
// rows in rv; each pivot column is broadcast through ldg's first row (free
// until the end) and read back as vectors; the L rows go to ldg (stride
// NB + 4) and the reciprocal diagonal to dgi.  NB: the panel width (32, or 16
// in the narrow-panel factor variant)
template <typename T, int NB = kNB>
__device__ inline void diag_block(T (&rv)[NB], int jb, int lane, T* ldg, T* dgi,
                                  int* __restrict__ info) {
  constexpr int LDG = NB + 4;
  T mydinv = T(0);
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    if (t < jb) {  // uniform
      if (lane < NB) ldg[lane] = rv[t];
      __builtin_amdgcn_wave_barrier();
      T d = lane_value(rv[t], t);
      if (!(d > T(0))) {  // uniform
        if (lane == 0) atomicMax(info, 1);
        d = T(1);
      }
      T sq, dinv;
      pivot(d, sq, dinv);
      // rank-1 update by column t of the rows l > t (unscaled); lanes
      // outside the update take g = 0, an exact no-op, so that the update
      // needs no per-entry select
      const bool below = lane > t && lane < jb;
      const T g = below ? rv[t] * (dinv * dinv) : T(0);
#pragma unroll
      for (int l = t + 1; l < NB; ++l) rv[l] -= g * ldg[l];
      __builtin_amdgcn_wave_barrier();
      rv[t] = lane == t ? sq : (below ? rv[t] * dinv : rv[t]);
      mydinv = lane == t ? dinv : mydinv;
    }
  }
  if (lane < jb) {
#pragma unroll
    for (int c = 0; c < NB; ++c) ldg[lane * LDG + c] = rv[c];
    dgi[lane] = mydinv;
  }
}

// a row below the diagonal block: L_r = A_r L_diag^{-T} (forward substitution
// on the row's registers with the block from LDS)
template <typename T, int NB = kNB>
__device__ inline void row_solve(T (&rv)[NB], int jb, const T* ldg, const T* dgi) {
  constexpr int LDG = NB + 4;
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    if (t < jb) {
      T s0 = rv[t], s1 = T(0);
#pragma unroll
      for (int l = 0; l < t; l += 2) {
        s0 -= rv[l] * ldg[t * LDG + l];
        if (l + 1 < t) s1 -= rv[l + 1] * ldg[t * LDG + l + 1];
      }
      rv[t] = (s0 + s1) * dgi[t];
    }
  }
}

// the sweep panel P (nf x ns, rows [0, ns) = L_JJ^{-1}, rows [ns, nf) = Q_R)
// and its transpose PT (ns x nf) from the eliminated front's E rows
template <typename T, int NT>
__device__ void panel_out(const SupDev& S, const T* __restrict__ F, T* __restrict__ P,
                          T* __restrict__ PT) {
  constexpr int kU = kGather;
  const int tid = threadIdx.x;
  const int ns = S.ns, nf = S.ns + S.rs, ld = nf + S.ns;
  // entry idx = c + r ns, with (c, r) carried along as in front_init
  const int np = nf * ns;  // in-LDS fronts: < 2^31
  const int sq = NT / ns, sr = NT % ns;
  const int bq = (NT * kU) / ns, br = (NT * kU) % ns;
  int cb = tid % ns, rb = tid / ns;
  for (int i0 = tid; i0 < np; i0 += NT * kU) {
    T pv[kU];
    int c = cb, r = rb;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      pv[u] = i0 + u * NT < np ? F[(nf + c) + static_cast<int64_t>(r) * ld] : T(0);
      if (r < ns && r < c) pv[u] = T(0);
      if (r >= ns) pv[u] = -pv[u];
      c += sr;
      r += sq;
      if (c >= ns) {
        c -= ns;
        ++r;
      }
    }
    c = cb;
    r = rb;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int idx = i0 + u * NT;
      if (idx < np) {
        PT[idx] = pv[u];
        P[r + static_cast<int64_t>(c) * nf] = pv[u];
      }
      c += sr;
      r += sq;
      if (c >= ns) {
        c -= ns;
        ++r;
      }
    }
    cb += br;
    rb += bq;
    if (cb >= ns) {
      cb -= ns;
      ++rb;
    }
  }
}

// NB: panel width (32; 16 in the narrow-panel variant: shorter register
// rows and diagonal-block steps, one more trailing pass per 16 columns).
// The FP32 256-thread variants are built for 6 wavefronts per SIMD (80
// VGPRs, no spills; 5 at their natural 81-83); the FP64 ones would spill.
template <typename T, int NT, int RPT, int NB>
__global__ __launch_bounds__(NT, (NT == 256 && sizeof(T) == 4 ? 6 : 1)) void k_factor_level(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl,
    const int* __restrict__ relpos, const int* __restrict__ cidx,
    const int* __restrict__ a_e, const int* __restrict__ a_row,
    const int* __restrict__ a_col, const double* __restrict__ vals, int64_t nnz,
    T* __restrict__ fcur, int64_t fcur_bs, const T* __restrict__ fprev,
    int64_t fprev_bs, T* __restrict__ panels, T* __restrict__ panelsT, int64_t panel_bs,
    int* __restrict__ info) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int LDP = NB + 4, LDG = NB + 4;
  const int s = lvl[blockIdx.x];
  const int b = blockIdx.y;
  const SupDev S = sup[s];
  const int ns = S.ns, rs = S.rs, nf = S.ns + S.rs, ld = nf + S.ns;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid / 64;
  constexpr int NW = NT / 64;
  T* ldg = reinterpret_cast<T*>(smem_raw);   // diagonal block rows, stride LDG
  T* dgi = ldg + NB * LDG;                  // reciprocal diagonal
  T* pan = dgi + NB;                         // panel rows, stride LDP
  T* F = fcur + b * fcur_bs + S.front_off;
  front_init<T, NT>(S, b, sup, relpos, cidx, a_e, a_row, a_col, vals, nnz, F, fprev, fprev_bs,
                    reinterpret_cast<int*>(pan));
  for (int j0 = 0; j0 < ns; j0 += NB) {
    const int jb = min(NB, ns - j0);
    // panel rows: J, R and E rows from j0 on.  E row e (front row nf + e)
    // has no entry left of column e until column e is eliminated, so the E
    // rows e >= j0 + jb are zero in this panel and take no part in it
    const int m = min(ld - j0, nf + jb);
    const int n1 = nf - j0;    // end of the J/R rows (panel-local)
    T rv[RPT][NB];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + q * NT;
#pragma unroll
      for (int c = 0; c < NB; ++c)
        rv[q][c] = (r < m && c < jb) ? F[(j0 + r) + static_cast<int64_t>(j0 + c) * ld] : T(0);
    }
    // (1) the jb x jb diagonal block, by wavefront 0 alone (lanes < jb hold
    // its rows)
    if (wave == 0) diag_block<T, NB>(rv[0], jb, lane, ldg, dgi, info + b);
    __syncthreads();
    // (2) every row below the block, on the thread's own registers
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + q * NT;
      if (r >= jb && r < m) row_solve<T, NB>(rv[q], jb, ldg, dgi);
    }
    // panel to F (its E rows) and to LDS (zero columns past jb; the 16 rows
    // past m are zeroed for the tile reads)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + q * NT;
      if (r < m) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          // only the E rows of a finished panel are read again (the sweep
          // panel comes from them); its J/R rows live on in LDS only
          if (c < jb && r >= n1) F[(j0 + r) + static_cast<int64_t>(j0 + c) * ld] = rv[q][c];
          pan[r * LDP + c] = c < jb ? rv[q][c] : T(0);
        }
      }
    }
    for (int idx = tid; idx < 16 * LDP; idx += NT) pan[m * LDP + idx] = T(0);
    __syncthreads();
    // trailing update, 16x16 tiles, computed transposed so that a lane's
    // output runs down a column of F:
    //   (a) rows [jb, n1) x cols [jb, n1), lower triangle
    //   (b) E rows [n1, m) x cols [jb, n1)
    const int w = n1 - jb;
    const int nt = (w + 15) >> 4;
    const int ntri = nt * (nt + 1) / 2;
    const int nte = ((m - n1 + 15) >> 4) * nt;
    constexpr int TB = 2;  // tiles per wavefront whose loads are in flight together
    for (int tt0 = wave; tt0 < ntri + nte; tt0 += NW * TB) {
      T cold[TB][4];
      int addr[TB][4];  // front offsets (< 2^31 entries)
      bool ok[TB][4];
      int i0[TB], l0[TB];
#pragma unroll
      for (int u = 0; u < TB; ++u) {
        const int tt = tt0 + u * NW;
        int imax = 0;
        bool diag = false;
        i0[u] = l0[u] = jb;
        if (tt < ntri) {
          int tj = 0, rem = tt;
          while (rem >= nt - tj) { rem -= nt - tj; ++tj; }
          i0[u] = jb + (tj + rem) * 16;
          l0[u] = jb + tj * 16;
          imax = n1;
          diag = rem == 0;
        } else if (tt < ntri + nte) {
          const int w2 = tt - ntri;
          i0[u] = n1 + (w2 / nt) * 16;
          l0[u] = jb + (w2 % nt) * 16;
          imax = m;
        }  // else: no tile, imax = 0 masks every entry
        const int i = i0[u] + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int l = l0[u] + Mma<T>::row(lane, e);
          ok[u][e] = i < imax && l < n1 && (!diag || i >= l);
          addr[u][e] = (j0 + i) + (j0 + l) * ld;
          cold[u][e] = ok[u][e] ? F[addr[u][e]] : T(0);
        }
      }
#pragma unroll
      for (int u = 0; u < TB; ++u) {
        typename Mma<T>::v4 acc = {T(0), T(0), T(0), T(0)};
        const T* pa = pan + (l0[u] + (lane & 15)) * LDP + (lane >> 4);
        const T* pb = pan + (i0[u] + (lane & 15)) * LDP + (lane >> 4);
#pragma unroll
        for (int k0 = 0; k0 < NB; k0 += 4) acc = Mma<T>::op(pa[k0], pb[k0], acc);
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (ok[u][e]) F[addr[u][e]] = cold[u][e] - acc[e];
      }
    }
    __syncthreads();
  }
  // the sweep panel and its transpose (the backward sweep's operand runs
  // along the rows of P)
  panel_out<T, NT>(S, F, panels + b * panel_bs + S.panel_off,
                   panelsT + b * panel_bs + S.panel_off);
  (void)rs;
}

// ---------------------------------------------------------------------------
// Fronts too tall for the in-LDS panel (3-D separators of thousands of rows):
// the same elimination as grid-wide launches, so that one front spreads over
// the whole chip --
//   k_front_zero_big   workgroup per 16 front columns: zero, E identity, and
//                      the A entries of those columns (sorted by column);
//   k_extend_big       per child index k, the k-th child's update block added
//                      in 64x64 tiles (one child's map is injective, so the
//                      tiles never collide and the sum order is fixed);
//   per 64-column panel:
//   k_panel_rows_big   workgroup per 256 panel rows: wavefront 0 factors the
//                      diagonal block (read only, so every workgroup can), then
//                      each thread solves one row in place in F;
//   k_trail_big        the trailing update as 64x64 MFMA tiles;
//   k_front_out_big    the sweep panels as 64x64 tiles, transposed via LDS.
// ---------------------------------------------------------------------------
constexpr int kZeroCols = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_front_zero_big(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl, const int* __restrict__ a_e,
    const int* __restrict__ a_row, const int* __restrict__ a_col,
    const double* __restrict__ vals, int64_t nnz, T* __restrict__ fcur, int64_t fcur_bs,
    int nbatch) {
  const int f = blockIdx.y / nbatch, b = blockIdx.y % nbatch;
  const SupDev S = sup[lvl[f]];
  const int nf = S.ns + S.rs, ld = nf + S.ns;
  const int c0 = blockIdx.x * kZeroCols;
  if (c0 >= nf) return;
  const int c1 = min(nf, c0 + kZeroCols);
  T* F = fcur + b * fcur_bs + S.front_off;
  for (int col = c0; col < c1; ++col) {
    T* Fcol = F + static_cast<int64_t>(col) * ld;
    for (int row = threadIdx.x; row < ld; row += 256)
      Fcol[row] = row == nf + col ? T(1) : T(0);  // E row nf + j holds the identity column j
  }
  // the entries of columns [c0, c1): a_col ascends within a supernode
  int lo = S.a_begin, hi = S.a_end;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a_col[mid] < c0) lo = mid + 1; else hi = mid;
  }
  int end = lo, hi2 = S.a_end;
  while (end < hi2) {
    const int mid = (end + hi2) >> 1;
    if (a_col[mid] < c1) end = mid + 1; else hi2 = mid;
  }
  __syncthreads();
  const double* v = vals + b * nnz;
  for (int q = lo + threadIdx.x; q < end; q += 256)
    F[a_row[q] + static_cast<int64_t>(a_col[q]) * ld] = static_cast<T>(v[a_e[q]]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_extend_big(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl, const int* __restrict__ relpos,
    const int* __restrict__ cidx, T* __restrict__ fcur, int64_t fcur_bs,
    const T* __restrict__ fprev, int64_t fprev_bs, int kchild, int nbatch) {
  __shared__ int rp[2][64];
  const int f = blockIdx.z / nbatch, b = blockIdx.z % nbatch;
  const SupDev S = sup[lvl[f]];
  if (S.c_begin + kchild >= S.c_end) return;
  const SupDev C = sup[cidx[S.c_begin + kchild]];
  const int rc = C.rs, ldc = C.ns + C.rs + C.ns, ld = S.ns + S.rs + S.ns;
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  if (i0 >= rc || j0 >= rc || i0 + 63 < j0) return;  // outside, or above the diagonal
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid < 128) {
    const int q = (tid < 64 ? i0 : j0) + lane;
    rp[tid >> 6][lane] = q < rc ? relpos[C.rptr + q] : 0;
  }
  __syncthreads();
  const T* Fc = fprev + b * fprev_bs + C.front_off + C.ns + static_cast<int64_t>(C.ns) * ldc;
  T* F = fcur + b * fcur_bs + S.front_off;
  const int i = i0 + lane;
  constexpr int kCols = 16;  // columns per wavefront, loads first
  T cv[kCols], fv[kCols];
  int64_t fa[kCols];
#pragma unroll
  for (int u = 0; u < kCols; ++u) {
    const int jj = w * kCols + u, j = j0 + jj;
    fa[u] = -1;
    if (i < rc && j < rc && i >= j) {
      fa[u] = rp[0][lane] + static_cast<int64_t>(rp[1][jj]) * ld;
      cv[u] = Fc[i + static_cast<int64_t>(j) * ldc];
      fv[u] = F[fa[u]];
    }
  }
#pragma unroll
  for (int u = 0; u < kCols; ++u)
    if (fa[u] >= 0) F[fa[u]] = fv[u] + cv[u];
}

// Big-front panels are kNB2 = 2 kNB columns wide.  Their diagonal block is
// factored as two kNB halves, D1 = L11 L11^T, L21 = A21 L11^{-T},
// D2 = A22 - L21 L21^T = L22 L22^T, by wavefront 0 of one workgroup
// (panel_diag), into the image [L11 | 1/diag L11 | L21 | L22 | 1/diag L22]
// (rows at stride kLDG) kept per front in global memory.  Panel 0's comes
// from k_diag_big; panel j0 + kNB2's from workgroup (0, 0) of panel j0's
// trailing update, which finishes that block (a one-panel lookahead), so the
// row solve of a panel never waits on the sequential diagonal factor.
constexpr int kRowsBig = 256;
constexpr int kNB2 = 2 * kNB;
constexpr int kOuter = 4 * kNB2;  // outer panel width of the large-front factor
constexpr int kDg = 3 * kNB * kLDG + 2 * kNB;

// called by every thread of the workgroup (barriers inside); Pn points at
// the block's top-left entry in F (column-major, leading dimension ld)
template <typename T>
__device__ __forceinline__ void panel_diag(const T* Pn, int ld, int jb, T* sm, int* __restrict__ info) {
  T* ldg0 = sm;
  T* dgi0 = ldg0 + kNB * kLDG;
  T* l21 = dgi0 + kNB;
  T* ldg1 = l21 + kNB * kLDG;
  T* dgi1 = ldg1 + kNB * kLDG;
  const int tid = threadIdx.x, lane = tid & 63;
  const int jb1 = min(kNB, jb), jb2 = jb - jb1;
  if (tid < 64) {
    T rv[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c)
      rv[c] = (lane < jb1 && c < jb1 && lane >= c) ? Pn[lane + static_cast<int64_t>(c) * ld]
                                                  : T(0);
    diag_block<T>(rv, jb1, lane, ldg0, dgi0, info);
  }
  __syncthreads();
  if (jb2 > 0) {
    if (tid < 64) {
      T rv[kNB];
#pragma unroll
      for (int c = 0; c < kNB; ++c)
        rv[c] = (lane < jb2 && c < jb1) ? Pn[(jb1 + lane) + static_cast<int64_t>(c) * ld] : T(0);
      row_solve<T>(rv, jb1, ldg0, dgi0);
      if (lane < jb2) {
#pragma unroll
        for (int c = 0; c < kNB; ++c) l21[lane * kLDG + c] = rv[c];
      }
    }
    __syncthreads();
    if (tid < 64) {
      T rv[kNB];
#pragma unroll
      for (int c = 0; c < kNB; ++c)
        rv[c] = (lane < jb2 && c < jb2 && lane >= c)
                    ? Pn[(jb1 + lane) + static_cast<int64_t>(jb1 + c) * ld]
                    : T(0);
      // L21 L21^T as 2 x 2 16x16 tiles on the matrix unit, staged in ldg1
      // (overwritten by the factor below)
      typename Mma<T>::v4 acc[2][2];
#pragma unroll
      for (int tr = 0; tr < 2; ++tr)
#pragma unroll
        for (int tc = 0; tc < 2; ++tc) acc[tr][tc] = {T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int q0 = 0; q0 < kNB; q0 += 4) {
        const int q = q0 + (lane >> 4);
        T op[2];
#pragma unroll
        for (int tr = 0; tr < 2; ++tr) {
          const int rr = 16 * tr + (lane & 15);
          op[tr] = rr < jb2 ? l21[rr * kLDG + q] : T(0);
        }
#pragma unroll
        for (int tr = 0; tr < 2; ++tr)
#pragma unroll
          for (int tc = 0; tc < 2; ++tc) acc[tr][tc] = Mma<T>::op(op[tr], op[tc], acc[tr][tc]);
      }
#pragma unroll
      for (int tr = 0; tr < 2; ++tr)
#pragma unroll
        for (int tc = 0; tc < 2; ++tc)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            ldg1[(16 * tr + Mma<T>::row(lane, e)) * kLDG + 16 * tc + (lane & 15)] = acc[tr][tc][e];
#pragma unroll
      for (int c = 0; c < kNB; ++c)
        if (c <= lane && lane < jb2 && c < jb2) rv[c] -= ldg1[lane * kLDG + c];
      diag_block<T>(rv, jb2, lane, ldg1, dgi1, info);
    }
  }
  __syncthreads();
}

// panel 0 of every big front: one wavefront per (front, batch member)
template <typename T>
__global__ __launch_bounds__(64) void k_diag_big(const SupDev* __restrict__ sup,
                                                const int* __restrict__ lvl,
                                                const T* __restrict__ fcur, int64_t fcur_bs,
                                                T* __restrict__ dg, int nbatch,
                                                int* __restrict__ info) {
  __shared__ T sm[kDg];
  const int f = blockIdx.x / nbatch, b = blockIdx.x % nbatch;
  const SupDev S = sup[lvl[f]];
  const int ld = 2 * S.ns + S.rs;
  panel_diag<T>(fcur + b * fcur_bs + S.front_off, ld, min(kNB2, S.ns), sm, info + b);
  T* o = dg + static_cast<int64_t>(blockIdx.x) * kDg;
  for (int i = threadIdx.x; i < kDg; i += 64) o[i] = sm[i];
}

// the rows below a panel's diagonal block: workgroup x takes rows
// [jb + 256 x, jb + 256 x + 256) and solves each against the factored block
template <typename T>
__global__ __launch_bounds__(kRowsBig) void k_panel_rows_big(const SupDev* __restrict__ sup,
                                                            const int* __restrict__ lvl,
                                                            T* __restrict__ fcur, int64_t fcur_bs,
                                                            const T* __restrict__ dg, int j0,
                                                            int nbatch) {
  __shared__ T sm[kDg];
  const int f = blockIdx.y / nbatch, b = blockIdx.y % nbatch;
  const SupDev S = sup[lvl[f]];
  const int ns = S.ns, nf = S.ns + S.rs, ld = nf + S.ns;
  if (j0 >= ns) return;  // this front has no panel j0 (uniform)
  // E rows e >= j0 + jb are zero in this panel's columns (see k_factor_level)
  const int jb = min(kNB2, ns - j0), jb1 = min(kNB, jb), jb2 = jb - jb1;
  const int m = min(ld - j0, nf + jb);
  const int rbase = jb + blockIdx.x * kRowsBig;
  if (rbase >= m) return;
  const int tid = threadIdx.x;
  T* Pn = fcur + b * fcur_bs + S.front_off + j0 + static_cast<int64_t>(j0) * ld;
  const int r = rbase + tid;
  T ra[kNB], rb[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    ra[c] = (r < m && c < jb1) ? Pn[r + static_cast<int64_t>(c) * ld] : T(0);
    rb[c] = (r < m && c < jb2) ? Pn[r + static_cast<int64_t>(jb1 + c) * ld] : T(0);
  }
  const T* dgb = dg + static_cast<int64_t>(blockIdx.y) * kDg;
  for (int i = tid; i < kDg; i += kRowsBig) sm[i] = dgb[i];
  __syncthreads();
  if (r >= m) return;
  const T* ldg0 = sm;
  const T* dgi0 = ldg0 + kNB * kLDG;
  const T* l21 = dgi0 + kNB;
  const T* ldg1 = l21 + kNB * kLDG;
  const T* dgi1 = ldg1 + kNB * kLDG;
  row_solve<T>(ra, jb1, ldg0, dgi0);
#pragma unroll
  for (int c = 0; c < kNB; ++c)
    if (c < jb1) Pn[r + static_cast<int64_t>(c) * ld] = ra[c];
  if (jb2 > 0) {
#pragma unroll
    for (int c = 0; c < kNB; ++c) {
      if (c < jb2) {
        T s = T(0);
#pragma unroll
        for (int q = 0; q < kNB; ++q) s += ra[q] * l21[c * kLDG + q];
        rb[c] -= s;
      }
    }
    row_solve<T>(rb, jb2, ldg1, dgi1);
#pragma unroll
    for (int c = 0; c < kNB; ++c)
      if (c < jb2) Pn[r + static_cast<int64_t>(jb1 + c) * ld] = rb[c];
  }
}

// The update after an outer panel (K up to kOuter): an LDS-tiled GEMM.
// Workgroup tile kTR rows (i) x kTC columns (l); K in chunks of kKC staged in
// LDS k-major (a chunk's rows are contiguous in the column-major front, so
// both the loads and the per-lane LDS reads are unit-stride), the next chunk
// held in registers while the current one is multiplied.  Each wavefront:
// kTR / 4 rows x kTC columns as (kTR / 64) x (kTC / 16) MFMA tiles, results
// transposed so a lane's values run down a column of F.  Same row, column
// and lower-triangle rules and the same lookahead tail as k_trail_big.
constexpr int kTR = 128, kTC = 64, kKC = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_trail_outer(const SupDev* __restrict__ sup,
                                                    const int* __restrict__ lvl,
                                                    T* __restrict__ fcur, int64_t fcur_bs,
                                                    int j0, int nbatch, T* __restrict__ dg,
                                                    int* __restrict__ info) {
  __shared__ T As[kKC][kTR + 4];  // rows i of the chunk
  __shared__ T Bs[kKC][kTC + 4];  // rows l of the chunk
  const int f = blockIdx.z / nbatch, b = blockIdx.z % nbatch;
  const SupDev S = sup[lvl[f]];
  const int ns = S.ns, nf = S.ns + S.rs, ld = nf + S.ns;
  if (j0 >= ns) return;
  const int jb = min(kOuter, ns - j0), m = min(ld - j0, nf + jb), n1 = nf - j0;
  const int rb = jb + blockIdx.x * kTR, cb = jb + blockIdx.y * kTC;
  if (rb >= m || cb >= n1 || (rb + kTR - 1 < n1 && rb + kTR - 1 < cb)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  T* Pn = fcur + b * fcur_bs + S.front_off + j0 + static_cast<int64_t>(j0) * ld;
  // chunk loads: A: kTR x kKC (16 per thread), B: kTC x kKC (8 per thread);
  // thread t takes row (t % kTR) [A] / (t % kTC) [B] of columns t / kTR + ...
  constexpr int kLA = kTR * kKC / 256, kLB = kTC * kKC / 256;
  T ra[kLA], rbv[kLB];
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < kLA; ++u) {
      const int idx = tid + 256 * u, r = idx % kTR, kk = k0 + idx / kTR;
      ra[u] = (rb + r < m && kk < jb) ? Pn[(rb + r) + static_cast<int64_t>(kk) * ld] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kLB; ++u) {
      const int idx = tid + 256 * u, r = idx % kTC, kk = k0 + idx / kTC;
      rbv[u] = (cb + r < n1 && kk < jb) ? Pn[(cb + r) + static_cast<int64_t>(kk) * ld] : T(0);
    }
  };
  constexpr int kMR = kTR / 64, kMC = kTC / 16;  // MFMA tiles per wavefront
  typename Mma<T>::v4 acc[kMR][kMC];
#pragma unroll
  for (int x = 0; x < kMR; ++x)
#pragma unroll
    for (int y = 0; y < kMC; ++y) acc[x][y] = {T(0), T(0), T(0), T(0)};
  const int kq = lane >> 4, l16 = lane & 15;
  load(0);
  for (int k0 = 0; k0 < jb; k0 += kKC) {
    __syncthreads();  // the previous chunk's reads are done
#pragma unroll
    for (int u = 0; u < kLA; ++u) {
      const int idx = tid + 256 * u;
      As[idx / kTR][idx % kTR] = ra[u];
    }
#pragma unroll
    for (int u = 0; u < kLB; ++u) {
      const int idx = tid + 256 * u;
      Bs[idx / kTC][idx % kTC] = rbv[u];
    }
    __syncthreads();
    if (k0 + kKC < jb) load(k0 + kKC);
#pragma unroll
    for (int kk = 0; kk < kKC; kk += 4) {
      T bi[kMR], al[kMC];
#pragma unroll
      for (int x = 0; x < kMR; ++x) bi[x] = As[kk + kq][wave * (kTR / 4) + 16 * x + l16];
#pragma unroll
      for (int y = 0; y < kMC; ++y) al[y] = Bs[kk + kq][16 * y + l16];
#pragma unroll
      for (int x = 0; x < kMR; ++x)
#pragma unroll
        for (int y = 0; y < kMC; ++y) acc[x][y] = Mma<T>::op(al[y], bi[x], acc[x][y]);
    }
  }
#pragma unroll
  for (int x = 0; x < kMR; ++x) {
    const int i = rb + wave * (kTR / 4) + 16 * x + l16;
    if (i >= m) continue;
#pragma unroll
    for (int y = 0; y < kMC; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int l = cb + 16 * y + Mma<T>::row(lane, e);
        if (l < n1 && (i >= n1 || i >= l)) Pn[i + static_cast<int64_t>(l) * ld] -= acc[x][y][e];
      }
  }
  if (blockIdx.x == 0 && blockIdx.y == 0 && j0 + jb < ns) {
    __shared__ T sm[kDg];
    __syncthreads();  // the tile's stores by every wavefront before wavefront 0 reads them
    panel_diag<T>(Pn + jb + static_cast<int64_t>(jb) * ld, ld, min(kNB2, ns - j0 - jb), sm,
                  info + b);
    T* o = dg + static_cast<int64_t>(blockIdx.z) * kDg;
    for (int q = threadIdx.x; q < kDg; q += 256) o[q] = sm[q];
  }
}

// C[i][l] -= sum_k Pan[i][k] Pan[l][k] over the trailing region of the
// solved panel columns [j0, j0 + jb), jb = min(KW, ns - j0): rows i in
// [jb, m) (J/R rows, then the E rows still live), columns l in [jb, ce),
// lower triangle only for the J/R rows.  ce = n1 (cend < 0: the whole
// trailing front) or, inside an outer panel, min(cend, ns) - j0 (the outer
// panel's own columns; the rest of the front takes one update of width KW =
// kOuter after the outer panel).  Workgroup (x, y) covers rows jb + 64x..,
// columns jb + 64y..; z = front x batch member.  Each wavefront: 16 rows x 64
// columns (four 16x16 MFMA tiles), computed transposed so a lane's results
// run down a column of F.
template <typename T, int KW>
__global__ __launch_bounds__(256) void k_trail_big(const SupDev* __restrict__ sup,
                                                  const int* __restrict__ lvl,
                                                  T* __restrict__ fcur, int64_t fcur_bs,
                                                  int j0, int cend, int nbatch,
                                                  T* __restrict__ dg, int* __restrict__ info) {
  const int f = blockIdx.z / nbatch, b = blockIdx.z % nbatch;
  const SupDev S = sup[lvl[f]];
  const int ns = S.ns, nf = S.ns + S.rs, ld = nf + S.ns;
  if (j0 >= ns) return;
  // E rows e >= j0 + jb are zero in these panel columns (see k_factor_level)
  const int jb = min(KW, ns - j0), m = min(ld - j0, nf + jb), n1 = nf - j0;
  const int ce = cend < 0 ? n1 : min(n1, min(cend, ns) - j0);
  const int rb = jb + blockIdx.x * 64, cb = jb + blockIdx.y * 64;
  if (rb >= m || cb >= ce || (rb + 63 < n1 && rb + 63 < cb)) return;  // empty or upper
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T* Pn = fcur + b * fcur_bs + S.front_off + j0 + static_cast<int64_t>(j0) * ld;
  const int i = rb + wave * 16 + (lane & 15);
  const int kq = lane >> 4;
  typename Mma<T>::v4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = {T(0), T(0), T(0), T(0)};
  constexpr int kStage = 8;  // k-steps whose operand loads are issued together
#pragma unroll(KW <= kNB2 ? KW / (4 * kStage) : 1)
  for (int k0 = 0; k0 < KW; k0 += 4 * kStage) {
    T bv[kStage], av[kStage][4];
#pragma unroll
    for (int j = 0; j < kStage; ++j) {
      const int kk = k0 + 4 * j + kq;
      const bool kok = kk < jb;
      bv[j] = (kok && i < m) ? Pn[i + static_cast<int64_t>(kk) * ld] : T(0);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int l = cb + 16 * t + (lane & 15);
        av[j][t] = (kok && l < ce) ? Pn[l + static_cast<int64_t>(kk) * ld] : T(0);
      }
    }
#pragma unroll
    for (int j = 0; j < kStage; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = Mma<T>::op(av[j][t], bv[j], acc[t]);
  }
  if (i < m) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int l = cb + 16 * t + Mma<T>::row(lane, e);
        if (l < ce && (i >= n1 || i >= l)) Pn[i + static_cast<int64_t>(l) * ld] -= acc[t][e];
      }
  }
  // lookahead: tile (0, 0) is the next panel's (columns j0 + jb on) diagonal
  // block, now final
  if (blockIdx.x == 0 && blockIdx.y == 0 && j0 + jb < ns) {
    __shared__ T sm[kDg];
    __syncthreads();  // the tile's stores by every wavefront before wavefront 0 reads them
    panel_diag<T>(Pn + jb + static_cast<int64_t>(jb) * ld, ld, min(kNB2, ns - j0 - jb), sm,
                  info + b);
    T* o = dg + static_cast<int64_t>(blockIdx.z) * kDg;
    for (int q = threadIdx.x; q < kDg; q += 256) o[q] = sm[q];
  }
}

// the sweep panels of a big front: tile (x, y) holds panel rows r in
// [64x, 64x+64) and columns c in [64y, 64y+64), read from the E rows (F row
// nf + c, column r; contiguous in c) straight into PT and, transposed through
// LDS, into P (contiguous in r)
template <typename T>
__global__ __launch_bounds__(256) void k_front_out_big(const SupDev* __restrict__ sup,
                                                      const int* __restrict__ lvl,
                                                      const T* __restrict__ fcur,
                                                      int64_t fcur_bs, T* __restrict__ panels,
                                                      T* __restrict__ panelsT, int64_t panel_bs,
                                                      int nbatch) {
  __shared__ T tile[64][65];
  const int f = blockIdx.z / nbatch, b = blockIdx.z % nbatch;
  const SupDev S = sup[lvl[f]];
  const int ns = S.ns, nf = S.ns + S.rs, ld = nf + S.ns;
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  if (r0 >= nf || c0 >= ns) return;
  const T* F = fcur + b * fcur_bs + S.front_off;
  T* P = panels + b * panel_bs + S.panel_off;
  T* PT = panelsT + b * panel_bs + S.panel_off;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int rr = w; rr < 64; rr += 4) {
    const int r = r0 + rr, c = c0 + lane;
    T v = T(0);
    if (r < nf && c < ns) {
      v = F[(nf + c) + static_cast<int64_t>(r) * ld];
      if (r < ns && r < c) v = T(0);
      if (r >= ns) v = -v;
      PT[static_cast<int64_t>(r) * ns + c] = v;
    }
    tile[rr][lane] = v;
  }
  __syncthreads();
  for (int cc = w; cc < 64; cc += 4) {
    const int c = c0 + cc, r = r0 + lane;
    if (r < nf && c < ns) P[r + static_cast<int64_t>(c) * nf] = tile[lane][cc];
  }
}

// ---------------------------------------------------------------------------
// multi-RHS sweeps on MFMA (Mma<T> above), one wavefront per 16x16 output
// tile at a time.
// ---------------------------------------------------------------------------

constexpr int kMaxTiles = 8;  // 16x16 register tiles a wave keeps across a barrier

// C tile = A(rows r0.., K) * B(K, cols c0..), A read through A(r, k) (global),
// B from LDS rows brow(k) with row stride TC.  Operands are fetched kPrefetch
// k-steps at a time so that the global loads of a chunk are in flight together.
constexpr int kPrefetch = 8;
template <typename T, int TC, typename AFn, typename BRow>
__device__ inline typename Mma<T>::v4 mma_tile(AFn A, int M, int K, const T* Bs, BRow brow,
                                               int r0, int c0, int lane) {
  typename Mma<T>::v4 acc = {T(0), T(0), T(0), T(0)};
  const int rr = r0 + (lane & 15);
  const int kq = lane >> 4;
  const int cc = c0 + (lane & 15);
  const bool rok = rr < M;
  for (int k0 = 0; k0 < K; k0 += 4 * kPrefetch) {
    T av[kPrefetch], bv[kPrefetch];
#pragma unroll
    for (int j = 0; j < kPrefetch; ++j) {
      const int kk = k0 + 4 * j + kq;
      const bool kok = kk < K;
      av[j] = (rok && kok) ? A(rr, kk) : T(0);
      bv[j] = kok ? Bs[brow(kk) * TC + cc] : T(0);
    }
#pragma unroll
    for (int j = 0; j < kPrefetch; ++j)
      if (k0 + 4 * j < K) acc = Mma<T>::op(av[j], bv[j], acc);
  }
  return acc;
}

struct RowId {
  __device__ inline int operator()(int k) const { return k; }
};

// Sweep grids are 1-D over (column tile, supernode, batch), column tile
// fastest, so that the workgroups that read one supernode's panel run
// together.  The hardware deals workgroups round-robin to the 8 XCDs; the
// remap gives each XCD a contiguous range of logical indices, so a panel is
// fetched into one L2 and re-read from there.  gridDim.x is a multiple of 8;
// indices at or past `total` are padding.
struct SweepBlock {
  int s, b, z;
  bool live;
};
__device__ inline SweepBlock sweep_block(const int* lvl, int cnt, int nzb, int total) {
  const int per = gridDim.x >> 3;
  const int L = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  SweepBlock o;
  o.live = L < total;
  const int Lc = o.live ? L : 0;
  o.z = Lc % nzb;
  const int rest = Lc / nzb;
  o.s = lvl[rest % cnt];
  o.b = rest / cnt;
  return o;
}

// forward sweep (deepest level first): f = [b_J; sum of the children's
// updates] in LDS, then [y_J; w] = P_s f_J: y_J to W, u_s = f_R - w to the
// update buffer of this depth.  PL: panel staged in LDS, `ct` column tiles
// per workgroup.
constexpr int kFwdKids = 8;  // children whose extend-add data k_fwd_level stages
template <typename T, int TC, bool PL>
__global__ __launch_bounds__(kThreads) void k_fwd_level(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl,
    const int* __restrict__ relpos, const int* __restrict__ cidx,
    const int* __restrict__ active, const T* __restrict__ panels,
    int64_t panel_bs, const double* __restrict__ Bsrc, int64_t b_bs,
    const int* __restrict__ brow, T* __restrict__ W, int64_t w_bs, int k,
    T* __restrict__ ucur, int64_t ucur_bs, const T* __restrict__ uprev,
    int64_t uprev_bs, const T* __restrict__ ufz, int64_t ufz_bs, int ct, int pmax, int cnt,
    int nzb, int total) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const SweepBlock sb = sweep_block(lvl, cnt, nzb, total);
  if (!sb.live) return;
  const int s = sb.s;
  const int b = sb.b;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const SupDev S = sup[s];
  const int ns = S.ns, nf = S.ns + S.rs;
  const T* Pg = panels + b * panel_bs + S.panel_off;
  T* Ps = reinterpret_cast<T*>(smem_raw);
  T* f = PL ? Ps + pmax : Ps;
  if (PL) {
    const int np = nf * ns;
    for (int idx = tid; idx < np; idx += kThreads) Ps[idx] = Pg[idx];
  }
  const T* P = PL ? Ps : Pg;
  auto Pa = [=](int r, int kk) { return P[r + kk * nf]; };
  T* Wb = W + b * w_bs;
  const double* Bb = Bsrc + b * b_bs;
  T* uo = ucur + b * ucur_bs + S.upd_off * k;
  // the first kFwdKids children: update-block offset (ufz-based ones stored
  // as -(offset + 1)), R rows (0: inactive, skipped) and relative positions,
  // read once per workgroup, so that each child's extend-add waits for one
  // memory round trip instead of a chain of index loads
  __shared__ int64_t kid_u[kFwdKids];
  __shared__ int kid_rs[kFwdKids], kid_rptr[kFwdKids];
  const int nkid = S.c_end - S.c_begin;
  if (tid < min(nkid, kFwdKids)) {
    const int c = cidx[S.c_begin + tid];
    const SupDev C = sup[c];
    kid_rs[tid] = (!active || active[c]) ? C.rs : 0;
    kid_rptr[tid] = C.rptr;
    kid_u[tid] = C.fz_uoff >= 0 ? -(C.fz_uoff * k + 1) : C.upd_off * k;
  }
  constexpr int NCT = TC / 16;
  const int nt = ((nf + 15) / 16) * NCT;
  for (int tile = 0; tile < ct; ++tile) {
    const int col0 = (sb.z * ct + tile) * TC;
    if (col0 >= k) break;
    __syncthreads();
    for (int idx = tid; idx < nf * TC; idx += kThreads) {
      const int r = idx / TC, c = idx % TC;
      T v = T(0);
      if (r < ns && col0 + c < k) {
        const int gr = brow ? brow[S.first + r] : S.first + r;
        v = static_cast<T>(Bb[static_cast<int64_t>(gr) * k + col0 + c]);
      }
      f[idx] = v;
    }
    __syncthreads();
    for (int q = 0; q < nkid; ++q) {
      int rsc, rptr;
      const T* u;
      if (q < kFwdKids) {
        rsc = kid_rs[q];
        rptr = kid_rptr[q];
        const int64_t uo_c = kid_u[q];
        u = uo_c < 0 ? ufz + b * ufz_bs + (-uo_c - 1) : uprev + b * uprev_bs + uo_c;
      } else {
        const int c = cidx[S.c_begin + q];
        const SupDev C = sup[c];
        rsc = (!active || active[c]) ? C.rs : 0;
        rptr = C.rptr;
        u = C.fz_uoff >= 0 ? ufz + b * ufz_bs + C.fz_uoff * k
                           : uprev + b * uprev_bs + C.upd_off * k;
      }
      if (!rsc) continue;  // inactive child (uniform)
      const int* rp = relpos + rptr;
      for (int idx = tid; idx < rsc * TC; idx += kThreads) {
        const int r = idx / TC, cl = idx % TC;
        if (col0 + cl < k) f[rp[r] * TC + cl] += u[static_cast<int64_t>(r) * k + col0 + cl];
      }
      __syncthreads();
    }
    for (int t = wave; t < nt; t += 4) {
      const int tr = (t / NCT) * 16, tc = (t % NCT) * 16;
      // the J rows of P_s hold L_JJ^{-1}, lower triangular: row tile tr needs
      // columns up to its last row only
      const int kt = tr < ns ? min(ns, tr + 16) : ns;
      typename Mma<T>::v4 a = mma_tile<T, TC>(Pa, nf, kt, f, RowId(), tr, tc, lane);
      const int c = tc + (lane & 15);
      if (col0 + c >= k) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = tr + Mma<T>::row(lane, i);
        if (r < ns)
          Wb[static_cast<int64_t>(S.first + r) * k + col0 + c] = a[i];
        else if (r < nf)
          uo[static_cast<int64_t>(r - ns) * k + col0 + c] = f[r * TC + c] - a[i];
      }
    }
  }
}

// backward sweep (root first): g = [y_J; -x_R] in LDS, x_J = P_s^T g to W
// (and the FP64 solution, xmode 1 store / 2 accumulate).  PL / ct as above.
template <typename T, int TC, bool PL>
__global__ __launch_bounds__(kThreads) void k_bwd_level(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl,
    const int* __restrict__ ridx, const int* __restrict__ active,
    const T* __restrict__ panels, int64_t panel_bs, T* __restrict__ W,
    int64_t w_bs, int k, double* __restrict__ X64, const int* __restrict__ xrow,
    int xmode, int ct, int pmax, int cnt, int nzb, int total) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const SweepBlock sb = sweep_block(lvl, cnt, nzb, total);
  if (!sb.live) return;
  const int s = sb.s;
  const int b = sb.b;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const SupDev S = sup[s];
  const int ns = S.ns, nf = S.ns + S.rs;
  const bool act = !active || active[s];
  const T* Pg = panels + b * panel_bs + S.panel_off;  // the transposed panel PT
  T* Ps = reinterpret_cast<T*>(smem_raw);
  T* g = PL ? Ps + pmax : Ps;
  if (PL) {
    const int np = nf * ns;
    for (int idx = tid; idx < np; idx += kThreads) Ps[idx] = Pg[idx];
  }
  const T* P = PL ? Ps : Pg;
  auto PaT = [=](int i, int kk) { return P[i + kk * ns]; };
  T* Wb = W + b * w_bs;
  constexpr int NCT = TC / 16;
  const int nt = ((ns + 15) / 16) * NCT;
  for (int tile = 0; tile < ct; ++tile) {
    const int col0 = (sb.z * ct + tile) * TC;
    if (col0 >= k) break;
    __syncthreads();
    for (int idx = tid; idx < nf * TC; idx += kThreads) {
      const int r = idx / TC, c = idx % TC;
      T val = T(0);
      if (col0 + c < k) {
        if (r < ns)
          val = act ? Wb[static_cast<int64_t>(S.first + r) * k + col0 + c] : T(0);
        else
          val = -Wb[static_cast<int64_t>(ridx[S.rptr + r - ns]) * k + col0 + c];
      }
      g[idx] = val;
    }
    __syncthreads();
    for (int t = wave; t < nt; t += 4) {
      const int tr = (t / NCT) * 16, tc = (t % NCT) * 16;
      typename Mma<T>::v4 a = mma_tile<T, TC>(PaT, ns, nf, g, RowId(), tr, tc, lane);
      const int c = tc + (lane & 15);
      if (col0 + c >= k) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = tr + Mma<T>::row(lane, i);
        if (r < ns) {
          Wb[static_cast<int64_t>(S.first + r) * k + col0 + c] = a[i];
          if (xmode) {
            const int gr = xrow ? xrow[S.first + r] : S.first + r;
            double* xp = X64 + b * w_bs + static_cast<int64_t>(gr) * k + col0 + c;
            *xp = (xmode == 2 ? *xp : 0.0) + static_cast<double>(a[i]);
          }
        }
      }
    }
  }
}

// backward sweep, wavefront tiles read straight from W into the matrix
// operands (no LDS staging, no barrier except for split-K):
//   x_J = P_s^T g,   g = [y_J; -x_R]  (rows of W: J rows, then R rows)
// A wavefront owns CW = 16 V columns: lane l holds the V adjacent columns
// c0 + V (l%16) + v of every row it touches, loaded and stored as one vector,
// and the column set {c0 + V j + v : j < 16} is the v-th 16-column matrix
// tile.  KS = 1: the four wavefronts of a workgroup take four adjacent column
// blocks; KS = 4: they split the K range (the front rows) of one column block
// and add their partial tiles through LDS (levels with few supernodes);
// KS = 0: they take the four 16-row tiles (MT = 1, ns <= 64) of one column
// block, each walking the whole K range, and store after a barrier (the other
// wavefronts still read y_J from those rows until then): four times the
// wavefronts of KS = 1 for the upper levels, where a level holds too few
// wavefronts to hide the latency of its K loop.
// MT: 16-row output tiles (ns <= 16 MT); PF: k-steps whose loads are in
// flight together.  Grid as sweep_block, nzb = column groups per workgroup.
template <typename T, int V>
struct Vec {
  typedef T v __attribute__((ext_vector_type(V)));
};
template <typename T>
struct Vec<T, 1> {
  typedef T v;
};

template <typename T, int V>
__device__ inline void load_cols(const T* p, int navail, T (&out)[V]) {
  if (navail >= V) {
    typename Vec<T, V>::v t;
    __builtin_memcpy(&t, p, sizeof(t));
    if constexpr (V == 1) {
      out[0] = t;
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) out[v] = t[v];
    }
  } else {
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = v < navail ? p[v] : T(0);
  }
}

// x with its sign flipped where m holds the sign bit (m is 0 or 0x80000000):
// one integer xor where the value is used.  A load negated in place would
// make the wavefront wait for it right there, before the loads that follow
// it are issued.
__device__ inline float flip_sign(float x, uint32_t m) {
  return __uint_as_float(__float_as_uint(x) ^ m);
}
__device__ inline double flip_sign(double x, uint32_t m) {
  return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(m) << 32));
}

// Wavefronts per SIMD asked of the register allocator (kBwdWaves): 5 for the
// FP32 split-K variants and the wide ones of at most 3 row tiles with 3
// k-steps loaded ahead (bwd_wave_mt), 4 for the wide 4-row-tile one; their
// accumulators then stay in VGPRs instead of adding AGPRs, with no spills
// (split-K MT 8: 87 + 32 -> 87 registers, 4 -> 5 wavefronts per SIMD; MT 6:
// 4 -> 6; wide MT 3, the leaf level: 62 + 52 -> 94, 4 -> 5; wide MT 2: 5 -> 6;
// wide MT 4: 78 + 68 -> 113, 3 -> 4).  The FP64 ones would spill at MT 6 and
// 8, the wide ones beyond these counts or with 4 k-steps ahead (all measured
// slower).
template <typename T, int MT, int V, int KS, int PF>
constexpr int kBwdWaves = sizeof(T) != 4 ? 1
                          : KS > 1 ? 5
                          : KS == 1 && MT <= 3 && PF <= 3 ? 5
                          : KS == 1 && MT == 4 && PF <= 3 ? 4
                          : 1;
template <typename T, int MT, int V, int KS, int PF>
__global__ __launch_bounds__(kThreads, (kBwdWaves<T, MT, V, KS, PF>)) void k_bwd_wave(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl,
    const int* __restrict__ ridx, const int* __restrict__ active,
    const T* __restrict__ panelsT, int64_t panel_bs, T* __restrict__ W, int64_t w_bs, int k,
    double* __restrict__ X64, const int* __restrict__ xrow, int xmode, int cnt, int nzb,
    int total, int nzc, T* __restrict__ xs, const int64_t* __restrict__ xs_off,
    int64_t xs_bs) {
  constexpr int CW = 16 * V;
  const SweepBlock sb = sweep_block(lvl, cnt, nzb, total);
  if (!sb.live) return;  // uniform over the workgroup
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const SupDev S = sup[sb.s];
  const int ns = S.ns, nf = S.ns + S.rs;
  // the R rows' positions, staged in LDS once per workgroup (every wavefront
  // of a workgroup works on the same supernode), so that the x_R loads of a
  // k-step wait for one memory round trip, not for an index load first
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr size_t kRedBytes = KS > 1 ? static_cast<size_t>(KS) * MT * V * 4 * 64 * sizeof(T) : 0;
  int* rl = reinterpret_cast<int*>(smem_raw + kRedBytes);
  for (int i = threadIdx.x; i < S.rs; i += kThreads) rl[i] = ridx[S.rptr + i];
  __syncthreads();
  // z = column group + nzc * row group; a row group is 16 MT output rows
  // (KS = 0: 64 rows, wavefront w taking rows [16 w, 16 w + 16))
  const int zc = sb.z % nzc;
  const int r0 = KS == 0 ? 16 * wave : (sb.z / nzc) * 16 * MT;
  const int col0 = (KS == 1 ? zc * 4 + wave : zc) * CW;
  if (col0 >= k) return;  // uniform over the workgroup when KS != 1
  // uniform over the workgroup, except KS = 0, whose idle wavefronts still
  // take part in the barrier below
  if (KS != 0 && r0 >= ns) return;
  const bool rows = r0 < ns;
  const bool act = !active || active[sb.s];
  const T* PT = panelsT + sb.b * panel_bs + S.panel_off;
  T* Wb = W + sb.b * w_bs;
  const int cl = col0 + V * (lane & 15);  // first column of this lane
  const int navail = k - cl;               // columns of the lane inside the block
  const int kq = lane >> 4;
  // K range of this wavefront
  int kbeg = 0, kend = nf;
  if (KS > 1) {
    const int chunk = ((nf + 4 * KS - 1) / (4 * KS)) * 4;
    kbeg = min(nf, wave * chunk);
    kend = min(nf, kbeg + chunk);
  }
  typename Mma<T>::v4 acc[MT][V];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int v = 0; v < V; ++v) acc[t][v] = {T(0), T(0), T(0), T(0)};
  for (int k0 = kbeg; k0 < (rows ? kend : 0); k0 += 4 * PF) {
    T bv[PF][V], av[PF][MT];
    uint32_t sg[PF];  // the R rows enter negated: sign flipped at the product
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      const int kk = k0 + 4 * j + kq;
      const bool kok = kk < kend;
      // one load path for J and R rows (row chosen by a select): loads into
      // the same registers on two branches would make the second branch wait
      // for the first one's loads
      const int kr = kk - ns;
      const int ri = kr >= 0 && kok ? rl[kr] : 0;
      const int row = kr >= 0 ? ri : S.first + kk;
      sg[j] = kr >= 0 ? 0x80000000u : 0u;
#pragma unroll
      for (int v = 0; v < V; ++v) bv[j][v] = T(0);
      if (kok && navail > 0 && (kr >= 0 || act))
        load_cols<T, V>(Wb + static_cast<int64_t>(row) * k + cl, navail, bv[j]);
      // L_JJ^{-1} (the J rows of P_s) is lower triangular: k-steps wholly
      // above a tile's first row multiply zeros and are skipped, and so are
      // tiles wholly past this supernode's ns rows (the level's tile count
      // fits its tallest supernode); not in the split-K variants, where the
      // branches cost more registers than the skipped work saves
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const int r = r0 + t * 16 + (lane & 15);
        const bool live = KS > 1 || (k0 + 4 * j + 4 > r0 + 16 * t && r0 + 16 * t < ns);
        av[j][t] = (kok && live && r < ns) ? PT[r + static_cast<int64_t>(kk) * ns] : T(0);
      }
    }
#pragma unroll
    for (int j = 0; j < PF; ++j)
      if (k0 + 4 * j < kend) {
        T b[V];
#pragma unroll
        for (int v = 0; v < V; ++v) b[v] = flip_sign(bv[j][v], sg[j]);
#pragma unroll
        for (int t = 0; t < MT; ++t)
          if (KS > 1 || (k0 + 4 * j + 4 > r0 + 16 * t && r0 + 16 * t < ns)) {
#pragma unroll
            for (int v = 0; v < V; ++v) acc[t][v] = Mma<T>::op(av[j][t], b[v], acc[t][v]);
          }
      }
  }
  if (KS > 1) {
    // partial tiles through LDS: wave w adds up and stores the tiles t with
    // t % KS == w
    T* red = reinterpret_cast<T*>(smem_raw);  // [KS][MT][V][4][64]
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int v = 0; v < V; ++v)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          red[(((wave * MT + t) * V + v) * 4 + i) * 64 + lane] = acc[t][v][i];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      if (t % KS != wave) continue;
#pragma unroll
      for (int v = 0; v < V; ++v)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          T a = T(0);
#pragma unroll
          for (int w = 0; w < KS; ++w) a += red[(((w * MT + t) * V + v) * 4 + i) * 64 + lane];
          acc[t][v][i] = a;
        }
    }
  }
  if (KS == 0) __syncthreads();  // every wavefront has read y_J
  if (navail <= 0 || !rows) return;
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    if (KS > 1 && t % KS != wave) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + t * 16 + Mma<T>::row(lane, i);
      if (r >= ns) continue;
      // with row groups the other groups still read y_J from W: the result
      // goes to the scratch block xs, copied into W by k_bwd_copy_big
      T* dst = xs ? xs + sb.b * xs_bs + (xs_off[sb.s] + r) * k + cl
                  : Wb + static_cast<int64_t>(S.first + r) * k + cl;
      if (navail >= V) {
        typename Vec<T, V>::v o;
        if constexpr (V == 1) {
          o = acc[t][0][i];
        } else {
#pragma unroll
          for (int v = 0; v < V; ++v) o[v] = acc[t][v][i];
        }
        __builtin_memcpy(dst, &o, sizeof(o));
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (v < navail) dst[v] = acc[t][v][i];
      }
      if (xmode && !xs) {
        const int gr = xrow ? xrow[S.first + r] : S.first + r;
        double* xp = X64 + sb.b * w_bs + static_cast<int64_t>(gr) * k + cl;
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (v < navail) xp[v] = (xmode == 2 ? xp[v] : 0.0) + static_cast<double>(acc[t][v][i]);
      }
    }
  }
}

// Forward sweep for fronts too tall for the in-LDS kernel: (1) f = [b_J; 0] +
// the children's updates, assembled in a global block per supernode (rows
// of the front, k columns; children one after the other); (2) the GEMM
// [y_J; w] = P_s f_J by wavefronts reading f and P_s straight into MFMA
// registers, 16 MT rows x 16 columns per wavefront: y_J to the solution
// rows, u_s = f_R - w to the update buffer.
// Grid of both: 1-D over (z, supernode of the level, batch member), z
// fastest, XCD-remapped as sweep_block.
struct LevelBlock {
  int q, b, z;
  bool live;
};
__device__ inline LevelBlock level_block(int cnt, int nzb, int total) {
  const int per = gridDim.x >> 3;
  const int L = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  LevelBlock o;
  o.live = L < total;
  const int Lc = o.live ? L : 0;
  o.z = Lc % nzb;
  o.q = (Lc / nzb) % cnt;
  o.b = Lc / nzb / cnt;
  return o;
}

// x_J of the row-grouped supernodes of a level: scratch -> W (and the FP64
// solution, xmode 1 store / 2 accumulate); workgroup (x, front x batch, z)
// copies rows [32x, 32x + 32) and columns [64z, 64z + 64)
constexpr int kCopyRows = 32;
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_copy_big(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl, const T* __restrict__ xs,
    const int64_t* __restrict__ xs_off, int64_t xs_bs,
    T* __restrict__ W, int64_t w_bs, int k, double* __restrict__ X64,
    const int* __restrict__ xrow, int xmode, int nbatch) {
  const int f = blockIdx.y / nbatch, b = blockIdx.y % nbatch;
  const int s = lvl[f];
  const SupDev S = sup[s];
  const int r0 = blockIdx.x * kCopyRows, c = blockIdx.z * 64 + (threadIdx.x & 63);
  if (r0 >= S.ns || c >= k) return;
  const T* x = xs + b * xs_bs + xs_off[s] * k;
  T* Wb = W + b * w_bs;
  constexpr int kU = kCopyRows / 4;
  T v[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int r = r0 + (threadIdx.x >> 6) + 4 * u;
    v[u] = r < S.ns ? x[static_cast<int64_t>(r) * k + c] : T(0);
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int r = r0 + (threadIdx.x >> 6) + 4 * u;
    if (r >= S.ns) continue;
    Wb[static_cast<int64_t>(S.first + r) * k + c] = v[u];
    if (xmode) {
      const int gr = xrow ? xrow[S.first + r] : S.first + r;
      double* xp = X64 + b * w_bs + static_cast<int64_t>(gr) * k + c;
      *xp = (xmode == 2 ? *xp : 0.0) + static_cast<double>(v[u]);
    }
  }
}

// forward-sweep right-hand side of a big front, f = [B_J; 0] plus the
// children's updates: k_fwd_init_big fills workgroup x's kFwdRows rows and
// 64 columns; k_fwd_extend_big adds the kchild-th child's update rows (one
// child's rows land on distinct front rows, so the workgroups never collide
// and the sum order is fixed)
constexpr int kFwdRows = 32;
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_init_big(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl, const double* __restrict__ Bsrc,
    int64_t b_bs, const int* __restrict__ brow, int k, T* __restrict__ fs,
    const int64_t* __restrict__ fs_off, int64_t fs_bs, int nbatch) {
  const int f = blockIdx.y / nbatch, b = blockIdx.y % nbatch;
  const int s = lvl[f];
  const SupDev S = sup[s];
  const int ns = S.ns, nf = S.ns + S.rs;
  const int r0 = blockIdx.x * kFwdRows, col0 = blockIdx.z * 64;
  if (r0 >= nf) return;
  T* fo = fs + b * fs_bs + fs_off[s] * k;
  const double* Bb = Bsrc + b * b_bs;
  const int c = col0 + (threadIdx.x & 63);
  if (c >= k) return;
  constexpr int kU = kFwdRows / 4;
  T v[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int r = r0 + (threadIdx.x >> 6) + 4 * u;
    v[u] = T(0);
    if (r < ns) {
      const int gr = brow ? brow[S.first + r] : S.first + r;
      v[u] = static_cast<T>(Bb[static_cast<int64_t>(gr) * k + c]);
    }
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int r = r0 + (threadIdx.x >> 6) + 4 * u;
    if (r < nf) fo[static_cast<int64_t>(r) * k + c] = v[u];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_extend_big(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl, const int* __restrict__ relpos,
    const int* __restrict__ cidx, const int* __restrict__ active, int k,
    const T* __restrict__ uprev, int64_t uprev_bs, const T* __restrict__ ufz, int64_t ufz_bs,
    T* __restrict__ fs, const int64_t* __restrict__ fs_off, int64_t fs_bs, int kchild,
    int nbatch) {
  const int f = blockIdx.y / nbatch, b = blockIdx.y % nbatch;
  const int s = lvl[f];
  const SupDev S = sup[s];
  if (S.c_begin + kchild >= S.c_end) return;
  const int ch = cidx[S.c_begin + kchild];
  if (active && !active[ch]) return;
  const SupDev C = sup[ch];
  const int r0 = blockIdx.x * kFwdRows, col0 = blockIdx.z * 64;
  if (r0 >= C.rs) return;
  const T* u = C.fz_uoff >= 0 ? ufz + b * ufz_bs + C.fz_uoff * k
                              : uprev + b * uprev_bs + C.upd_off * k;
  const int* rp = relpos + C.rptr;
  T* fo = fs + b * fs_bs + fs_off[s] * k;
  const int c = col0 + (threadIdx.x & 63);
  if (c >= k) return;
  constexpr int kU = kFwdRows / 4;
  T uv[kU], fv[kU];
  int64_t fa[kU];
#pragma unroll
  for (int q = 0; q < kU; ++q) {
    const int r = r0 + (threadIdx.x >> 6) + 4 * q;
    fa[q] = -1;
    if (r < C.rs) {
      fa[q] = static_cast<int64_t>(rp[r]) * k + c;
      uv[q] = u[static_cast<int64_t>(r) * k + c];
      fv[q] = fo[fa[q]];
    }
  }
#pragma unroll
  for (int q = 0; q < kU; ++q)
    if (fa[q] >= 0) fo[fa[q]] = fv[q] + uv[q];
}

template <typename T, int MT>
__global__ __launch_bounds__(256) void k_fwd_wave_big(
    const SupDev* __restrict__ sup, const int* __restrict__ lvl, int cnt,
    const T* __restrict__ panels, int64_t panel_bs, const T* __restrict__ fs,
    const int64_t* __restrict__ fs_off, int64_t fs_bs, T* __restrict__ W, int64_t w_bs, int k,
    T* __restrict__ ucur, int64_t ucur_bs, int nzc, int nzb, int total) {
  constexpr int PF = 4;
  const LevelBlock lb = level_block(cnt, nzb, total);
  if (!lb.live) return;
  const int s = lvl[lb.q];
  const SupDev S = sup[s];
  const int ns = S.ns, nf = S.ns + S.rs;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int zc = lb.z % nzc, r0 = (lb.z / nzc) * 16 * MT;
  const int col0 = (zc * 4 + wave) * 16;
  if (r0 >= nf || col0 >= k) return;
  const int c = col0 + (lane & 15);
  const bool cok = c < k;
  const int kq = lane >> 4;
  const T* P = panels + lb.b * panel_bs + S.panel_off;  // nf x ns, column-major
  const T* f = fs + lb.b * fs_bs + fs_off[s] * k;
  typename Mma<T>::v4 acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = {T(0), T(0), T(0), T(0)};
  // the J rows of P_s hold L_JJ^{-1}, lower triangular: rows below ns need
  // columns up to their own index only
  const int kend = r0 + 16 * MT <= ns ? r0 + 16 * MT : ns;
  for (int k0 = 0; k0 < kend; k0 += 4 * PF) {
    T bv[PF], av[PF][MT];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      const int kk = k0 + 4 * j + kq;
      const bool kok = kk < kend;
      bv[j] = (kok && cok) ? f[static_cast<int64_t>(kk) * k + c] : T(0);
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const int r = r0 + t * 16 + (lane & 15);
        av[j][t] = (kok && r < nf) ? P[r + static_cast<int64_t>(kk) * nf] : T(0);
      }
    }
#pragma unroll
    for (int j = 0; j < PF; ++j)
      if (k0 + 4 * j < kend) {
#pragma unroll
        for (int t = 0; t < MT; ++t) acc[t] = Mma<T>::op(av[j][t], bv[j], acc[t]);
      }
  }
  if (!cok) return;
  T* Wb = W + lb.b * w_bs;
  T* uo = ucur + lb.b * ucur_bs + S.upd_off * k;
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + t * 16 + Mma<T>::row(lane, i);
      if (r < ns)
        Wb[static_cast<int64_t>(S.first + r) * k + c] = acc[t][i];
      else if (r < nf)
        uo[static_cast<int64_t>(r - ns) * k + c] = f[static_cast<int64_t>(r) * k + c] - acc[t][i];
    }
}

struct LRow {
  const int* l;
  __device__ inline int operator()(int k) const { return l[k]; }
};
struct OffRow {
  int o;
  __device__ inline int operator()(int k) const { return o + k; }
};
// rows of [y_J; x_R] inside a fused subtree block: J rows from jl, R rows via l
struct SplitRow {
  int jl, ns;
  const int* l;
  __device__ inline int operator()(int k) const { return k < ns ? jl + k : l[k - ns]; }
};

// Fused bottom subtree, backward sweep.  The LDS block X holds the subtree's
// rows [lo, hi) and, below them, the root's off-diagonal rows R_T (ancestors:
// final x).  Sublevels root first: t = y_J - L_RJ^T x_R goes to the scratch
// Tb, then x_J = L_JJ^{-T} t back into X; one HBM read and write per row.
template <typename T, int TC>
__global__ __launch_bounds__(kThreads) void k_bwd_fused(
    const SupDev* __restrict__ sup, const FusedDev* __restrict__ fz,
    const int* __restrict__ flv, const int* __restrict__ fmem,
    const int* __restrict__ toff, const int* __restrict__ lridx,
    const int* __restrict__ ridx, const int* __restrict__ active,
    const T* __restrict__ panels, int64_t panel_bs, T* __restrict__ W, int64_t w_bs,
    int k, double* __restrict__ X64, const int* __restrict__ xrow, int xmode) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* X = reinterpret_cast<T*>(smem_raw);
  const FusedDev F = fz[blockIdx.x];
  const int b = blockIdx.y;
  const int col0 = blockIdx.z * TC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nloc = F.hi - F.lo, nrows = nloc + F.rsT;
  T* Tb = X + nrows * TC;
  const bool act = !active || active[F.root];
  T* Wb = W + b * w_bs;
  for (int idx = tid; idx < nrows * TC; idx += kThreads) {
    const int r = idx / TC, c = idx % TC;
    T v = T(0);
    if (col0 + c < k) {
      if (r < nloc)
        v = act ? Wb[static_cast<int64_t>(F.lo + r) * k + col0 + c] : T(0);
      else
        v = Wb[static_cast<int64_t>(ridx[F.rptrT + r - nloc]) * k + col0 + c];
    }
    X[idx] = v;
  }
  __syncthreads();
  constexpr int NCT = TC / 16;
  for (int L = F.lv_begin; L < F.lv_end; ++L) {
    const int m0 = flv[L], m1 = flv[L + 1];
    int base = 0;
    for (int mi = m0; mi < m1; ++mi) {  // x_J = P^T [y_J; -x_R] into the scratch
      const int s = fmem[mi];
      const SupDev S = sup[s];
      const int ns = S.ns, nf = S.ns + S.rs;
      const T* P = panels + b * panel_bs + S.panel_off;  // the transposed panel PT
      auto PaT = [=](int i, int kk) {
        const T v = P[i + static_cast<int64_t>(kk) * ns];
        return kk < ns ? v : -v;
      };
      const int nt = ((ns + 15) / 16) * NCT;
      const int to = toff[s], jl = S.first - F.lo;
      const SplitRow br{jl, ns, lridx + S.rptr};
      for (int t = ((wave - base) % 4 + 4) % 4; t < nt; t += 4) {
        const int tr = (t / NCT) * 16, tc = (t % NCT) * 16;
        typename Mma<T>::v4 acc = mma_tile<T, TC>(PaT, ns, nf, X, br, tr, tc, lane);
        const int c = tc + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = tr + Mma<T>::row(lane, i);
          if (r < ns) Tb[(to + r) * TC + c] = acc[i];
        }
      }
      base += nt;
    }
    __syncthreads();
    for (int mi = m0; mi < m1; ++mi) {
      const int s = fmem[mi];
      const SupDev S = sup[s];
      const int to = toff[s], jl = S.first - F.lo;
      for (int idx = tid; idx < S.ns * TC; idx += kThreads) X[jl * TC + idx] = Tb[to * TC + idx];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < nloc * TC; idx += kThreads) {
    const int r = idx / TC, c = idx % TC;
    if (col0 + c >= k) continue;
    const T v = X[idx];
    Wb[static_cast<int64_t>(F.lo + r) * k + col0 + c] = v;
    if (xmode) {
      const int gr = xrow ? xrow[F.lo + r] : F.lo + r;
      double* xp = X64 + b * w_bs + static_cast<int64_t>(gr) * k + col0 + c;
      *xp = (xmode == 2 ? *xp : 0.0) + static_cast<double>(v);
    }
  }
}

// Fused bottom subtree, forward sweep.  X holds b on the subtree rows and
// accumulates the updates of R_T below them.  Sublevels deepest first:
// y_J = L_JJ^{-1} f_J (to Tb, then into X), then X[R rows] -= L_RJ y_J, one
// colour group at a time so that no two members add into the same row at
// once.  The root's accumulated R_T rows are its update block.
template <typename T, int TC>
__global__ __launch_bounds__(kThreads) void k_fwd_fused(
    const SupDev* __restrict__ sup, const FusedDev* __restrict__ fz,
    const int* __restrict__ flv, const int* __restrict__ fmem,
    const int* __restrict__ fgptr, const int* __restrict__ fgb,
    const int* __restrict__ toff, const int* __restrict__ lridx,
    const int* __restrict__ active, const T* __restrict__ panels, int64_t panel_bs,
    const double* __restrict__ Bsrc, int64_t b_bs, const int* __restrict__ brow,
    T* __restrict__ W, int64_t w_bs, int k, T* __restrict__ ufz, int64_t ufz_bs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* X = reinterpret_cast<T*>(smem_raw);
  const FusedDev F = fz[blockIdx.x];
  if (active && !active[F.root]) return;  // zero right-hand side below this root
  const int b = blockIdx.y;
  const int col0 = blockIdx.z * TC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nloc = F.hi - F.lo, nrows = nloc + F.rsT;
  T* Tb = X + nrows * TC;
  const double* Bb = Bsrc + b * b_bs;
  for (int idx = tid; idx < nrows * TC; idx += kThreads) {
    const int r = idx / TC, c = idx % TC;
    T v = T(0);
    if (r < nloc && col0 + c < k) {
      const int gr = brow ? brow[F.lo + r] : F.lo + r;
      v = static_cast<T>(Bb[static_cast<int64_t>(gr) * k + col0 + c]);
    }
    X[idx] = v;
  }
  __syncthreads();
  constexpr int NCT = TC / 16;
  for (int L = F.lv_end - 1; L >= F.lv_begin; --L) {
    const int m0 = flv[L], m1 = flv[L + 1];
    int base = 0;
    for (int mi = m0; mi < m1; ++mi) {  // y_J = L_JJ^{-1} f_J into the scratch
      const int s = fmem[mi];
      const SupDev S = sup[s];
      const int ns = S.ns, nf = S.ns + S.rs;
      const T* P = panels + b * panel_bs + S.panel_off;
      auto Linv = [=](int r, int kk) { return P[r + static_cast<int64_t>(kk) * nf]; };
      const int nt = ((ns + 15) / 16) * NCT;
      const int to = toff[s], jl = S.first - F.lo;
      for (int t = ((wave - base) % 4 + 4) % 4; t < nt; t += 4) {
        const int tr = (t / NCT) * 16, tc = (t % NCT) * 16;
        typename Mma<T>::v4 acc = mma_tile<T, TC>(Linv, ns, ns, X, OffRow{jl}, tr, tc, lane);
        const int c = tc + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = tr + Mma<T>::row(lane, i);
          if (r < ns) Tb[(to + r) * TC + c] = acc[i];
        }
      }
      base += nt;
    }
    // X[R rows] -= Q_R f_J, one colour group at a time (f_J still in X)
    for (int g = fgptr[L]; g < fgptr[L + 1]; ++g) {
      base = 0;
      for (int mi = fgb[g]; mi < fgb[g + 1]; ++mi) {
        const int s = fmem[mi];
        const SupDev S = sup[s];
        const int ns = S.ns, rs = S.rs, nf = S.ns + S.rs;
        const T* P = panels + b * panel_bs + S.panel_off;
        auto Qr = [=](int r, int kk) { return P[(ns + r) + static_cast<int64_t>(kk) * nf]; };
        const int nt = ((rs + 15) / 16) * NCT;
        const int jl = S.first - F.lo;
        const int* lr = lridx + S.rptr;
        for (int t = ((wave - base) % 4 + 4) % 4; t < nt; t += 4) {
          const int tr = (t / NCT) * 16, tc = (t % NCT) * 16;
          typename Mma<T>::v4 acc = mma_tile<T, TC>(Qr, rs, ns, X, OffRow{jl}, tr, tc, lane);
          const int c = tc + (lane & 15);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int r = tr + Mma<T>::row(lane, i);
            if (r < rs) X[lr[r] * TC + c] -= acc[i];
          }
        }
        base += nt;
      }
      __syncthreads();
    }
    for (int mi = m0; mi < m1; ++mi) {  // y into the subtree rows
      const int s = fmem[mi];
      const SupDev S = sup[s];
      const int to = toff[s], jl = S.first - F.lo;
      for (int idx = tid; idx < S.ns * TC; idx += kThreads) X[jl * TC + idx] = Tb[to * TC + idx];
    }
    __syncthreads();
  }
  T* Wb = W + b * w_bs;
  T* uo = ufz + b * ufz_bs + F.uoff * k;
  for (int idx = tid; idx < nrows * TC; idx += kThreads) {
    const int r = idx / TC, c = idx % TC;
    if (col0 + c >= k) continue;
    if (r < nloc)
      Wb[static_cast<int64_t>(F.lo + r) * k + col0 + c] = X[idx];
    else
      uo[static_cast<int64_t>(r - nloc) * k + col0 + c] = X[idx];
  }
}

// lane-layout probe of the 16x16x4 matrix instruction: C = A(16x4) B(4x16)
template <typename T>
__global__ void k_mma_probe(const T* __restrict__ A, const T* __restrict__ B, T* __restrict__ C) {
  const int lane = threadIdx.x;
  typename Mma<T>::v4 acc = {T(0), T(0), T(0), T(0)};
  acc = Mma<T>::op(A[(lane & 15) * 4 + (lane >> 4)], B[(lane >> 4) * 16 + (lane & 15)], acc);
#pragma unroll
  for (int i = 0; i < 4; ++i) C[Mma<T>::row(lane, i) * 16 + (lane & 15)] = acc[i];
}

// ---------------------------------------------------------------------------
// elementwise / gather helpers (one wavefront per row, lanes over columns)
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ void k_permute(const TI* __restrict__ in, TO* __restrict__ out,
                          const int* __restrict__ perm, int n, int k,
                          int64_t in_bs, int64_t out_bs, int to_elim) {
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int b = blockIdx.y;
  if (row >= n) return;
  const int src = to_elim ? perm[row] : row;
  const int dst = to_elim ? row : perm[row];
  const TI* pi = in + b * in_bs + static_cast<int64_t>(src) * k;
  TO* po = out + b * out_bs + static_cast<int64_t>(dst) * k;
  for (int c = threadIdx.x & 63; c < k; c += 64) po[c] = static_cast<TO>(pi[c]);
}

template <typename TI, typename TO>
__global__ void k_convert(const TI* __restrict__ in, TO* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[i] = static_cast<TO>(in[i]);
}

__global__ void k_axpy_f32(const float* __restrict__ c, double* __restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[i] += static_cast<double>(c[i]);
}

__global__ void k_zero_int(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// fv[b] = vals[b] * (b == batch ? scale : 1): the values the factorisation
// reads when a fault is injected into one batch member's factor
__global__ void k_scale_vals(const double* __restrict__ vals, double* __restrict__ fv,
                             int64_t nnz, int nb, int batch, double scale) {
  const int64_t total = nnz * nb;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    fv[i] = i / nnz == batch ? vals[i] * scale : vals[i];
}

__global__ void k_fill(double* __restrict__ x, int64_t n, double v) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[i] = v;
}

// r[b] = B[b or shared] - A_b x[b]  (FP64, elimination order)
__global__ void k_residual(const int* __restrict__ crow, const int* __restrict__ ccol,
                           const int* __restrict__ vidx, const double* __restrict__ vals,
                           int64_t nnz, const double* __restrict__ B, int64_t b_bs,
                           const double* __restrict__ X, double* __restrict__ R,
                           int n, int k) {
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int b = blockIdx.y;
  if (row >= n) return;
  const int64_t xb = static_cast<int64_t>(b) * n * k;
  const double* vb = vals + b * nnz;
  const int p0 = crow[row], p1 = crow[row + 1];
  for (int c = threadIdx.x & 63; c < k; c += 64) {
    double acc = B[b * b_bs + static_cast<int64_t>(row) * k + c];
    for (int p = p0; p < p1; ++p)
      acc -= vb[vidx ? vidx[p] : p] * X[xb + static_cast<int64_t>(ccol[p]) * k + c];
    R[xb + static_cast<int64_t>(row) * k + c] = acc;
  }
}

// Backward-error check of a solve on a sample of its columns: for batch
// member b and the first cs columns (cs <= LPR), sums of r^2 (r = B - A_b X
// in FP64), X^2 and B^2 over every row, and of A_b's squared entries.  LPR
// lanes per row (32 or 64), kChkRows rows per thread; the first ME entries of
// each row are loaded in three rounds for all of a thread's rows together
// (indices, then values and fields, then the arithmetic), so the loads of
// different rows overlap instead of chaining; entries past ME (longer rows
// than the stencil's) take a plain loop.  Each workgroup writes its four
// partial sums to part[b][blockIdx.x], summed in a fixed order by
// k_check_sum.
constexpr int kChkRows = 2;
template <typename TX, int LPR, int ME>
__global__ __launch_bounds__(256) void k_check(const int* __restrict__ crow,
                                               const int* __restrict__ ccol,
                                               const int* __restrict__ vidx,
                                               const double* __restrict__ vals, int64_t nnz,
                                               const double* __restrict__ B, int64_t b_bs,
                                               const TX* __restrict__ X, int n, int k, int cs,
                                               double* __restrict__ part) {
  constexpr int RPB = 256 / LPR;  // rows per pass of the workgroup
  constexpr int R = kChkRows;
  __shared__ double red[4][4];
  const int b = blockIdx.y, lane = threadIdx.x % LPR, sub = threadIdx.x / LPR;
  const int64_t xb = static_cast<int64_t>(b) * n * k;
  const double* vb = vals + b * nnz;
  const int c = lane < cs ? lane : 0;   // lanes past the sample read column 0, count zero
  int row[R], p0[R], p1[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    row[i] = (blockIdx.x * R + i) * RPB + sub;
    const bool ok = row[i] < n;
    p0[i] = ok ? crow[row[i]] : 0;
    p1[i] = ok ? crow[row[i] + 1] : 0;
  }
  int ci[R][ME], vi[R][ME];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int e = 0; e < ME; ++e) {
      const int p = p0[i] + e;
      const bool live = p < p1[i];
      ci[i][e] = live ? ccol[p] : -1;
      vi[i][e] = live ? (vidx ? vidx[p] : p) : -1;
    }
  double av[R][ME], bv[R];
  TX xv[R][ME], xd[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const bool ok = row[i] < n;
    bv[i] = ok ? B[b * b_bs + static_cast<int64_t>(row[i]) * k + c] : 0.0;
    xd[i] = ok ? X[xb + static_cast<int64_t>(row[i]) * k + c] : TX(0);
#pragma unroll
    for (int e = 0; e < ME; ++e) {
      av[i][e] = vi[i][e] >= 0 ? vb[vi[i][e]] : 0.0;
      xv[i][e] = ci[i][e] >= 0 ? X[xb + static_cast<int64_t>(ci[i][e]) * k + c] : TX(0);
    }
  }
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const bool in = lane < cs;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (row[i] >= n) continue;
    double acc = bv[i];
#pragma unroll
    for (int e = 0; e < ME; ++e) acc -= av[i][e] * static_cast<double>(xv[i][e]);
    double a2 = 0.0;
#pragma unroll
    for (int e = 0; e < ME; ++e) a2 += av[i][e] * av[i][e];
    for (int p = p0[i] + ME; p < p1[i]; ++p) {  // rows longer than ME
      const double a = vb[vidx ? vidx[p] : p];
      acc -= a * static_cast<double>(X[xb + static_cast<int64_t>(ccol[p]) * k + c]);
      a2 += a * a;
    }
    const double xr = static_cast<double>(xd[i]);
    if (in) {
      v[0] += acc * acc;
      v[1] += xr * xr;
      v[2] += bv[i] * bv[i];
    }
    if (lane == 0) v[3] += a2;
  }
  const int wl = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_down(v[i], o, 64);
  if (wl == 0)
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][i] = v[i];
  __syncthreads();
  if (threadIdx.x < 4)
    part[(static_cast<int64_t>(b) * gridDim.x + blockIdx.x) * 4 + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// out[b][i] = sum over the nblk partial sums part[b][.][i] (one workgroup per
// batch member; thread t sums blocks t, t + 256, ... then a fixed tree)
__global__ __launch_bounds__(256) void k_check_sum(const double* __restrict__ part, int nblk,
                                                   double* __restrict__ out) {
  __shared__ double red[256][4];
  const int b = blockIdx.x, t = threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int q = t; q < nblk; q += 256)
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] += part[(static_cast<int64_t>(b) * nblk + q) * 4 + i];
#pragma unroll
  for (int i = 0; i < 4; ++i) red[t][i] = v[i];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h)
#pragma unroll
      for (int i = 0; i < 4; ++i) red[t][i] += red[t + h][i];
    __syncthreads();
  }
  if (t < 4) out[4 * b + t] = red[0][t];
}

// Y[b] = alpha A_b X[b] (FP64; A_b from vals[b], CSR rows in the storage
// order of X and Y when vidx maps their entries to the caller's order)
__global__ void k_spmm(const int* __restrict__ crow, const int* __restrict__ ccol,
                       const int* __restrict__ vidx, const double* __restrict__ vals,
                       int64_t nnz, double alpha, const double* __restrict__ X,
                       double* __restrict__ Y, int n, int k) {
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int b = blockIdx.y;
  if (row >= n) return;
  const int64_t xb = static_cast<int64_t>(b) * n * k;
  const double* vb = vals + b * nnz;
  const int p0 = crow[row], p1 = crow[row + 1];
  for (int c = threadIdx.x & 63; c < k; c += 64) {
    double acc = 0.0;
    for (int p = p0; p < p1; ++p)
      acc += vb[vidx ? vidx[p] : p] * X[xb + static_cast<int64_t>(ccol[p]) * k + c];
    Y[xb + static_cast<int64_t>(row) * k + c] = alpha * acc;
  }
}

// Rows of the Jacobian, r data at once (datum j: adjoint field Y[.][.][j],
// forward field of its source U[.][.][src[j]]):
//   gv[b][e][j] = -wq[b] Y[b][row_e][j] U[b][col_e][src[j]]
// reduced over the wavenumbers straight away: ge = sum_b gv, gn = sum_b ky_b^2 gv,
// and for each boundary entry t of e, gb[t] = sum_b bcoef[b][t] gv.  One
// wavefront per row of the caller's pattern, lanes over the data.  iperm
// (may be null) maps rows to the storage order of Y and U.
constexpr int kMaxBcPerEntry = 4;
__global__ void k_jpair(int n, const int* __restrict__ crow, const int* __restrict__ ccol,
                        const int* __restrict__ iperm, const double* __restrict__ Y, int r,
                        const double* __restrict__ U, int nS, const int* __restrict__ src,
                        const double* __restrict__ wq, int nb, const double* __restrict__ ky2,
                        const int* __restrict__ bptr, const double* __restrict__ bcoef, int nbc,
                        double* __restrict__ ge, double* __restrict__ gn,
                        double* __restrict__ gb) {
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const int prow = iperm ? iperm[row] : row;
  for (int j = lane; j < r; j += 64) {
    const int sj = src[j];
    for (int e = crow[row]; e < crow[row + 1]; ++e) {
      const int pc = iperm ? iperm[ccol[e]] : ccol[e];
      const int t0 = bptr[e], nt = min(bptr[e + 1] - t0, kMaxBcPerEntry);
      double a = 0.0, c = 0.0, bq[kMaxBcPerEntry] = {0.0, 0.0, 0.0, 0.0};
      for (int b = 0; b < nb; ++b) {
        const double p = -wq[b] * Y[(static_cast<int64_t>(b) * n + prow) * r + j] *
                         U[(static_cast<int64_t>(b) * n + pc) * nS + sj];
        a += p;
        c += ky2[b] * p;
#pragma unroll
        for (int q = 0; q < kMaxBcPerEntry; ++q)
          if (q < nt) bq[q] += bcoef[static_cast<int64_t>(b) * nbc + t0 + q] * p;
      }
      ge[static_cast<int64_t>(e) * r + j] = a;
      gn[static_cast<int64_t>(e) * r + j] = c;
#pragma unroll
      for (int q = 0; q < kMaxBcPerEntry; ++q)
        if (q < nt) gb[static_cast<int64_t>(t0 + q) * r + j] = bq[q];
    }
  }
}

// J[c][j] = -sigma[c] (sum over the cell's map entries of le ge + ln gn, plus
// its boundary entries' gb); one wavefront per cell, lanes over the data
__global__ void k_jmap(int ncell, int r, const int* __restrict__ tptr,
                       const int* __restrict__ tent, const int* __restrict__ te,
                       const double* __restrict__ le, const double* __restrict__ ln,
                       const int* __restrict__ tbptr, const int* __restrict__ tbq,
                       const double* __restrict__ ge, const double* __restrict__ gn,
                       const double* __restrict__ gb, const double* __restrict__ sigma,
                       double* __restrict__ J) {
  const int c = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (c >= ncell) return;
  for (int j = lane; j < r; j += 64) {
    double acc = 0.0;
    for (int t = tptr[c]; t < tptr[c + 1]; ++t) {
      const int64_t e = te[t];
      acc += le[tent[t]] * ge[e * r + j] + ln[tent[t]] * gn[e * r + j];
    }
    for (int t = tbptr[c]; t < tbptr[c + 1]; ++t) acc += gb[static_cast<int64_t>(tbq[t]) * r + j];
    J[static_cast<int64_t>(c) * r + j] = -sigma[c] * acc;
  }
}

// gv[b][e] = -scale[b] sum_c Y[b][row][c] X[b][col_e][c]; LPR lanes per row
// (64: one wavefront per row; 32 or 16 for rows of at most 128 or 64
// columns, so that one wavefront serves 64 / LPR rows through the same chain
// of dependent loads: row pointers, columns, fields, output index) of the given
// pattern: the caller's, or, when the fields are stored in elimination order,
// the permuted one with oidx mapping its entries back to the caller's order
// (rows then run in elimination order: consecutive wavefronts read
// consecutive rows of Y, and a row's neighbours mostly lie in its own
// supernode).  The hardware deals workgroups round-robin to the 8 XCDs; the
// block index is remapped so that each XCD walks one contiguous range of rows
// (its own L2).  gridDim.x must be a multiple of 8.  Up to MAXE entries of a
// row at a time; each lane loads 4 columns of the Y row and of every X row
// before it multiplies, so that 4 (1 + MAXE) loads per lane are in flight,
// the last partial block of columns included; FP64 sums.  With pad_col /
// pad_out (rows of at most MAXE entries, padded with -1) a row's columns and
// output positions come from one load each, with no row-pointer load first.
constexpr size_t kGatherLds = 25 * 1024;  // occupancy cap of k_grad_values (see its launch)
// FP32 rows of several per wavefront (LPR < 64) and up to 5 entries: built
// for 6 wavefronts per SIMD (88 -> 76 VGPRs, no spills), the launch's LDS cap:
// 16-line batch 9.62 -> 9.51 ms, its electrode basis 8.66 -> 8.54 ms (8-entry
// rows would spill)
template <typename TI, int MAXE, int LPR, bool PAD>
__global__ __launch_bounds__(256, (LPR < 64 && sizeof(TI) == 4 && MAXE <= 5 ? 6 : 1))
void k_grad_values(const int* __restrict__ crow, const int* __restrict__ ccol,
                              const int* __restrict__ oidx, const int* __restrict__ pad_col,
                              const int* __restrict__ pad_out, const TI* __restrict__ Y,
                              const TI* __restrict__ X, int n, int k,
                              const double* __restrict__ scale, double* __restrict__ gv,
                              int64_t nnz) {
  constexpr int U = 4;
  constexpr int RPW = 64 / LPR;  // rows per wavefront
  const int per = gridDim.x >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int lane = threadIdx.x & 63;
  const int sl = lane % LPR;  // lane within the row's segment
  int row = (lb * (blockDim.x / 64) + threadIdx.x / 64) * RPW + lane / LPR;
  // one row per wavefront: the row and its neighbours' rows are the same in
  // every lane; said so, their bases live in scalar registers and each field
  // load takes one 32-bit lane offset (61 instead of 88 VGPRs; the launch
  // caps the wavefronts per SIMD at 6, see kGatherLds)
  if constexpr (LPR == 64) row = __builtin_amdgcn_readfirstlane(row);
  const int b = blockIdx.y;
  // a row's segment is wholly active or wholly gone, and the reductions
  // below stay inside one segment
  if (row >= n) return;
  const int64_t xb = static_cast<int64_t>(b) * n * k;
  const TI* y = Y + xb + static_cast<int64_t>(row) * k;
  const double sc = scale ? scale[b] : 1.0;
  const int pbeg = PAD ? 0 : crow[row], pend = PAD ? MAXE : crow[row + 1];
  for (int p0 = pbeg; p0 < pend; p0 += MAXE) {
    const int ne = min(MAXE, pend - p0);
    const TI* xr[MAXE];
#pragma unroll
    for (int q = 0; q < MAXE; ++q) {
      int cq;
      if constexpr (PAD) {
        cq = pad_col[static_cast<int64_t>(row) * MAXE + q];
        if (cq < 0) cq = row;  // empty slot: any valid row, its sum is dropped
      } else {
        cq = ccol[p0 + min(q, ne - 1)];  // surplus slots repeat the last entry
      }
      if constexpr (LPR == 64) cq = __builtin_amdgcn_readfirstlane(cq);
      xr[q] = X + xb + static_cast<int64_t>(cq) * k;
    }
    double acc[MAXE];
#pragma unroll
    for (int q = 0; q < MAXE; ++q) acc[q] = 0.0;
    int c = sl;
    for (; c + LPR * (U - 1) < k; c += LPR * U) {
      TI yv[U], xv[MAXE][U];
#pragma unroll
      for (int u = 0; u < U; ++u) yv[u] = y[static_cast<uint32_t>(c + LPR * u)];
#pragma unroll
      for (int q = 0; q < MAXE; ++q)
#pragma unroll
        for (int u = 0; u < U; ++u) xv[q][u] = xr[q][static_cast<uint32_t>(c + LPR * u)];
#pragma unroll
      for (int q = 0; q < MAXE; ++q)
#pragma unroll
        for (int u = 0; u < U; ++u)
          acc[q] += static_cast<double>(yv[u]) * static_cast<double>(xv[q][u]);
    }
    if constexpr (sizeof(TI) == 8 && LPR == 64) {
      // FP64 whole-wavefront rows: the remainder chunk by chunk (the block
      // form below holds 4 (1 + MAXE) doubles per lane and measured slower)
      for (; c < k; c += 64) {
        const double yv = static_cast<double>(y[c]);
#pragma unroll
        for (int q = 0; q < MAXE; ++q) acc[q] += yv * static_cast<double>(xr[q][c]);
      }
    } else if (c < k) {  // the remaining < LPR U columns: one block, every load
      // issued before the first use; columns past k read column k - 1 (in
      // bounds, same cache line) and count zero, so no load sits behind a branch
      TI yv[U], xv[MAXE][U];
      int cu[U];
#pragma unroll
      for (int u = 0; u < U; ++u) cu[u] = min(c + LPR * u, k - 1);
#pragma unroll
      for (int u = 0; u < U; ++u) yv[u] = y[static_cast<uint32_t>(cu[u])];
#pragma unroll
      for (int q = 0; q < MAXE; ++q)
#pragma unroll
        for (int u = 0; u < U; ++u) xv[q][u] = xr[q][static_cast<uint32_t>(cu[u])];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (c + LPR * u >= k) yv[u] = TI(0);
#pragma unroll
      for (int q = 0; q < MAXE; ++q)
#pragma unroll
        for (int u = 0; u < U; ++u)
          acc[q] += static_cast<double>(yv[u]) * static_cast<double>(xv[q][u]);
    }
#pragma unroll
    for (int q = 0; q < MAXE; ++q) {
      double a = acc[q];
#pragma unroll
      for (int off = LPR / 2; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
      if constexpr (PAD) {
        if (sl == 0) {
          const int o = pad_out[static_cast<int64_t>(row) * MAXE + q];
          if (o >= 0) gv[b * nnz + o] = -sc * a;
        }
      } else {
        if (sl == 0 && q < ne) gv[b * nnz + (oidx ? oidx[p0 + q] : p0 + q)] = -sc * a;
      }
    }
  }
}

// The gather for long rows (padded entries) with a whole pass of a row's loads
// in flight at once: one wavefront per row, each lane loading 16-byte vectors
// (VW = 4 floats or 2 doubles) at columns c0 + VW (64 j + lane), j < NV, of the
// Y row and of each of the row's MAXE X rows -- (1 + MAXE) NV loads per lane
// issued before the first product, so that a row of up to 64 VW NV columns
// costs one memory round trip (k_grad_values' lanes walk a row in four).
// Every load is unconditional: a vector that would cross the end of the row
// is read at the last full vector of the row instead and its columns already
// counted are masked, so no load sits behind a branch.  k >= VW.  FP64 sums.
template <typename TI, int MAXE, int NV>
__global__ __launch_bounds__(256) void k_grad_rows(const int* __restrict__ pad_col,
                                                   const int* __restrict__ pad_out,
                                                   const TI* __restrict__ Y,
                                                   const TI* __restrict__ X, int n, int k,
                                                   const double* __restrict__ scale,
                                                   double* __restrict__ gv, int64_t nnz) {
  constexpr int VW = 16 / sizeof(TI);
  constexpr int CP = 64 * VW * NV;  // columns per pass
  const int per = gridDim.x >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);  // XCD remap as k_grad_values
  const int lane = threadIdx.x & 63;
  const int row = lb * 4 + (threadIdx.x >> 6);
  const int b = blockIdx.y;
  if (row >= n) return;  // uniform over the wavefront
  const int64_t xb = static_cast<int64_t>(b) * n * k;
  const TI* y = Y + xb + static_cast<int64_t>(row) * k;
  const TI* xr[MAXE];
#pragma unroll
  for (int q = 0; q < MAXE; ++q) {
    int cq = pad_col[static_cast<int64_t>(row) * MAXE + q];
    if (cq < 0) cq = row;  // empty slot: any valid row, its sum is dropped
    xr[q] = X + xb + static_cast<int64_t>(cq) * k;
  }
  double acc[MAXE];
#pragma unroll
  for (int q = 0; q < MAXE; ++q) acc[q] = 0.0;
  for (int c0 = 0; c0 < k; c0 += CP) {
    int ce[NV];
    TI yv[NV][VW], xv[MAXE][NV][VW];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const int c = c0 + VW * (64 * j + lane);
      ce[j] = min(c, k - VW);
      typename Vec<TI, VW>::v t;
      __builtin_memcpy(&t, y + ce[j], sizeof(t));
#pragma unroll
      for (int v = 0; v < VW; ++v) {
        // column ce + v counts here if no earlier vector of this lane's
        // pass or of another lane covered it: ce + v >= c
        yv[j][v] = ce[j] + v >= c ? t[v] : TI(0);
      }
#pragma unroll
      for (int q = 0; q < MAXE; ++q) {
        typename Vec<TI, VW>::v u;
        __builtin_memcpy(&u, xr[q] + ce[j], sizeof(u));
#pragma unroll
        for (int v = 0; v < VW; ++v) xv[q][j][v] = u[v];
      }
    }
#pragma unroll
    for (int q = 0; q < MAXE; ++q)
#pragma unroll
      for (int j = 0; j < NV; ++j)
#pragma unroll
        for (int v = 0; v < VW; ++v)
          acc[q] += static_cast<double>(yv[j][v]) * static_cast<double>(xv[q][j][v]);
  }
  const double sc = scale ? scale[b] : 1.0;
#pragma unroll
  for (int q = 0; q < MAXE; ++q) {
    double a = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) {
      const int o = pad_out[static_cast<int64_t>(row) * MAXE + q];
      if (o >= 0) gv[b * nnz + o] = -sc * a;
    }
  }
}

// ---------------------------------------------------------------------------
// assembly map
// ---------------------------------------------------------------------------
// ky2 holds nky x ngroup squared wavenumbers, entry e taking column grp[e]
// (grp null: one group)
__global__ void k_assemble(int nnz, int nky, const double* __restrict__ ky2,
                           const int* __restrict__ grp, int ngroup,
                           const int* __restrict__ ptr, const int* __restrict__ cell,
                           const double* __restrict__ le, const double* __restrict__ ln,
                           const int* __restrict__ bptr, const int* __restrict__ bcell,
                           const double* __restrict__ bcoef, int nbc,
                           const double* __restrict__ sigma, double* __restrict__ vals) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  double ae = 0.0, an = 0.0;
  for (int p = ptr[e]; p < ptr[e + 1]; ++p) {
    const double s = sigma[cell[p]];
    ae += le[p] * s;
    an += ln[p] * s;
  }
  const int q0 = bptr[e], q1 = bptr[e + 1];
  const int g = grp ? grp[e] : 0;
  for (int k = 0; k < nky; ++k) {
    double v = ae + ky2[k * ngroup + g] * an;
    for (int q = q0; q < q1; ++q) v += bcoef[static_cast<int64_t>(k) * nbc + q] * sigma[bcell[q]];
    vals[static_cast<int64_t>(k) * nnz + e] = v;
  }
}

// per entry: ge = sum_k gv[k][e], gn = sum_k ky_k^2 gv[k][e]
__global__ void k_reduce_ky(int nnz, int nky, const double* __restrict__ ky2,
                            const int* __restrict__ grp, int ngroup,
                            const double* __restrict__ gv, double* __restrict__ ge,
                            double* __restrict__ gn) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const int gr = grp ? grp[e] : 0;
  double a = 0.0, c = 0.0;
  for (int k = 0; k < nky; ++k) {
    const double g = gv[static_cast<int64_t>(k) * nnz + e];
    a += g;
    c += ky2[k * ngroup + gr] * g;
  }
  ge[e] = a;
  gn[e] = c;
}

// grad_m[c] = -sigma[c] * sum over the cell's map entries (transposed CSR;
// tle / tln: the map coefficients in the same cell-major order, read
// contiguously)
// four lanes per cell: lane j takes the cell's entries j, j + 4, ... (and the
// boundary terms' wavenumbers j, j + 4, ...), so that a cell's loads are in
// flight together instead of a chain of dependent loads; the lanes' sums are
// added in a fixed order
constexpr int kLmapLanes = 4;
__global__ void k_lmap_t(int ncell, int nnz, int nky, const int* __restrict__ tptr,
                         const int* __restrict__ te,
                         const double* __restrict__ tle, const double* __restrict__ tln,
                         const int* __restrict__ tbptr, const int* __restrict__ tbq,
                         const int* __restrict__ bc_e, const double* __restrict__ bcoef,
                         int nbc, const double* __restrict__ ge, const double* __restrict__ gn,
                         const double* __restrict__ gv, const double* __restrict__ sigma,
                         double* __restrict__ grad_m) {
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) / kLmapLanes;
  const int j = threadIdx.x % kLmapLanes;
  if (c >= ncell) return;  // uniform over the cell's lanes
  double acc = 0.0;
  for (int t = tptr[c] + j; t < tptr[c + 1]; t += kLmapLanes) {
    const int e = te[t];
    acc += tle[t] * ge[e] + tln[t] * gn[e];
  }
  for (int t = tbptr[c]; t < tbptr[c + 1]; ++t) {
    const int q = tbq[t];
    const int e = bc_e[q];
    for (int k = j; k < nky; k += kLmapLanes)
      acc += bcoef[static_cast<int64_t>(k) * nbc + q] * gv[static_cast<int64_t>(k) * nnz + e];
  }
#pragma unroll
  for (int off = 1; off < kLmapLanes; off <<= 1) acc += __shfl_xor(acc, off, kLmapLanes);
  if (j == 0) grad_m[c] = -sigma[c] * acc;
}

// ---------------------------------------------------------------------------
// survey: projection and adjoint sources
// ---------------------------------------------------------------------------
// 16 lanes per datum: lane j takes the wavenumbers j, j + 16, ..., so that the
// field loads of every wavenumber are in flight together; the lanes' sums are
// added by a fixed butterfly (deterministic).
constexpr int kDpredLanes = 16;
template <typename TU>
__global__ void k_dpred(int nD, const int* __restrict__ ptr, const int* __restrict__ rows,
                        const double* __restrict__ val, const int* __restrict__ sidx,
                        const TU* __restrict__ U, int64_t bs, int k, int nky,
                        const double* __restrict__ w, int s0, int s1, double* __restrict__ d) {
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kDpredLanes;
  const int j = threadIdx.x % kDpredLanes;
  if (i >= nD) return;  // uniform over the datum's lanes
  const int s = sidx[i] - s0;
  if (sidx[i] < s0 || sidx[i] >= s1) return;  // uniform over the datum's lanes
  double acc = 0.0;
  for (int kk = j; kk < nky; kk += kDpredLanes) {
    double a = 0.0;
    for (int p = ptr[i]; p < ptr[i + 1]; ++p)
      a += val[p] * static_cast<double>(U[kk * bs + static_cast<int64_t>(rows[p]) * k + s]);
    acc += w[kk] * a;
  }
#pragma unroll
  for (int off = kDpredLanes / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kDpredLanes);
  if (j == 0) d[i] = acc;
}

// G[row][src - s0] = sum coef * v[datum] for each (row, source) target
__global__ void k_adjoint_rhs(int ntgt, const int* __restrict__ tptr,
                              const int* __restrict__ trow, const int* __restrict__ tsrc,
                              const int* __restrict__ tdat, const double* __restrict__ tcoef,
                              const double* __restrict__ v, int k, int s0, int s1,
                              double* __restrict__ G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntgt) return;
  const int src = tsrc[t];
  if (src < s0 || src >= s1) return;
  double acc = 0.0;
  for (int q = tptr[t]; q < tptr[t + 1]; ++q) acc += tcoef[q] * v[tdat[q]];
  G[static_cast<int64_t>(trow[t]) * k + (src - s0)] = acc;
}

inline int grid1(int64_t n, int bs) { return static_cast<int>((n + bs - 1) / bs); }
inline int grid_stride(int64_t n) {
  return static_cast<int>(std::min<int64_t>((n + 255) / 256, 4096));
}

}  // namespace

// ===========================================================================
// plan
// ===========================================================================
// buffers and context of one solve check: the device partial sums and totals,
// their pinned host copy (nbatch x 4 sums, then nbatch pivot flags), the
// events of a deferred check, and the solve it belongs to
struct CheckSlot {
  double* d_chk = nullptr;
  void* h_chk = nullptr;
  hipEvent_t ev = nullptr;   // the results are on the host
  hipEvent_t src = nullptr;  // the solve is done (the check stream waits on it)
  const double* B = nullptr;
  int64_t bbs = 0;
  void* xv = nullptr;
  int k = 0, flags = 0;
  hipStream_t st = nullptr;
};

struct dcx_plan {
  dcx::Symbolic S;
  int device = 0;
  int nbatch = 1;
  int precision = DCX_FP64;
  int refine = 1;
  size_t tsize = 8;
  // device maps
  SupDev* d_sup = nullptr;
  int *d_lvl = nullptr, *d_relpos = nullptr, *d_cidx = nullptr, *d_ridx = nullptr;
  // kernel-variant thresholds (dcx_plan_set_tuning; defaults from the
  // environment at plan creation)
  int tc_cap = 32;           // widest column tile of the in-LDS sweeps
  int splitk_below = 4096;   // backward levels with fewer wavefronts take split-K
  int rowsplit_below = 0;    // ... and those with ns <= 64 the row-split variant
  int wide_below = 321;      // factor levels with fewer workgroups take 1024 threads
  int panel_width = 32;      // columns per panel of the in-LDS factor (16 or 32)
  // factor launch order: each level's supernodes by front height (ld)
  // descending, cut into buckets launched with their own LDS size
  int* d_flvl = nullptr;
  std::vector<int> flvl_ld;                  // ld of each entry of that order
  std::vector<int> fbkt_ptr, fbkt;           // level d: bucket starts fbkt[fbkt_ptr[d]..]
  int *d_a_e = nullptr, *d_a_row = nullptr, *d_a_col = nullptr;
  int *d_perm = nullptr, *d_iperm = nullptr, *d_pcrow = nullptr, *d_pcol = nullptr,
      *d_pvidx = nullptr;
  int *d_crow = nullptr, *d_col = nullptr;
  int* d_info = nullptr;  // per batch member: a pivot was not positive
  // backward-error check of every solve (dcx_plan_set_check)
  double chk_tol = 0.0;     // <= 0: off
  int chk_passes = 2;       // refinement passes a failing check may add (FP32 plans)
  int chk_cols = 8;         // sample columns (<= 64)
  // check buffers: slots 0 and 1 hold up to two deferred checks (a FIFO),
  // slot 2 the immediate one
  CheckSlot slot[3];
  int slot_head = 0, slot_count = 0;
  std::vector<double> chk_eta;   // normwise backward error |R| / (|A| |X| + |B|)
  std::vector<double> chk_rho;   // relative residual |R| / |B| (the test)
  std::vector<int> chk_pivot;
  int chk_passes_done = 0;
  int chk_status = 0;       // 0 not run, 1 passed, 2 failed, 3 skipped (stream capture)
  hipStream_t chk_st = nullptr;    // stream of deferred checks (default priority)
  size_t chk_tot_off = 0;          // offset of the totals in a slot's d_chk
  // split backward sweep (DCX_SPLIT_BWD): the supernodes outside the support
  // run on the side stream; rest_ev marks their end
  int* d_inact_lvl = nullptr;
  std::vector<int> inact_lvl_ptr;
  hipEvent_t rest_ev = nullptr, split_ev = nullptr;
  bool rest_pending = false;
  // fault injection (tests): the factorisation reads d_fvals, in which batch
  // member inj_batch is scaled by 1 + inj_rel
  int inj_batch = -1;
  double inj_rel = 0.0;
  double* d_fvals = nullptr;
  // incremented whenever a workspace buffer is reallocated (kernels recorded
  // in a graph before that point hold stale addresses)
  int64_t buf_gen = 0;
  // factor
  void* d_panels = nullptr;
  void* d_panelsT = nullptr;
  void* d_front[2] = {nullptr, nullptr};
  double* d_vals = nullptr;
  bool factored = false;
  // fused bottom subtrees
  int fuse_rows = 0;
  FusedDev* d_fz = nullptr;
  int *d_flv = nullptr, *d_fmem = nullptr, *d_fgptr = nullptr, *d_fgb = nullptr,
      *d_toff = nullptr, *d_lridx = nullptr, *d_ulvl = nullptr;
  void* d_ufz = nullptr;
  size_t ufz_bytes = 0;
  // sparse right-hand sides
  int* d_active = nullptr;
  std::vector<int> act_lvl_ptr;
  int* d_act_lvl = nullptr;
  // workspace (grown on demand)
  void* d_upd[2] = {nullptr, nullptr};
  int upd_k = 0;
  void* d_w = nullptr;       // working block in factor precision
  size_t w_bytes = 0;
  double* d_r = nullptr;     // FP64 residual (mixed)
  double* d_b = nullptr;     // permuted FP64 rhs
  double* d_x = nullptr;     // permuted FP64 iterate (mixed)
  size_t r_bytes = 0, b_bytes = 0, x_bytes = 0;
  int64_t device_bytes = 0;
  int max_row_nnz = 0;  // longest row of the caller's pattern
  // the gather's padded rows (gpad_me = 5 or 8 entries per row, 0: rows too
  // long): columns and output positions, in elimination order (p) and in the
  // caller's order (u)
  int gpad_me = 0;
  int *d_gpad_col_p = nullptr, *d_gpad_out_p = nullptr, *d_gpad_col_u = nullptr,
      *d_gpad_out_u = nullptr;
  // forward-sweep scratch of the levels whose fronts exceed the in-LDS kernel
  int64_t* d_fs_off = nullptr;  // row offset of each supernode within its level
  void* d_fs = nullptr;
  size_t fs_bytes = 0;
  void* d_dg = nullptr;  // factored diagonal blocks of the big fronts' panels, kDg per front
  // second stream and level events of dcx_factor_solve
  hipStream_t side = nullptr;
  std::vector<hipEvent_t> lev_ev;
  // optional event timing of the sweeps (forward, backward) of every solve
  bool timing = false;
  // the plan's own streams beside the caller's (dcx_plan_set_overlap): the
  // factorisation overlapped with the first forward sweep, the split backward
  // sweep and the deferred checks; off, a plan's work stays on the caller's
  // stream in order
  bool overlap = true;
  std::vector<hipEvent_t> ev;  // triples per sweep pair: start, mid, end

  ~dcx_plan() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : lev_ev) (void)hipEventDestroy(e);
    if (side) (void)hipStreamDestroy(side);
    void* ptrs[] = {d_sup, d_lvl, d_flvl, d_relpos, d_cidx, d_ridx, d_a_e, d_a_row, d_a_col,
                    d_perm, d_iperm, d_pcrow, d_pcol, d_pvidx, d_crow, d_col, d_info, d_panels,
                    d_front[0], d_front[1], d_vals, d_active, d_act_lvl, d_upd[0],
                    d_upd[1], d_w, d_r, d_b, d_x, d_fz, d_flv, d_fmem, d_fgptr, d_fgb,
                    d_toff, d_lridx, d_ulvl, d_ufz, d_panelsT, d_fs_off, d_fs, d_dg,
                    d_fvals, d_inact_lvl, d_gpad_col_p, d_gpad_out_p, d_gpad_col_u,
                    d_gpad_out_u};
    for (void* p : ptrs)
      if (p) (void)dev_free(p);
    for (auto& c : slot) {
      if (c.d_chk) (void)dev_free(c.d_chk);
      if (c.h_chk) (void)hipHostFree(c.h_chk);
      if (c.ev) (void)hipEventDestroy(c.ev);
      if (c.src) (void)hipEventDestroy(c.src);
    }
    if (rest_ev) (void)hipEventDestroy(rest_ev);
    if (split_ev) (void)hipEventDestroy(split_ev);
    if (chk_st) (void)hipStreamDestroy(chk_st);
  }
};

static int complete_pending(dcx_plan* P);
static int complete_all(dcx_plan* P);

namespace {

int env_int(const char* name, int dflt) {
  const char* e = std::getenv(name);
  return e ? std::atoi(e) : dflt;
}

int default_fuse_rows(int precision) {
  const char* e = std::getenv("DCX_FUSE_ROWS");
  if (e) return std::atoi(e);
  (void)precision;
  return 0;  // fused subtrees off by default (slower than the level kernels at this occupancy)
}

// a workspace buffer may not grow while the stream is being captured (the
// synchronisation and the free are not capturable, and a graph recorded
// before would keep the freed address)
int refuse_in_capture(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone)
    return fail(DCX_ERR_STATE,
                "a workspace buffer would grow during stream capture: run the same call "
                "once before capturing it");
  return DCX_OK;
}

bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

int ensure(dcx_plan* P, void** p, size_t* cap, size_t bytes, hipStream_t st) {
  if (*cap >= bytes && *p) return DCX_OK;
  if (int rc = refuse_in_capture(st)) return rc;
  ++P->buf_gen;
  HIP_TRY(hipStreamSynchronize(st));
  if (*p) HIP_TRY(dev_free(*p));
  *p = nullptr;
  *cap = 0;
  HIP_TRY(dev_malloc(p, std::max<size_t>(bytes, 256)));
  *cap = bytes;
  return DCX_OK;
}

// (re)upload the fused-subtree structure and the supernode table
int upload_fused(dcx_plan* P) {
  auto& S = P->S;
  void* olds[] = {P->d_fz, P->d_flv, P->d_fmem, P->d_fgptr, P->d_fgb, P->d_toff,
                  P->d_lridx, P->d_ulvl, P->d_sup};
  for (void* q : olds)
    if (q) HIP_TRY(dev_free(q));
  P->d_fz = nullptr;
  P->d_flv = P->d_fmem = P->d_fgptr = P->d_fgb = P->d_toff = P->d_lridx = P->d_ulvl = nullptr;
  P->d_sup = nullptr;
  std::vector<SupDev> sd(S.nsup);
  for (int s = 0; s < S.nsup; ++s) {
    const int t = S.fused_of_root[s];
    sd[s] = SupDev{S.first[s], S.ns[s], S.rs[s], S.rptr[s], S.cptr[s], S.cptr[s + 1],
                   S.aptr[s], S.aptr[s + 1], S.panel_off[s], S.front_off[s], S.upd_off[s],
                   t >= 0 ? S.f_uoff[t] : int64_t(-1)};
  }
  HIP_TRY(upload(&P->d_sup, sd));
  std::vector<FusedDev> fd(S.nfused);
  for (int t = 0; t < S.nfused; ++t) {
    const int r = S.f_root[t];
    fd[t] = FusedDev{r, S.f_lo[t], S.f_hi[t], S.rs[r], S.rptr[r], S.f_lv_begin[t],
                     S.f_lv_end[t], S.f_tb_rows[t], S.f_uoff[t]};
  }
  HIP_TRY(upload(&P->d_fz, fd));
  HIP_TRY(upload(&P->d_flv, S.f_lv));
  HIP_TRY(upload(&P->d_fmem, S.f_mem));
  HIP_TRY(upload(&P->d_fgptr, S.f_gptr));
  HIP_TRY(upload(&P->d_fgb, S.f_gb));
  HIP_TRY(upload(&P->d_toff, S.toff));
  HIP_TRY(upload(&P->d_lridx, S.lridx));
  HIP_TRY(upload(&P->d_ulvl, S.ulvl_sup));
  return DCX_OK;
}

// default tolerances of the solve check on the relative residual |R| / |B|
// (DCX_CHECK_TOL overrides): FP32 factors stay below about 1e-5 on the
// benchmark problems, FP64 ones below 1e-12; a failing check means a factor
// that broke down or lost accuracy far beyond rounding
constexpr double kCheckTol32 = 1e-3;
constexpr double kCheckTol64 = 1e-8;

constexpr size_t kLdsBudget = 64 * 1024;
constexpr size_t kFusedLdsBudget = 80 * 1024;  // two workgroups per CU

// levels whose tallest panel does not fit k_factor_level's LDS or threads
// take the grid-wide large-front kernels
inline int level_max_ld(const dcx::Symbolic& S, int d) {
  int maxld = 0;
  for (int q = S.lvl_ptr[d]; q < S.lvl_ptr[d + 1]; ++q) {
    const int s = S.lvl_sup[q];
    maxld = std::max(maxld, 2 * S.ns[s] + S.rs[s]);
  }
  return maxld;
}
inline size_t factor_lds(int maxld, size_t tsize, int nb = kNB) {
  return (static_cast<size_t>(nb) * (nb + 4) + nb + static_cast<size_t>(maxld + 16) * (nb + 4)) *
         tsize;
}
// also levels with fewer than DCX_FACTOR_GRID_BELOW (front x batch)
// workgroups, where one workgroup per front leaves the chip idle
inline bool level_is_big(const dcx::Symbolic& S, int d, size_t tsize, int nbatch) {
  static const int grid_below = env_int("DCX_FACTOR_GRID_BELOW", 0);
  const int maxld = level_max_ld(S, d);
  const int64_t wgs = static_cast<int64_t>(S.lvl_ptr[d + 1] - S.lvl_ptr[d]) * nbatch;
  return factor_lds(maxld, tsize) > 160 * 1024 || maxld > 1024 || wgs < grid_below;
}

// fronts of one LDS size a CU holds at once (8: the wavefront limit of
// 256-thread workgroups)
inline int factor_occupancy(int ld, size_t tsize) {
  return static_cast<int>(std::min<size_t>(8, (160 * 1024) / factor_lds(ld, tsize)));
}

// the factor launch order (returned, for upload) and its buckets: within a
// level, supernodes by ld descending; a new bucket starts where the
// fronts-per-CU class rises by 2 or more and the rest of the level holds at
// least 2048 workgroups (a smaller gain does not pay for the extra launch on
// the critical path)
std::vector<int> build_factor_order(dcx_plan* P) {
  const auto& S = P->S;
  std::vector<int> order(S.lvl_sup);
  P->flvl_ld.assign(S.nsup, 0);
  P->fbkt_ptr.assign(S.ndepth + 1, 0);
  P->fbkt.clear();
  auto ld_of = [&](int s) { return 2 * S.ns[s] + S.rs[s]; };
  for (int d = 0; d < S.ndepth; ++d) {
    const int q0 = S.lvl_ptr[d], q1 = S.lvl_ptr[d + 1];
    std::stable_sort(order.begin() + q0, order.begin() + q1,
                     [&](int a, int b) { return ld_of(a) > ld_of(b); });
    for (int q = q0; q < q1; ++q) P->flvl_ld[q] = ld_of(order[q]);
    if (q1 > q0) P->fbkt.push_back(q0);
    int occ = q1 > q0 ? factor_occupancy(P->flvl_ld[q0], P->tsize) : 0;
    for (int q = q0 + 1; q < q1; ++q) {
      const int o = factor_occupancy(P->flvl_ld[q], P->tsize);
      if (o >= occ + 2 && static_cast<int64_t>(q1 - q) * P->nbatch >= 2048) {
        P->fbkt.push_back(q);
        occ = o;
      }
    }
    P->fbkt_ptr[d + 1] = static_cast<int>(P->fbkt.size());
  }
  P->fbkt.push_back(S.nsup);  // end sentinel
  return order;
}

// the values the factorisation reads: the caller's, or the fault-injected copy
inline const double* factor_vals(const dcx_plan* P) {
  return P->d_fvals && P->inj_batch >= 0 ? P->d_fvals : P->d_vals;
}

template <typename T, int NT, int RPT, int NB = kNB>
void factor_launch(dcx_plan* P, int d, const int* lvl, int cnt, size_t lds, hipStream_t st,
                   hipEvent_t done = nullptr) {
  const auto& S = P->S;
  T* fr[2] = {static_cast<T*>(P->d_front[0]), static_cast<T*>(P->d_front[1])};
  const int64_t fbs[2] = {S.front_size[0], S.front_size[1]};
  dim3 grid(cnt, P->nbatch);
  // `done`: an event that the launch itself signals when the kernel ends (no
  // separate marker between this level and the next)
  hipExtLaunchKernelGGL((k_factor_level<T, NT, RPT, NB>), grid, dim3(NT), lds, st, nullptr, done,
                        0, P->d_sup, lvl, P->d_relpos, P->d_cidx, P->d_a_e, P->d_a_row,
                        P->d_a_col, factor_vals(P), static_cast<int64_t>(S.nnz), fr[d & 1],
                        fbs[d & 1], fr[(d + 1) & 1], fbs[(d + 1) & 1],
                        static_cast<T*>(P->d_panels), static_cast<T*>(P->d_panelsT),
                        S.panel_size, P->d_info);
}

// a level with fronts too tall for k_factor_level: grid-wide launches
template <typename T>
int factor_level_big(dcx_plan* P, int d, int cnt, hipStream_t st) {
  const auto& S = P->S;
  T* fr[2] = {static_cast<T*>(P->d_front[0]), static_cast<T*>(P->d_front[1])};
  const int64_t fbs[2] = {S.front_size[0], S.front_size[1]};
  const int* lvl = P->d_lvl + S.lvl_ptr[d];
  const int nb = P->nbatch;
  int maxns = 0, maxnf = 0, maxkids = 0;
  std::vector<int> maxrc;
  for (int q = S.lvl_ptr[d]; q < S.lvl_ptr[d + 1]; ++q) {
    const int s = S.lvl_sup[q];
    maxns = std::max(maxns, S.ns[s]);
    maxnf = std::max(maxnf, S.ns[s] + S.rs[s]);
    const int nk = S.cptr[s + 1] - S.cptr[s];
    if (nk > maxkids) maxrc.resize(nk, 0);
    maxkids = std::max(maxkids, nk);
    for (int c = 0; c < nk; ++c) maxrc[c] = std::max(maxrc[c], S.rs[S.cidx[S.cptr[s] + c]]);
  }
  if (static_cast<int64_t>(cnt) * nb > 65535)
    return fail(DCX_ERR_ARG, "too many large fronts in one level");
  hipLaunchKernelGGL(k_front_zero_big<T>, dim3((maxnf + kZeroCols - 1) / kZeroCols, cnt * nb),
                     dim3(256), 0, st, P->d_sup, lvl, P->d_a_e, P->d_a_row, P->d_a_col, factor_vals(P),
                     static_cast<int64_t>(S.nnz), fr[d & 1], fbs[d & 1], nb);
  for (int c = 0; c < maxkids; ++c) {
    if (!maxrc[c]) continue;
    const int nt = (maxrc[c] + 63) / 64;
    hipLaunchKernelGGL(k_extend_big<T>, dim3(nt, nt, cnt * nb), dim3(256), 0, st, P->d_sup, lvl,
                       P->d_relpos, P->d_cidx, fr[d & 1], fbs[d & 1], fr[(d + 1) & 1],
                       fbs[(d + 1) & 1], c, nb);
  }
  T* dg = static_cast<T*>(P->d_dg);
  hipLaunchKernelGGL(k_diag_big<T>, dim3(cnt * nb), dim3(64), 0, st, P->d_sup, lvl, fr[d & 1],
                     fbs[d & 1], dg, nb, P->d_info);
  // outer panels of kOuter columns: inside one, 64-column panels whose
  // trailing updates stay within its columns; then one update of width kOuter
  // for the rest of the front
  for (int J0 = 0; J0 < maxns; J0 += kOuter) {
    const int Jend = J0 + kOuter;
    for (int j0 = J0; j0 < std::min(Jend, maxns); j0 += kNB2) {
      // tallest live row range (J, R and the E rows e < j0 + kNB2)
      const int rows = std::min(maxnf + maxns - j0, maxnf + kNB2);
      hipLaunchKernelGGL(k_panel_rows_big<T>, dim3((rows + kRowsBig - 1) / kRowsBig, cnt * nb),
                         dim3(kRowsBig), 0, st, P->d_sup, lvl, fr[d & 1], fbs[d & 1], dg, j0, nb);
      if (j0 + kNB2 < std::min(Jend, maxns)) {
        dim3 tg((rows + 63) / 64, (Jend - j0 - kNB2 + 63) / 64, cnt * nb);
        hipLaunchKernelGGL((k_trail_big<T, kNB2>), tg, dim3(256), 0, st, P->d_sup, lvl,
                           fr[d & 1], fbs[d & 1], j0, Jend, nb, dg, P->d_info);
      }
    }
    const int rows = std::min(maxnf + maxns - J0, maxnf + kOuter), cols = maxnf - J0;
    dim3 tg((rows + kTR - 1) / kTR, (cols + kTC - 1) / kTC, cnt * nb);
    hipLaunchKernelGGL(k_trail_outer<T>, tg, dim3(256), 0, st, P->d_sup, lvl, fr[d & 1],
                       fbs[d & 1], J0, nb, dg, P->d_info);
  }
  hipLaunchKernelGGL(k_front_out_big<T>, dim3((maxnf + 63) / 64, (maxns + 63) / 64, cnt * nb),
                     dim3(256), 0, st, P->d_sup, lvl, fr[d & 1], fbs[d & 1],
                     static_cast<T*>(P->d_panels), static_cast<T*>(P->d_panelsT), S.panel_size,
                     nb);
  return DCX_OK;
}

// one launch per level; the threads (x rows per thread) cover the tallest
// panel of the level
// done (optional): an event the level's last launch signals at its end;
// *signalled tells whether it did (levels through the grid-wide kernels and
// empty levels do not: the caller records the event itself)
template <typename T>
int factor_level(dcx_plan* P, int d, hipStream_t st, hipEvent_t done = nullptr,
                 bool* signalled = nullptr) {
  const auto& S = P->S;
  if (signalled) *signalled = false;
  const int cnt = S.lvl_ptr[d + 1] - S.lvl_ptr[d];
  if (!cnt) return DCX_OK;
  const int maxld = level_max_ld(S, d);
  const size_t lds = factor_lds(maxld, sizeof(T));
  if (level_is_big(S, d, sizeof(T), P->nbatch)) return factor_level_big<T>(P, d, cnt, st);
  // levels with few workgroups (the top of the tree) take more wavefronts per
  // front for the trailing update
  static const int buckets = env_int("DCX_FACTOR_BUCKETS", 1);
  const bool wide = static_cast<int64_t>(cnt) * P->nbatch < P->wide_below;
  if (signalled) *signalled = done != nullptr;
  if (!buckets) {
    const int* lvl = P->d_lvl + S.lvl_ptr[d];
    if (maxld <= 256 && !wide)
      factor_launch<T, 256, 1>(P, d, lvl, cnt, lds, st, done);
    else
      factor_launch<T, 1024, 1>(P, d, lvl, cnt, lds, st, done);
    return DCX_OK;
  }
  // one launch per bucket of the level (fronts of one occupancy class), each
  // with the LDS of its tallest front (the first: descending order)
  for (int t = P->fbkt_ptr[d]; t < P->fbkt_ptr[d + 1]; ++t) {
    const int b0 = P->fbkt[t];
    const int b1 = t + 1 < P->fbkt_ptr[d + 1] ? P->fbkt[t + 1] : S.lvl_ptr[d + 1];
    const int bld = P->flvl_ld[b0];
    const int* lvl = P->d_flvl + b0;
    hipEvent_t ev = t + 1 == P->fbkt_ptr[d + 1] ? done : nullptr;  // the level's last launch
    if (P->panel_width == 16) {
      const size_t l16 = factor_lds(bld, sizeof(T), 16);
      if (bld <= 256 && !wide) factor_launch<T, 256, 1, 16>(P, d, lvl, b1 - b0, l16, st, ev);
      else factor_launch<T, 1024, 1, 16>(P, d, lvl, b1 - b0, l16, st, ev);
      continue;
    }
    const size_t blds = factor_lds(bld, sizeof(T));
    if (bld <= 256 && !wide)
      factor_launch<T, 256, 1>(P, d, lvl, b1 - b0, blds, st, ev);
    else
      factor_launch<T, 1024, 1>(P, d, lvl, b1 - b0, blds, st, ev);
  }
  return DCX_OK;
}

template <typename T>
int launch_factor(dcx_plan* P, hipStream_t st) {
  for (int d = P->S.ndepth - 1; d >= 0; --d) {
    const int rc = factor_level<T>(P, d, st);
    if (rc) return rc;
  }
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

// where the forward sweep reads its right-hand side and where the backward
// sweep leaves the FP64 solution
struct SweepIO {
  const double* B = nullptr;  // FP64 right-hand side (batch stride b_bs)
  int64_t b_bs = 0;
  const int* brow = nullptr;  // B row of each elimination row (null: same)
  double* X64 = nullptr;      // FP64 solution (null: only the working block)
  const int* xrow = nullptr;  // X64 row of each elimination row (null: same)
  int xmode = 0;              // 0 none, 1 store, 2 accumulate
};

template <typename T, int TC, bool PL>
void sweep_launch(dcx_plan* P, T* W, int k, int maxnf, int maxns, int pmax, int ct, bool fwd, int d,
                  const int* lvl, int cnt, const int* active, const SweepIO& io,
                  hipStream_t st) {
  const auto& S = P->S;
  const int nz = (k + TC - 1) / TC;
  const int nzb = (nz + ct - 1) / ct;
  const int total = cnt * P->nbatch * nzb;
  dim3 grid(((total + 7) / 8) * 8);
  const size_t lds = (static_cast<size_t>(PL ? pmax : 0) + static_cast<size_t>(maxnf) * TC) * sizeof(T);
  T* panels = static_cast<T*>(P->d_panels);
  const int64_t w_bs = static_cast<int64_t>(S.n) * k;
  if (fwd) {
    T* u[2] = {static_cast<T*>(P->d_upd[0]), static_cast<T*>(P->d_upd[1])};
    const int64_t ubs[2] = {S.upd_rows[0] * k, S.upd_rows[1] * k};
    hipLaunchKernelGGL((k_fwd_level<T, TC, PL>), grid, dim3(kThreads), lds, st, P->d_sup, lvl,
                       P->d_relpos, P->d_cidx, active, panels, S.panel_size, io.B, io.b_bs,
                       io.brow, W, w_bs, k, u[d & 1], ubs[d & 1], u[(d + 1) & 1],
                       ubs[(d + 1) & 1], static_cast<const T*>(P->d_ufz), P->S.f_upd_rows * k,
                       ct, pmax, cnt, nzb, total);
  } else {
    hipLaunchKernelGGL((k_bwd_level<T, TC, PL>), grid, dim3(kThreads), lds, st, P->d_sup, lvl,
                       P->d_ridx, active, static_cast<const T*>(P->d_panelsT), S.panel_size, W,
                       w_bs, k, io.X64, io.xrow, io.xmode, ct, pmax, cnt, nzb, total);
  }
}

// one level: panel staged in LDS when it fits; column tiles per workgroup so
// that the level still has about kTargetBlocks workgroups
constexpr size_t kSweepLds = 160 * 1024;


int pick_tc(int maxnf, int maxns, size_t tsize, int cap);

// rows of the widest level whose forward sweep takes the large-front path or
// whose backward sweep is row-grouped (both use the scratch block d_fs)
inline int64_t big_fwd_rows(const dcx_plan* P, size_t tsize) {
  const auto& S = P->S;
  int64_t best = 0;
  for (int d = 0; d < S.ndepth; ++d) {
    int maxnf = 0, maxns = 0;
    int64_t rows = 0;
    for (int q = S.lvl_ptr[d]; q < S.lvl_ptr[d + 1]; ++q) {
      const int s = S.lvl_sup[q];
      maxnf = std::max(maxnf, S.ns[s] + S.rs[s]);
      maxns = std::max(maxns, S.ns[s]);
      rows += S.ns[s] + S.rs[s];
    }
    if (pick_tc(maxnf, maxns, tsize, P->tc_cap) == 0 || maxns > 128) best = std::max(best, rows);
  }
  return best;
}

template <typename T, int MT, int V, int KS, int PF>
void bwd_wave_launch(dcx_plan* P, T* W, int k, const int* lvl, int cnt, const int* active,
                     const SweepIO& io, hipStream_t st, int nrg, int maxrs) {
  constexpr int CW = 16 * V;
  const int ncb = (k + CW - 1) / CW;              // column blocks
  const int nzc = KS == 1 ? (ncb + 3) / 4 : ncb;  // column groups per supernode and batch member
  static_assert(KS != 0 || MT == 1, "row-split tiles are 16 rows per wavefront");
  const int nzb = nzc * nrg;                      // x row groups of 16 MT rows
  const int total = cnt * P->nbatch * nzb;
  dim3 grid(((total + 7) / 8) * 8);
  // split-K partial tiles, then the R rows of the level's widest front
  const size_t lds = (KS > 1 ? static_cast<size_t>(KS) * MT * V * 4 * 64 * sizeof(T) : 0) +
                     static_cast<size_t>(std::max(maxrs, 1)) * sizeof(int);
  const int64_t w_bs = static_cast<int64_t>(P->S.n) * k;
  T* xs = nrg > 1 ? static_cast<T*>(P->d_fs) : nullptr;
  const int64_t xs_bs = big_fwd_rows(P, sizeof(T)) * k;
  hipLaunchKernelGGL((k_bwd_wave<T, MT, V, KS, PF>), grid, dim3(kThreads), lds, st, P->d_sup,
                     lvl, P->d_ridx, active, static_cast<const T*>(P->d_panelsT),
                     P->S.panel_size, W, w_bs, k, io.X64, io.xrow, io.xmode, cnt, nzb, total,
                     nzc, xs, P->d_fs_off, xs_bs);
  if (xs) {
    const int maxns = nrg * 16 * MT;  // bounds ns over the level
    hipLaunchKernelGGL(k_bwd_copy_big<T>,
                       dim3((maxns + kCopyRows - 1) / kCopyRows, cnt * P->nbatch, (k + 63) / 64),
                       dim3(256), 0, st, P->d_sup, lvl, xs, P->d_fs_off, xs_bs, W, w_bs,
                       k, io.X64, io.xrow, io.xmode, P->nbatch);
  }
}

template <typename T, int V, int KS, int PF>
bool bwd_wave_mt(dcx_plan* P, T* W, int k, int maxns, int maxrs, const int* lvl, int cnt,
                 const int* active, const SweepIO& io, hipStream_t st) {
  const int mt = (maxns + 15) / 16;
  // FP32 wide tiles of up to 4 row tiles: 3 k-steps ahead, built for 5 (4 for
  // MT 4) wavefronts per SIMD (kBwdWaves): mixed eval+grad 3.337 -> 3.318 ms
  // with MT <= 3, 3.33 -> 3.28 ms with MT 4 added; the 16-line batch 9.88 ->
  // 9.60 ms (2 k-steps ahead: within 0.01 ms of 3)
  constexpr int PFN = (KS == 1 && sizeof(T) == 4) ? 3 : PF;
  if (mt <= 1) bwd_wave_launch<T, 1, V, KS, PFN>(P, W, k, lvl, cnt, active, io, st, 1, maxrs);
  else if (mt <= 2) bwd_wave_launch<T, 2, V, KS, PFN>(P, W, k, lvl, cnt, active, io, st, 1, maxrs);
  else if (mt <= 3) bwd_wave_launch<T, 3, V, KS, PFN>(P, W, k, lvl, cnt, active, io, st, 1, maxrs);
  else if (mt <= 4) bwd_wave_launch<T, 4, V, KS, PFN>(P, W, k, lvl, cnt, active, io, st, 1, maxrs);
  else if (mt <= 6) bwd_wave_launch<T, 6, V, KS, PF>(P, W, k, lvl, cnt, active, io, st, 1, maxrs);
  else  // taller fronts: row groups of 128 rows
    bwd_wave_launch<T, 8, V, KS, PF>(P, W, k, lvl, cnt, active, io, st, (mt + 7) / 8, maxrs);
  return true;
}

// backward level through the wavefront kernel when its output tiles fit:
// wide vector columns where the level has enough wavefronts to fill the
// chip, split-K otherwise
template <typename T>
bool bwd_wave(dcx_plan* P, T* W, int k, int maxns, int maxrs, const int* lvl, int cnt,
              const int* active, const SweepIO& io, hipStream_t st) {
  static const int use = env_int("DCX_BWD_WAVE", 1);
  const int min_waves = P->splitk_below;
  if (!use) return false;
  constexpr int VW = 16 / sizeof(T);  // 16-byte vectors
  const int64_t waves = static_cast<int64_t>(cnt) * P->nbatch * ((k + 16 * VW - 1) / (16 * VW));
  if (maxns <= 64 && waves < P->rowsplit_below) {  // four wavefronts per column block
    bwd_wave_launch<T, 1, VW, 0, 4>(P, W, k, lvl, cnt, active, io, st, 1, maxrs);
    return true;
  }
  if (waves < min_waves)
    return bwd_wave_mt<T, 1, 4, 4>(P, W, k, maxns, maxrs, lvl, cnt, active, io, st);
  return bwd_wave_mt<T, VW, 1, 4>(P, W, k, maxns, maxrs, lvl, cnt, active, io, st);
}

template <typename T, int TC>
void sweep_tc(dcx_plan* P, T* W, int k, int maxnf, int maxns, int maxpanel, bool fwd, int d,
              const int* lvl, int cnt, const int* active, const SweepIO& io, hipStream_t st) {
  static const int use_pl = env_int("DCX_PANEL_LDS", 0);
  static const int ct_max = env_int("DCX_CT_MAX", 1);
  static const int target = env_int("DCX_TARGET_BLOCKS", 8192);
  const int pmax = (maxpanel + 3) & ~3;
  const bool pl = use_pl &&
      (static_cast<size_t>(pmax) + static_cast<size_t>(maxnf) * TC) * sizeof(T) <= kSweepLds;
  const int nz = (k + TC - 1) / TC;
  int ct = 1;
  if (pl) {
    const int64_t blocks = static_cast<int64_t>(cnt) * P->nbatch * nz;
    ct = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>({int64_t(nz), int64_t(ct_max),
                                                                    blocks / target})));
    sweep_launch<T, TC, true>(P, W, k, maxnf, maxns, pmax, ct, fwd, d, lvl, cnt, active, io, st);
  } else {
    sweep_launch<T, TC, false>(P, W, k, maxnf, maxns, 0, 1, fwd, d, lvl, cnt, active, io, st);
  }
}

// widest column tile whose LDS block fits and whose register tiles fit
int pick_tc(int maxnf, int maxns, size_t tsize, int cap) {
  for (int tc : {64, 32, 16}) {
    if (tc > cap) continue;
    if (static_cast<size_t>(maxnf) * tc * tsize <= kLdsBudget) return tc;
  }
  return 0;
}

// full solve on the working block W (elimination order, factor precision)
template <typename T, int TC>
void fused_tc(dcx_plan* P, T* W, int k, bool fwd, const int* active, const SweepIO& io,
              hipStream_t st) {
  const auto& S = P->S;
  dim3 grid(S.nfused, P->nbatch, (k + TC - 1) / TC);
  const size_t lds = static_cast<size_t>(S.f_max_lds_rows) * TC * sizeof(T);
  T* panels = static_cast<T*>(P->d_panels);
  const int64_t w_bs = static_cast<int64_t>(S.n) * k;
  T* ufz = static_cast<T*>(P->d_ufz);
  const int64_t ufz_bs = S.f_upd_rows * k;
  if (fwd)
    hipLaunchKernelGGL((k_fwd_fused<T, TC>), grid, dim3(kThreads), lds, st, P->d_sup, P->d_fz,
                       P->d_flv, P->d_fmem, P->d_fgptr, P->d_fgb, P->d_toff, P->d_lridx, active,
                       panels, S.panel_size, io.B, io.b_bs, io.brow, W, w_bs, k, ufz, ufz_bs);
  else
    hipLaunchKernelGGL((k_bwd_fused<T, TC>), grid, dim3(kThreads), lds, st, P->d_sup, P->d_fz,
                       P->d_flv, P->d_fmem, P->d_toff, P->d_lridx, P->d_ridx, active,
                       static_cast<const T*>(P->d_panelsT), S.panel_size, W, w_bs, k, io.X64,
                       io.xrow, io.xmode);
}

template <typename T>
int launch_fused(dcx_plan* P, T* W, int k, bool fwd, const int* active, const SweepIO& io,
                 hipStream_t st) {
  const size_t rows = static_cast<size_t>(P->S.f_max_lds_rows);
  if (rows * 64 * sizeof(T) <= kFusedLdsBudget)
    fused_tc<T, 64>(P, W, k, fwd, active, io, st);
  else if (rows * 32 * sizeof(T) <= kFusedLdsBudget)
    fused_tc<T, 32>(P, W, k, fwd, active, io, st);
  else if (rows * 16 * sizeof(T) <= 160 * 1024)
    fused_tc<T, 16>(P, W, k, fwd, active, io, st);
  else
    return fail(DCX_ERR_ARG, "fused subtree too large for LDS");
  return DCX_OK;
}

// update buffers of the forward sweep for k right-hand sides (grown on
// demand; growing synchronises the stream once)
template <typename T>
int ensure_upd(dcx_plan* P, int k, hipStream_t st) {
  const auto& S = P->S;
  const size_t fsb = static_cast<size_t>(P->nbatch) * big_fwd_rows(P, sizeof(T)) * k * sizeof(T);
  if (fsb > P->fs_bytes) {
    if (int rc = refuse_in_capture(st)) return rc;
    ++P->buf_gen;
    HIP_TRY(hipStreamSynchronize(st));
    if (P->d_fs) HIP_TRY(dev_free(P->d_fs));
    P->d_fs = nullptr;
    P->fs_bytes = 0;
    HIP_TRY(dev_malloc(&P->d_fs, fsb));
    P->fs_bytes = fsb;
  }
  if (P->upd_k >= k) return DCX_OK;
  const size_t ub0 = static_cast<size_t>(P->nbatch) * S.upd_rows[0] * k * sizeof(T);
  const size_t ub1 = static_cast<size_t>(P->nbatch) * S.upd_rows[1] * k * sizeof(T);
  const size_t ubf = static_cast<size_t>(P->nbatch) * S.f_upd_rows * k * sizeof(T);
  if (int rc = refuse_in_capture(st)) return rc;
  ++P->buf_gen;
  HIP_TRY(hipStreamSynchronize(st));
  void* olds[] = {P->d_upd[0], P->d_upd[1], P->d_ufz};
  for (void* q : olds)
    if (q) HIP_TRY(dev_free(q));
  P->d_upd[0] = P->d_upd[1] = P->d_ufz = nullptr;
  HIP_TRY(dev_malloc(&P->d_upd[0], std::max<size_t>(ub0, 256)));
  HIP_TRY(dev_malloc(&P->d_upd[1], std::max<size_t>(ub1, 256)));
  HIP_TRY(dev_malloc(&P->d_ufz, std::max<size_t>(ubf, 256)));
  P->upd_k = k;
  return DCX_OK;
}

// forward sweep of a level with fronts too tall for the in-LDS kernel
template <typename T>
int fwd_big(dcx_plan* P, T* W, int k, int d, const int* lvl, int cnt, int maxnf,
            const int* act, const SweepIO& io, hipStream_t st) {
  const auto& S = P->S;
  const int64_t rows = big_fwd_rows(P, sizeof(T));
  const int64_t fs_bs = rows * k;
  T* fs = static_cast<T*>(P->d_fs);
  T* u[2] = {static_cast<T*>(P->d_upd[0]), static_cast<T*>(P->d_upd[1])};
  const int64_t ubs[2] = {S.upd_rows[0] * k, S.upd_rows[1] * k};
  const int nzg = (k + 63) / 64, nb = P->nbatch;
  if (static_cast<int64_t>(cnt) * nb > 65535 || nzg > 65535)
    return fail(DCX_ERR_ARG, "too many large fronts or right-hand sides in one level");
  int maxkids = 0;
  std::vector<int> maxrc;
  for (int q = S.ulvl_ptr[d]; q < S.ulvl_ptr[d + 1]; ++q) {  // bounds the active subset too
    const int s = S.ulvl_sup[q];
    const int nk = S.cptr[s + 1] - S.cptr[s];
    if (nk > maxkids) maxrc.resize(nk, 0);
    maxkids = std::max(maxkids, nk);
    for (int c = 0; c < nk; ++c) maxrc[c] = std::max(maxrc[c], S.rs[S.cidx[S.cptr[s] + c]]);
  }
  hipLaunchKernelGGL(k_fwd_init_big<T>, dim3((maxnf + kFwdRows - 1) / kFwdRows, cnt * nb, nzg),
                     dim3(256), 0, st, P->d_sup, lvl, io.B, io.b_bs, io.brow, k, fs, P->d_fs_off,
                     fs_bs, nb);
  for (int c = 0; c < maxkids; ++c) {
    if (!maxrc[c]) continue;
    hipLaunchKernelGGL(k_fwd_extend_big<T>,
                       dim3((maxrc[c] + kFwdRows - 1) / kFwdRows, cnt * nb, nzg), dim3(256), 0,
                       st, P->d_sup, lvl, P->d_relpos, P->d_cidx, act, k, u[(d + 1) & 1],
                       ubs[(d + 1) & 1], static_cast<const T*>(P->d_ufz), P->S.f_upd_rows * k,
                       fs, P->d_fs_off, fs_bs, c, nb);
  }
  constexpr int MT = 8;
  const int nzc = ((k + 15) / 16 + 3) / 4;
  const int nrg = (maxnf + 16 * MT - 1) / (16 * MT);
  const int tw = cnt * P->nbatch * nzc * nrg;
  hipLaunchKernelGGL((k_fwd_wave_big<T, MT>), dim3(((tw + 7) / 8) * 8), dim3(256), 0, st,
                     P->d_sup, lvl, cnt, static_cast<const T*>(P->d_panels), S.panel_size, fs,
                     P->d_fs_off, fs_bs, W, static_cast<int64_t>(S.n) * k, k, u[d & 1],
                     ubs[d & 1], nzc, nzc * nrg, tw);
  return DCX_OK;
}

// one level of one sweep direction (sparse forward: active supernodes only)
// subset (backward sweep only): 0 every supernode of the level, 1 those of
// the right-hand side's support, 2 the others
template <typename T>
int sweep_level(dcx_plan* P, T* W, int k, bool fwd, int d, const int* act, const SweepIO& io,
                hipStream_t st, int subset = 0) {
  const auto& S = P->S;
  const int* lvl;
  int cnt;
  if ((fwd && act) || subset == 1) {
    lvl = P->d_act_lvl + P->act_lvl_ptr[d];
    cnt = P->act_lvl_ptr[d + 1] - P->act_lvl_ptr[d];
  } else if (subset == 2) {
    lvl = P->d_inact_lvl + P->inact_lvl_ptr[d];
    cnt = P->inact_lvl_ptr[d + 1] - P->inact_lvl_ptr[d];
  } else {
    lvl = P->d_ulvl + S.ulvl_ptr[d];
    cnt = S.ulvl_ptr[d + 1] - S.ulvl_ptr[d];
  }
  if (!cnt) return DCX_OK;
  int maxnf = 0, maxns = 0, maxrs = 0, maxpanel = 0;
  for (int q = S.ulvl_ptr[d]; q < S.ulvl_ptr[d + 1]; ++q) {
    int s = S.ulvl_sup[q];
    maxnf = std::max(maxnf, S.ns[s] + S.rs[s]);
    maxns = std::max(maxns, S.ns[s]);
    maxrs = std::max(maxrs, S.rs[s]);
    maxpanel = std::max(maxpanel, (S.ns[s] + S.rs[s]) * S.ns[s]);
  }
  if (!fwd && bwd_wave<T>(P, W, k, maxns, maxrs, lvl, cnt, act, io, st)) return DCX_OK;
  const int tc = pick_tc(maxnf, maxns, sizeof(T), P->tc_cap);
  if (!tc) {
    if (fwd) return fwd_big<T>(P, W, k, d, lvl, cnt, maxnf, act, io, st);
    return fail(DCX_ERR_ARG, "front too large for the in-LDS sweep");
  }
  switch (tc) {
    case 64: sweep_tc<T, 64>(P, W, k, maxnf, maxns, maxpanel, fwd, d, lvl, cnt, act, io, st); break;
    case 32: sweep_tc<T, 32>(P, W, k, maxnf, maxns, maxpanel, fwd, d, lvl, cnt, act, io, st); break;
    default: sweep_tc<T, 16>(P, W, k, maxnf, maxns, maxpanel, fwd, d, lvl, cnt, act, io, st); break;
  }
  return DCX_OK;
}

// forward then backward sweep on W; skip_fwd: the forward sweep was already
// run (overlapped with the factorisation), only the backward one remains
template <typename T>
int sweeps(dcx_plan* P, T* W, int k, bool sparse_rhs, const SweepIO& io, hipStream_t st,
           bool skip_fwd = false, bool split = false) {
  const auto& S = P->S;
  int rc = ensure_upd<T>(P, k, st);
  if (rc) return rc;
  const bool use_act = sparse_rhs && P->d_active;
  const int* act = use_act ? P->d_active : nullptr;
  const bool timed = P->timing && !skip_fwd;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  if (timed) {
    for (auto& e : ev) {
      HIP_TRY(hipEventCreate(&e));
      P->ev.push_back(e);
    }
    HIP_TRY(hipEventRecord(ev[0], st));
  }
  for (int pass = skip_fwd ? 1 : 0; pass < 2; ++pass) {
    const bool fwd = pass == 0;
    if (!fwd && timed) HIP_TRY(hipEventRecord(ev[1], st));
    if (fwd && S.nfused) {
      rc = launch_fused<T>(P, W, k, true, act, io, st);
      if (rc) return rc;
    }
    const bool sp = split && !fwd;
    for (int dd = 0; dd < S.ndepth; ++dd) {
      const int d = fwd ? S.ndepth - 1 - dd : dd;
      rc = sweep_level<T>(P, W, k, fwd, d, act, io, st, sp ? 1 : 0);
      if (rc) return rc;
    }
    if (sp) {
      // the supernodes outside the support, root first, on the side stream
      // after the support's ones; rest_ev marks their end (dcx_plan_join)
      HIP_TRY(hipEventRecord(P->split_ev, st));
      HIP_TRY(hipStreamWaitEvent(P->side, P->split_ev, 0));
      for (int d = 0; d < S.ndepth; ++d) {
        rc = sweep_level<T>(P, W, k, false, d, act, io, P->side, 2);
        if (rc) return rc;
      }
      HIP_TRY(hipEventRecord(P->rest_ev, P->side));
      P->rest_pending = true;
    }
    if (!fwd && S.nfused) {
      rc = launch_fused<T>(P, W, k, false, act, io, st);
      if (rc) return rc;
    }
  }
  if (timed) HIP_TRY(hipEventRecord(ev[2], st));
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

// the plan's second stream (default priority) and its level events, made once
int ensure_side(dcx_plan* P) {
  if (P->side) return DCX_OK;
  HIP_TRY(hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking));
  P->lev_ev.assign(P->S.ndepth + 2, nullptr);
  for (auto& e : P->lev_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&P->rest_ev, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&P->split_ev, hipEventDisableTiming));
  return DCX_OK;
}

// Factorisation and the first solve's forward sweep, overlapped: the factor
// levels run on `st`; the forward sweep of level d (deepest first) runs on the
// plan's second stream as soon as level d is factored, using the compute
// units the narrow upper factor levels leave idle; the backward sweep then
// runs on `st` after both.  Without the overlap (fused subtrees) it is the
// plain sequence.
template <typename T>
int factor_and_sweeps(dcx_plan* P, T* W, int k, bool sparse_rhs, const SweepIO& io,
                      hipStream_t st, bool split = false) {
  const auto& S = P->S;
  if (S.nfused || !P->overlap) {
    int rc = launch_factor<T>(P, st);
    if (rc) return rc;
    return sweeps<T>(P, W, k, sparse_rhs, io, st);
  }
  int rc = ensure_upd<T>(P, k, st);
  if (rc) return rc;
  rc = ensure_side(P);
  if (rc) return rc;
  const int* act = sparse_rhs && P->d_active ? P->d_active : nullptr;
  // the side stream's first wait (the deepest level's event) orders it after
  // everything queued on `st` before this call.  Outside stream capture the
  // level events are signalled by the factor launches themselves; under
  // capture they are recorded (a launch-signalled event is no capture
  // dependency: the side stream's sweep would be left out of the graph)
  const bool cap = capturing(st);
  hipEvent_t join = P->lev_ev[S.ndepth + 1];
  for (int d = S.ndepth - 1; d >= 1; --d) {
    bool signalled = false;
    rc = factor_level<T>(P, d, st, cap ? nullptr : P->lev_ev[d], &signalled);
    if (rc) return rc;
    if (!signalled) HIP_TRY(hipEventRecord(P->lev_ev[d], st));
    HIP_TRY(hipStreamWaitEvent(P->side, P->lev_ev[d], 0));
    rc = sweep_level<T>(P, W, k, true, d, act, io, P->side);
    if (rc) return rc;
  }
  // the root level: factored, then its forward sweep on `st` after the side
  // stream's levels (finished by then, while the root was factored) -- no
  // event between the root's factorisation and its sweep
  HIP_TRY(hipEventRecord(join, P->side));
  rc = factor_level<T>(P, 0, st);
  if (rc) return rc;
  HIP_TRY(hipStreamWaitEvent(st, join, 0));
  rc = sweep_level<T>(P, W, k, true, 0, act, io, st);
  if (rc) return rc;
  return sweeps<T>(P, W, k, sparse_rhs, io, st, true, split);
}

void launch_permute_d2T(dcx_plan* P, const double* in, void* out, int k, int nblocks,
                        int64_t in_bs, int64_t out_bs, bool f32, hipStream_t st) {
  const int n = P->S.n;
  dim3 grid(grid1(n, 4), nblocks);
  if (f32)
    hipLaunchKernelGGL((k_permute<double, float>), grid, dim3(256), 0, st, in,
                       static_cast<float*>(out), P->d_perm, n, k, in_bs, out_bs, 1);
  else
    hipLaunchKernelGGL((k_permute<double, double>), grid, dim3(256), 0, st, in,
                       static_cast<double*>(out), P->d_perm, n, k, in_bs, out_bs, 1);
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
namespace {

constexpr int kMixMaxSeg = 256;  // segments (lines) per call
constexpr int kMixRows = 64;     // rows per workgroup, 16 per wave
constexpr int kMixMaxK = 128;    // columns: one or two per lane

struct MixSegs {
  int nseg;
  int row[kMixMaxSeg + 1];   // segment l = rows [row[l], row[l+1])
  int tile[kMixMaxSeg + 1];  // first workgroup of segment l
};

template <typename T>
__device__ inline void load4(const T* p, T (&v)[4]) {
  v[0] = p[0]; v[1] = p[1]; v[2] = p[2]; v[3] = p[3];
}
template <>
__device__ inline void load4<float>(const float* p, float (&v)[4]) {
  const float4 q = *reinterpret_cast<const float4*>(p);
  v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
template <>
__device__ inline void load4<double>(const double* p, double (&v)[4]) {
  const double2 a = *reinterpret_cast<const double2*>(p);
  const double2 b = *reinterpret_cast<const double2*>(p + 2);
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}

// H[b][r][:] = G[b][r][:] W[b][l] for the rows r of segment l (W[b][l]: k x k,
// row-major).  A workgroup takes 64 rows of one segment of one system: the
// rows are staged in LDS (k padded to kp, a multiple of 4, with zeros), lane
// j owns columns j and j + 64 (NG = 2) for 16 rows of its wave, and W's row a
// is read once per wave from L2 and used for all 16 rows.  A wavefront stages
// its own 16 rows only (no workgroup barrier).
template <typename T, int NG>
__global__ __launch_bounds__(256) void k_rows_mix(const T* __restrict__ G,
                                                  const T* __restrict__ W, T* __restrict__ H,
                                                  int n, int k, int kp, MixSegs segs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* g = reinterpret_cast<T*>(smem_raw);
  const int t = blockIdx.x, b = blockIdx.y;
  int l = 0;
  while (l + 1 < segs.nseg && segs.tile[l + 1] <= t) ++l;
  const int r0 = segs.row[l] + (t - segs.tile[l]) * kMixRows;
  const int nr = min(kMixRows, segs.row[l + 1] - r0);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // each wavefront stages and then reads only its own 16 rows: a wavefront
  // fence orders its LDS writes before its reads, and no workgroup barrier
  // couples the wavefronts, so one's loads overlap another's FMAs
  const T* Gb = G + (static_cast<size_t>(b) * n + r0 + wv * 16) * k;
  T* gw_st = g + wv * 16 * kp;
  const int nrw = nr - wv * 16;  // rows of this wavefront inside the tile
  if (nrw <= 0) return;          // past a short tile's end: no barrier waits for it
  {
    // all 32 loads issued before any LDS write (a runtime loop waited on each
    // load in turn); columns lane and lane + 64 cover kp <= 128
    T v[16][2];
#pragma unroll
    for (int rr = 0; rr < 16; ++rr)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int a = lane + 64 * c;
        v[rr][c] = (rr < nrw && a < k) ? Gb[static_cast<size_t>(rr) * k + a] : T(0);
      }
#pragma unroll
    for (int rr = 0; rr < 16; ++rr)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int a = lane + 64 * c;
        if (a < kp) gw_st[rr * kp + a] = v[rr][c];
      }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const T* Wl = W + (static_cast<size_t>(b) * segs.nseg + l) * k * k;
  const T* gw = gw_st;
  T acc[NG][16];
#pragma unroll
  for (int c = 0; c < NG; ++c)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[c][i] = T(0);
  // W's rows a..a+3 for the next step are loaded one step ahead, so the FMAs
  // of a step do not wait on that step's L2 reads
  T wn[NG][4];
  auto load_w = [&](int a0) {
#pragma unroll
    for (int c = 0; c < NG; ++c) {
      const int j = lane + 64 * c;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        wn[c][q] = (j < k && a0 + q < k) ? Wl[static_cast<size_t>(a0 + q) * k + j] : T(0);
    }
  };
  load_w(0);
  for (int a = 0; a < k; a += 4) {
    T w[NG][4];
#pragma unroll
    for (int c = 0; c < NG; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q) w[c][q] = wn[c][q];
    load_w(a + 4);  // rows past k load nothing
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      T gv[4];
      load4(gw + i * kp + a, gv);
#pragma unroll
      for (int c = 0; c < NG; ++c)
        acc[c][i] += gv[0] * w[c][0] + gv[1] * w[c][1] + gv[2] * w[c][2] + gv[3] * w[c][3];
    }
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int rr = wv * 16 + i;
    if (rr >= nr) break;
    T* Hr = H + (static_cast<size_t>(b) * n + r0 + rr) * k;
#pragma unroll
    for (int c = 0; c < NG; ++c)
      if (lane + 64 * c < k) Hr[lane + 64 * c] = acc[c][i];
  }
}

template <typename T>
void launch_rows_mix(const void* g, const void* w, void* h, int nbatch, int n, int k,
                     const MixSegs& segs, hipStream_t st) {
  const int kp = (k + 3) & ~3;
  const dim3 grid(segs.tile[segs.nseg], nbatch);
  const size_t lds = static_cast<size_t>(kMixRows) * kp * sizeof(T);
  const T* G = static_cast<const T*>(g);
  const T* W = static_cast<const T*>(w);
  T* H = static_cast<T*>(h);
  if (k <= 64)
    hipLaunchKernelGGL((k_rows_mix<T, 1>), grid, dim3(256), lds, st, G, W, H, n, k, kp, segs);
  else
    hipLaunchKernelGGL((k_rows_mix<T, 2>), grid, dim3(256), lds, st, G, W, H, n, k, kp, segs);
}

}  // namespace

extern "C" {

const char* dcx_last_error(void) { return g_err.c_str(); }

int dcx_device_count(int* out) {
  if (!out) return fail(DCX_ERR_ARG, "null output");
  HIP_TRY(hipGetDeviceCount(out));
  return DCX_OK;
}

int dcx_set_allocator(dcx_alloc_fn alloc, dcx_free_fn free_fn, void* ctx) {
  if ((alloc == nullptr) != (free_fn == nullptr))
    return fail(DCX_ERR_ARG, "give both an allocation and a free function, or neither");
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_alloc_fn = alloc;
  g_free_fn = free_fn;
  g_alloc_ctx = ctx;
  return DCX_OK;
}

int dcx_plan_create(int n, const int32_t* crow, const int32_t* col, int nbatch,
                    const double* coords, int dim, int leaf_max, int precision,
                    dcx_plan** out) {
  try {
    if (!out || !crow || !col) return fail(DCX_ERR_ARG, "null argument");
    *out = nullptr;
    if (nbatch < 1) return fail(DCX_ERR_ARG, "nbatch must be >= 1");
    if (precision != DCX_FP64 && precision != DCX_FP32)
      return fail(DCX_ERR_ARG, "precision must be DCX_FP64 or DCX_FP32");
    if (coords && (dim < 1 || dim > 3)) return fail(DCX_ERR_ARG, "dim must be 1..3");
    std::unique_ptr<dcx_plan> P(new dcx_plan());
    P->fuse_rows = default_fuse_rows(precision);
    std::string err = dcx::analyse(n, crow, col, coords, dim, leaf_max, P->fuse_rows, P->S);
    if (!err.empty()) return fail(DCX_ERR_ARG, err);
    auto& S = P->S;
    HIP_TRY(hipGetDevice(&P->device));
    P->nbatch = nbatch;
    P->precision = precision;
    P->tsize = precision == DCX_FP32 ? 4 : 8;
    P->tc_cap = env_int("DCX_SWEEP_TC", 32);
    // split-K backward levels below 1024 wavefronts: with tile tuning off the
    // 25k-cell line's eval+grad takes 3.29 ms (FP64 5.51) against 3.37 (5.57)
    // at 4096; the sweep column tile (16 or 32) measured equal there
    P->splitk_below = env_int("DCX_BWD_SPLITK_BELOW", 1024);
    P->rowsplit_below = env_int("DCX_BWD_ROWSPLIT_BELOW", 0);
    P->wide_below = env_int("DCX_FACTOR_WIDE_BELOW", 321);
    P->panel_width = env_int("DCX_FACTOR_PANEL_WIDTH", 32) == 16 ? 16 : 32;
    {
      int rc = upload_fused(P.get());
      if (rc) return rc;
    }
    HIP_TRY(upload(&P->d_lvl, S.lvl_sup));
    HIP_TRY(upload(&P->d_flvl, build_factor_order(P.get())));
    HIP_TRY(upload(&P->d_relpos, S.relpos));
    HIP_TRY(upload(&P->d_cidx, S.cidx));
    HIP_TRY(upload(&P->d_ridx, S.ridx));
    HIP_TRY(upload(&P->d_a_e, S.a_e));
    HIP_TRY(upload(&P->d_a_row, S.a_row));
    HIP_TRY(upload(&P->d_a_col, S.a_col));
    HIP_TRY(upload(&P->d_perm, S.perm));
    HIP_TRY(upload(&P->d_iperm, S.iperm));
    HIP_TRY(upload(&P->d_pcrow, S.pcrow));
    HIP_TRY(upload(&P->d_pcol, S.pcol));
    HIP_TRY(upload(&P->d_pvidx, S.pval_idx));
    {
      std::vector<int64_t> off(S.nsup, 0);
      for (int d = 0; d < S.ndepth; ++d) {
        int64_t rows = 0;
        for (int q = S.lvl_ptr[d]; q < S.lvl_ptr[d + 1]; ++q) {
          const int sp = S.lvl_sup[q];
          off[sp] = rows;
          rows += S.ns[sp] + S.rs[sp];
        }
      }
      HIP_TRY(upload(&P->d_fs_off, off));
    }
    std::vector<int> vcrow(crow, crow + n + 1), vcol(col, col + crow[n]);
    for (int i = 0; i < n; ++i) P->max_row_nnz = std::max(P->max_row_nnz, crow[i + 1] - crow[i]);
    HIP_TRY(upload(&P->d_crow, vcrow));
    HIP_TRY(upload(&P->d_col, vcol));
    P->gpad_me = P->max_row_nnz <= 5 ? 5 : P->max_row_nnz <= 8 ? 8 : 0;
    if (P->gpad_me) {
      const int me = P->gpad_me;
      std::vector<int> pc(static_cast<size_t>(n) * me, -1), po(pc.size(), -1);
      std::vector<int> uc(pc.size(), -1), uo(pc.size(), -1);
      for (int i = 0; i < n; ++i) {
        for (int p = S.pcrow[i]; p < S.pcrow[i + 1]; ++p) {
          pc[static_cast<size_t>(i) * me + (p - S.pcrow[i])] = S.pcol[p];
          po[static_cast<size_t>(i) * me + (p - S.pcrow[i])] = S.pval_idx[p];
        }
        for (int p = crow[i]; p < crow[i + 1]; ++p) {
          uc[static_cast<size_t>(i) * me + (p - crow[i])] = col[p];
          uo[static_cast<size_t>(i) * me + (p - crow[i])] = p;
        }
      }
      HIP_TRY(upload(&P->d_gpad_col_p, pc));
      HIP_TRY(upload(&P->d_gpad_out_p, po));
      HIP_TRY(upload(&P->d_gpad_col_u, uc));
      HIP_TRY(upload(&P->d_gpad_out_u, uo));
    }
    HIP_TRY(dev_malloc(&P->d_info, sizeof(int) * nbatch));
    {
      const char* e = std::getenv("DCX_CHECK_TOL");
      P->chk_tol = e ? std::atof(e) : (precision == DCX_FP32 ? kCheckTol32 : kCheckTol64);
    }
    const size_t pb = static_cast<size_t>(nbatch) * S.panel_size * P->tsize;
    const size_t f0 = static_cast<size_t>(nbatch) * S.front_size[0] * P->tsize;
    const size_t f1 = static_cast<size_t>(nbatch) * S.front_size[1] * P->tsize;
    const size_t vb = static_cast<size_t>(nbatch) * S.nnz * sizeof(double);
    HIP_TRY(dev_malloc(&P->d_panels, std::max<size_t>(pb, 256)));
    HIP_TRY(dev_malloc(&P->d_panelsT, std::max<size_t>(pb, 256)));
    HIP_TRY(dev_malloc(&P->d_front[0], std::max<size_t>(f0, 256)));
    HIP_TRY(dev_malloc(&P->d_front[1], std::max<size_t>(f1, 256)));
    HIP_TRY(dev_malloc(&P->d_vals, std::max<size_t>(vb, 256)));
    int maxbig = 0;
    for (int d = 0; d < S.ndepth; ++d)
      if (level_is_big(S, d, P->tsize, nbatch)) maxbig = std::max(maxbig, S.lvl_ptr[d + 1] - S.lvl_ptr[d]);
    const size_t dgb = static_cast<size_t>(nbatch) * maxbig * kDg * P->tsize;
    HIP_TRY(dev_malloc(&P->d_dg, std::max<size_t>(dgb, 256)));
    P->device_bytes = static_cast<int64_t>(2 * pb + f0 + f1 + vb + dgb);
    *out = P.release();
    return DCX_OK;
  } catch (const std::exception& e) {
    return fail(DCX_ERR_INTERNAL, e.what());
  }
}

int dcx_plan_set_check(dcx_plan* P, double tol, int max_passes, int sample_cols) {
  if (!P) return fail(DCX_ERR_ARG, "null plan");
  if (max_passes < 0 || max_passes > 16) return fail(DCX_ERR_ARG, "max_passes must be 0..16");
  if (sample_cols < 1 || sample_cols > 64) return fail(DCX_ERR_ARG, "sample_cols must be 1..64");
  P->chk_tol = tol;
  P->chk_passes = max_passes;
  P->chk_cols = sample_cols;
  return DCX_OK;
}

int dcx_plan_get_check(const dcx_plan* P, double* rho, double* eta, int* pivot_fail,
                       int* passes, int* status) {
  if (!P) return fail(DCX_ERR_ARG, "null plan");
  for (int b = 0; b < P->nbatch; ++b) {
    const bool run = static_cast<int>(P->chk_eta.size()) == P->nbatch;
    if (rho) rho[b] = run ? P->chk_rho[b] : 0.0;
    if (eta) eta[b] = run ? P->chk_eta[b] : 0.0;
    if (pivot_fail) pivot_fail[b] = run ? P->chk_pivot[b] : 0;
  }
  if (passes) *passes = P->chk_passes_done;
  if (status) *status = P->chk_status;
  return DCX_OK;
}

int dcx_plan_join(dcx_plan* P, void* stream) {
  if (!P) return fail(DCX_ERR_ARG, "null plan");
  if (P->rest_pending)
    HIP_TRY(hipStreamWaitEvent(static_cast<hipStream_t>(stream), P->rest_ev, 0));
  return DCX_OK;
}

int dcx_plan_check_wait(dcx_plan* P) {
  if (!P) return fail(DCX_ERR_ARG, "null plan");
  HIP_TRY(hipSetDevice(P->device));
  return complete_pending(P);
}

int dcx_plan_inject_fault(dcx_plan* P, int batch, double rel) {
  if (!P) return fail(DCX_ERR_ARG, "null plan");
  if (batch >= P->nbatch) return fail(DCX_ERR_ARG, "batch member out of range");
  P->inj_batch = rel != 0.0 ? batch : -1;
  P->inj_rel = rel;
  return DCX_OK;
}

int dcx_plan_destroy(dcx_plan* plan) {
  delete plan;
  return DCX_OK;
}

int dcx_plan_get_info(const dcx_plan* P, dcx_plan_info* o) {
  if (!P || !o) return fail(DCX_ERR_ARG, "null argument");
  const auto& S = P->S;
  o->n = S.n;
  o->nnz = S.nnz;
  o->nbatch = P->nbatch;
  o->nsup = S.nsup;
  o->ndepth = S.ndepth;
  o->max_front = S.max_front;
  o->max_ns = S.max_ns;
  o->precision = P->precision;
  o->nnz_L = S.nnz_L;
  o->factor_flops = S.factor_flops;
  o->solve_flops_per_rhs = S.solve_flops_per_rhs;
  o->device_bytes = P->device_bytes;
  int64_t snf = 0, sns = 0;
  int nl = 0;
  for (int d = 0; d < S.ndepth; ++d) {
    if (S.ulvl_ptr[d + 1] > S.ulvl_ptr[d]) ++nl;
    for (int q = S.ulvl_ptr[d]; q < S.ulvl_ptr[d + 1]; ++q) {
      snf += S.ns[S.ulvl_sup[q]] + S.rs[S.ulvl_sup[q]];
      sns += S.ns[S.ulvl_sup[q]];
    }
  }
  o->sweep_rows_nf = snf;
  o->sweep_rows_ns = sns;
  o->panel_size = S.panel_size;
  o->level_launches = nl;
  o->nfused = S.nfused;
  o->buffer_generation = P->buf_gen;
  return DCX_OK;
}

int dcx_plan_get_perm(const dcx_plan* P, int32_t* perm) {
  if (!P || !perm) return fail(DCX_ERR_ARG, "null argument");
  std::memcpy(perm, P->S.perm.data(), sizeof(int32_t) * P->S.n);
  return DCX_OK;
}

int dcx_plan_set_refinement(dcx_plan* P, int passes) {
  if (!P || passes < 0 || passes > 16) return fail(DCX_ERR_ARG, "passes must be 0..16");
  P->refine = passes;
  return DCX_OK;
}

int dcx_plan_set_tuning(dcx_plan* P, int key, int value) {
  if (!P) return fail(DCX_ERR_ARG, "null argument");
  switch (key) {
    case DCX_TUNE_SWEEP_TC:
      if (value != 16 && value != 32 && value != 64)
        return fail(DCX_ERR_ARG, "sweep column tile must be 16, 32 or 64");
      P->tc_cap = value;
      break;
    case DCX_TUNE_SPLITK_BELOW:
      if (value < 0) return fail(DCX_ERR_ARG, "split-K threshold must be >= 0");
      P->splitk_below = value;
      break;
    case DCX_TUNE_WIDE_BELOW:
      if (value < 0) return fail(DCX_ERR_ARG, "wide-factor threshold must be >= 0");
      P->wide_below = value;
      break;
    case DCX_TUNE_PANEL_WIDTH:
      if (value != 16 && value != 32) return fail(DCX_ERR_ARG, "panel width must be 16 or 32");
      P->panel_width = value;
      break;
    case DCX_TUNE_ROWSPLIT_BELOW:
      if (value < 0) return fail(DCX_ERR_ARG, "row-split threshold must be >= 0");
      P->rowsplit_below = value;
      break;
    default:
      return fail(DCX_ERR_ARG, "unknown tuning key");
  }
  return DCX_OK;
}

int dcx_plan_get_tuning(const dcx_plan* P, int key, int* value) {
  if (!P || !value) return fail(DCX_ERR_ARG, "null argument");
  switch (key) {
    case DCX_TUNE_SWEEP_TC: *value = P->tc_cap; break;
    case DCX_TUNE_SPLITK_BELOW: *value = P->splitk_below; break;
    case DCX_TUNE_WIDE_BELOW: *value = P->wide_below; break;
    case DCX_TUNE_PANEL_WIDTH: *value = P->panel_width; break;
    case DCX_TUNE_ROWSPLIT_BELOW: *value = P->rowsplit_below; break;
    default: return fail(DCX_ERR_ARG, "unknown tuning key");
  }
  return DCX_OK;
}

int dcx_plan_set_timing(dcx_plan* P, int enable) {
  if (!P) return fail(DCX_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(P->device));
  for (hipEvent_t e : P->ev) HIP_TRY(hipEventDestroy(e));
  P->ev.clear();
  P->timing = enable != 0;
  return DCX_OK;
}

int dcx_plan_set_overlap(dcx_plan* P, int enable) {
  if (!P) return fail(DCX_ERR_ARG, "null argument");
  // pending checks keep their events, and the rest of a split sweep stays
  // joined through rest_ev: nothing to wait for here
  P->overlap = enable != 0;
  return DCX_OK;
}

int dcx_plan_sweep_times(dcx_plan* P, double* fwd_ms, double* bwd_ms, int* nsweeps) {
  if (!P || !fwd_ms || !bwd_ms || !nsweeps) return fail(DCX_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(P->device));
  double f = 0.0, b = 0.0;
  const int n = static_cast<int>(P->ev.size() / 3);
  for (int i = 0; i < n; ++i) {
    float a = 0.f, c = 0.f;
    HIP_TRY(hipEventSynchronize(P->ev[3 * i + 2]));
    HIP_TRY(hipEventElapsedTime(&a, P->ev[3 * i], P->ev[3 * i + 1]));
    HIP_TRY(hipEventElapsedTime(&c, P->ev[3 * i + 1], P->ev[3 * i + 2]));
    f += a;
    b += c;
  }
  *fwd_ms = f;
  *bwd_ms = b;
  *nsweeps = n;
  return DCX_OK;
}

int dcx_plan_set_fusion(dcx_plan* P, int fuse_rows) {
  if (!P || fuse_rows < 0) return fail(DCX_ERR_ARG, "bad arguments");
  HIP_TRY(hipSetDevice(P->device));
  HIP_TRY(hipDeviceSynchronize());
  P->fuse_rows = fuse_rows;
  dcx::build_fused(P->S, fuse_rows);
  int rc = upload_fused(P);
  if (rc) return rc;
  P->upd_k = 0;  // the fused update buffer is resized on the next solve
  if (P->d_active) {  // active lists follow the supernodes above the subtrees
    std::vector<int> act(P->S.nsup);
    HIP_TRY(hipMemcpy(act.data(), P->d_active, sizeof(int) * act.size(), hipMemcpyDeviceToHost));
    std::vector<int> ptr(P->S.ndepth + 1, 0), lst;
    for (int d = 0; d < P->S.ndepth; ++d) {
      for (int q = P->S.ulvl_ptr[d]; q < P->S.ulvl_ptr[d + 1]; ++q)
        if (act[P->S.ulvl_sup[q]]) lst.push_back(P->S.ulvl_sup[q]);
      ptr[d + 1] = static_cast<int>(lst.size());
    }
    if (P->d_act_lvl) HIP_TRY(dev_free(P->d_act_lvl));
    P->d_act_lvl = nullptr;
    if (lst.empty()) lst.push_back(0);
    HIP_TRY(upload(&P->d_act_lvl, lst));
    P->act_lvl_ptr = ptr;
  }
  return DCX_OK;
}

int dcx_set_rhs_support(dcx_plan* P, const int32_t* rows, int nrows) {
  if (!P || (nrows > 0 && !rows)) return fail(DCX_ERR_ARG, "null argument");
  const auto& S = P->S;
  std::vector<int> act(S.nsup, 0);
  for (int i = 0; i < nrows; ++i) {
    if (rows[i] < 0 || rows[i] >= S.n) return fail(DCX_ERR_ARG, "support row out of range");
    for (int s = S.snode_of[S.iperm[rows[i]]]; s >= 0 && !act[s]; s = S.parent[s]) act[s] = 1;
  }
  std::vector<int> ptr(S.ndepth + 1, 0), lst, iptr(S.ndepth + 1, 0), ilst;
  for (int d = 0; d < S.ndepth; ++d) {
    for (int q = S.ulvl_ptr[d]; q < S.ulvl_ptr[d + 1]; ++q)
      (act[S.ulvl_sup[q]] ? lst : ilst).push_back(S.ulvl_sup[q]);
    ptr[d + 1] = static_cast<int>(lst.size());
    iptr[d + 1] = static_cast<int>(ilst.size());
  }
  HIP_TRY(hipDeviceSynchronize());  // the old tables may still be read
  if (P->d_active) HIP_TRY(dev_free(P->d_active));
  if (P->d_act_lvl) HIP_TRY(dev_free(P->d_act_lvl));
  if (P->d_inact_lvl) HIP_TRY(dev_free(P->d_inact_lvl));
  P->d_active = P->d_act_lvl = P->d_inact_lvl = nullptr;
  HIP_TRY(upload(&P->d_active, act));
  if (lst.empty()) lst.push_back(0);
  HIP_TRY(upload(&P->d_act_lvl, lst));
  if (ilst.empty()) ilst.push_back(0);
  HIP_TRY(upload(&P->d_inact_lvl, ilst));
  P->act_lvl_ptr = ptr;
  P->inact_lvl_ptr = iptr;
  return DCX_OK;
}

namespace {
// copy the caller's values, clear the pivot flags, and prepare the values the
// factorisation reads (a fault-injected copy when one is armed)
static int load_values(dcx_plan* P, const double* values, hipStream_t st) {
  if (P->rest_pending) {  // the rest of a split sweep still reads the factor
    HIP_TRY(hipStreamWaitEvent(st, P->rest_ev, 0));
    P->rest_pending = false;
  }
  const size_t vb = static_cast<size_t>(P->nbatch) * P->S.nnz * sizeof(double);
  HIP_TRY(hipMemcpyAsync(P->d_vals, values, vb, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(256), 0, st, P->d_info, P->nbatch);
  if (P->inj_batch >= 0) {
    if (!P->d_fvals) {
      if (int rc = refuse_in_capture(st)) return rc;
      HIP_TRY(dev_malloc(&P->d_fvals, std::max<size_t>(vb, 256)));
    }
    hipLaunchKernelGGL(k_scale_vals, dim3(grid_stride(static_cast<int64_t>(P->nbatch) * P->S.nnz)),
                       dim3(256), 0, st, P->d_vals, P->d_fvals, static_cast<int64_t>(P->S.nnz),
                       P->nbatch, P->inj_batch, 1.0 + P->inj_rel);
  }
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

// The backward-error check of a finished solve (X: its solution, float or
// double, rows in elimination order when `permuted`): k_check on the first
// chk_cols columns, the sums and the pivot flags to pinned host memory, one
// synchronisation.  eta_b = |R| / (|A_b| |X| + |B|) (Frobenius norms over the
// sample) for every batch member; returns the number failing (eta > tol or a
// pivot that was not positive).
static int launch_check(dcx_plan* P, CheckSlot& C, const double* B, int64_t b_bs,
                        const void* X, bool xf32, bool permuted, int k, hipStream_t st) {
  const auto& S = P->S;
  const int nb = P->nbatch;
  const int cs = std::min(k, std::min(P->chk_cols, 64));
  const int lpr = cs > 32 ? 64 : cs > 16 ? 32 : cs > 8 ? 16 : 8;
  const int nblk = grid1(S.n, (256 / lpr) * kChkRows);
  if (!C.d_chk) {
    // partial sums (nb x nblk x 4, nblk at its largest: 64 lanes per row)
    // followed by the totals (nb x 4)
    const size_t mx = grid1(S.n, 4 * kChkRows);
    HIP_TRY(dev_malloc(&C.d_chk, sizeof(double) * 4 * nb * (mx + 1)));
    HIP_TRY(hipHostMalloc(&C.h_chk, (sizeof(double) * 4 + sizeof(int)) * nb));
    P->chk_tot_off = static_cast<size_t>(4) * nb * mx;
  }
  double* part = C.d_chk;
  double* tot = C.d_chk + P->chk_tot_off;
  const int* rp = permuted ? P->d_pcrow : P->d_crow;
  const int* cl = permuted ? P->d_pcol : P->d_col;
  const int* vi = permuted ? P->d_pvidx : nullptr;
  dim3 grid(nblk, nb);
  auto go = [&](auto kern, const auto* Xp) {
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, rp, cl, vi, P->d_vals,
                       static_cast<int64_t>(S.nnz), B, b_bs, Xp, S.n, k, cs, part);
  };
  const float* Xf = static_cast<const float*>(X);
  const double* Xd = static_cast<const double*>(X);
  // entries per row loaded together: 5 (2-D stencils), else 8
  const bool wide = P->max_row_nnz > 5;
  auto pick = [&](auto tx, const auto* Xp) {
    using TX = decltype(tx);
    switch (lpr) {
      case 8: wide ? go(k_check<TX, 8, 8>, Xp) : go(k_check<TX, 8, 5>, Xp); break;
      case 16: wide ? go(k_check<TX, 16, 8>, Xp) : go(k_check<TX, 16, 5>, Xp); break;
      case 32: wide ? go(k_check<TX, 32, 8>, Xp) : go(k_check<TX, 32, 5>, Xp); break;
      default: wide ? go(k_check<TX, 64, 8>, Xp) : go(k_check<TX, 64, 5>, Xp); break;
    }
  };
  if (xf32) pick(0.0f, Xf); else pick(0.0, Xd);
  hipLaunchKernelGGL(k_check_sum, dim3(nb), dim3(256), 0, st, part, nblk, tot);
  HIP_TRY(hipGetLastError());
  double* hs = static_cast<double*>(C.h_chk);
  int* hi = reinterpret_cast<int*>(hs + 4 * nb);
  HIP_TRY(hipMemcpyAsync(hs, tot, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hi, P->d_info, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
  return DCX_OK;
}

// the check's verdict from the host copies (after the stream or the check
// event has been waited on); returns the number of failing systems
static int eval_check(dcx_plan* P, const CheckSlot& C) {
  const int nb = P->nbatch;
  const double* hs = static_cast<const double*>(C.h_chk);
  const int* hi = reinterpret_cast<const int*>(hs + 4 * nb);
  P->chk_eta.assign(nb, 0.0);
  P->chk_rho.assign(nb, 0.0);
  P->chk_pivot.assign(nb, 0);
  int bad = 0;
  auto ratio = [](double num, double den) {
    const double q = den > 0.0 ? num / den : (num > 0.0 ? INFINITY : 0.0);
    return std::isfinite(q) ? q : INFINITY;
  };
  for (int b = 0; b < nb; ++b) {
    const double r = std::sqrt(hs[4 * b]), x = std::sqrt(hs[4 * b + 1]),
                 bn = std::sqrt(hs[4 * b + 2]), a = std::sqrt(hs[4 * b + 3]);
    P->chk_rho[b] = ratio(r, bn);
    P->chk_eta[b] = ratio(r, a * x + bn);
    P->chk_pivot[b] = hi[b] != 0;
    if (P->chk_pivot[b] || !(P->chk_rho[b] <= P->chk_tol)) ++bad;
  }
  return bad;
}

// launch_check, one synchronisation, eval_check; a negative value is a
// status (HIP error)
static int run_check(dcx_plan* P, const double* B, int64_t b_bs, const void* X, bool xf32,
                     bool permuted, int k, hipStream_t st) {
  if (int rc = launch_check(P, P->slot[2], B, b_bs, X, xf32, permuted, k, st)) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  return eval_check(P, P->slot[2]);
}

static std::string check_message(const dcx_plan* P) {
  std::string m;
  for (int b = 0; b < P->nbatch; ++b) {
    if (!P->chk_pivot[b] && P->chk_rho[b] <= P->chk_tol) continue;
    char buf[112];
    std::snprintf(buf, sizeof(buf), "%s system %d: %s relative residual %.2e", m.empty() ? "" : ";",
                  b, P->chk_pivot[b] ? "non-positive pivot," : "", P->chk_rho[b]);
    m += buf;
  }
  return m;
}
}  // namespace

int dcx_factorize(dcx_plan* P, const double* values, void* stream) {
  if (!P || !values) return fail(DCX_ERR_ARG, "null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(P->device));
  if (int rc0 = complete_all(P)) return rc0;
  if (int rc = load_values(P, values, st)) return rc;
  int rc = P->precision == DCX_FP32 ? launch_factor<float>(P, st) : launch_factor<double>(P, st);
  if (rc) return rc;
  P->factored = true;
  return DCX_OK;
}

// one refinement pass of an FP32 plan: r = B - A x (FP64), W = A^{-1} r
// (FP32 sweeps), x += W (FP64)
static int mixed_refine_pass(dcx_plan* P, const double* B, int64_t bbs, double* x, float* W,
                             int k, bool permuted, hipStream_t st) {
  const auto& S = P->S;
  const int nb = P->nbatch, n = S.n;
  const int64_t blk = static_cast<int64_t>(n) * k;
  int rc = ensure(P, reinterpret_cast<void**>(&P->d_r), &P->r_bytes,
                  static_cast<size_t>(nb) * blk * 8, st);
  if (rc) return rc;
  dim3 rgrid(grid1(n, 4), nb);
  if (permuted)
    hipLaunchKernelGGL(k_residual, rgrid, dim3(256), 0, st, P->d_pcrow, P->d_pcol, P->d_pvidx,
                       P->d_vals, static_cast<int64_t>(S.nnz), B, bbs, x, P->d_r, n, k);
  else
    hipLaunchKernelGGL(k_residual, rgrid, dim3(256), 0, st, P->d_crow, P->d_col,
                       static_cast<const int*>(nullptr), P->d_vals,
                       static_cast<int64_t>(S.nnz), B, bbs, x, P->d_r, n, k);
  SweepIO ir;
  ir.B = P->d_r;
  ir.b_bs = blk;
  ir.brow = permuted ? nullptr : P->d_perm;
  ir.xrow = permuted ? nullptr : P->d_perm;
  ir.X64 = x;
  ir.xmode = 2;
  return sweeps<float>(P, W, k, false, ir, st);
}

// After a check that found `bad` failing systems (the verdict in chk_*): an
// FP64 plan or a non-positive pivot fails at once; an FP32 plan gets up to
// chk_passes refinement passes, each followed by a check.  B, xv, k, flags:
// the solve's right-hand side and solution as the caller passed them (B the
// copy when b and x alias).
static int finish_check(dcx_plan* P, int bad, const double* B, int64_t bbs, void* xv, int k,
                        int flags, hipStream_t st) {
  const bool permuted = flags & DCX_PERMUTED, f32x = (flags & DCX_F32_FIELDS) != 0;
  const int nb = P->nbatch;
  const int64_t blk = static_cast<int64_t>(P->S.n) * k;
  double* x = f32x ? (P->refine > 0 ? P->d_x : nullptr) : static_cast<double*>(xv);
  float* W = f32x ? static_cast<float*>(xv) : static_cast<float*>(P->d_w);
  for (int extra = 0;; ++extra) {
    if (bad < 0) return bad;
    P->chk_passes_done = (P->precision == DCX_FP32 ? P->refine : 0) + extra;
    if (bad == 0) {
      P->chk_status = 1;
      return DCX_OK;
    }
    bool pivot = false;
    for (int v : P->chk_pivot) pivot = pivot || v;
    if (P->precision == DCX_FP64) {
      P->chk_status = 2;
      return fail(DCX_ERR_NUMERIC,
                  "FP64 solve failed its check (matrix not positive definite?):" +
                      check_message(P));
    }
    if (pivot || extra >= P->chk_passes) {
      P->chk_status = 2;
      return fail(DCX_ERR_NUMERIC, "FP32 solve failed its check after " +
                                       std::to_string(P->chk_passes_done) +
                                       " refinement pass(es):" + check_message(P));
    }
    if (P->rest_pending) HIP_TRY(hipStreamWaitEvent(st, P->rest_ev, 0));  // all rows of x
    if (f32x && !x) {  // first pass on FP32 fields: an FP64 iterate from them
      int rc = ensure(P, reinterpret_cast<void**>(&P->d_x), &P->x_bytes,
                      static_cast<size_t>(nb) * blk * 8, st);
      if (rc) return rc;
      x = P->d_x;
      hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_stride(nb * blk)), dim3(256), 0,
                         st, W, x, nb * blk);
    }
    int rc = mixed_refine_pass(P, B, bbs, x, W, k, permuted, st);
    if (rc) return rc;
    if (f32x)
      hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_stride(nb * blk)), dim3(256), 0,
                         st, x, W, nb * blk);
    HIP_TRY(hipGetLastError());
    bad = run_check(P, B, bbs, f32x ? static_cast<const void*>(W) : x, f32x, permuted, k, st);
  }
}

// the check of a finished solve: now (one synchronisation) or, with
// DCX_CHECK_DEFER, launched with its verdict left to dcx_plan_check_wait
static int start_check(dcx_plan* P, const double* B, int64_t bbs, void* xv, int k, int flags,
                       hipStream_t st) {
  const bool permuted = flags & DCX_PERMUTED, f32x = (flags & DCX_F32_FIELDS) != 0;
  P->chk_status = 0;
  if (flags & DCX_CHECK_DEFER) {
    // on the plan's check stream, after the solve (and the rest of a split
    // backward sweep): the work the caller queues next on `st` does not wait
    // for it.  Without the plan's own streams (dcx_plan_set_overlap) the
    // check runs on `st` in order.  At most two checks are pending; a third
    // completes the oldest.
    if (P->slot_count == 2)
      if (int rc = complete_pending(P)) return rc;
    if (P->overlap && !P->chk_st)
      HIP_TRY(hipStreamCreateWithFlags(&P->chk_st, hipStreamNonBlocking));
    hipStream_t cs = P->overlap ? P->chk_st : st;
    CheckSlot& C = P->slot[(P->slot_head + P->slot_count) % 2];
    if (!C.ev) {
      HIP_TRY(hipEventCreateWithFlags(&C.ev, hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&C.src, hipEventDisableTiming));
    }
    if (cs != st) {
      HIP_TRY(hipEventRecord(C.src, st));
      HIP_TRY(hipStreamWaitEvent(cs, C.src, 0));
    }
    if (P->rest_pending) HIP_TRY(hipStreamWaitEvent(cs, P->rest_ev, 0));
    if (int rc = launch_check(P, C, B, bbs, xv, f32x, permuted, k, cs)) return rc;
    HIP_TRY(hipEventRecord(C.ev, cs));
    C.B = B;
    C.bbs = bbs;
    C.xv = xv;
    C.k = k;
    C.flags = flags;
    C.st = st;
    ++P->slot_count;
    return DCX_OK;
  }
  if (P->rest_pending) HIP_TRY(hipStreamWaitEvent(st, P->rest_ev, 0));
  return finish_check(P, run_check(P, B, bbs, xv, f32x, permuted, k, st), B, bbs, xv, k, flags,
                      st);
}

// dcx_solve and dcx_factor_solve (with_factor: the factorisation runs here,
// overlapped with the first forward sweep).  With the check on (and the
// stream not being captured) every solve ends with run_check; an FP32 plan
// whose check fails adds refinement passes (FP64 residual, FP32 correction)
// up to chk_passes, then reports DCX_ERR_NUMERIC; an FP64 plan reports at
// once.  A non-positive pivot is reported without refinement.
static int solve_impl(dcx_plan* P, const double* b, void* xv, int k, int flags, hipStream_t st,
               bool with_factor) {
  const auto& S = P->S;
  const int nb = P->nbatch;
  const bool shared = flags & DCX_SHARED_RHS;
  const bool permuted = flags & DCX_PERMUTED;
  const bool sparse = (flags & DCX_SPARSE_RHS) != 0;
  const bool f32x = (flags & DCX_F32_FIELDS) != 0;
  if (f32x && (P->precision != DCX_FP32 || !permuted))
    return fail(DCX_ERR_ARG, "DCX_F32_FIELDS needs an FP32 plan and DCX_PERMUTED");
  if (f32x && static_cast<const void*>(b) == xv)
    return fail(DCX_ERR_ARG, "b and a float x must not alias");
  // a pending check's refinement would find the FP64 iterate of fixed passes
  // overwritten: with fixed passes, complete them first
  if (P->refine > 0 && P->precision == DCX_FP32)
    if (int rc0 = complete_all(P)) return rc0;
  const bool check = P->chk_tol > 0.0 && !capturing(st);
  P->chk_status = check ? 0 : 3;
  // split backward sweep: only where it is safe and useful (see DCX_SPLIT_BWD):
  // a sparse right-hand side, the caller's block as the working block (FP64,
  // or FP32 fields without fixed refinement), no fused subtrees, no
  // row-grouped backward levels (their scratch is shared), no sweep timing
  const bool split = (flags & DCX_SPLIT_BWD) && sparse && P->d_inact_lvl && !S.nfused &&
                     S.max_ns <= 128 && !P->timing && P->overlap &&
                     (P->precision == DCX_FP64 ? permuted : (f32x && P->refine == 0));
  double* x = f32x ? nullptr : static_cast<double*>(xv);
  const int64_t blk = static_cast<int64_t>(S.n) * k;
  const int n = S.n;
  SweepIO io;
  io.B = b;
  io.b_bs = shared ? 0 : blk;
  io.brow = permuted ? nullptr : P->d_perm;
  io.xrow = permuted ? nullptr : P->d_perm;
  int rc;
  // the residual (refinement, check) needs the right-hand side intact
  if ((check || (P->precision == DCX_FP32 && P->refine > 0)) && b == x) {
    rc = ensure(P, reinterpret_cast<void**>(&P->d_b), &P->b_bytes,
                static_cast<size_t>(shared ? 1 : nb) * blk * 8, st);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(P->d_b, b, (shared ? 1 : nb) * blk * 8, hipMemcpyDeviceToDevice, st));
    io.B = P->d_b;
  }
  const double* Borig = io.B;
  const int64_t Bbs = io.b_bs;
  if (P->precision == DCX_FP64) {
    double* W = x;
    if (!permuted) {
      rc = ensure(P, &P->d_w, &P->w_bytes, static_cast<size_t>(nb) * blk * 8, st);
      if (rc) return rc;
      W = static_cast<double*>(P->d_w);
      io.X64 = x;
      io.xmode = 1;
    }
    if (split && (rc = ensure_side(P))) return rc;
    rc = with_factor ? factor_and_sweeps<double>(P, W, k, sparse, io, st, split)
                     : sweeps<double>(P, W, k, sparse, io, st, false, split);
    if (rc || !check) return rc;
    return start_check(P, Borig, Bbs, xv, k, flags, st);
  }
  // mixed: FP32 sweeps, FP64 residual refinement; the solution accumulates in
  // x (double) or, for a float x, in a plan buffer rounded into x at the end
  float* W;
  if (f32x) {
    W = static_cast<float*>(xv);
    if (P->refine > 0) {
      rc = ensure(P, reinterpret_cast<void**>(&P->d_x), &P->x_bytes,
                  static_cast<size_t>(nb) * blk * 8, st);
      if (rc) return rc;
      x = P->d_x;
    }
  } else {
    rc = ensure(P, &P->d_w, &P->w_bytes, static_cast<size_t>(nb) * blk * 4, st);
    if (rc) return rc;
    W = static_cast<float*>(P->d_w);
  }
  io.X64 = x;
  io.xmode = x ? 1 : 0;
  if (split && (rc = ensure_side(P))) return rc;
  rc = with_factor ? factor_and_sweeps<float>(P, W, k, sparse, io, st, split)
                   : sweeps<float>(P, W, k, sparse, io, st, false, split);
  if (rc) return rc;
  for (int it = 0; it < P->refine; ++it) {
    rc = mixed_refine_pass(P, Borig, Bbs, x, W, k, permuted, st);
    if (rc) return rc;
  }
  if (f32x && x)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_stride(nb * blk)), dim3(256), 0, st,
                       x, W, nb * blk);
  HIP_TRY(hipGetLastError());
  if (!check) return DCX_OK;
  return start_check(P, Borig, Bbs, xv, k, flags, st);
}

// complete a deferred check (its solve's verdict, refinement, status)
static int complete_all(dcx_plan* P) {
  while (P->slot_count > 0)
    if (int rc = complete_pending(P)) {
      while (P->slot_count > 0) {  // drop the rest: their solves are superseded
        P->slot_head = (P->slot_head + 1) % 2;
        --P->slot_count;
      }
      return rc;
    }
  return DCX_OK;
}

static int complete_pending(dcx_plan* P) {
  if (P->slot_count == 0) return DCX_OK;
  CheckSlot& C = P->slot[P->slot_head];
  P->slot_head = (P->slot_head + 1) % 2;
  --P->slot_count;
  HIP_TRY(hipEventSynchronize(C.ev));
  return finish_check(P, eval_check(P, C), C.B, C.bbs, C.xv, C.k, C.flags, C.st);
}

int dcx_solve(dcx_plan* P, const double* b, void* xv, int k, int flags, void* stream) {
  if (!P || !b || !xv) return fail(DCX_ERR_ARG, "null argument");
  if (!P->factored) return fail(DCX_ERR_STATE, "dcx_solve before dcx_factorize");
  if (k < 1) return fail(DCX_ERR_ARG, "k must be >= 1");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(P->device));
  return solve_impl(P, b, xv, k, flags, st, false);
}

int dcx_factor_solve(dcx_plan* P, const double* values, const double* b, void* xv, int k,
                     int flags, void* stream) {
  if (!P || !values || !b || !xv) return fail(DCX_ERR_ARG, "null argument");
  if (k < 1) return fail(DCX_ERR_ARG, "k must be >= 1");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(P->device));
  if (int rc0 = complete_all(P)) return rc0;
  if (int rc = load_values(P, values, st)) return rc;
  P->factored = true;
  return solve_impl(P, b, xv, k, flags, st, true);
}

int dcx_grad_values(dcx_plan* P, const void* y, const void* x, int k,
                    const double* scale, double* gv, int flags, void* stream) {
  if (!P || !y || !x || !gv) return fail(DCX_ERR_ARG, "null argument");
  if (k < 1) return fail(DCX_ERR_ARG, "k must be >= 1");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(P->device));
  const auto& S = P->S;
  // lanes per row: whole wavefronts for long rows, 32 or 16 lanes (2 or 4
  // rows per wavefront) when 4 columns per lane cover the row
  const int lpr = k <= 64 ? 16 : k <= 128 ? 32 : 64;
  const int rows_per_block = 4 * (64 / lpr);  // 256 threads
  const int blocks = ((grid1(S.n, rows_per_block) + 7) / 8) * 8;  // XCD remap: multiple of 8
  dim3 grid(blocks, P->nbatch);
  // permuted fields: rows walked in elimination order through the permuted
  // pattern (consecutive waves read consecutive rows of y, and the neighbours
  // mostly lie in the same supernode), results scattered to the caller's
  // entry order
  const bool perm = (flags & DCX_PERMUTED) != 0;
  const int* rp = perm ? P->d_pcrow : P->d_crow;
  const int* cl = perm ? P->d_pcol : P->d_col;
  const int* oi = perm ? P->d_pvidx : nullptr;
  const int* pcol = perm ? P->d_gpad_col_p : P->d_gpad_col_u;
  const int* pout = perm ? P->d_gpad_out_p : P->d_gpad_out_u;
  // entries per pass: the longest row up to 5 (2-D stencils), else 8
  const bool wide = P->max_row_nnz > 5;
  auto go = [&](auto tag, auto me) {
    using TI = decltype(tag);
    constexpr int ME = decltype(me)::value;
    // padded rows when every row fits the pass width
    const bool padded = P->gpad_me == ME;
    // k_grad_values reserves kGatherLds bytes of LDS it does not use: at most
    // 6 workgroups (wavefronts per SIMD) per CU.  With its row bases in
    // scalar registers the one-row-per-wavefront form needs 61 VGPRs, i.e. 8
    // wavefronts per SIMD, and at 8 the neighbour rows that several rows
    // re-read are evicted before the re-reads: mixed eval+grad of the 25k-cell
    // line 3.54 ms at 8, 3.40 at 7, 3.33 at 6 and 5, 3.35 at 4, 3.40 at 3
    // (3.37 for the earlier per-lane-address form, 88 VGPRs, 5 per SIMD)
    auto launch = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid, dim3(256), kGatherLds, st, rp, cl, oi,
                         padded ? pcol : nullptr,
                         padded ? pout : nullptr, static_cast<const TI*>(y),
                         static_cast<const TI*>(x), S.n, k, scale, gv,
                         static_cast<int64_t>(S.nnz));
    };
    // k_grad_rows for long padded FP64 rows, four vectors per lane and pass
    // (two passes on the 25k-cell line, 849 columns): eval+grad 6.20 ->
    // 6.10 ms there.  FP32 rows keep k_grad_values: k_grad_rows (two vectors
    // per lane) timed faster alone but made the mixed eval+grad slower,
    // 3.78 -> 3.94 ms.  DCX_GATHER_NV (vectors per lane, 0: k_grad_values)
    // overrides both.
    static const int nv_env = env_int("DCX_GATHER_NV", -1);
    const int nv = nv_env >= 0 ? nv_env : (sizeof(TI) == 4 ? 0 : 4);
    constexpr int VWI = 16 / static_cast<int>(sizeof(TI));
    if (padded && lpr == 64 && nv > 0 && k >= VWI) {
      // whole rows in flight (k_grad_rows), NV vectors of each row per lane
      auto rl = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, pcol, pout, static_cast<const TI*>(y),
                           static_cast<const TI*>(x), S.n, k, scale, gv,
                           static_cast<int64_t>(S.nnz));
      };
      if (nv <= 2) rl(k_grad_rows<TI, ME, 2>);
      else if (nv <= 4) rl(k_grad_rows<TI, ME, 4>);
      else rl(k_grad_rows<TI, ME, 8>);
    } else if (padded) {
      if (lpr == 16) launch(k_grad_values<TI, ME, 16, true>);
      else if (lpr == 32) launch(k_grad_values<TI, ME, 32, true>);
      else launch(k_grad_values<TI, ME, 64, true>);
    } else {
      if (lpr == 16) launch(k_grad_values<TI, ME, 16, false>);
      else if (lpr == 32) launch(k_grad_values<TI, ME, 32, false>);
      else launch(k_grad_values<TI, ME, 64, false>);
    }
  };
  using E5 = std::integral_constant<int, 5>;
  using E8 = std::integral_constant<int, 8>;
  if (flags & DCX_F32_FIELDS) {
    if (wide) go(0.0f, E8()); else go(0.0f, E5());
  } else {
    if (wide) go(0.0, E8()); else go(0.0, E5());
  }
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

int dcx_spmm(dcx_plan* P, const double* vals, const double* x, double* y, int k, double alpha,
             int flags, void* stream) {
  if (!P || !vals || !x || !y) return fail(DCX_ERR_ARG, "null argument");
  if (k < 1) return fail(DCX_ERR_ARG, "k must be >= 1");
  if (x == y) return fail(DCX_ERR_ARG, "x and y must not alias");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(P->device));
  const auto& S = P->S;
  dim3 grid(grid1(S.n, 4), P->nbatch);
  if (flags & DCX_PERMUTED)
    hipLaunchKernelGGL(k_spmm, grid, dim3(256), 0, st, P->d_pcrow, P->d_pcol, P->d_pvidx, vals,
                       static_cast<int64_t>(S.nnz), alpha, x, y, S.n, k);
  else
    hipLaunchKernelGGL(k_spmm, grid, dim3(256), 0, st, P->d_crow, P->d_col,
                       static_cast<const int*>(nullptr), vals, static_cast<int64_t>(S.nnz), alpha,
                       x, y, S.n, k);
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

int dcx_mma_probe(int fp32, const void* A, const void* B, void* C, void* stream) {
  if (!A || !B || !C) return fail(DCX_ERR_ARG, "null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (fp32)
    hipLaunchKernelGGL(k_mma_probe<float>, dim3(1), dim3(64), 0, st,
                       static_cast<const float*>(A), static_cast<const float*>(B),
                       static_cast<float*>(C));
  else
    hipLaunchKernelGGL(k_mma_probe<double>, dim3(1), dim3(64), 0, st,
                       static_cast<const double*>(A), static_cast<const double*>(B),
                       static_cast<double*>(C));
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

int dcx_permute_rows(dcx_plan* P, const double* in, double* out, int k, int nblocks,
                     int to_elim, void* stream) {
  if (!P || !in || !out || in == out) return fail(DCX_ERR_ARG, "bad arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(P->device));
  const int64_t blk = static_cast<int64_t>(P->S.n) * k;
  dim3 grid(grid1(P->S.n, 4), nblocks);
  hipLaunchKernelGGL((k_permute<double, double>), grid, dim3(256), 0, st, in, out, P->d_perm,
                     P->S.n, k, blk, blk, to_elim);
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

}  // extern "C"

// ===========================================================================
// assembly map
// ===========================================================================
struct dcx_lmap {
  int nnz = 0, ncell = 0, nky = 0, nbc = 0, device = 0;
  int ngroup = 1;           // wavenumber groups (dcx_lmap_set_groups)
  double* d_ky2 = nullptr;  // nky x ngroup
  int* d_grp = nullptr;     // group of each value entry, null for one group
  int *d_ptr = nullptr, *d_cell = nullptr;
  double *d_le = nullptr, *d_ln = nullptr;
  int *d_bptr = nullptr, *d_bcell = nullptr, *d_bc_e = nullptr;
  double* d_bcoef = nullptr;
  int *d_tptr = nullptr, *d_tent = nullptr, *d_te = nullptr, *d_tbptr = nullptr, *d_tbq = nullptr;
  double *d_tle = nullptr, *d_tln = nullptr;  // le, ln in transposed (cell-major) order
  double *d_ge = nullptr, *d_gn = nullptr;
  int max_bc_per_entry = 0;
  double* d_jws = nullptr;  // Jacobian-rows workspace, grown on demand
  size_t jws_bytes = 0;
  ~dcx_lmap() {
    void* ptrs[] = {d_ky2, d_grp, d_ptr, d_cell, d_le, d_ln, d_bptr, d_bcell, d_bc_e, d_bcoef,
                    d_tptr, d_tent, d_te, d_tbptr, d_tbq, d_ge, d_gn, d_jws, d_tle, d_tln};
    for (void* p : ptrs)
      if (p) (void)dev_free(p);
  }
};

extern "C" {

int dcx_lmap_create(int nnz, int ncell, int nky, const double* ky2, const int32_t* ptr,
                    const int32_t* cell, const double* le, const double* ln, int nbc,
                    const int32_t* bc_e, const int32_t* bc_cell, const double* bc_coef,
                    dcx_lmap** out) {
  try {
    if (!out || !ky2 || !ptr || !cell || !le || !ln) return fail(DCX_ERR_ARG, "null argument");
    if (nbc > 0 && (!bc_e || !bc_cell || !bc_coef)) return fail(DCX_ERR_ARG, "null bc arrays");
    *out = nullptr;
    std::unique_ptr<dcx_lmap> M(new dcx_lmap());
    M->nnz = nnz;
    M->ncell = ncell;
    M->nky = nky;
    M->nbc = nbc;
    HIP_TRY(hipGetDevice(&M->device));
    const int np = ptr[nnz];
    for (int p = 0; p < np; ++p)
      if (cell[p] < 0 || cell[p] >= ncell) return fail(DCX_ERR_ARG, "map cell out of range");
    // bc entries grouped by value entry (CSR over entries) and by cell
    std::vector<int> bptr(nnz + 1, 0), order(nbc);
    for (int q = 0; q < nbc; ++q) {
      if (bc_e[q] < 0 || bc_e[q] >= nnz || bc_cell[q] < 0 || bc_cell[q] >= ncell)
        return fail(DCX_ERR_ARG, "bc entry out of range");
      bptr[bc_e[q] + 1]++;
    }
    for (int e = 0; e < nnz; ++e) M->max_bc_per_entry = std::max(M->max_bc_per_entry, bptr[e + 1]);
    for (int e = 0; e < nnz; ++e) bptr[e + 1] += bptr[e];
    {
      std::vector<int> fill(bptr.begin(), bptr.end() - 1);
      for (int q = 0; q < nbc; ++q) order[fill[bc_e[q]]++] = q;
    }
    std::vector<int> bcell(nbc), be(nbc);
    std::vector<double> bcoef(static_cast<size_t>(nky) * nbc);
    for (int t = 0; t < nbc; ++t) {
      const int q = order[t];
      bcell[t] = bc_cell[q];
      be[t] = bc_e[q];
      for (int k = 0; k < nky; ++k) bcoef[static_cast<size_t>(k) * nbc + t] = bc_coef[static_cast<size_t>(k) * nbc + q];
    }
    // transposed CSR by cell
    std::vector<int> tptr(ncell + 1, 0), tent(np), te(np);
    std::vector<int> ent_of(np);
    for (int e = 0; e < nnz; ++e)
      for (int p = ptr[e]; p < ptr[e + 1]; ++p) ent_of[p] = e;
    for (int p = 0; p < np; ++p) tptr[cell[p] + 1]++;
    for (int c = 0; c < ncell; ++c) tptr[c + 1] += tptr[c];
    {
      std::vector<int> fill(tptr.begin(), tptr.end() - 1);
      for (int p = 0; p < np; ++p) {
        int t = fill[cell[p]]++;
        tent[t] = p;
        te[t] = ent_of[p];
      }
    }
    std::vector<int> tbptr(ncell + 1, 0), tbq(nbc);
    for (int t = 0; t < nbc; ++t) tbptr[bcell[t] + 1]++;
    for (int c = 0; c < ncell; ++c) tbptr[c + 1] += tbptr[c];
    {
      std::vector<int> fill(tbptr.begin(), tbptr.end() - 1);
      for (int t = 0; t < nbc; ++t) tbq[fill[bcell[t]]++] = t;
    }
    HIP_TRY(upload(&M->d_ky2, std::vector<double>(ky2, ky2 + nky)));
    HIP_TRY(upload(&M->d_ptr, std::vector<int>(ptr, ptr + nnz + 1)));
    HIP_TRY(upload(&M->d_cell, std::vector<int>(cell, cell + np)));
    HIP_TRY(upload(&M->d_le, std::vector<double>(le, le + np)));
    HIP_TRY(upload(&M->d_ln, std::vector<double>(ln, ln + np)));
    HIP_TRY(upload(&M->d_bptr, bptr));
    if (nbc) {
      HIP_TRY(upload(&M->d_bcell, bcell));
      HIP_TRY(upload(&M->d_bc_e, be));
      HIP_TRY(upload(&M->d_bcoef, bcoef));
      HIP_TRY(upload(&M->d_tbq, tbq));
    }
    HIP_TRY(upload(&M->d_tptr, tptr));
    {
      std::vector<double> tle(np), tln(np);
      for (int t = 0; t < np; ++t) {
        tle[t] = le[tent[t]];
        tln[t] = ln[tent[t]];
      }
      HIP_TRY(upload(&M->d_tle, tle));
      HIP_TRY(upload(&M->d_tln, tln));
    }
    HIP_TRY(upload(&M->d_tent, tent));
    HIP_TRY(upload(&M->d_te, te));
    HIP_TRY(upload(&M->d_tbptr, tbptr));
    HIP_TRY(dev_malloc(&M->d_ge, std::max<size_t>(8, nnz * sizeof(double))));
    HIP_TRY(dev_malloc(&M->d_gn, std::max<size_t>(8, nnz * sizeof(double))));
    *out = M.release();
    return DCX_OK;
  } catch (const std::exception& e) {
    return fail(DCX_ERR_INTERNAL, e.what());
  }
}

int dcx_lmap_destroy(dcx_lmap* m) {
  delete m;
  return DCX_OK;
}

int dcx_lmap_set_groups(dcx_lmap* M, int ngroup, const int32_t* group, const double* ky2) {
  try {
    if (!M || !group || !ky2 || ngroup < 1) return fail(DCX_ERR_ARG, "bad arguments");
    for (int e = 0; e < M->nnz; ++e)
      if (group[e] < 0 || group[e] >= ngroup) return fail(DCX_ERR_ARG, "group out of range");
    HIP_TRY(hipSetDevice(M->device));
    HIP_TRY(hipDeviceSynchronize());  // earlier launches may still read the old table
    if (M->d_ky2) HIP_TRY(dev_free(M->d_ky2));
    if (M->d_grp) HIP_TRY(dev_free(M->d_grp));
    M->d_ky2 = nullptr;
    M->d_grp = nullptr;
    HIP_TRY(upload(&M->d_ky2,
                   std::vector<double>(ky2, ky2 + static_cast<size_t>(M->nky) * ngroup)));
    HIP_TRY(upload(&M->d_grp, std::vector<int>(group, group + M->nnz)));
    M->ngroup = ngroup;
    return DCX_OK;
  } catch (const std::exception& e) {
    return fail(DCX_ERR_INTERNAL, e.what());
  }
}

int dcx_assemble(dcx_lmap* M, const double* sigma, double* values, void* stream) {
  if (!M || !sigma || !values) return fail(DCX_ERR_ARG, "null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(M->device));
  hipLaunchKernelGGL(k_assemble, dim3(grid1(M->nnz, 256)), dim3(256), 0, st, M->nnz, M->nky,
                     M->d_ky2, M->d_grp, M->ngroup, M->d_ptr, M->d_cell, M->d_le, M->d_ln, M->d_bptr, M->d_bcell,
                     M->d_bcoef, M->nbc, sigma, values);
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

int dcx_lmap_transpose(dcx_lmap* M, const double* gv, const double* sigma, double* grad_m,
                       void* stream) {
  if (!M || !gv || !sigma || !grad_m) return fail(DCX_ERR_ARG, "null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(M->device));
  hipLaunchKernelGGL(k_reduce_ky, dim3(grid1(M->nnz, 256)), dim3(256), 0, st, M->nnz, M->nky,
                     M->d_ky2, M->d_grp, M->ngroup, gv, M->d_ge, M->d_gn);
  hipLaunchKernelGGL(k_lmap_t, dim3(grid1(static_cast<int64_t>(M->ncell) * kLmapLanes, 256)),
                     dim3(256), 0, st, M->ncell, M->nnz,
                     M->nky, M->d_tptr, M->d_te, M->d_tle, M->d_tln, M->d_tbptr,
                     M->d_tbq, M->d_bc_e, M->d_bcoef, M->nbc, M->d_ge, M->d_gn, gv, sigma,
                     grad_m);
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

int dcx_rows_mix(const void* g, const void* w, void* h, int nbatch, int n, int k, int nseg,
                 const int32_t* seg_rows, int flags, void* stream) {
  if (!g || !w || !h || !seg_rows) return fail(DCX_ERR_ARG, "null argument");
  if (nbatch < 1 || n < 1) return fail(DCX_ERR_ARG, "nbatch and n must be >= 1");
  if (k < 1 || k > kMixMaxK) return fail(DCX_ERR_ARG, "k must be 1..128");
  if (nseg < 1 || nseg > kMixMaxSeg) return fail(DCX_ERR_ARG, "nseg must be 1..256");
  if (nbatch > 65535) return fail(DCX_ERR_ARG, "nbatch must be <= 65535");
  MixSegs segs{};
  segs.nseg = nseg;
  int64_t tiles = 0;
  for (int l = 0; l <= nseg; ++l) {
    const int r = seg_rows[l];
    if (r < 0 || r > n || (l > 0 && r < seg_rows[l - 1]))
      return fail(DCX_ERR_ARG, "segment rows must ascend within 0..n");
    segs.row[l] = r;
    segs.tile[l] = static_cast<int>(tiles);
    if (l < nseg) tiles += (static_cast<int64_t>(seg_rows[l + 1]) - r + kMixRows - 1) / kMixRows;
  }
  if (tiles == 0) return DCX_OK;
  if (tiles > INT32_MAX) return fail(DCX_ERR_ARG, "too many rows");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (flags & DCX_F32_FIELDS)
    launch_rows_mix<float>(g, w, h, nbatch, n, k, segs, st);
  else
    launch_rows_mix<double>(g, w, h, nbatch, n, k, segs, st);
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

}  // extern "C"

namespace {
// checks shared by dcx_jacobian_rows and its allocation-free variant; the
// workspace they need, in bytes, for r data
int jacobian_rows_args(const dcx_lmap* M, const dcx_plan* P, const void* Y, int r,
                       const void* U, int nS, const void* src, const void* wq,
                       const void* sigma, const void* J, size_t* need) {
  if (!M || !P || !Y || !U || !src || !wq || !sigma || !J)
    return fail(DCX_ERR_ARG, "null argument");
  if (r < 1 || nS < 1) return fail(DCX_ERR_ARG, "r and nS must be >= 1");
  if (P->nbatch != M->nky || P->S.nnz != M->nnz)
    return fail(DCX_ERR_ARG, "map and plan do not describe the same systems");
  if (M->max_bc_per_entry > kMaxBcPerEntry)
    return fail(DCX_ERR_ARG, "more boundary terms per entry than k_jpair holds");
  if (M->ngroup != 1)
    return fail(DCX_ERR_ARG, "Jacobian rows of a map with wavenumber groups are not supported");
  *need = (2 * static_cast<size_t>(M->nnz) + std::max(M->nbc, 1)) * r * sizeof(double);
  return DCX_OK;
}

// the two kernels of a block of Jacobian rows, with the pairwise sums in the
// workspace ws (need bytes; every entry of it is written before it is read)
void jacobian_rows_launch(const dcx_lmap* M, const dcx_plan* P, const double* Y, int r,
                          const double* U, int nS, const int32_t* src, const double* wq,
                          const double* sigma, double* J, int flags, double* ws,
                          hipStream_t st) {
  double* ge = ws;
  double* gn = ge + static_cast<size_t>(M->nnz) * r;
  double* gb = gn + static_cast<size_t>(M->nnz) * r;
  const int n = P->S.n;
  hipLaunchKernelGGL(k_jpair, dim3(grid1(n, 4)), dim3(256), 0, st, n, P->d_crow, P->d_col,
                     (flags & DCX_PERMUTED) ? static_cast<const int*>(P->d_iperm) : nullptr, Y,
                     r, U, nS, src, wq, M->nky, M->d_ky2, M->d_bptr, M->d_bcoef, M->nbc, ge, gn,
                     gb);
  hipLaunchKernelGGL(k_jmap, dim3(grid1(M->ncell, 4)), dim3(256), 0, st, M->ncell, r, M->d_tptr,
                     M->d_tent, M->d_te, M->d_le, M->d_ln, M->d_tbptr, M->d_tbq, ge, gn, gb,
                     sigma, J);
}
}  // namespace

// The allocation-free variant's workspace: one file-scope device array per
// GPU, held until the process ends.  In host code `scratch` names a host
// stand-in, not a device address: its device address is taken once per
// device with hipGetSymbolAddress (scratch_ptr) and only that address is
// passed to kernels.
constexpr size_t kScratchBytes = static_cast<size_t>(64) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};

// the device address of `scratch` on the current device `dev`
int scratch_ptr(int dev, double** out) {
  if (dev < 0 || dev >= kMaxScratchDevices) return fail(DCX_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(DCX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                   hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  *out = static_cast<double*>(g_scratch_dev[dev]);
  return DCX_OK;
}
}  // namespace

extern "C" {

int dcx_jacobian_rows(dcx_lmap* M, dcx_plan* P, const double* Y, int r, const double* U, int nS,
                      const int32_t* src, const double* wq, const double* sigma, double* J,
                      int flags, void* stream) {
  size_t need = 0;
  if (int rc = jacobian_rows_args(M, P, Y, r, U, nS, src, wq, sigma, J, &need)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(M->device));
  if (M->jws_bytes < need) {
    HIP_TRY(hipStreamSynchronize(st));
    if (M->d_jws) HIP_TRY(dev_free(M->d_jws));
    M->d_jws = nullptr;
    M->jws_bytes = 0;
    HIP_TRY(dev_malloc(&M->d_jws, need));
    M->jws_bytes = need;
  }
  jacobian_rows_launch(M, P, Y, r, U, nS, src, wq, sigma, J, flags, M->d_jws, st);
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

int dcx_jacobian_rows_noalloc(dcx_lmap* M, dcx_plan* P, const double* Y, int r,
                              const double* U, int nS, const int32_t* src, const double* wq,
                              const double* sigma, double* J, int flags, void* stream) {
  size_t need = 0;
  if (int rc = jacobian_rows_args(M, P, Y, r, U, nS, src, wq, sigma, J, &need)) return rc;
  if (need > kScratchBytes) {
    const size_t per_row = need / static_cast<size_t>(r);
    return fail(DCX_ERR_ARG, std::to_string(r) + " rows need " + std::to_string(need >> 20) +
                                 " MiB of workspace; the allocation-free variant holds " +
                                 std::to_string(kScratchBytes >> 20) + " MiB: at most " +
                                 std::to_string(kScratchBytes / per_row) + " rows per call");
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(M->device));
  double* ws = nullptr;
  if (int rc = scratch_ptr(M->device, &ws)) return rc;
  jacobian_rows_launch(M, P, Y, r, U, nS, src, wq, sigma, J, flags, ws, st);
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

}  // extern "C"

// ===========================================================================
// survey
// ===========================================================================
struct dcx_survey {
  int nD = 0, nS = 0, n = 0, device = 0;
  int *d_ptr = nullptr, *d_rows = nullptr, *d_rows_p = nullptr, *d_sidx = nullptr;
  double* d_val = nullptr;
  // adjoint targets, per numbering (0 original, 1 elimination)
  int ntgt = 0;
  int *d_tptr = nullptr, *d_tdat = nullptr, *d_trow = nullptr, *d_trow_p = nullptr,
      *d_tsrc = nullptr;
  double* d_tcoef = nullptr;
  std::vector<int> t_row, t_src;
  ~dcx_survey() {
    void* ptrs[] = {d_ptr, d_rows, d_rows_p, d_sidx, d_val, d_tptr, d_tdat, d_tcoef,
                    d_trow, d_trow_p, d_tsrc};
    for (void* p : ptrs)
      if (p) (void)dev_free(p);
  }
};

extern "C" {

int dcx_survey_create(dcx_plan* P, int nD, int nS, const int32_t* pd_ptr, const int32_t* pd_col,
                      const double* pd_val, const int32_t* sidx, dcx_survey** out) {
  try {
    if (!P || !out || !pd_ptr || !pd_col || !pd_val || !sidx) return fail(DCX_ERR_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<dcx_survey> V(new dcx_survey());
    V->nD = nD;
    V->nS = nS;
    V->n = P->S.n;
    HIP_TRY(hipGetDevice(&V->device));
    const int np = pd_ptr[nD];
    std::vector<int> rows(pd_col, pd_col + np), rows_p(np);
    for (int p = 0; p < np; ++p) {
      if (rows[p] < 0 || rows[p] >= V->n) return fail(DCX_ERR_ARG, "Pd column out of range");
      rows_p[p] = P->S.iperm[rows[p]];
    }
    for (int i = 0; i < nD; ++i)
      if (sidx[i] < 0 || sidx[i] >= nS) return fail(DCX_ERR_ARG, "source index out of range");
    // targets (row, source) with their contributing data, sorted by (row, src)
    std::vector<std::tuple<int, int, int, double>> trip;  // row, src, datum, coef
    for (int i = 0; i < nD; ++i)
      for (int p = pd_ptr[i]; p < pd_ptr[i + 1]; ++p) trip.emplace_back(rows[p], sidx[i], i, pd_val[p]);
    std::sort(trip.begin(), trip.end(), [](const auto& a, const auto& b) {
      if (std::get<0>(a) != std::get<0>(b)) return std::get<0>(a) < std::get<0>(b);
      if (std::get<1>(a) != std::get<1>(b)) return std::get<1>(a) < std::get<1>(b);
      return std::get<2>(a) < std::get<2>(b);
    });
    std::vector<int> tptr, tdat;
    std::vector<double> tcoef;
    for (size_t q = 0; q < trip.size(); ++q) {
      const bool fresh = q == 0 || std::get<0>(trip[q]) != std::get<0>(trip[q - 1]) ||
                         std::get<1>(trip[q]) != std::get<1>(trip[q - 1]);
      if (fresh) {
        tptr.push_back(static_cast<int>(q));
        V->t_row.push_back(std::get<0>(trip[q]));
        V->t_src.push_back(std::get<1>(trip[q]));
      }
      tdat.push_back(std::get<2>(trip[q]));
      tcoef.push_back(std::get<3>(trip[q]));
    }
    tptr.push_back(static_cast<int>(trip.size()));
    V->ntgt = static_cast<int>(V->t_row.size());
    if (static_cast<int>(tptr.size()) != V->ntgt + 1) return fail(DCX_ERR_INTERNAL, "target table");
    HIP_TRY(upload(&V->d_ptr, std::vector<int>(pd_ptr, pd_ptr + nD + 1)));
    HIP_TRY(upload(&V->d_rows, rows));
    HIP_TRY(upload(&V->d_rows_p, rows_p));
    HIP_TRY(upload(&V->d_val, std::vector<double>(pd_val, pd_val + np)));
    HIP_TRY(upload(&V->d_sidx, std::vector<int>(sidx, sidx + nD)));
    HIP_TRY(upload(&V->d_tptr, tptr));
    HIP_TRY(upload(&V->d_tdat, tdat));
    HIP_TRY(upload(&V->d_tcoef, tcoef));
    std::vector<int> trow_p(V->ntgt);
    for (int t = 0; t < V->ntgt; ++t) trow_p[t] = P->S.iperm[V->t_row[t]];
    HIP_TRY(upload(&V->d_trow, V->t_row));
    HIP_TRY(upload(&V->d_trow_p, trow_p));
    HIP_TRY(upload(&V->d_tsrc, V->t_src));
    *out = V.release();
    return DCX_OK;
  } catch (const std::exception& e) {
    return fail(DCX_ERR_INTERNAL, e.what());
  }
}

int dcx_survey_destroy(dcx_survey* v) {
  delete v;
  return DCX_OK;
}

int dcx_survey_dpred(dcx_survey* V, const void* U, int k, int nky, const double* w, int s0,
                     int s1, double* d, int flags, void* stream) {
  if (!V || !U || !w || !d) return fail(DCX_ERR_ARG, "null argument");
  if (s0 < 0 || s1 > V->nS || s0 > s1 || s1 - s0 > k)
    return fail(DCX_ERR_ARG, "source window does not fit the block width");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(V->device));
  const int64_t bs = static_cast<int64_t>(V->n) * k;
  const int* rows = (flags & DCX_PERMUTED) ? V->d_rows_p : V->d_rows;
  const int blocks = grid1(static_cast<int64_t>(V->nD) * kDpredLanes, 256);
  if (flags & DCX_F32_FIELDS)
    hipLaunchKernelGGL(k_dpred<float>, dim3(blocks), dim3(256), 0, st, V->nD,
                       V->d_ptr, rows, V->d_val, V->d_sidx, static_cast<const float*>(U), bs, k,
                       nky, w, s0, s1, d);
  else
    hipLaunchKernelGGL(k_dpred<double>, dim3(blocks), dim3(256), 0, st, V->nD,
                       V->d_ptr, rows, V->d_val, V->d_sidx, static_cast<const double*>(U), bs,
                       k, nky, w, s0, s1, d);
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

int dcx_survey_adjoint_rhs(dcx_survey* V, const double* v, double* G, int k, int s0, int s1,
                           int flags, void* stream) {
  if (!V || !v || !G) return fail(DCX_ERR_ARG, "null argument");
  if (s0 < 0 || s1 > V->nS || s0 > s1 || s1 - s0 > k)
    return fail(DCX_ERR_ARG, "source window does not fit the block width");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipSetDevice(V->device));
  const int64_t tot = static_cast<int64_t>(V->n) * k;
  // DCX_SPARSE_RHS: G is already zero outside the target entries
  if (!(flags & DCX_SPARSE_RHS))
    hipLaunchKernelGGL(k_fill, dim3(grid_stride(tot)), dim3(256), 0, st, G, tot, 0.0);
  if (V->ntgt)
    hipLaunchKernelGGL(k_adjoint_rhs, dim3(grid1(V->ntgt, 256)), dim3(256), 0, st, V->ntgt,
                       V->d_tptr, (flags & DCX_PERMUTED) ? V->d_trow_p : V->d_trow, V->d_tsrc,
                       V->d_tdat, V->d_tcoef, v, k, s0, s1, G);
  HIP_TRY(hipGetLastError());
  return DCX_OK;
}

}  // extern "C"
// ===========================================================================
// host-only symbolic analysis
// ===========================================================================
struct dcx_symbolic {
  dcx::Symbolic S;
};

extern "C" {

int dcx_symbolic_analyse(int n, const int32_t* crow, const int32_t* col, const double* coords,
                         int dim, int leaf_max, int fuse_rows, dcx_symbolic** out) {
  try {
    if (!out || !crow || !col) return fail(DCX_ERR_ARG, "null argument");
    *out = nullptr;
    if (coords && (dim < 1 || dim > 3)) return fail(DCX_ERR_ARG, "dim must be 1..3");
    std::unique_ptr<dcx_symbolic> Y(new dcx_symbolic());
    std::string err = dcx::analyse(n, crow, col, coords, dim, leaf_max, fuse_rows, Y->S);
    if (!err.empty()) return fail(DCX_ERR_ARG, err);
    *out = Y.release();
    return DCX_OK;
  } catch (const std::exception& e) {
    return fail(DCX_ERR_INTERNAL, e.what());
  }
}

int dcx_symbolic_view_get(const dcx_symbolic* Y, dcx_symbolic_view* o) {
  if (!Y || !o) return fail(DCX_ERR_ARG, "null argument");
  const auto& S = Y->S;
  o->n = S.n;
  o->nnz = S.nnz;
  o->nsup = S.nsup;
  o->ndepth = S.ndepth;
  o->max_front = S.max_front;
  o->max_ns = S.max_ns;
  o->perm = S.perm.data();
  o->first = S.first.data();
  o->ns = S.ns.data();
  o->rs = S.rs.data();
  o->parent = S.parent.data();
  o->depth = S.depth.data();
  o->rptr = S.rptr.data();
  o->ridx = S.ridx.data();
  o->cptr = S.cptr.data();
  o->cidx = S.cidx.data();
  o->relpos = S.relpos.data();
  o->aptr = S.aptr.data();
  o->a_e = S.a_e.data();
  o->a_row = S.a_row.data();
  o->a_col = S.a_col.data();
  o->lvl_ptr = S.lvl_ptr.data();
  o->lvl_sup = S.lvl_sup.data();
  o->nnz_L = S.nnz_L;
  o->factor_flops = S.factor_flops;
  o->solve_flops_per_rhs = S.solve_flops_per_rhs;
  o->nfused = S.nfused;
  o->nfused_lv = static_cast<int>(S.f_lv.size()) - 1;
  o->nfused_mem = static_cast<int>(S.f_mem.size());
  o->nfused_grp = static_cast<int>(S.f_gb.size()) - 1;
  o->f_max_lds_rows = S.f_max_lds_rows;
  o->f_root = S.f_root.data();
  o->f_lo = S.f_lo.data();
  o->f_hi = S.f_hi.data();
  o->f_lv_begin = S.f_lv_begin.data();
  o->f_lv_end = S.f_lv_end.data();
  o->f_tb_rows = S.f_tb_rows.data();
  o->f_lv = S.f_lv.data();
  o->f_mem = S.f_mem.data();
  o->f_gptr = S.f_gptr.data();
  o->f_gb = S.f_gb.data();
  o->toff = S.toff.data();
  o->lridx = S.lridx.data();
  o->ulvl_ptr = S.ulvl_ptr.data();
  o->ulvl_sup = S.ulvl_sup.data();
  return DCX_OK;
}

int dcx_symbolic_destroy(dcx_symbolic* Y) {
  delete Y;
  return DCX_OK;
}

}  // extern "C"
