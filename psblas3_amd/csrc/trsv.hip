// Sparse triangular solves on the device:
//  - BJAC / ILU(0) preconditioner application:
//    y = alpha * U^-1 (d .* L^-1 x) + beta * y with L unit lower, U unit upper
//    (psb_d_bjac_apply_vect, prec/impl/psb_d_bjacprec_impl.f90:166-173:
//    psb_spsm with scale 'L' and the inverse pivots, then psb_spsm on U);
//  - the CSR matrix's own solve y = alpha * op(T)^-1 x + beta * y, T lower or
//    upper, unit or not, op 'N', 'T' or 'C' (psb_d_cuda_csrg_inner_vect_sv,
//    cuda/impl/psb_d_cuda_csrg_inner_vect_sv.F90; the CPU algorithm it must
//    equal is psb_d_csr_cssv, base/serial/impl/psb_d_csr_impl.F90:782-1039),
//    FP64 and FP32.
//
// Each triangle is solved by levels: the host analysis gives every row the
// length of its longest dependency chain.  A row's sum runs in its CSR order,
// y(i) = x(i) - sum_j val(j) y(ja(j)) with separate multiply and add, as the
// reference's inner_csrsv (base/serial/impl/psb_d_csr_impl.F90:890-941), so
// the result equals the sequential CPU solve bit for bit.
//
// A non-unit row divides by its diagonal, (x(i) - sum) / a(i,i), as the CPU
// does.  The transposed solve (op 'T' / 'C') is solved as the triangle of the
// transpose, whose row c lists the entries (i, c) of T in the order the CPU
// applies them (rows in solve order, CSR order inside a row); its rows are
// computed as the CPU updates them, by running subtraction
// y(c) = ((x(c) - v1 y(i1)) - v2 y(i2)) ..., then the division.
//
// The solve runs in level order: each triangle is renumbered so that level l
// holds positions lev_ptr[l] .. lev_ptr[l+1]-1 (rows ascending inside a
// level), x is gathered into that order by one full-grid kernel and the
// solution scattered back by another (the last one also applies alpha and
// beta), so every level reads and writes a contiguous range.
//
// A launch per level costs a few microseconds of dispatch, more than the work
// of a small level, so runs of consecutive small levels (each at most 2048
// rows of at most 4 entries, or 1024 rows of at most 8) are solved by ONE
// workgroup of 1024 threads that walks the levels with an LDS barrier
// between them.  Everything a run reads of the solution lives in LDS
// (128 KB): its own rows (first holding x, then overwritten by the solution)
// and the "halo" of positions solved before the run that it depends on,
// loaded once at the start.  A level therefore waits only for the LDS traffic
// of the level before it; the run's solution goes to global memory once, at
// its end.  A run's entries (16-bit LDS slot, value) are laid out per level,
// column-major and padded to the run's width, the smallest of 2, 3, 4, 8 that
// fits (entry k of the level's t-th row at off + k*count + t; a padding slot
// adds value 0 times a zero LDS slot, which leaves the sum's bits unchanged),
// and are loaded two levels ahead.  Larger or wider levels get a launch of
// their own (one thread per row, level-ordered CSR).  The permutations, both
// triangles and the pivot scaling are captured once into a graph that works
// on internal buffers.
#pragma clang fp contract(off)

#include <string.h>

#include <algorithm>
#include <vector>

#include "psb_hip_internal.hpp"

namespace psbhip {

// A run of levels solved by one workgroup: levels [lev0, lev1), entry width
// W, nhalo positions from before the run at halo[hoff ..].
struct TriSegment {
  int lev0, lev1, width, hoff, nhalo;
};

constexpr int kSegThreads = 1024;
constexpr int kSegRows = 16384;          // LDS slots of a run: its rows, its halo, two spare
constexpr int kSinkSlot = kSegRows - 2;  // written by lanes beyond their level, never read
constexpr int kZeroSlot = kSegRows - 1;  // holds 0, read by padding entries
constexpr int kSegLevels = 1024;         // levels of a run (their descriptors sit in LDS)

template <typename T>
struct TriLevels {
  int n = 0, nnz = 0, nlev = 0;
  int* perm = nullptr;          // level position -> row
  int* irp = nullptr;           // level-ordered CSR, 0-based; columns are level positions
  int* ja = nullptr;
  T* val = nullptr;
  std::vector<int> lev_ptr;     // host: level l = positions lev_ptr[l] .. lev_ptr[l+1]-1
  // per-level column-major copies for the workgroup-solved runs
  int4* lev_info = nullptr;     // {count, width, entry offset, first position} per level
  uint16_t* eja = nullptr;      // LDS slot of each dependency (kSegRows <= 65536)
  T* eval = nullptr;
  int* halo = nullptr;          // positions before each run that the run reads
  std::vector<TriSegment> segs; // host: the runs, in level order
  std::vector<int> seg_of;      // host: segment index of a level, -1 if launched alone
};

// 0: runs of small levels by one workgroup (default); 1: one launch per level.
static int g_trsv_variant = 0;
// Register stages of a run's entry pipeline: entries are loaded two levels
// ahead (five stages measured 5% slower at L=40 and 64).
constexpr int kSegStages = 3;

// Rows per thread and entry width a workgroup run supports.
__host__ __device__ constexpr int seg_rpt(int W) { return W <= 4 ? 2 : 1; }
constexpr int kSegWidths[] = {2, 3, 4, 8};   // the entry widths a run can be built with

struct IluDev {
  int n = 0;
  TriLevels<double> L, U;
  double* d = nullptr;          // inverse pivots
  double* xa = nullptr;         // right-hand side of the triangle being solved (level order)
  double* ya = nullptr;         // its solution (level order)
  double* t1 = nullptr;         // L^-1 x (row order)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// A CSR matrix's triangle prepared for op(T)^-1 x: the strict part (or, for
// op 'T' / 'C', the strict part of the transpose) in level order, and the
// diagonal in the same order when the triangle is not unit.
template <typename T>
struct TriSolve {
  int n = 0, key = -1;          // key: lower | unit << 1 | transposed << 2
  bool run = false;             // transposed: rows by running subtraction
  TriLevels<T> tl;
  T* dv = nullptr;              // diagonal in level order (non-unit), else null
  T* xa = nullptr;              // right-hand side (level order)
  T* ya = nullptr;              // solution (level order)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// dst[p] = src[perm[p]], times d[perm[p]] when d is given (the pivot scaling
// y(i) * d(i) of the reference's 'L' scale)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_perm_gather(int n, const int* __restrict__ perm,
                                                        const T* __restrict__ src,
                                                        const T* __restrict__ d,
                                                        T* __restrict__ dst) {
  for (int p = blockIdx.x * kBlock + threadIdx.x; p < n; p += gridDim.x * kBlock) {
    const int i = perm[p];
    dst[p] = d ? src[i] * d[i] : src[i];
  }
}

// dst[perm[p]] = src[p]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_perm_scatter(int n, const int* __restrict__ perm,
                                                         const T* __restrict__ src,
                                                         T* __restrict__ dst) {
  for (int p = blockIdx.x * kBlock + threadIdx.x; p < n; p += gridDim.x * kBlock)
    dst[perm[p]] = src[p];
}

// y[perm[p]] = alpha * src[p] + beta * y[perm[p]] with the special cases of
// the library's axpby (alpha or beta 0 or 1), separately rounded; beta == 0
// and alpha == -1 negate, as the CPU solve's scaling does
template <typename T>
__global__ __launch_bounds__(kBlock) void k_perm_scatter_axpby(int n, const int* __restrict__ perm,
                                                               const T* __restrict__ src,
                                                               T alpha, T beta, T* y) {
  for (int p = blockIdx.x * kBlock + threadIdx.x; p < n; p += gridDim.x * kBlock) {
    const int i = perm[p];
    T r;
    if (alpha == T(0)) r = (beta == T(0)) ? T(0) : (beta == T(1) ? y[i] : beta * y[i]);
    else if (beta == T(0)) r = (alpha == T(1)) ? src[p] : alpha * src[p];
    else r = alpha * src[p] + beta * y[i];
    y[i] = r;
  }
}

// One level: positions p0 .. p0+count-1, one thread each.  RUN: running
// subtraction from x (the transposed solve); DIV: divide by the diagonal.
template <typename T, bool DIV, bool RUN>
__global__ __launch_bounds__(kBlock) void k_trsv_level(int p0, int count,
                                                       const int* __restrict__ irp,
                                                       const int* __restrict__ ja,
                                                       const T* __restrict__ val,
                                                       const T* __restrict__ dv,
                                                       const T* __restrict__ x, T* y) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int p = p0 + t;
  T r;
  if (RUN) {
    r = x[p];
    for (int e = irp[p]; e < irp[p + 1]; e++) r = r - val[e] * y[ja[e]];
  } else {
    T acc = T(0);
    for (int e = irp[p]; e < irp[p + 1]; e++) acc = acc + val[e] * y[ja[e]];
    r = x[p] - acc;
  }
  y[p] = DIV ? r / dv[p] : r;
}

// Levels [lev0, lev1) by one workgroup.  Row r of thread t in a level is
// level row t + r*kSegThreads; lanes beyond the level work on the level's last
// row and write the sink slot.  Loads are unconditional and neither the sum
// nor the stores are behind a branch, so a level's LDS reads are in flight
// together.  The sum runs over the row's entries in CSR order, then the
// padding adds +0 to it (never -0: the sum starts at +0), which leaves it
// unchanged; in the running form the padding subtracts +0, which leaves
// every value, -0 included, unchanged.
template <typename T, int W, bool DIV>
struct SegStage {
  static constexpr int R = seg_rpt(W);
  int dep[R][W];
  T v[R][W];
  T dv[DIV ? R : 1];
};

template <typename T, int W, bool DIV>
__device__ __forceinline__ void seg_fetch(SegStage<T, W, DIV>& S, const int4 L,
                                          const uint16_t* __restrict__ eja,
                                          const T* __restrict__ eval,
                                          const T* __restrict__ dv) {
#pragma unroll
  for (int r = 0; r < SegStage<T, W, DIV>::R; r++) {
    const int tt = min((int)threadIdx.x + r * kSegThreads, L.x - 1);
#pragma unroll
    for (int k = 0; k < W; k++) {
      const int64_t e = (int64_t)L.z + (int64_t)k * L.x + tt;
      S.dep[r][k] = eja[e];
      S.v[r][k] = eval[e];
    }
    if (DIV) S.dv[r] = dv[L.w + tt];
  }
}

__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T, int W, bool DIV, bool RUN>
__device__ __forceinline__ void seg_solve(const SegStage<T, W, DIV>& S, const int4 L, int base,
                                          T* ys) {
  constexpr int R = SegStage<T, W, DIV>::R;
  lds_barrier();   // the LDS slots of the levels before this one are written
  T yi[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int tt = min((int)threadIdx.x + r * kSegThreads, L.x - 1);
    T yk[W];
#pragma unroll
    for (int k = 0; k < W; k++) yk[k] = ys[S.dep[r][k]];
    const T xi = ys[L.w - base + tt];
    T v;
    if (RUN) {
      v = xi;
#pragma unroll
      for (int k = 0; k < W; k++) v = v - S.v[r][k] * yk[k];
    } else {
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < W; k++) acc = acc + S.v[r][k] * yk[k];
      v = xi - acc;
    }
    yi[r] = DIV ? v / S.dv[r] : v;
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int t = threadIdx.x + r * kSegThreads;
    ys[t < L.x ? L.w - base + t : kSinkSlot] = yi[r];
  }
}

// x and y in level order; the run's positions are base .. base+nrun-1; dv
// the diagonal in level order when DIV.
template <typename T, int W, int NS, bool DIV, bool RUN>
__global__ __launch_bounds__(kSegThreads) void k_trsv_seg(int lev0, int lev1, int base, int nrun,
                                                          const int* __restrict__ halo, int nhalo,
                                                          const int4* __restrict__ info,
                                                          const uint16_t* __restrict__ eja,
                                                          const T* __restrict__ eval,
                                                          const T* __restrict__ dv,
                                                          const T* __restrict__ x, T* y) {
  __shared__ T ys[kSegRows];
  __shared__ int4 linfo[kSegLevels];   // the run's level descriptors: no scalar load per level
  for (int q = threadIdx.x; q < lev1 - lev0; q += kSegThreads) linfo[q] = info[lev0 + q];
  for (int p = threadIdx.x; p < nrun; p += kSegThreads) ys[p] = x[base + p];
  for (int h = threadIdx.x; h < nhalo; h += kSegThreads) ys[nrun + h] = y[halo[h]];
  if (threadIdx.x == 0) ys[kZeroSlot] = T(0);
  lds_barrier();
  // Entries NS-1 levels ahead: NS register stages rotate, the one solved
  // holds level l.  The loop runs whole rounds of NS (the stage indices are
  // compile-time constants once unrolled) and the fetches are unconditional
  // (past the run's end they repeat its last level), so every path leaves the
  // same loads in flight and the compiler's waits stay exact.
  SegStage<T, W, DIV> st[NS];
  int4 Ls[NS];
  auto level = [&](int l) { return linfo[min(l, lev1 - 1) - lev0]; };
#pragma unroll
  for (int q = 0; q < NS - 1; q++) {
    Ls[q] = level(lev0 + q);
    seg_fetch<T, W, DIV>(st[q], Ls[q], eja, eval, dv);
  }
  for (int l = lev0; l < lev1; l += NS) {
#pragma unroll
    for (int j = 0; j < NS; j++) {
      const int ahead = (j + NS - 1) % NS;
      Ls[ahead] = level(l + j + NS - 1);
      seg_fetch<T, W, DIV>(st[ahead], Ls[ahead], eja, eval, dv);
      if (l + j < lev1) seg_solve<T, W, DIV, RUN>(st[j], Ls[j], base, ys);   // uniform
    }
  }
  lds_barrier();
  for (int p = threadIdx.x; p < nrun; p += kSegThreads) y[base + p] = ys[p];
}

// Level of every row (longest chain of dependencies through the triangle),
// rows grouped by level in ascending row order within a level.
static void analyse(int n, bool lower, const int* irp, const int* ja, std::vector<int>& rows,
                    std::vector<int>& lev_ptr, std::vector<int>& lev) {
  lev.assign(n, 0);
  int nlev = 0;
  for (int t = 0; t < n; t++) {
    const int i = lower ? t : n - 1 - t;
    int l = 0;
    for (int p = irp[i] - 1; p < irp[i + 1] - 1; p++) l = std::max(l, lev[ja[p] - 1] + 1);
    lev[i] = l;
    nlev = std::max(nlev, l + 1);
  }
  lev_ptr.assign(nlev + 1, 0);
  for (int i = 0; i < n; i++) lev_ptr[lev[i] + 1]++;
  for (int l = 0; l < nlev; l++) lev_ptr[l + 1] += lev_ptr[l];
  rows.assign(n, 0);
  std::vector<int> fill(lev_ptr.begin(), lev_ptr.end() - 1);
  for (int i = 0; i < n; i++) rows[fill[lev[i]]++] = i;
}

// Runs of consecutive levels a workgroup can take: every level of a run has
// at most seg_rpt(W) * kSegThreads rows of at most W entries, W one of
// kSegWidths; the run's rows plus its halo fill at most kSegRows - 2 LDS
// slots, and it has at most kSegLevels levels.
template <typename V>
static void plan_segments(TriLevels<V>& T, const int* irp, const int* ja,
                          const std::vector<int>& rows, const std::vector<int>& lev) {
  T.segs.clear();
  T.seg_of.assign(T.nlev, -1);
  std::vector<int> width(T.nlev, 0);
  for (int l = 0; l < T.nlev; l++)
    for (int p = T.lev_ptr[l]; p < T.lev_ptr[l + 1]; p++)
      width[l] = std::max(width[l], irp[rows[p] + 1] - irp[rows[p]]);
  auto fits = [&](int l, int W) {
    return width[l] <= W && T.lev_ptr[l + 1] - T.lev_ptr[l] <= seg_rpt(W) * kSegThreads;
  };
  auto first_fit = [&](int q, int from) {   // smallest width >= from that level q fits
    for (int W : kSegWidths)
      if (W >= from && fits(q, W)) return W;
    return 0;
  };
  std::vector<int> stamp(T.n > 0 ? T.n : 1, -1);
  int cand = 0;
  int l = 0;
  while (l < T.nlev) {
    int W = first_fit(l, 1);
    if (!W) {
      l++;
      continue;
    }
    cand++;
    int nh = 0;
    auto add_halo = [&](int q) {   // the rows before the run that level q reads
      for (int p = T.lev_ptr[q]; p < T.lev_ptr[q + 1]; p++)
        for (int e = irp[rows[p]] - 1; e < irp[rows[p] + 1] - 1; e++) {
          const int j = ja[e] - 1;
          if (lev[j] < l && stamp[j] != cand) {
            stamp[j] = cand;
            nh++;
          }
        }
    };
    add_halo(l);
    int e = l + 1;
    for (; e < T.nlev; e++) {
      bool ok = fits(e, W);
      for (int W2 = first_fit(e, W + 1); !ok && W2; W2 = first_fit(e, W2 + 1)) {
        bool all = true;   // every level so far must also fit the wider limits
        for (int q = l; q < e && all; q++) all = fits(q, W2);
        if (all) {
          W = W2;
          ok = true;
        }
      }
      if (!ok) break;
      add_halo(e);
      if (T.lev_ptr[e + 1] - T.lev_ptr[l] + nh > kSegRows - 2 || e + 1 - l > kSegLevels) break;
    }
    if (e - l >= 2) {   // a lone level is as cheap as its own launch
      for (int q = l; q < e; q++) T.seg_of[q] = (int)T.segs.size();
      T.segs.push_back({l, e, W, 0, 0});
    }
    l = e;
  }
}

template <typename V>
static int upload_segments(TriLevels<V>& T, const int* irp, const int* ja, const V* val,
                           const std::vector<int>& rows, const std::vector<int>& pos,
                           const std::vector<int>& lev) {
  if (T.segs.empty()) return PSB_HIP_SUCCESS;
  std::vector<int4> info(T.nlev, make_int4(0, 0, 0, 0));
  int64_t ent = 0;
  for (const TriSegment& S : T.segs)
    for (int l = S.lev0; l < S.lev1; l++) {
      const int c = T.lev_ptr[l + 1] - T.lev_ptr[l];
      info[l] = make_int4(c, S.width, (int)std::min<int64_t>(ent, INT32_MAX), T.lev_ptr[l]);
      ent += (int64_t)c * S.width;
    }
  if (ent > INT32_MAX) return set_error(PSB_HIP_UNSUPPORTED, "trsv", "triangle too large");
  std::vector<uint16_t> eja(ent, kZeroSlot);   // padding: the zero slot, value 0
  std::vector<V> eval(ent, V(0));
  std::vector<int> halo, slot(T.n > 0 ? T.n : 1, -1);
  for (TriSegment& S : T.segs) {
    const int base = T.lev_ptr[S.lev0], nrun = T.lev_ptr[S.lev1] - base;
    S.hoff = (int)halo.size();
    for (int l = S.lev0; l < S.lev1; l++) {
      const int4 L = info[l];
      for (int t = 0; t < L.x; t++) {
        const int i = rows[L.w + t];
        const int b = irp[i] - 1, c = irp[i + 1] - 1 - b;
        for (int k = 0; k < c; k++) {
          const int j = ja[b + k] - 1;
          int sl;
          if (lev[j] >= S.lev0) {
            sl = pos[j] - base;   // solved inside the run
          } else {
            if (slot[j] < 0) {     // first use of this position of the halo
              slot[j] = nrun + (int)halo.size() - S.hoff;
              halo.push_back(pos[j]);
            }
            sl = slot[j];
          }
          eja[(size_t)L.z + (size_t)k * L.x + t] = (uint16_t)sl;
          eval[(size_t)L.z + (size_t)k * L.x + t] = val[b + k];
        }
      }
    }
    S.nhalo = (int)halo.size() - S.hoff;
    for (int h = S.hoff; h < (int)halo.size(); h++) slot[rows[halo[h]]] = -1;
    if (nrun + S.nhalo > kSegRows - 2)
      return set_error(PSB_HIP_RUNTIME, "trsv", "run larger than its LDS slots");
  }
  PSB_HIP_CHECK(hipMalloc(&T.lev_info, sizeof(int4) * (size_t)T.nlev));
  PSB_HIP_CHECK(hipMalloc(&T.eja, sizeof(uint16_t) * eja.size()));
  PSB_HIP_CHECK(hipMalloc(&T.eval, sizeof(V) * eval.size()));
  PSB_HIP_CHECK(hipMalloc(&T.halo, sizeof(int) * std::max<size_t>(halo.size(), 1)));
  hipStream_t s = ctx().stream;
  PSB_HIP_CHECK(hipMemcpyAsync(T.lev_info, info.data(), sizeof(int4) * info.size(),
                               hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(T.eja, eja.data(), sizeof(uint16_t) * eja.size(),
                               hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(T.eval, eval.data(), sizeof(V) * eval.size(),
                               hipMemcpyHostToDevice, s));
  if (!halo.empty())
    PSB_HIP_CHECK(hipMemcpyAsync(T.halo, halo.data(), sizeof(int) * halo.size(),
                                 hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));   // the host vectors go out of scope
  return PSB_HIP_SUCCESS;
}

template <typename V>
static int upload_tri(TriLevels<V>& T, int n, bool lower, const int* irp, const int* ja,
                      const V* val, bool segments) {
  T.n = n;
  T.nnz = irp[n] - 1;
  for (int i = 0; i < n; i++)
    for (int p = irp[i] - 1; p < irp[i + 1] - 1; p++) {
      const int c = ja[p];
      if (c < 1 || c > n || (lower ? c > i : c < i + 2))
        return set_error(PSB_HIP_INVALID, "trsv", "entry outside its triangle");
    }
  std::vector<int> rows, lev;
  analyse(n, lower, irp, ja, rows, T.lev_ptr, lev);
  T.nlev = (int)T.lev_ptr.size() - 1;
  std::vector<int> pos(n > 0 ? n : 1, 0);
  for (int p = 0; p < n; p++) pos[rows[p]] = p;
  // the triangle in level order: row rows[p] becomes position p, its entries
  // keep their CSR order and point at the positions of their columns
  std::vector<int> pirp(n + 1, 0), pja(T.nnz > 0 ? T.nnz : 1, 0);
  std::vector<V> pval(T.nnz > 0 ? T.nnz : 1, V(0));
  for (int p = 0; p < n; p++) {
    const int i = rows[p], b = irp[i] - 1, c = irp[i + 1] - 1 - b;
    for (int k = 0; k < c; k++) {
      pja[pirp[p] + k] = pos[ja[b + k] - 1];
      pval[pirp[p] + k] = val[b + k];
    }
    pirp[p + 1] = pirp[p] + c;
  }
  PSB_HIP_CHECK(hipMalloc(&T.irp, sizeof(int) * pirp.size()));
  PSB_HIP_CHECK(hipMalloc(&T.ja, sizeof(int) * pja.size()));
  PSB_HIP_CHECK(hipMalloc(&T.val, sizeof(V) * pval.size()));
  PSB_HIP_CHECK(hipMalloc(&T.perm, sizeof(int) * (size_t)(n > 0 ? n : 1)));
  hipStream_t s = ctx().stream;
  PSB_HIP_CHECK(hipMemcpyAsync(T.irp, pirp.data(), sizeof(int) * pirp.size(),
                               hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(T.ja, pja.data(), sizeof(int) * pja.size(), hipMemcpyHostToDevice,
                               s));
  PSB_HIP_CHECK(hipMemcpyAsync(T.val, pval.data(), sizeof(V) * pval.size(),
                               hipMemcpyHostToDevice, s));
  if (n > 0)
    PSB_HIP_CHECK(hipMemcpyAsync(T.perm, rows.data(), sizeof(int) * (size_t)n,
                                 hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));   // the host vectors go out of scope
  if (segments) {
    plan_segments(T, irp, ja, rows, lev);
  } else {
    T.segs.clear();
    T.seg_of.assign(T.nlev, -1);   // every level launched alone
  }
  return upload_segments(T, irp, ja, val, rows, pos, lev);
}

// x and y in level order; dv the diagonal in level order (null: unit);
// run: rows by running subtraction (the transposed solve)
template <typename V, bool DIV, bool RUN>
static void enqueue_levels_t(const TriLevels<V>& T, const V* dv, const V* x, V* y,
                             hipStream_t s) {
  for (int l = 0; l < T.nlev;) {
    const int g = T.seg_of[l];
    if (g >= 0) {
      const TriSegment& S = T.segs[g];
      const int base = T.lev_ptr[S.lev0], nrun = T.lev_ptr[S.lev1] - base;
#define PSB_SEG_LAUNCH(W)                                                                     \
  hipLaunchKernelGGL((k_trsv_seg<V, W, kSegStages, DIV, RUN>), dim3(1), dim3(kSegThreads), 0, s, \
                     S.lev0, S.lev1, base, nrun, T.halo + S.hoff, S.nhalo, T.lev_info, T.eja,   \
                     T.eval, dv, x, y)
      switch (S.width) {
        case 2: PSB_SEG_LAUNCH(2); break;
        case 3: PSB_SEG_LAUNCH(3); break;
        case 4: PSB_SEG_LAUNCH(4); break;
        default: PSB_SEG_LAUNCH(8); break;
      }
#undef PSB_SEG_LAUNCH
      l = S.lev1;
      continue;
    }
    const int count = T.lev_ptr[l + 1] - T.lev_ptr[l];
    hipLaunchKernelGGL((k_trsv_level<V, DIV, RUN>), dim3((count + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, T.lev_ptr[l], count, T.irp, T.ja, T.val, dv, x, y);
    l++;
  }
}

template <typename V>
static void enqueue_levels(const TriLevels<V>& T, const V* dv, bool run, const V* x, V* y,
                           hipStream_t s) {
  if (dv && run) enqueue_levels_t<V, true, true>(T, dv, x, y, s);
  else if (dv) enqueue_levels_t<V, true, false>(T, dv, x, y, s);
  else if (run) enqueue_levels_t<V, false, true>(T, dv, x, y, s);
  else enqueue_levels_t<V, false, false>(T, dv, x, y, s);
}

// The captured part of an application: x in L level order (xa) -> U^-1
// (d .* L^-1 x) in U level order (ya).  The gather of the caller's x and the
// scatter into the caller's y (with alpha and beta) run around it.
static void enqueue_apply(IluDev* F, hipStream_t s) {
  const int n = F->n, g = stream_grid(n);
  enqueue_levels<double>(F->L, nullptr, false, F->xa, F->ya, s);
  hipLaunchKernelGGL(k_perm_scatter<double>, dim3(g), dim3(kBlock), 0, s, n, F->L.perm, F->ya,
                     F->t1);
  hipLaunchKernelGGL(k_perm_gather<double>, dim3(g), dim3(kBlock), 0, s, n, F->U.perm, F->t1,
                     F->d, F->xa);
  enqueue_levels<double>(F->U, nullptr, false, F->xa, F->ya, s);
}

template <typename V>
static void free_tri(TriLevels<V>& T) {
  void* p[] = {T.perm, T.irp, T.ja, T.val, T.lev_info, T.eja, T.eval, T.halo};
  for (void* q : p)
    if (q) (void)hipFree(q);
  T = TriLevels<V>{};
}

static void destroy_ilu(IluDev* F) {
  if (!F) return;
  if (ctx().initialized) (void)hipStreamSynchronize(ctx().stream);
  if (F->exec) (void)hipGraphExecDestroy(F->exec);
  if (F->graph) (void)hipGraphDestroy(F->graph);
  free_tri(F->L);
  free_tri(F->U);
  void* p[] = {F->d, F->xa, F->ya, F->t1};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete F;
}

template <typename V>
static void destroy_sv(TriSolve<V>* S) {
  if (!S) return;
  if (ctx().initialized) (void)hipStreamSynchronize(ctx().stream);
  if (S->exec) (void)hipGraphExecDestroy(S->exec);
  if (S->graph) (void)hipGraphDestroy(S->graph);
  free_tri(S->tl);
  void* p[] = {S->dv, S->xa, S->ya};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete S;
}

// Prepares op(T)^-1 for the CSR triangle (irp, ja, val), 1-based host
// arrays: the diagonal of a non-unit row (its last entry if lower, first if
// upper, the entries psb_d_csr_cssv divides by) is taken out, the rest must
// lie strictly inside the triangle.  For op 'T' / 'C' the strict part is
// transposed with row c's entries in the CPU's update order: the rows i of T
// in solve order (descending for a lower triangle), CSR order within a row.
template <typename V>
static int build_sv(TriSolve<V>* S, int n, const int* irp, const int* ja, const V* val,
                    bool lower, bool unit, bool trans) {
  S->n = n;
  S->run = trans;
  const int nnz = irp[n] - 1;
  std::vector<int> sirp(n + 1, 1), sja;
  std::vector<V> sval, dv;
  sja.reserve(nnz > 0 ? nnz : 1);
  sval.reserve(nnz > 0 ? nnz : 1);
  if (!unit) dv.resize(n > 0 ? n : 1);
  for (int i = 0; i < n; i++) {
    int b = irp[i] - 1, e = irp[i + 1] - 1;
    if (!unit) {
      if (e <= b) return set_error(PSB_HIP_INVALID, "trsv", "non-unit row without entries");
      if (lower) dv[i] = val[--e];
      else dv[i] = val[b++];
    }
    for (int k = b; k < e; k++) {
      sja.push_back(ja[k]);
      sval.push_back(val[k]);
    }
    sirp[i + 1] = (int)sja.size() + 1;
  }
  bool tlower = lower;
  if (trans) {
    for (int i = 0; i < n; i++)   // the strict part must be a triangle before it is turned
      for (int k = sirp[i] - 1; k < sirp[i + 1] - 1; k++)
        if (sja[k] < 1 || sja[k] > n || (lower ? sja[k] > i : sja[k] < i + 2))
          return set_error(PSB_HIP_INVALID, "trsv", "entry outside its triangle");
    const int snz = sirp[n] - 1;
    std::vector<int> cnt(n + 1, 0);
    for (int k = 0; k < snz; k++) cnt[sja[k]]++;
    std::vector<int> tirp(n + 1, 1);
    for (int c = 0; c < n; c++) tirp[c + 1] = tirp[c] + cnt[c + 1];
    std::vector<int> fill(tirp.begin(), tirp.end() - 1), tja(snz > 0 ? snz : 1);
    std::vector<V> tval(snz > 0 ? snz : 1);
    for (int t = 0; t < n; t++) {
      const int i = lower ? n - 1 - t : t;   // the CPU's order of the rows
      for (int k = sirp[i] - 1; k < sirp[i + 1] - 1; k++) {
        const int q = fill[sja[k] - 1]++ - 1;
        tja[q] = i + 1;
        tval[q] = sval[k];
      }
    }
    sirp.swap(tirp);
    sja.swap(tja);
    sval.swap(tval);
    tlower = !lower;
  }
  int rc = upload_tri(S->tl, n, tlower, sirp.data(), sja.data(), sval.data(),
                      g_trsv_variant != 1);
  if (rc) return rc;
  const size_t nb = sizeof(V) * (size_t)(n > 0 ? n : 1);
  hipError_t e = hipSuccess;
  V** bufs[] = {&S->xa, &S->ya};
  for (V** b : bufs)
    if (e == hipSuccess) e = hipMalloc((void**)b, nb);
  hipStream_t s = ctx().stream;
  if (e == hipSuccess && !unit && n > 0) {   // the diagonal, gathered into level order
    e = hipMalloc((void**)&S->dv, nb);
    if (e == hipSuccess) e = hipMemcpy(S->xa, dv.data(), nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_perm_gather<V>, dim3(stream_grid(n)), dim3(kBlock), 0, s, n,
                         S->tl.perm, S->xa, (const V*)nullptr, S->dv);
      e = hipStreamSynchronize(s);
    }
  }
  if (e == hipSuccess && n > 0) {
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      enqueue_levels<V>(S->tl, S->dv, S->run, S->xa, S->ya, s);
      e = hipStreamEndCapture(s, &S->graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&S->exec, S->graph, nullptr, nullptr, 0);
  }
  if (e != hipSuccess)
    return set_error(e == hipErrorOutOfMemory ? PSB_HIP_OUTOFMEM : PSB_HIP_RUNTIME, "trsv",
                     hipGetErrorString(e));
  return PSB_HIP_SUCCESS;
}

// The CSR matrix's prepared solve for key (lower | unit << 1 | trans << 2),
// built from the device copy of the matrix on first use and kept until the
// matrix is rewritten, freed or asked for another triangle or op.
template <typename V>
static int csr_sv_get(CsrDev* m, int key, TriSolve<V>** out) {
  TriSolve<V>* S = (TriSolve<V>*)m->sv;
  if (S && S->key == key) {
    *out = S;
    return PSB_HIP_SUCCESS;
  }
  release_csr_sv(m);
  std::vector<int> irp(m->rows + 1), ja(m->nnz > 0 ? m->nnz : 1);
  std::vector<V> val(m->nnz > 0 ? m->nnz : 1);
  hipStream_t s = ctx().stream;
  PSB_HIP_CHECK(hipMemcpyAsync(irp.data(), m->irp, sizeof(int) * irp.size(),
                               hipMemcpyDeviceToHost, s));
  if (m->nnz > 0) {
    PSB_HIP_CHECK(hipMemcpyAsync(ja.data(), m->ja, sizeof(int) * (size_t)m->nnz,
                                 hipMemcpyDeviceToHost, s));
    PSB_HIP_CHECK(hipMemcpyAsync(val.data(), m->val, sizeof(V) * (size_t)m->nnz,
                                 hipMemcpyDeviceToHost, s));
  }
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  S = new TriSolve<V>();
  const int rc = build_sv<V>(S, m->rows, irp.data(), ja.data(), val.data(), (key & 1) != 0,
                             (key & 2) != 0, (key & 4) != 0);
  if (rc) {
    destroy_sv(S);
    return rc;
  }
  S->key = key;
  m->sv = S;
  *out = S;
  return PSB_HIP_SUCCESS;
}

template <typename V>
static int csr_sv_apply(CsrDev* m, int key, V alpha, const MultiVec* x, V beta, MultiVec* y) {
  TriSolve<V>* S = nullptr;
  int rc = csr_sv_get<V>(m, key, &S);
  if (rc) return rc;
  hipStream_t s = ctx().stream;
  const int n = S->n, g = stream_grid(n);
  for (int c = 0; c < x->count; c++) {
    const V* xc = (const V*)x->v + (size_t)c * x->pitch;
    V* yc = (V*)y->v + (size_t)c * y->pitch;
    hipLaunchKernelGGL(k_perm_gather<V>, dim3(g), dim3(kBlock), 0, s, n, S->tl.perm, xc,
                       (const V*)nullptr, S->xa);
    PSB_HIP_LAUNCH_CHECK();
    PSB_HIP_CHECK(hipGraphLaunch(S->exec, s));
    hipLaunchKernelGGL(k_perm_scatter_axpby<V>, dim3(g), dim3(kBlock), 0, s, n, S->tl.perm,
                       S->ya, alpha, beta, yc);
    PSB_HIP_LAUNCH_CHECK();
  }
  return PSB_HIP_SUCCESS;
}

void release_csr_sv(CsrDev* m) {
  if (!m || !m->sv) return;
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE) destroy_sv((TriSolve<double>*)m->sv);
  else destroy_sv((TriSolve<float>*)m->sv);
  m->sv = nullptr;
}

}  // namespace psbhip

using namespace psbhip;

extern "C" {

int psb_hip_ilu_alloc(void** ilu, int n, const int* lirp, const int* lja, const double* lval,
                      const int* uirp, const int* uja, const double* uval, const double* d) {
  PSB_ENSURE_INIT();
  if (!ilu || n < 0 || !lirp || !uirp || !d) return PSB_HIP_INVALID;
  *ilu = nullptr;
  IluDev* F = new IluDev();
  F->n = n;
  const bool seg = g_trsv_variant != 1;
  int rc = upload_tri(F->L, n, true, lirp, lja, lval, seg);
  if (!rc) rc = upload_tri(F->U, n, false, uirp, uja, uval, seg);
  if (rc) {
    destroy_ilu(F);
    return rc;
  }
  const size_t nb = sizeof(double) * (size_t)(n > 0 ? n : 1);
  hipError_t e = hipSuccess;
  double** bufs[] = {&F->d, &F->xa, &F->ya, &F->t1};
  for (double** b : bufs)
    if (e == hipSuccess) e = hipMalloc((void**)b, nb);
  if (e == hipSuccess && n > 0)
    e = hipMemcpy(F->d, d, sizeof(double) * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess && n > 0) {
    hipStream_t s = ctx().stream;
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      enqueue_apply(F, s);
      e = hipStreamEndCapture(s, &F->graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&F->exec, F->graph, nullptr, nullptr, 0);
  }
  if (e != hipSuccess) {
    destroy_ilu(F);
    return set_error(e == hipErrorOutOfMemory ? PSB_HIP_OUTOFMEM : PSB_HIP_RUNTIME, __func__,
                     hipGetErrorString(e));
  }
  *ilu = F;
  return PSB_HIP_SUCCESS;
}

int psb_hip_ilu_set_variant(int variant) {
  const int old = g_trsv_variant;
  if (variant == 0 || variant == 1) g_trsv_variant = variant;
  return old;
}

int psb_hip_ilu_segments(void* ilu, int* lower_segments, int* upper_segments) {
  IluDev* F = (IluDev*)ilu;
  if (!F) return PSB_HIP_INVALID;
  if (lower_segments) *lower_segments = (int)F->L.segs.size();
  if (upper_segments) *upper_segments = (int)F->U.segs.size();
  return PSB_HIP_SUCCESS;
}

int psb_hip_ilu_free(void* ilu) {
  destroy_ilu((IluDev*)ilu);
  return PSB_HIP_SUCCESS;
}

int psb_hip_ilu_levels(void* ilu, int* lower_levels, int* upper_levels) {
  IluDev* F = (IluDev*)ilu;
  if (!F) return PSB_HIP_INVALID;
  if (lower_levels) *lower_levels = F->L.nlev;
  if (upper_levels) *upper_levels = F->U.nlev;
  return PSB_HIP_SUCCESS;
}

int psb_hip_ilu_apply(void* ilu, double alpha, void* xv, double beta, void* yv) {
  PSB_ENSURE_INIT();
  IluDev* F = (IluDev*)ilu;
  MultiVec *x = (MultiVec*)xv, *y = (MultiVec*)yv;
  if (!F || !x || !y || x->elem_type != PSB_HIP_TYPE_DOUBLE || y->elem_type != PSB_HIP_TYPE_DOUBLE ||
      x->size < F->n || y->size < F->n || x->count != 1 || y->count != 1)
    return set_error(PSB_HIP_INVALID, __func__, "x/y type or size");
  if (F->n == 0) return PSB_HIP_SUCCESS;
  hipStream_t s = ctx().stream;
  const int n = F->n, g = stream_grid(n);
  hipLaunchKernelGGL(k_perm_gather<double>, dim3(g), dim3(kBlock), 0, s, n, F->L.perm,
                     (const double*)x->v, (const double*)nullptr, F->xa);
  PSB_HIP_LAUNCH_CHECK();
  PSB_HIP_CHECK(hipGraphLaunch(F->exec, s));
  hipLaunchKernelGGL(k_perm_scatter_axpby<double>, dim3(g), dim3(kBlock), 0, s, n, F->U.perm,
                     F->ya, alpha, beta, (double*)y->v);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_csr_spsv(void* dmat, int lower, int unit, int trans, double alpha, void* xv,
                     double beta, void* yv) {
  PSB_ENSURE_INIT();
  CsrDev* m = (CsrDev*)dmat;
  MultiVec *x = (MultiVec*)xv, *y = (MultiVec*)yv;
  if (!m || !x || !y) return set_error(PSB_HIP_INVALID, __func__, "null argument");
  if (m->rows != m->cols) return set_error(PSB_HIP_INVALID, __func__, "matrix not square");
  if (trans < 0 || trans > 2) return set_error(PSB_HIP_INVALID, __func__, "trans");
  if (x->elem_type != m->elem_type || y->elem_type != m->elem_type || x->size < m->rows ||
      y->size < m->rows || x->count != y->count)
    return set_error(PSB_HIP_INVALID, __func__, "x/y type, size or count");
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  const int key = (lower ? 1 : 0) | (unit ? 2 : 0) | (trans ? 4 : 0);
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
    return csr_sv_apply<double>(m, key, alpha, x, beta, y);
  return csr_sv_apply<float>(m, key, (float)alpha, x, (float)beta, y);
}

int psb_hip_csr_spsv_levels(void* dmat, int* levels, int* segments) {
  CsrDev* m = (CsrDev*)dmat;
  if (!m) return PSB_HIP_INVALID;
  int nl = 0, ns = 0;
  if (m->sv && m->elem_type == PSB_HIP_TYPE_DOUBLE) {
    nl = ((TriSolve<double>*)m->sv)->tl.nlev;
    ns = (int)((TriSolve<double>*)m->sv)->tl.segs.size();
  } else if (m->sv) {
    nl = ((TriSolve<float>*)m->sv)->tl.nlev;
    ns = (int)((TriSolve<float>*)m->sv)->tl.segs.size();
  }
  if (levels) *levels = nl;
  if (segments) *segments = ns;
  return PSB_HIP_SUCCESS;
}

}  // extern "C"
