// Device-resident Krylov solvers: CG and BiCGSTAB with the DIAG (Jacobi)
// preconditioner, and FP32-inner / FP64-outer iterative refinement.
//
// The operation sequence is the reference's (krylov/psb_dcg.F90:241-294,
// krylov/psb_dcgstab.f90:248-364, convergence krylov/psb_d_krylov_conv_mod.f90:
// 279-356), regrouped into a few fused kernels per iteration.  Every scalar
// (rho, sigma, alpha, beta, omega, the residual norm, the stop flag) lives
// in device memory, so iterations are queued without host round trips; the
// host reads the stop flag once every `check_every` iterations.  Kernels
// queued after convergence see the flag and return at once, so the iterate
// and the iteration count are exactly those of the stopping iteration.
//
// Reductions: each block of a fused kernel stores its partial sums; a
// one-block finalisation kernel then adds the partials in index order and
// updates the scalars (deterministic run to run).  The kernel boundary makes
// the partials of every XCD visible; a device-scope fence per block would
// write back the XCD's whole L2 after the fused kernel's stores, which
// measured 1.4x slower on MI355X.
//
// CG per iteration:        K_p   p = z | z + beta*p
//                          K_q   q = A*p, partial p.q        F_q  sigma, alpha
//                          K_x   x += alpha p, r -= alpha q, z = d.*r, partial r.r, r.z
//                                                            F_x  ||r||, stop test, rho, beta
// BiCGSTAB per iteration:  K_a   p = r | r + beta*(p - omega v), f = d.*p
//                          K_v   v = A*f, partial q.v        F_v  alpha
//                          K_s   s = r - alpha v, z = d.*s
//                          K_t   t = A*z, partial t.t, t.s   F_t  omega
//                          K_x   x = (x + alpha f) + omega z, r = s - omega t,
//                                partial r.r, q.r            F_x  ||r||, stop test, rho, beta
#include <math.h>
#include <string.h>

#include <vector>

#include "psb_hip_internal.hpp"

namespace psbhip {

constexpr int kMaxGrid = 32768;  // blocks of a fused kernel (partials per reduction)

struct KState {
  // scalars of the recurrence
  double rho, rho_old, alpha, beta, omega, sigma;
  double errnum, errden, eps;
  double bnorm;
  double rref;    // reliable updates: residual norm at the last replacement
  double delta;   // reliable updates: replacement factor
  int it;         // iterations since the last restart
  int itx;        // total iterations
  int itmax;
  int done;       // stop test satisfied (or itmax reached)
  int breakdown;  // 1: restart requested; 2: residual replacement requested
  int istop;
};

// --------------------------------------------------------------- views
// rows<R>(row0, x, ax): ax[k] = (A x)[row0 + k] for R = 1, or for R = 2 the
// adjacent pair (row0 even) sharing 8-byte index and 2-element value loads.
// Rows past nr are clamped to stored rows (the results are not used).  With
// xd / fd (R entries each, fd zeroed by the caller), xd[k] receives x(row0 +
// k) where the row stores its diagonal (the entry the product gathers at the
// row's own column) and fd[k] is set.
// HK > 0: the hack size as a compile-time constant (the row's hack and lane
// by shift and mask instead of an integer division; measured 4% faster on the
// FP32 inner solves at L = 256); HK == 0: any hksz.
template <typename T, int HK = 0>
struct HllView {
  int nr, hksz;
  const int* hkoffs;
  const int* ja;
  const T* val;
  const short* res16;   // 16-bit column offsets (HllDev::res16 / base16) or null
  const int* base16;
  template <int R>
  __device__ __forceinline__ void rows(int row0, const T* __restrict__ x, T (&ax)[R],
                                       T* xd = nullptr, int* fd = nullptr) const {
    const int last = R == 1 ? nr - 1 : ((nr - 1) & ~1);
    const int r = row0 < last ? row0 : last;
    const int stride = HK > 0 ? HK : hksz;
    int hk = r / stride;
    // one row per lane and hack size = wavefront size: the wavefront is one
    // hack (tiles are whole hacks), so its offsets come through scalar loads
    if (R == 1 && HK == kWave) hk = __builtin_amdgcn_readfirstlane(hk);
    const int lane = r - hk * stride;
    int off0, w;
    if constexpr (R == 2) {
      hll_pair_hack<HK>(hkoffs, nr, hk, stride, off0, w);
    } else {
      off0 = hkoffs[hk];
      w = (hkoffs[hk + 1] - off0) / stride;
    }
    if (res16) {
      const short* pe = res16 + off0 + lane;
      const int* pb = base16 + off0 / stride;
      if constexpr (R == 1)
        ax[0] = slot_dot16<T, T>(pe, pb, val + off0 + lane, stride, w, r, x, xd, fd);
      else
        slot_dot16_pair<T, T>(pe, pb, val + off0 + lane, stride, w, r, x, ax[0], ax[1], xd, fd);
    } else {
      if constexpr (R == 1)
        ax[0] = slot_dot<T, T>(ja + off0 + lane, val + off0 + lane, stride, w, x, r, xd, fd);
      else
        slot_dot_pair<T, T>(ja + off0 + lane, val + off0 + lane, stride, w, x, ax[0], ax[1], r,
                            xd, fd);
    }
  }
};
template <typename T>
using HllView64 = HllView<T, 64>;
template <typename T>
inline HllView64<T> as64(const HllView<T>& v) {
  return HllView64<T>{v.nr, v.hksz, v.hkoffs, v.ja, v.val, v.res16, v.base16};
}

template <typename T>
struct EllView {
  int nr, pitch, nzm;
  const int* ja;
  const T* val;
  const short* res16;   // 16-bit column offsets (EllDev::res16 / base16) or null
  const int* base16;
  template <int R>
  __device__ __forceinline__ void rows(int row0, const T* __restrict__ x, T (&ax)[R],
                                       T* xd = nullptr, int* fd = nullptr) const {
    const int last = R == 1 ? nr - 1 : ((nr - 1) & ~1);
    const int r = row0 < last ? row0 : last;
    if (res16) {
      const int* pb = base16 + (size_t)(r >> 6) * nzm;
      if constexpr (R == 1)
        ax[0] = slot_dot16<T, T>(res16 + r, pb, val + r, pitch, nzm, r, x, xd, fd);
      else
        slot_dot16_pair<T, T>(res16 + r, pb, val + r, pitch, nzm, r, x, ax[0], ax[1], xd, fd);
    } else {
      if constexpr (R == 1) ax[0] = slot_dot<T, T>(ja + r, val + r, pitch, nzm, x, r, xd, fd);
      else slot_dot_pair<T, T>(ja + r, val + r, pitch, nzm, x, ax[0], ax[1], r, xd, fd);
    }
  }
};

// Grids.  Products take one tile of kBlock*R rows per block up to kMaxGrid
// tiles, grid-stride beyond; vector kernels walk 16-byte vectors grid-stride.
inline int spmv_grid(int n, int rpl) {
  const int per = kBlock * rpl;
  int tiles = (n + per - 1) / per;
  return tiles < 1 ? 1 : (tiles < kMaxGrid ? tiles : kMaxGrid);
}
template <typename T>
inline int vec_grid(int n) {
  const int V = 16 / (int)sizeof(T);
  int g = (n / V + kBlock - 1) / kBlock;
  if (g > 2048) g = 2048;
  return g < 1 ? 1 : g;
}

// BODY runs for each row < nr of the thread's tiles with `row` and `ax` = (A X)[row]
#define PSB_TILE_ROWS(R, M, X, BODY)                                            \
  for (int tile_ = blockIdx.x; tile_ * (kBlock * R) < (M).nr; tile_ += gridDim.x) { \
    const int row0_ = tile_ * (kBlock * R) + R * (int)threadIdx.x;              \
    T axr_[R];                                                                  \
    (M).template rows<R>(row0_, (X), axr_);                                     \
    _Pragma("unroll") for (int k_ = 0; k_ < R; k_++) {                          \
      const int row = row0_ + k_;                                               \
      const T ax = axr_[k_];                                                    \
      if (row < (M).nr) { BODY }                                                \
    }                                                                           \
  }

// The same tiles with xrow = X[row] beside ax: the x entry the product
// gathered at the row's diagonal, or a load of X[row] for a row that stores
// no diagonal (the same value either way).
#define PSB_TILE_ROWS_X(R, M, X, BODY)                                          \
  for (int tile_ = blockIdx.x; tile_ * (kBlock * R) < (M).nr; tile_ += gridDim.x) { \
    const int row0_ = tile_ * (kBlock * R) + R * (int)threadIdx.x;              \
    T axr_[R], xd_[R];                                                          \
    int fd_[R];                                                                 \
    _Pragma("unroll") for (int k_ = 0; k_ < R; k_++) {                          \
      xd_[k_] = T(0);                                                           \
      fd_[k_] = 0;                                                              \
    }                                                                           \
    (M).template rows<R>(row0_, (X), axr_, xd_, fd_);                           \
    _Pragma("unroll") for (int k_ = 0; k_ < R; k_++) {                          \
      const int row = row0_ + k_;                                               \
      const T ax = axr_[k_];                                                    \
      if (row < (M).nr) {                                                       \
        const T xrow = fd_[k_] ? xd_[k_] : (X)[row];                            \
        BODY                                                                    \
      }                                                                         \
    }                                                                           \
  }

// vector loop over n entries: BODY_V sees i (first of 16/sizeof(T) entries),
// BODY_S handles the scalar tail (block 0)
#define PSB_VEC_LOOP(N, BODY_V, BODY_S)                                         \
  {                                                                             \
    constexpr int V = 16 / sizeof(T);                                           \
    const int nv = (N) / V;                                                     \
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < nv; k += gridDim.x * kBlock) { \
      const int i = V * k;                                                      \
      BODY_V                                                                    \
    }                                                                           \
    if (blockIdx.x == 0)                                                        \
      for (int i = nv * V + threadIdx.x; i < (N); i += kBlock) { BODY_S }       \
  }

// Each block stores its NP partial sums (partials[p*kMaxGrid + block]).
template <int NP>
__device__ __forceinline__ void store_partials(double (&v)[NP], double* partials) {
  __shared__ double lds[kBlock / kWave];
#pragma unroll
  for (int p = 0; p < NP; p++) {
    double s = block_sum(v[p], lds);
    if (threadIdx.x == 0) partials[p * kMaxGrid + blockIdx.x] = s;
  }
}

__device__ __forceinline__ bool halted(const KState* st) {
  return st->done != 0 || st->breakdown != 0;
}

// --------------------------------------------------------------- finalisation
enum FinOp {
  FIN_RESTART_CG, FIN_RESTART_BI, FIN_BNORM, FIN_CG_Q, FIN_CG_X, FIN_BI_V, FIN_BI_T, FIN_BI_X,
  FIN_RU_X, FIN_RU_RESET, FIN_RU_RESET_FIRST
};

// stop test of psb_d_check_conv_vect for istop 2 and 3 (errden = ||b|| or ||r0||)
__device__ __forceinline__ void stop_test(KState* st, double rr, double* hist) {
  st->errnum = sqrt(rr);
  bool res;
  if (st->errden == 0.0) res = st->errnum <= st->eps;
  else res = st->errnum <= st->eps * st->errden;
  if (hist) hist[st->itx] = st->errden > 0.0 ? st->errnum / st->errden : st->errnum;
  if (res || st->itmax <= st->itx) st->done = 1;
}

constexpr int kFin = 1024;  // threads of the finalisation kernel

// Sum of partials[0..nb) over the kFin threads: eight independent loads in
// flight per thread, fixed order (deterministic for a given nb).
__device__ __forceinline__ double fin_sum(const double* __restrict__ partials, int nb,
                                          double* lds /* kFin / kWave */) {
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int b = threadIdx.x;
  for (; b + 7 * kFin < nb; b += 8 * kFin) {
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] += partials[b + k * kFin];
  }
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (b + k * kFin < nb) a[k] += partials[b + k * kFin];
  double t = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  t = wave_sum(t);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = t;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < kFin / kWave; k++) r += lds[k];
  __syncthreads();
  return r;
}

template <int OP, int NP>
__global__ __launch_bounds__(kFin) void k_fin(KState* st, const double* __restrict__ partials,
                                              int nb, double* hist,
                                              const double* __restrict__ aux) {
  if (OP == FIN_RESTART_CG || OP == FIN_RESTART_BI) {
    if (st->done) return;
  } else if (OP != FIN_BNORM && OP != FIN_RU_RESET && OP != FIN_RU_RESET_FIRST) {
    if (halted(st)) return;
  }
  __shared__ double lds[kFin / kWave];
  double v[NP];
#pragma unroll
  for (int p = 0; p < NP; p++) v[p] = fin_sum(partials + p * kMaxGrid, nb, lds);
  if (threadIdx.x != 0) return;
  const double v1 = v[NP - 1];  // second partial (NP == 2)
  if (OP == FIN_BNORM) {
    st->bnorm = sqrt(v[0]);
    if (st->istop == 2) st->errden = st->bnorm;
  } else if (OP == FIN_RESTART_CG || OP == FIN_RESTART_BI) {
    const bool bicg = OP == FIN_RESTART_BI;
    st->breakdown = 0;
    st->it = 0;
    st->errnum = sqrt(v[0]);
    // ||r0||: CG re-initialises at every restart (psb_dcg.F90:231), BiCGSTAB
    // once before its restart loop (psb_dcgstab.f90:220)
    if (st->istop == 3 && (!bicg || st->itx == 0)) st->errden = st->errnum;
    if (bicg) {
      // BiCGSTAB checks convergence before the first iteration
      bool res = st->errden == 0.0 ? st->errnum <= st->eps : st->errnum <= st->eps * st->errden;
      if (res || st->itmax <= st->itx) st->done = 1;
      st->rho = v[0];  // q = r: rho = q.r
    } else {
      st->rho = v1;
    }
    st->rho_old = 0.0;
    if (!st->done && bicg && st->rho == 0.0) {  // first-iteration breakdown test
      st->itx += 1;
      st->breakdown = 1;
    }
  } else if (OP == FIN_CG_Q || OP == FIN_BI_V) {
    // a zero or non-finite denominator or scalar is a breakdown, caught
    // before x takes it (FP32 inner solves can overflow near a breakdown):
    // the solve restarts from b - A x
    st->sigma = v[0];
    const double alpha = v[0] != 0.0 ? st->rho / v[0] : 0.0;
    if (v[0] == 0.0 || !isfinite(v[0]) || !isfinite(alpha)) {
      st->itx += 1;  // the reference counted this iteration before breaking
      st->breakdown = 1;
    } else {
      st->alpha = alpha;
    }
  } else if (OP == FIN_BI_T) {
    if (v[0] == 0.0 || !isfinite(v[0]) || !isfinite(v1)) {
      st->itx += 1;
      st->breakdown = 1;
      return;
    }
    st->omega = v1 / v[0];
    if (st->omega == 0.0 || !isfinite(st->omega)) {
      st->itx += 1;
      st->breakdown = 1;
    }
  } else if (OP == FIN_CG_X) {
    st->itx += 1;
    st->it += 1;
    stop_test(st, v[0], hist);
    if (!st->done) {
      st->rho_old = st->rho;
      st->rho = v1;
      if (st->rho_old == 0.0) {
        st->itx += 1;  // next iteration counted, then breaks (psb_dcg.F90:253-258)
        st->breakdown = 1;
      } else {
        st->beta = st->rho / st->rho_old;
      }
    }
  } else if (OP == FIN_BI_X) {
    st->itx += 1;
    st->it += 1;
    stop_test(st, v[0], hist);
    if (!st->done) {
      st->rho_old = st->rho;
      st->rho = v1;
      const double beta = (st->rho / st->rho_old) * (st->alpha / st->omega);
      if (st->rho == 0.0 || !isfinite(beta)) {
        st->itx += 1;  // next iteration counted, then breaks (psb_dcgstab.f90:267-271)
        st->breakdown = 1;
      } else {
        st->beta = beta;
      }
    }
  } else if (OP == FIN_RU_X) {
    st->itx += 1;
    st->it += 1;
    const double rn = sqrt(v[0]);
    st->errnum = rn;
    // rref: norm at the last replacement; eps*errden: the absolute target
    if (rn <= st->delta * st->rref || rn <= st->eps * st->errden || st->itx >= st->itmax) {
      st->rho_old = st->rho;
      st->breakdown = 2;
      return;
    }
    st->rho_old = st->rho;
    st->rho = v1;
    if (st->rho_old == 0.0) {
      st->itx += 1;
      st->breakdown = 1;
    } else {
      st->beta = st->rho / st->rho_old;
    }
  } else if (OP == FIN_RU_RESET || OP == FIN_RU_RESET_FIRST) {
    st->rref = *aux;
    st->breakdown = 0;
    st->rho = v[0];
    if (OP == FIN_RU_RESET_FIRST) {
      st->it = 0;
    } else if (st->rho_old == 0.0) {
      st->breakdown = 1;
    } else {
      st->beta = st->rho / st->rho_old;
    }
  }
}

// --------------------------------------------------------------- restart / norms
// r = b - A*x, then (CG) z = d.*r with partials r.r, r.z or (BiCGSTAB) q = r
// with partial r.r.
template <typename T, class M, int R>
__global__ __launch_bounds__(kBlock) void k_restart(M m, const T* b,
                                                    const T* __restrict__ x, T* r, T* q,
                                                    const T* __restrict__ d, T* z, KState* st,
                                                    double* partials, int bicg) {
  if (st->done) return;
  double v[2] = {0.0, 0.0};
  PSB_TILE_ROWS(R, m, x, {
    T ri = b[row] - ax;
    r[row] = ri;
    v[0] += (double)ri * (double)ri;
    if (bicg) {
      q[row] = ri;
    } else {
      T zi = d ? d[row] * ri : ri;
      if (z != r) z[row] = zi;
      v[1] += (double)ri * (double)zi;
    }
  })
  store_partials<2>(v, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sumsq(int n, const T* __restrict__ b,
                                                  double* partials) {
  double v[1] = {0.0};
  PSB_VEC_LOOP(n, { v[0] += vdot<T>(ldv(b + i), ldv(b + i)); },
               { v[0] += (double)b[i] * (double)b[i]; })
  store_partials<1>(v, partials);
}

// A product's output row: FP32 outputs by non-temporal store, FP64 by a
// plain store.  At L = 256 the non-temporal store made the FP32 inner solves
// 5% faster (reliable 238-240 -> 227-228 ms, restarted 297 -> 282 ms) and the
// FP64 CG 5% slower (414 -> 434 ms), where the plain store leaves part of
// the 134 MB output in the Infinity Cache for the next kernel.
template <typename T>
__device__ __forceinline__ void store_prod(T* __restrict__ out, int row, T ax) {
  if constexpr (sizeof(T) == 4) __builtin_nontemporal_store(ax, out + row);
  else out[row] = ax;
}

// --------------------------------------------------------------- CG kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_p(int n, const T* __restrict__ z, T* __restrict__ p,
                                                 const KState* st) {
  if (halted(st)) return;
  const bool first = st->it == 0;
  const T beta = (T)st->beta;
  PSB_VEC_LOOP(n,
               {
                 v16<T> zi = ldv(z + i);
                 stv(p + i, first ? zi : zi + beta * ldv(p + i));
               },
               { p[i] = first ? z[i] : z[i] + beta * p[i]; })
}

template <typename T, class M, int R>
__global__ __launch_bounds__(kBlock) void k_cg_q(M m, const T* __restrict__ p, T* __restrict__ q,
                                                 const KState* st, double* partials) {
  if (halted(st)) return;
  double v[1] = {0.0};
  // p(row) for the p.q partial comes from the product's own gather at the
  // diagonal: one read of p fewer per iteration (8n bytes FP64, 4n FP32)
  PSB_TILE_ROWS_X(R, m, p, {
    store_prod(q, row, ax);
    v[0] += (double)xrow * (double)ax;
  })
  store_partials<1>(v, partials);
}

// a.b partial (the CSR path: the product runs through the CSR kernels)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(int n, const T* __restrict__ a,
                                                const T* __restrict__ b, const KState* st,
                                                double* partials) {
  if (halted(st)) return;
  double v[1] = {0.0};
  PSB_VEC_LOOP(n, { v[0] += vdot<T>(ldv(a + i), ldv(b + i)); },
               { v[0] += (double)a[i] * (double)b[i]; })
  store_partials<1>(v, partials);
}

// t.t and t.s partials (the CSR path of BiCGSTAB)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot2(int n, const T* __restrict__ t,
                                                 const T* __restrict__ s, const KState* st,
                                                 double* partials) {
  if (halted(st)) return;
  double v[2] = {0.0, 0.0};
  PSB_VEC_LOOP(n,
               {
                 v16<T> ti = ldv(t + i);
                 v[0] += vdot<T>(ti, ti);
                 v[1] += vdot<T>(ti, ldv(s + i));
               },
               {
                 v[0] += (double)t[i] * (double)t[i];
                 v[1] += (double)t[i] * (double)s[i];
               })
  store_partials<2>(v, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_x(int n, T* __restrict__ x, T* r,
                                                 const T* __restrict__ p, const T* __restrict__ q,
                                                 const T* __restrict__ d, T* z, const KState* st,
                                                 double* partials) {
  if (halted(st)) return;
  const T alpha = (T)st->alpha;
  double v[2] = {0.0, 0.0};
  PSB_VEC_LOOP(n,
               {
                 v16<T> ri = ldv(r + i) - alpha * ldv(q + i);
                 stv(x + i, ldv(x + i) + alpha * ldv(p + i));
                 stv(r + i, ri);
                 v16<T> zi = ri;
                 if (d) {
                   zi = ldv(d + i) * ri;
                   stv(z + i, zi);
                 }
                 v[0] += vdot<T>(ri, ri);
                 v[1] += vdot<T>(ri, zi);
               },
               {
                 x[i] = x[i] + alpha * p[i];
                 T ri = r[i] - alpha * q[i];
                 r[i] = ri;
                 T zi = d ? d[i] * ri : ri;
                 if (d) z[i] = zi;
                 v[0] += (double)ri * (double)ri;
                 v[1] += (double)ri * (double)zi;
               })
  store_partials<2>(v, partials);
}

// --------------------------------------------------------------- BiCGSTAB kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_a(int n, const T* __restrict__ r, T* p,
                                                 const T* __restrict__ v, const T* __restrict__ d,
                                                 T* f, const KState* st) {
  if (halted(st)) return;
  const bool first = st->it == 0;
  const T beta = (T)st->beta, momega = (T)(-st->omega);
  PSB_VEC_LOOP(n,
               {
                 v16<T> pi = ldv(r + i);
                 if (!first) pi = pi + beta * (momega * ldv(v + i) + ldv(p + i));
                 stv(p + i, pi);
                 if (d) stv(f + i, ldv(d + i) * pi);
               },
               {
                 T pi = first ? r[i] : r[i] + beta * (momega * v[i] + p[i]);
                 p[i] = pi;
                 if (d) f[i] = d[i] * pi;
               })
}

template <typename T, class M, int R>
__global__ __launch_bounds__(kBlock) void k_bi_v(M m, const T* __restrict__ f, T* __restrict__ v,
                                                 const T* __restrict__ q, const KState* st,
                                                 double* partials) {
  if (halted(st)) return;
  double s[1] = {0.0};
  PSB_TILE_ROWS(R, m, f, {
    store_prod(v, row, ax);
    s[0] += (double)q[row] * (double)ax;
  })
  store_partials<1>(s, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_s(int n, const T* __restrict__ r,
                                                 const T* __restrict__ v, const T* __restrict__ d,
                                                 T* s, T* z, const KState* st) {
  if (halted(st)) return;
  const T malpha = (T)(-st->alpha);
  PSB_VEC_LOOP(n,
               {
                 v16<T> si = malpha * ldv(v + i) + ldv(r + i);
                 stv(s + i, si);
                 if (d) stv(z + i, ldv(d + i) * si);
               },
               {
                 T si = malpha * v[i] + r[i];
                 s[i] = si;
                 if (d) z[i] = d[i] * si;
               })
}

template <typename T, class M, int R>
__global__ __launch_bounds__(kBlock) void k_bi_t(M m, const T* __restrict__ z, T* __restrict__ t,
                                                 const T* __restrict__ s, const KState* st,
                                                 double* partials) {
  if (halted(st)) return;
  double v[2] = {0.0, 0.0};
  PSB_TILE_ROWS(R, m, z, {
    store_prod(t, row, ax);
    v[0] += (double)ax * (double)ax;
    v[1] += (double)ax * (double)s[row];
  })
  store_partials<2>(v, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_x(int n, T* __restrict__ x, T* __restrict__ r,
                                                 const T* __restrict__ f, const T* __restrict__ z,
                                                 const T* __restrict__ s, const T* __restrict__ t,
                                                 const T* __restrict__ q, const KState* st,
                                                 double* partials) {
  if (halted(st)) return;
  const T alpha = (T)st->alpha, omega = (T)st->omega;
  double v[2] = {0.0, 0.0};
  PSB_VEC_LOOP(n,
               {
                 stv(x + i, omega * ldv(z + i) + (alpha * ldv(f + i) + ldv(x + i)));
                 v16<T> ri = -omega * ldv(t + i) + ldv(s + i);
                 stv(r + i, ri);
                 v[0] += vdot<T>(ri, ri);
                 v[1] += vdot<T>(ldv(q + i), ri);
               },
               {
                 T xi = alpha * f[i] + x[i];
                 x[i] = omega * z[i] + xi;
                 T ri = -omega * t[i] + s[i];
                 r[i] = ri;
                 v[0] += (double)ri * (double)ri;
                 v[1] += (double)q[i] * (double)ri;
               })
  store_partials<2>(v, partials);
}

// --------------------------------------------------------------- mixed precision helpers
// r32 = (float)(r * inv), inv = 1/||r|| read from device
__global__ __launch_bounds__(kBlock) void k_to_single(int n, const double* __restrict__ r,
                                                      const double* __restrict__ scale,
                                                      float* __restrict__ r32) {
  const double inv = 1.0 / *scale;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    r32[i] = (float)(r[i] * inv);
}

// x += scale * (double)d32
__global__ __launch_bounds__(kBlock) void k_update_double(int n, double* __restrict__ x,
                                                          const float* __restrict__ d32,
                                                          const double* __restrict__ scale) {
  const double s = *scale;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    x[i] = x[i] + s * (double)d32[i];
}

__global__ __launch_bounds__(kBlock) void k_d2s_vec(int n, const double* __restrict__ a,
                                                    float* __restrict__ b) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) b[i] = (float)a[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero(int n, T* __restrict__ a) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) a[i] = T(0);
}

// Reliable updates: the FP32 CG step is k_cg_x; its finalisation (FIN_RU_X)
// requests a residual replacement once the recursive residual has dropped
// by delta since the last one.
// x64 += x32; x32 = 0
__global__ __launch_bounds__(kBlock) void k_ru_accum(int n, double* __restrict__ x,
                                                     float* __restrict__ x32) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] = x[i] + (double)x32[i];
    x32[i] = 0.0f;
  }
}

// r32 = (float)r64; z = d.*r32; partial r32.z
__global__ __launch_bounds__(kBlock) void k_ru_reset(int n, const double* __restrict__ r64,
                                                     float* r, const float* __restrict__ d,
                                                     float* z, double* partials) {
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    float ri = (float)r64[i];
    r[i] = ri;
    float zi = d ? d[i] * ri : ri;
    if (d) z[i] = zi;
    v[0] += (double)ri * (double)zi;
  }
  store_partials<1>(v, partials);
}

// --------------------------------------------------------------- host driver
enum { FMT_HLL = 0, FMT_ELL = 1, FMT_CSR = 2 };

struct Solver {
  int method = 0, fmt = 0, n = 0, etype = 0;
  void* mat = nullptr;
  bool use_single = false;       // HLL FP32 copy of a double matrix
  int rpl = 1;                   // rows per lane of the product kernels (1 or 2)
  KState* st = nullptr;          // device
  KState* hst = nullptr;         // pinned host mirror
  KState* ring = nullptr;        // pinned, two state copies of the pipelined poll
  hipEvent_t ev[2] = {nullptr, nullptr};
  int posted = 0;                // state copies queued so far
  double* partials = nullptr;
  double* hist = nullptr;
  std::vector<void*> bufs;
  hipStream_t stream = nullptr;
  // allocation-free mode: device buffers are carved from the library's
  // scratch array and the host mirror and events are the context's
  char* arena = nullptr;
  size_t arena_left = 0;
  bool borrowed = false;

  int alloc(void** p, size_t bytes) {
    if (borrowed) {
      const size_t b = (bytes + 255) & ~(size_t)255;
      if (b > arena_left) return set_error(PSB_HIP_INVALID, "krylov", "scratch exhausted");
      *p = arena;
      arena += b;
      arena_left -= b;
      return PSB_HIP_SUCCESS;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return set_error(PSB_HIP_OUTOFMEM, "krylov", hipGetErrorString(e));
    bufs.push_back(*p);
    return PSB_HIP_SUCCESS;
  }
  ~Solver() {
    for (void* p : bufs) (void)hipFree(p);
    if (borrowed) return;
    if (hst) (void)hipHostFree(hst);
    if (ring) (void)hipHostFree(ring);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device bytes one fused solve takes: state, partials, history and the work
// vectors (CG: r p q z; BiCGSTAB: r q p v f s z t).
static size_t solve_bytes(int method, int n, int etype, int itmax) {
  const size_t vec = align256(elem_size(etype) * (size_t)(n > 0 ? n : 1));
  return align256(sizeof(KState)) + align256(sizeof(double) * 2 * kMaxGrid) +
         align256(sizeof(double) * ((size_t)itmax + 2)) + (method == 0 ? 4 : 8) * vec;
}
static_assert(3 * sizeof(KState) <= kScratchHostBytes, "host mirror + ring in the pinned block");

template <typename T>
struct Mats {
  HllView<T> hll;
  EllView<T> ell;
};

template <typename T>
static Mats<T> make_views(Solver& s) {
  Mats<T> m{};
  if (s.fmt == FMT_HLL) {
    HllDev* h = (HllDev*)s.mat;
    m.hll.nr = h->rows;
    m.hll.hksz = h->hksz;
    m.hll.hkoffs = h->hkoffs;
    m.hll.ja = h->ja;
    m.hll.val = s.use_single ? (const T*)h->val32 : (const T*)h->val;
    const bool i16 = hll_use_idx16(h);
    m.hll.res16 = i16 ? h->res16 : nullptr;
    m.hll.base16 = i16 ? h->base16 : nullptr;
    s.rpl = sizeof(T) == 8 && h->hksz % 2 == 0 ? 2 : 1;
  } else if (s.fmt == FMT_ELL) {
    EllDev* e = (EllDev*)s.mat;
    m.ell.nr = e->rows;
    m.ell.pitch = e->pitch;
    m.ell.nzm = e->nzm;
    m.ell.ja = e->ja;
    m.ell.val = (const T*)e->val;
    const bool i16 = ell_use_idx16(e);
    m.ell.res16 = i16 ? e->res16 : nullptr;
    m.ell.base16 = i16 ? e->base16 : nullptr;
    s.rpl = sizeof(T) == 8 && e->pitch % 2 == 0 ? 2 : 1;
  } else {
    s.rpl = 1;
  }
  return m;
}

// Product kernel KERN<T, VT, R> with the solver's rows per lane R.
#define PSB_PRODUCT(KERN, T, VT, s, g, ...)                                                  \
  do {                                                                                       \
    if ((s).rpl == 2)                                                                        \
      hipLaunchKernelGGL((KERN<T, VT, 2>), dim3(g), dim3(kBlock), 0, (s).stream, __VA_ARGS__); \
    else                                                                                     \
      hipLaunchKernelGGL((KERN<T, VT, 1>), dim3(g), dim3(kBlock), 0, (s).stream, __VA_ARGS__); \
  } while (0)

// The same over an HLL view, with the hack size 64 as a constant when it is.
#define PSB_HLL_PRODUCT(KERN, T, s, g, view, ...)                                          \
  do {                                                                                     \
    if ((view).hksz == 64) PSB_PRODUCT(KERN, T, HllView64<T>, s, g, as64(view), __VA_ARGS__); \
    else PSB_PRODUCT(KERN, T, HllView<T>, s, g, view, __VA_ARGS__);                         \
  } while (0)

// One-block finalisation after a fused kernel launched on nb blocks.
template <int OP, int NP>
static void fin(Solver& s, int nb, const double* aux = nullptr) {
  hipLaunchKernelGGL((k_fin<OP, NP>), dim3(1), dim3(kFin), 0, s.stream, s.st, s.partials, nb,
                     s.hist, aux);
}

static int csr_product(Solver& s, const void* x, void* y) {
  // y = A*x through the CSR entry point (vector CSR + long rows)
  MultiVec mx{1, s.n, s.n, s.etype, const_cast<void*>(x)};
  MultiVec my{1, s.n, s.n, s.etype, y};
  return psb_hip_csr_spmv(s.mat, 1.0, &mx, 0.0, &my);
}

// r = b - A x and the restart sums; CSR computes the residual through the
// product entry point first and passes it as b to an empty-matrix sweep.
template <typename T>
static int restart(Solver& s, const Mats<T>& m, const T* b, const T* x, T* r, T* q, const T* d,
                   T* z, int bicg) {
  const int n = s.n;
  int g = spmv_grid(n, s.rpl);
  dim3 blk(kBlock);
  if (s.fmt == FMT_HLL) {
    PSB_HLL_PRODUCT(k_restart, T, s, g, m.hll, b, x, r, q, d, z, s.st, s.partials, bicg);
  } else if (s.fmt == FMT_ELL) {
    PSB_PRODUCT(k_restart, T, EllView<T>, s, g, m.ell, b, x, r, q, d, z, s.st, s.partials, bicg);
  } else {
    MultiVec mx{1, n, n, s.etype, (void*)x}, mr{1, n, n, s.etype, r}, mb{1, n, n, s.etype, (void*)b};
    int rc = psb_hip_mv_axpby(n, 1.0, &mb, 0.0, &mr);
    if (rc) return rc;
    rc = psb_hip_csr_spmv(s.mat, -1.0, &mx, 1.0, &mr);
    if (rc) return rc;
    EllView<T> zero{n, n, 0, nullptr, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL((k_restart<T, EllView<T>, 1>), dim3(g), blk, 0, s.stream, zero,
                       (const T*)r, x, r, q, d, z, s.st, s.partials, bicg);
  }
  PSB_HIP_LAUNCH_CHECK();
  if (bicg) fin<FIN_RESTART_BI, 2>(s, g);
  else fin<FIN_RESTART_CG, 2>(s, g);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

// One CG iteration (queued, no host synchronisation).
template <typename T>
static int cg_iter(Solver& s, const Mats<T>& m, T* x, T* r, T* p, T* q, const T* d, T* z) {
  const int n = s.n;
  dim3 blk(kBlock), gv(vec_grid<T>(n));
  const int gs = spmv_grid(n, s.rpl);
  hipLaunchKernelGGL((k_cg_p<T>), gv, blk, 0, s.stream, n, (const T*)z, p, s.st);
  if (s.fmt == FMT_HLL) {
    PSB_HLL_PRODUCT(k_cg_q, T, s, gs, m.hll, (const T*)p, q, s.st, s.partials);
    fin<FIN_CG_Q, 1>(s, gs);
  } else if (s.fmt == FMT_ELL) {
    PSB_PRODUCT(k_cg_q, T, EllView<T>, s, gs, m.ell, (const T*)p, q, s.st, s.partials);
    fin<FIN_CG_Q, 1>(s, gs);
  } else {
    int rc = csr_product(s, p, q);
    if (rc) return rc;
    hipLaunchKernelGGL((k_dot<T>), gv, blk, 0, s.stream, n, (const T*)p, (const T*)q, s.st,
                       s.partials);
    fin<FIN_CG_Q, 1>(s, gv.x);
  }
  hipLaunchKernelGGL((k_cg_x<T>), gv, blk, 0, s.stream, n, x, r, (const T*)p, (const T*)q, d, z,
                     s.st, s.partials);
  fin<FIN_CG_X, 2>(s, gv.x);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

template <typename T>
static int bi_iter(Solver& s, const Mats<T>& m, T* x, T* r, T* q, T* p, T* v, T* f, T* sv, T* z,
                   T* t, const T* d) {
  const int n = s.n;
  dim3 blk(kBlock), gv(vec_grid<T>(n));
  const int gs = spmv_grid(n, s.rpl);
  hipLaunchKernelGGL((k_bi_a<T>), gv, blk, 0, s.stream, n, (const T*)r, p, (const T*)v, d, f,
                     s.st);
  if (s.fmt == FMT_HLL) {
    PSB_HLL_PRODUCT(k_bi_v, T, s, gs, m.hll, (const T*)f, v, (const T*)q, s.st,
                s.partials);
    fin<FIN_BI_V, 1>(s, gs);
  } else if (s.fmt == FMT_ELL) {
    PSB_PRODUCT(k_bi_v, T, EllView<T>, s, gs, m.ell, (const T*)f, v, (const T*)q, s.st,
                s.partials);
    fin<FIN_BI_V, 1>(s, gs);
  } else {
    int rc = csr_product(s, f, v);
    if (rc) return rc;
    hipLaunchKernelGGL((k_dot<T>), gv, blk, 0, s.stream, n, (const T*)q, (const T*)v, s.st,
                       s.partials);
    fin<FIN_BI_V, 1>(s, gv.x);
  }
  hipLaunchKernelGGL((k_bi_s<T>), gv, blk, 0, s.stream, n, (const T*)r, (const T*)v, d, sv, z,
                     s.st);
  if (s.fmt == FMT_HLL) {
    PSB_HLL_PRODUCT(k_bi_t, T, s, gs, m.hll, (const T*)z, t, (const T*)sv, s.st,
                s.partials);
    fin<FIN_BI_T, 2>(s, gs);
  } else if (s.fmt == FMT_ELL) {
    PSB_PRODUCT(k_bi_t, T, EllView<T>, s, gs, m.ell, (const T*)z, t, (const T*)sv, s.st,
                s.partials);
    fin<FIN_BI_T, 2>(s, gs);
  } else {
    int rc = csr_product(s, z, t);
    if (rc) return rc;
    hipLaunchKernelGGL((k_dot2<T>), gv, blk, 0, s.stream, n, (const T*)t, (const T*)sv, s.st,
                       s.partials);
    fin<FIN_BI_T, 2>(s, gv.x);
  }
  hipLaunchKernelGGL((k_bi_x<T>), gv, blk, 0, s.stream, n, x, r, (const T*)f, (const T*)z,
                     (const T*)sv, (const T*)t, (const T*)q, s.st, s.partials);
  fin<FIN_BI_X, 2>(s, gv.x);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

static int read_state(Solver& s) {
  PSB_HIP_CHECK(hipMemcpyAsync(s.hst, s.st, sizeof(KState), hipMemcpyDeviceToHost, s.stream));
  PSB_HIP_CHECK(hipStreamSynchronize(s.stream));
  return PSB_HIP_SUCCESS;
}

// Pipelined poll.  After queueing a batch of iterations the host queues a
// copy of the state (post_state) and then waits only for the copy queued one
// batch earlier (prev_halted), so the next batch is already queued while the
// host looks at the previous one.  Batches queued after the stop are no-ops
// (every kernel returns on the stop flag); read_state then gives the final
// state.
static int post_state(Solver& s) {
  const int k = s.posted & 1;
  PSB_HIP_CHECK(hipMemcpyAsync(&s.ring[k], s.st, sizeof(KState), hipMemcpyDeviceToHost, s.stream));
  PSB_HIP_CHECK(hipEventRecord(s.ev[k], s.stream));
  s.posted++;
  return PSB_HIP_SUCCESS;
}

// Waits for the state copy queued before the last one; *halted = its stop or
// breakdown flag.  Needs two copies posted since the last read_state.
static int prev_halted(Solver& s, bool* halted) {
  const int k = (s.posted - 2) & 1;
  PSB_HIP_CHECK(hipEventSynchronize(s.ev[k]));
  *halted = s.ring[k].done != 0 || s.ring[k].breakdown != 0;
  return PSB_HIP_SUCCESS;
}

// ||b|| into the state (errden for istop 2)
template <typename T>
static int bnorm(Solver& s, const T* b) {
  dim3 gv(vec_grid<T>(s.n));
  hipLaunchKernelGGL((k_sumsq<T>), gv, dim3(kBlock), 0, s.stream, s.n, b, s.partials);
  fin<FIN_BNORM, 1>(s, gv.x);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

// The solve loop: restart, then batches of check_every queued iterations.
template <typename T>
static int solve(Solver& s, const T* b, T* x, const T* d, double eps, int itmax, int istop,
                 int check_every, int* iter, double* err) {
  const int n = s.n;
  int nvec = s.method == 0 ? 4 : 8;  // CG: r p q z ; BiCGSTAB: r q p v f s z t
  T* w[8] = {};
  for (int k = 0; k < nvec; k++) {
    int rc = s.alloc((void**)&w[k], sizeof(T) * (size_t)(n > 0 ? n : 1));
    if (rc) return rc;
  }
  // the initial state goes through the pinned mirror (an async copy must not
  // read a stack object that may be gone when the copy runs)
  *s.hst = KState{};
  s.hst->eps = eps;
  s.hst->itmax = itmax;
  s.hst->istop = istop;
  s.hst->errden = 1.0;
  PSB_HIP_CHECK(hipMemcpyAsync(s.st, s.hst, sizeof(KState), hipMemcpyHostToDevice, s.stream));
  Mats<T> m = make_views<T>(s);
  int rc = bnorm<T>(s, b);
  if (rc) return rc;
  if (check_every < 1) check_every = 8;
  for (;;) {
    // restart: r = b - A x (and the first z / rho); without a preconditioner
    // the vectors it would only copy are aliased (z = r, f = p, z = s)
    if (s.method == 0) {
      T* z = d ? w[3] : w[0];
      rc = restart<T>(s, m, b, x, w[0], nullptr, d, z, 0);
    } else {
      rc = restart<T>(s, m, b, x, w[0], w[1], d, nullptr, 1);
    }
    if (rc) return rc;
    auto batch = [&]() -> int {
      for (int k = 0; k < check_every; k++) {
        int rk;
        if (s.method == 0) {
          T* z = d ? w[3] : w[0];
          rk = cg_iter<T>(s, m, x, w[0], w[1], w[2], d, z);
        } else {
          T* f = d ? w[4] : w[2];
          T* z = d ? w[6] : w[5];
          rk = bi_iter<T>(s, m, x, w[0], w[1], w[2], w[3], f, w[5], z, w[7], d);
        }
        if (rk) return rk;
      }
      return post_state(s);
    };
    s.posted = 0;
    if ((rc = batch())) return rc;
    bool stop = false;
    while (!stop) {
      if ((rc = batch())) return rc;
      if ((rc = prev_halted(s, &stop))) return rc;
    }
    if ((rc = read_state(s))) return rc;
    if (s.hst->done) break;
    // breakdown: restart unless the iteration budget is spent
    if (s.hst->itx >= itmax) break;
    s.hst->breakdown = 0;
    PSB_HIP_CHECK(hipMemcpyAsync(s.st, s.hst, sizeof(KState), hipMemcpyHostToDevice, s.stream));
  }
  *iter = s.hst->itx;
  *err = s.hst->errden > 0.0 ? s.hst->errnum / s.hst->errden : s.hst->errnum;
  return PSB_HIP_SUCCESS;
}

// The matrix side of setup: format, size and element type.
static int describe(Solver& s, int method, int fmt, void* dmat) {
  if (!dmat) return set_error(PSB_HIP_INVALID, "krylov", "no matrix");
  s.method = method;
  s.fmt = fmt;
  s.mat = dmat;
  s.stream = ctx().stream;
  if (fmt == FMT_HLL) {
    HllDev* h = (HllDev*)dmat;
    if (h->perm)
      return set_error(PSB_HIP_UNSUPPORTED, "krylov",
                       "row-sorted HLL: the fused solvers take rows in order");
    s.n = h->rows;
    s.etype = h->elem_type;
  } else if (fmt == FMT_ELL) {
    EllDev* e = (EllDev*)dmat;
    s.n = e->rows;
    s.etype = e->elem_type;
  } else if (fmt == FMT_CSR) {
    CsrDev* c = (CsrDev*)dmat;
    s.n = c->rows;
    s.etype = c->elem_type;
  } else {
    return set_error(PSB_HIP_UNSUPPORTED, "krylov", "format");
  }
  return PSB_HIP_SUCCESS;
}

// State, partials and history buffers plus the host mirror and events
// (carved from the scratch array and borrowed from the context when
// s.borrowed is set).
static int setup_state(Solver& s, int itmax) {
  int rc = s.alloc((void**)&s.st, sizeof(KState));
  if (rc) return rc;
  rc = s.alloc((void**)&s.partials, sizeof(double) * 2 * kMaxGrid);
  if (rc) return rc;
  rc = s.alloc((void**)&s.hist, sizeof(double) * ((size_t)itmax + 2));
  if (rc) return rc;
  if (s.borrowed) {
    s.hst = (KState*)ctx().scratch_host;
    s.ring = s.hst + 1;
    s.ev[0] = ctx().scratch_ev[0];
    s.ev[1] = ctx().scratch_ev[1];
    return PSB_HIP_SUCCESS;
  }
  PSB_HIP_CHECK(hipHostMalloc((void**)&s.hst, sizeof(KState), hipHostMallocDefault));
  PSB_HIP_CHECK(hipHostMalloc((void**)&s.ring, 2 * sizeof(KState), hipHostMallocDefault));
  for (hipEvent_t& e : s.ev) PSB_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return PSB_HIP_SUCCESS;
}

static int setup(Solver& s, int method, int fmt, void* dmat, int itmax) {
  int rc = describe(s, method, fmt, dmat);
  if (rc) return rc;
  return setup_state(s, itmax);
}

static int mv_check(const MultiVec* v, int etype, int n) {
  if (!v || v->elem_type != etype || v->size < n || v->count != 1) return PSB_HIP_INVALID;
  return PSB_HIP_SUCCESS;
}

// FP64 residual of the refinement loops: r64 = b - A64 x (||r64|| into the
// outer state's errnum), read back.
static int fp64_residual(Solver& outer, const Mats<double>& m64, const double* b, const double* x,
                         double* r64) {
  int rc = restart<double>(outer, m64, b, x, r64, nullptr, nullptr, r64, 0);
  if (rc) return rc;
  return read_state(outer);
}

// FP32 CG with FP64 reliable updates (the Krylov recurrence is not restarted
// at a residual replacement).  x (double) holds the initial guess.
static int mixed_reliable(HllDev* h, const double* b, double* x, const double* d64, double eps,
                          int outer_itmax, double delta, int inner_itmax, int check_every,
                          int* outer_its, int* inner_its, double* err) {
  const int n = h->rows;
  Solver outer;
  int rc = setup(outer, 0, FMT_HLL, h, 1);
  if (rc) return rc;
  Solver inner;
  rc = setup(inner, 0, FMT_HLL, h, inner_itmax);
  if (rc) return rc;
  inner.use_single = true;
  inner.etype = PSB_HIP_TYPE_FLOAT;
  double* r64 = nullptr;
  float *x32 = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *d = nullptr;
  size_t fb = sizeof(float) * (size_t)n;
  if ((rc = outer.alloc((void**)&r64, sizeof(double) * (size_t)n))) return rc;
  if ((rc = outer.alloc((void**)&x32, fb))) return rc;
  if ((rc = outer.alloc((void**)&r, fb))) return rc;
  if ((rc = outer.alloc((void**)&p, fb))) return rc;
  if ((rc = outer.alloc((void**)&q, fb))) return rc;
  if (d64) {
    if ((rc = outer.alloc((void**)&z, fb))) return rc;
    if ((rc = outer.alloc((void**)&d, fb))) return rc;
  } else {
    z = r;
  }
  hipStream_t stream = outer.stream;
  dim3 blk(kBlock), gs(stream_grid(n));
  if (d64) hipLaunchKernelGGL(k_d2s_vec, gs, blk, 0, stream, n, d64, d);
  hipLaunchKernelGGL((k_zero<float>), gs, blk, 0, stream, n, x32);
  // outer state: ||b||, and the FP64 residuals
  *outer.hst = KState{};
  outer.hst->itmax = 1 << 30;
  outer.hst->istop = 2;
  outer.hst->errden = 1.0;
  PSB_HIP_CHECK(hipMemcpyAsync(outer.st, outer.hst, sizeof(KState), hipMemcpyHostToDevice, stream));
  if ((rc = bnorm<double>(outer, b))) return rc;
  Mats<double> m64 = make_views<double>(outer);
  Mats<float> m32 = make_views<float>(inner);
  if ((rc = fp64_residual(outer, m64, b, x, r64))) return rc;
  const double bnrm = outer.hst->bnorm;
  // inner state: eps*errden is the absolute target, delta the replacement factor
  *inner.hst = KState{};
  inner.hst->eps = eps;
  inner.hst->errden = bnrm;
  inner.hst->delta = delta;
  inner.hst->itmax = inner_itmax;
  inner.hst->istop = 2;
  PSB_HIP_CHECK(hipMemcpyAsync(inner.st, inner.hst, sizeof(KState), hipMemcpyHostToDevice, stream));
  auto reset = [&](bool first) {
    hipLaunchKernelGGL(k_ru_reset, gs, blk, 0, stream, n, (const double*)r64, r,
                       (const float*)d, z, inner.partials);
    if (first) fin<FIN_RU_RESET_FIRST, 1>(inner, gs.x, &outer.st->errnum);
    else fin<FIN_RU_RESET, 1>(inner, gs.x, &outer.st->errnum);
  };
  reset(true);
  PSB_HIP_LAUNCH_CHECK();
  int updates = 0;
  double rnorm = outer.hst->errnum;
  bool conv = bnrm == 0.0 ? rnorm <= eps : rnorm <= eps * bnrm;
  if (check_every < 1) check_every = 8;
  dim3 gv(vec_grid<float>(n));
  const int gsp = spmv_grid(n, inner.rpl);
  while (!conv) {
    // queue FP32 iterations until a replacement (or breakdown) is requested
    auto batch = [&]() -> int {
      for (int k = 0; k < check_every; k++) {
        hipLaunchKernelGGL((k_cg_p<float>), gv, blk, 0, stream, n, (const float*)z, p, inner.st);
        if (m32.hll.hksz == 64)
          hipLaunchKernelGGL((k_cg_q<float, HllView64<float>, 1>), dim3(gsp), blk, 0, stream,
                             as64(m32.hll), (const float*)p, q, inner.st, inner.partials);
        else
          hipLaunchKernelGGL((k_cg_q<float, HllView<float>, 1>), dim3(gsp), blk, 0, stream,
                             m32.hll, (const float*)p, q, inner.st, inner.partials);
        fin<FIN_CG_Q, 1>(inner, gsp);
        hipLaunchKernelGGL((k_cg_x<float>), gv, blk, 0, stream, n, x32, r, (const float*)p,
                           (const float*)q, (const float*)d, z, inner.st, inner.partials);
        fin<FIN_RU_X, 2>(inner, gv.x);
      }
      PSB_HIP_LAUNCH_CHECK();
      return post_state(inner);
    };
    inner.posted = 0;
    if ((rc = batch())) return rc;
    bool stop = false;
    while (!stop) {
      if ((rc = batch())) return rc;
      if ((rc = prev_halted(inner, &stop))) return rc;
    }
    if ((rc = read_state(inner))) return rc;
    // replacement: x += x32 in FP64, r = b - A x in FP64
    hipLaunchKernelGGL(k_ru_accum, gs, blk, 0, stream, n, x, x32);
    PSB_HIP_LAUNCH_CHECK();
    if ((rc = fp64_residual(outer, m64, b, x, r64))) return rc;
    updates++;
    rnorm = outer.hst->errnum;
    conv = bnrm == 0.0 ? rnorm <= eps : rnorm <= eps * bnrm;
    if (conv || updates >= outer_itmax || inner.hst->itx >= inner_itmax) break;
    // a true breakdown (rho_old == 0) restarts the recurrence from the new residual
    reset(inner.hst->breakdown == 1);
    PSB_HIP_LAUNCH_CHECK();
  }
  *outer_its = updates;
  *inner_its = inner.hst->itx;
  *err = bnrm > 0.0 ? rnorm / bnrm : rnorm;
  return PSB_HIP_SUCCESS;
}

// Restarted refinement: each cycle solves A32 d = r/||r|| from zero.
static int mixed_restart(int method, HllDev* h, const double* b, double* x, const double* d64,
                         double eps, int outer_itmax, double inner_eps, int inner_itmax,
                         int check_every, int* outer_its, int* inner_its, double* err) {
  const int n = h->rows;
  hipStream_t stream = ctx().stream;
  Solver outer;
  int rc = setup(outer, 0, FMT_HLL, h, 1);
  if (rc) return rc;
  double* r = nullptr;
  float *r32 = nullptr, *d32 = nullptr, *x32 = nullptr;
  if ((rc = outer.alloc((void**)&r, sizeof(double) * n))) return rc;
  if ((rc = outer.alloc((void**)&r32, sizeof(float) * n))) return rc;
  if ((rc = outer.alloc((void**)&x32, sizeof(float) * n))) return rc;
  if (d64) {
    if ((rc = outer.alloc((void**)&d32, sizeof(float) * n))) return rc;
    hipLaunchKernelGGL(k_d2s_vec, dim3(stream_grid(n)), dim3(kBlock), 0, stream, n, d64, d32);
    PSB_HIP_LAUNCH_CHECK();
  }
  Mats<double> m64 = make_views<double>(outer);
  *outer.hst = KState{};
  outer.hst->itmax = 1 << 30;
  outer.hst->istop = 2;
  outer.hst->errden = 1.0;
  PSB_HIP_CHECK(hipMemcpyAsync(outer.st, outer.hst, sizeof(KState), hipMemcpyHostToDevice, stream));
  if ((rc = bnorm<double>(outer, b))) return rc;
  int total_inner = 0, outer_k = 0;
  double bnrm = 0.0, rnorm = 0.0;
  for (;;) {
    if ((rc = fp64_residual(outer, m64, b, x, r))) return rc;
    bnrm = outer.hst->bnorm;
    rnorm = outer.hst->errnum;
    bool conv = bnrm == 0.0 ? rnorm <= eps : rnorm <= eps * bnrm;
    if (conv || outer_k >= outer_itmax || rnorm == 0.0) break;
    outer_k++;
    // inner tolerance: inner_eps, no tighter than the outer target needs
    double ie = inner_eps;
    if (bnrm > 0.0) {
      double need = 0.5 * eps * bnrm / rnorm;
      if (need > ie) ie = need < 0.5 ? need : 0.5;
    }
    hipLaunchKernelGGL(k_to_single, dim3(stream_grid(n)), dim3(kBlock), 0, stream, n,
                       (const double*)r, (const double*)&outer.st->errnum, r32);
    hipLaunchKernelGGL((k_zero<float>), dim3(stream_grid(n)), dim3(kBlock), 0, stream, n, x32);
    PSB_HIP_LAUNCH_CHECK();
    Solver inner;
    if ((rc = setup(inner, method == 0 ? 0 : 1, FMT_HLL, h, inner_itmax))) return rc;
    inner.use_single = true;
    inner.etype = PSB_HIP_TYPE_FLOAT;
    int it = 0;
    double ierr = 0.0;
    rc = solve<float>(inner, r32, x32, d64 ? d32 : nullptr, ie, inner_itmax, 2, check_every, &it,
                      &ierr);
    if (rc) return rc;
    total_inner += it;
    // x += ||r|| * d  (FP64)
    hipLaunchKernelGGL(k_update_double, dim3(stream_grid(n)), dim3(kBlock), 0, stream, n, x,
                       (const float*)x32, (const double*)&outer.st->errnum);
    PSB_HIP_LAUNCH_CHECK();
  }
  *outer_its = outer_k;
  *inner_its = total_inner;
  *err = bnrm > 0.0 ? rnorm / bnrm : rnorm;
  return PSB_HIP_SUCCESS;
}

static int fused_args(const char* who, int istop, int itmax) {
  if (istop != 2 && istop != 3)
    return set_error(PSB_HIP_UNSUPPORTED, who, "istop must be 2 or 3 in the fused solver");
  if (itmax < 0) return set_error(PSB_HIP_INVALID, who, "itmax");
  return PSB_HIP_SUCCESS;
}

// The body of both fused entry points once the solver's buffers are set up.
static int fused_run(const char* who, Solver& s, void* diag, void* bv, void* xv, double eps,
                     int itmax, int istop, int check_every, int* iter, double* err,
                     double* hist) {
  int rc;
  MultiVec *b = (MultiVec*)bv, *x = (MultiVec*)xv, *d = (MultiVec*)diag;
  if (mv_check(b, s.etype, s.n) || mv_check(x, s.etype, s.n) || (d && mv_check(d, s.etype, s.n)))
    return set_error(PSB_HIP_INVALID, who, "b/x/diag type or size");
  if (s.n == 0) {
    *iter = 0;
    *err = 0.0;
    return PSB_HIP_SUCCESS;
  }
  if (s.etype == PSB_HIP_TYPE_DOUBLE)
    rc = solve<double>(s, (const double*)b->v, (double*)x->v, d ? (const double*)d->v : nullptr,
                       eps, itmax, istop, check_every, iter, err);
  else
    rc = solve<float>(s, (const float*)b->v, (float*)x->v, d ? (const float*)d->v : nullptr, eps,
                      itmax, istop, check_every, iter, err);
  if (rc) return rc;
  if (hist && *iter > 0) {
    PSB_HIP_CHECK(hipMemcpyAsync(hist, s.hist + 1, sizeof(double) * (size_t)*iter,
                                 hipMemcpyDeviceToHost, s.stream));
    PSB_HIP_CHECK(hipStreamSynchronize(s.stream));
  }
  return PSB_HIP_SUCCESS;
}

}  // namespace psbhip

using namespace psbhip;

extern "C" {

int psb_hip_krylov_fused(int method, int fmt, void* dmat, void* diag, void* bv, void* xv,
                         double eps, int itmax, int istop, int check_every, int* iter,
                         double* err, double* hist) {
  int rc = ensure_init();
  if (rc) return rc;
  if ((rc = fused_args(__func__, istop, itmax))) return rc;
  Solver s;
  rc = setup(s, method == 0 ? 0 : 1, fmt, dmat, itmax);
  if (rc) return rc;
  return fused_run(__func__, s, diag, bv, xv, eps, itmax, istop, check_every, iter, err, hist);
}

int psb_hip_krylov_fused_noalloc(int method, int fmt, void* dmat, void* diag, void* bv, void* xv,
                                 double eps, int itmax, int istop, int check_every, int* iter,
                                 double* err, double* hist) {
  Context& c = ctx();
  if (!c.initialized || !c.scratch)
    return set_error(PSB_HIP_INVALID, __func__,
                     "the library is not initialised (call psb_hip_init first; this entry point "
                     "allocates nothing, initialisation included)");
  int rc = fused_args(__func__, istop, itmax);
  if (rc) return rc;
  Solver s;
  s.borrowed = true;
  if ((rc = describe(s, method == 0 ? 0 : 1, fmt, dmat))) return rc;
  const size_t need = solve_bytes(s.method, s.n, s.etype, itmax);
  if (need > c.scratch_bytes) {
    char msg[256];
    snprintf(msg, sizeof(msg),
             "needs %zu bytes of scratch (n = %d, itmax = %d), over the %zu-byte maximum; "
             "use psb_hip_krylov_fused",
             need, s.n, itmax, c.scratch_bytes);
    return set_error(PSB_HIP_INVALID, __func__, msg);
  }
  s.arena = c.scratch;
  s.arena_left = c.scratch_bytes;
  if ((rc = setup_state(s, itmax))) return rc;
  return fused_run(__func__, s, diag, bv, xv, eps, itmax, istop, check_every, iter, err, hist);
}

int psb_hip_krylov_mixed(int method, void* dmat, void* diag, void* bv, void* xv, double eps,
                         int outer_itmax, double inner_eps, int inner_itmax, int check_every,
                         int variant, int* outer_its, int* inner_its, double* err) {
  int rc = ensure_init();
  if (rc) return rc;
  HllDev* h = (HllDev*)dmat;
  if (h->elem_type != PSB_HIP_TYPE_DOUBLE || !h->val32)
    return set_error(PSB_HIP_INVALID, __func__, "needs a double HLL matrix with a single copy");
  if (h->perm)
    return set_error(PSB_HIP_UNSUPPORTED, __func__, "row-sorted HLL: rows must be in order");
  const int n = h->rows;
  MultiVec *b = (MultiVec*)bv, *x = (MultiVec*)xv, *d = (MultiVec*)diag;
  if (mv_check(b, PSB_HIP_TYPE_DOUBLE, n) || mv_check(x, PSB_HIP_TYPE_DOUBLE, n) ||
      (d && mv_check(d, PSB_HIP_TYPE_DOUBLE, n)))
    return set_error(PSB_HIP_INVALID, __func__, "b/x/diag type or size");
  if (n == 0) {
    *outer_its = *inner_its = 0;
    *err = 0.0;
    return PSB_HIP_SUCCESS;
  }
  const double* d64 = d ? (const double*)d->v : nullptr;
  if (variant == 1) {
    if (method != 0)
      return set_error(PSB_HIP_UNSUPPORTED, __func__, "reliable updates are implemented for CG");
    return mixed_reliable(h, (const double*)b->v, (double*)x->v, d64, eps, outer_itmax, inner_eps,
                          inner_itmax, check_every, outer_its, inner_its, err);
  }
  return mixed_restart(method, h, (const double*)b->v, (double*)x->v, d64, eps, outer_itmax,
                       inner_eps, inner_itmax, check_every, outer_its, inner_its, err);
}

}  // extern "C"
