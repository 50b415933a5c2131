// Device-resident Krylov solvers: CG and BiCGSTAB with the DIAG (Jacobi)
// preconditioner, and FP32-inner / FP64-outer iterative refinement.
//
// The operation sequence is the reference's (krylov/psb_dcg.F90:241-294,
// krylov/psb_dcgstab.f90:248-364, convergence krylov/psb_d_krylov_conv_mod.f90:
// 279-356), regrouped into a few fused kernels per iteration.  Every scalar
// (rho, sigma, alpha, beta, omega, the residual norm, the stop flag) lives
// in device memory and is produced by the last block of the reduction that
// defines it, so iterations are queued without host round trips; the host
// reads the stop flag once every `check_every` iterations.  Kernels queued
// after convergence see the flag and return at once, so the iterate and the
// iteration count are exactly those of the stopping iteration.
//
// CG per iteration (3 kernels):
//   K_p   p = z (first) | p = z + beta*p
//   K_q   q = A*p with the epilogue sum p.q -> sigma, alpha = rho/sigma
//   K_x   x += alpha*p; r -= alpha*q; z = d.*r; sums r.r, r.z ->
//         ||r||, stop test, rho_old = rho, rho = r.z, beta = rho/rho_old
// BiCGSTAB per iteration (5 kernels):
//   K_a   p = r (first) | p = r + beta*(p - omega*v); f = d.*p
//   K_v   v = A*f, sum q.v -> alpha = rho/(q.v)
//   K_s   s = r - alpha*v; z = d.*s
//   K_t   t = A*z, sums t.t, t.s -> omega
//   K_x   x = (x + alpha*f) + omega*z; r = s - omega*t; sums r.r, q.r ->
//         ||r||, stop test, rho, beta
// Reduction partials are per block of a bounded grid (<= kMaxGrid blocks, a
// fixed split of the rows), summed by the last block in index order:
// results are deterministic run to run.
#include <math.h>
#include <string.h>

#include <vector>

#include "psb_hip_internal.hpp"

namespace psbhip {

constexpr int kMaxGrid = 2048;   // blocks of every fused kernel (multiple of 8)

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
  int breakdown;  // restart requested
  int istop;
  int pad[2];
  unsigned int counter[8];   // grid-reduction arrival counters (kept at 0)
};

// --------------------------------------------------------------- views
template <typename T>
struct HllView {
  int nr, hksz;
  const int* hkoffs;
  const int* ja;
  const T* val;
  __device__ __forceinline__ T row(int r, const T* __restrict__ x) const {
    const int hk = r / hksz;
    const int lane = r - hk * hksz;
    const int off0 = hkoffs[hk];
    const int w = (hkoffs[hk + 1] - off0) / hksz;
    return slot_dot<T, T>(ja + off0 + lane, val + off0 + lane, hksz, w, x);
  }
  __device__ __forceinline__ void row2(int r0, int r1, const T* __restrict__ x, T& a0,
                                       T& a1) const {
    const int h0 = r0 / hksz, h1 = r1 / hksz;
    const int o0 = hkoffs[h0], o1 = hkoffs[h1];
    const int w0 = (hkoffs[h0 + 1] - o0) / hksz, w1 = (hkoffs[h1 + 1] - o1) / hksz;
    const int l0 = r0 - h0 * hksz, l1 = r1 - h1 * hksz;
    slot_dot2<T, T>(ja + o0 + l0, val + o0 + l0, w0, ja + o1 + l1, val + o1 + l1, w1, hksz, x,
                    a0, a1);
  }
};

template <typename T>
struct EllView {
  int nr, pitch, nzm;
  const int* ja;
  const T* val;
  __device__ __forceinline__ T row(int r, const T* __restrict__ x) const {
    return slot_dot<T, T>(ja + r, val + r, pitch, nzm, x);
  }
  __device__ __forceinline__ void row2(int r0, int r1, const T* __restrict__ x, T& a0,
                                       T& a1) const {
    slot_dot2<T, T>(ja + r0, val + r0, nzm, ja + r1, val + r1, nzm, pitch, x, a0, a1);
  }
};

// Tile walk of the fused products: two 256-row tiles per step (rows row0 and
// row0 + 256 of this thread), clamped to the chunk so that every lane
// computes (the values of clamped rows are discarded).
#define PSB_FUSED_ROWS(M, X, BODY)                                              \
  for (int base = c.first; base < c.last; base += 2 * kBlock) {                 \
    const int row0 = base + threadIdx.x, row1 = row0 + kBlock;                  \
    const int rc0 = row0 < c.last ? row0 : c.last - 1;                          \
    const int rc1 = row1 < c.last ? row1 : c.last - 1;                          \
    T acc0, acc1;                                                               \
    (M).row2(rc0, rc1, (X), acc0, acc1);                                        \
    { const int row = row0; const T ax = acc0; if (row < c.last) { BODY } }     \
    { const int row = row1; const T ax = acc1; if (row < c.last) { BODY } }     \
  }

// Rows [first, last) owned by this block: the grid is bounded, each block
// walks a contiguous chunk of 256-row tiles (XCD-aware block order).
struct Chunk {
  int first, last;
};
__device__ __forceinline__ Chunk block_rows(int n) {
  const int ntiles = (n + kBlock - 1) / kBlock;
  const int nb = gridDim.x;
  const int lb = xcd_block(nb);
  const int per = (ntiles + nb - 1) / nb;
  int t0 = lb * per, t1 = t0 + per;
  if (t1 > ntiles) t1 = ntiles;
  if (t0 > ntiles) t0 = ntiles;
  Chunk c;
  c.first = t0 * kBlock;
  c.last = t1 * kBlock < n ? t1 * kBlock : n;
  return c;
}

inline int fused_grid(int n) {
  int tiles = (n + kBlock - 1) / kBlock;
  int g = tiles < kMaxGrid ? tiles : kMaxGrid;
  return xcd_grid(g < 1 ? 1 : g);
}

// Block reduction + last-block finalisation of NP sums.  partials holds
// NP * kMaxGrid doubles.  Returns true in the last block (all threads), with
// v[] holding the grid totals.
template <int NP>
__device__ __forceinline__ bool reduce_last(double (&v)[NP], double* partials, unsigned int* counter) {
  __shared__ double lds[kBlock / kWave];
  __shared__ bool am_last;
#pragma unroll
  for (int p = 0; p < NP; p++) {
    double s = block_sum(v[p], lds);
    if (threadIdx.x == 0) partials[p * kMaxGrid + blockIdx.x] = s;
  }
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned int prev = atomicAdd(counter, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
#pragma unroll
  for (int p = 0; p < NP; p++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
      s += __builtin_nontemporal_load(&partials[p * kMaxGrid + b]);
    v[p] = block_sum(s, lds);
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

__device__ __forceinline__ bool halted(const KState* st) {
  return st->done != 0 || st->breakdown != 0;
}

// stop test of psb_d_check_conv_vect for istop 2 and 3 (errden = ||b|| or ||r0||)
__device__ __forceinline__ void stop_test(KState* st, double rr, double* hist) {
  st->errnum = sqrt(rr);
  bool res;
  if (st->errden == 0.0) res = st->errnum <= st->eps;
  else res = st->errnum <= st->eps * st->errden;
  if (hist) hist[st->itx] = st->errden > 0.0 ? st->errnum / st->errden : st->errnum;
  if (res || st->itmax <= st->itx) st->done = 1;
}

// --------------------------------------------------------------- restart kernels
// r = b - A*x, optional copy q = r (BiCGSTAB shadow), z = d.*r; sums
// r.r and r.z (CG) or q.r (BiCGSTAB, = r.r).
template <typename T, class M>
__global__ __launch_bounds__(kBlock) void k_restart(M m, const T* __restrict__ b,
                                                    const T* __restrict__ x, T* __restrict__ r,
                                                    T* __restrict__ q, const T* __restrict__ d,
                                                    T* __restrict__ z, KState* st,
                                                    double* partials, int bicg) {
  if (st->done) return;
  const Chunk c = block_rows(m.nr);
  double rr = 0.0, rz = 0.0;
  PSB_FUSED_ROWS(m, x, {
    T ri = b[row] - ax;
    r[row] = ri;
    rr += (double)ri * (double)ri;
    if (bicg) {
      q[row] = ri;
    } else {
      T zi = d ? d[row] * ri : ri;
      if (z != r) z[row] = zi;
      rz += (double)ri * (double)zi;
    }
  })
  double v[2] = {rr, rz};
  if (reduce_last<2>(v, partials, &st->counter[0]) && threadIdx.x == 0) {
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
      st->rho = v[1];
    }
    st->rho_old = 0.0;
    // the first iteration's breakdown test (rho == 0 in BiCGSTAB)
    if (!st->done && bicg && st->rho == 0.0) {
      st->itx += 1;
      st->breakdown = 1;
    }
  }
}

// ||b|| for istop = 2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bnorm(int n, const T* __restrict__ b, KState* st,
                                                  double* partials) {
  const Chunk c = block_rows(n);
  double s = 0.0;
  for (int i = c.first + threadIdx.x; i < c.last; i += kBlock) s += (double)b[i] * (double)b[i];
  double v[1] = {s};
  if (reduce_last<1>(v, partials, &st->counter[1]) && threadIdx.x == 0) {
    st->bnorm = sqrt(v[0]);
    if (st->istop == 2) st->errden = st->bnorm;
  }
}

// --------------------------------------------------------------- CG kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_p(int n, const T* __restrict__ z, T* __restrict__ p,
                                                 KState* st) {
  if (halted(st)) return;
  const Chunk c = block_rows(n);
  const bool first = st->it == 0;
  const T beta = (T)st->beta;
  constexpr int V = 16 / sizeof(T);
  const int nv = (c.last - c.first) / V;
  const T* zb = z + c.first;
  T* pb = p + c.first;
  for (int k = threadIdx.x; k < nv; k += 2 * kBlock) {
    const int k1 = k + kBlock;
    v16<T> z0 = ldv(zb + V * k), p0 = ldv(pb + V * k);
    v16<T> z1 = z0, p1 = p0;
    if (k1 < nv) {
      z1 = ldv(zb + V * k1);
      p1 = ldv(pb + V * k1);
    }
    stv(pb + V * k, first ? z0 : z0 + beta * p0);
    if (k1 < nv) stv(pb + V * k1, first ? z1 : z1 + beta * p1);
  }
  for (int i = c.first + nv * V + threadIdx.x; i < c.last; i += kBlock)
    p[i] = first ? z[i] : z[i] + beta * p[i];
}

template <typename T, class M>
__global__ __launch_bounds__(kBlock) void k_cg_q(M m, const T* __restrict__ p, T* __restrict__ q,
                                                 KState* st, double* partials) {
  if (halted(st)) return;
  const Chunk c = block_rows(m.nr);
  double pq = 0.0;
  PSB_FUSED_ROWS(m, p, {
    q[row] = ax;
    pq += (double)p[row] * (double)ax;
  })
  double v[1] = {pq};
  if (reduce_last<1>(v, partials, &st->counter[2]) && threadIdx.x == 0) {
    st->sigma = v[0];
    if (v[0] == 0.0) {
      st->itx += 1;   // the reference counted this iteration before breaking
      st->breakdown = 1;
    } else {
      st->alpha = st->rho / v[0];
    }
  }
}

// q = A*p computed by a separate product (CSR): sigma from a dot kernel
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_sigma(int n, const T* __restrict__ p,
                                                     const T* __restrict__ q, KState* st,
                                                     double* partials) {
  if (halted(st)) return;
  const Chunk c = block_rows(n);
  double pq = 0.0;
  for (int i = c.first + threadIdx.x; i < c.last; i += kBlock) pq += (double)p[i] * (double)q[i];
  double v[1] = {pq};
  if (reduce_last<1>(v, partials, &st->counter[2]) && threadIdx.x == 0) {
    st->sigma = v[0];
    if (v[0] == 0.0) {
      st->itx += 1;
      st->breakdown = 1;
    } else {
      st->alpha = st->rho / v[0];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_x(int n, T* __restrict__ x, T* __restrict__ r,
                                                 const T* __restrict__ p, const T* __restrict__ q,
                                                 const T* __restrict__ d, T* __restrict__ z,
                                                 KState* st, double* partials, double* hist) {
  if (halted(st)) return;
  const Chunk c = block_rows(n);
  const T alpha = (T)st->alpha;
  double rr = 0.0, rz = 0.0;
  constexpr int V = 16 / sizeof(T);
  const int nv = (c.last - c.first) / V;
  const int o = c.first;
  for (int k = threadIdx.x; k < nv; k += 2 * kBlock) {
    const int k1 = k + kBlock;
    const bool two = k1 < nv;
    const int i0 = o + V * k, i1 = o + V * (two ? k1 : k);
    v16<T> x0 = ldv(x + i0), p0 = ldv(p + i0), r0 = ldv(r + i0), q0 = ldv(q + i0);
    v16<T> x1 = ldv(x + i1), p1 = ldv(p + i1), r1 = ldv(r + i1), q1 = ldv(q + i1);
    v16<T> d0 = r0, d1 = r1;
    if (d) {
      d0 = ldv(d + i0);
      d1 = ldv(d + i1);
    }
    x0 = x0 + alpha * p0;
    r0 = r0 - alpha * q0;
    x1 = x1 + alpha * p1;
    r1 = r1 - alpha * q1;
    v16<T> z0 = d ? d0 * r0 : r0, z1 = d ? d1 * r1 : r1;
    stv(x + i0, x0);
    stv(r + i0, r0);
    if (z != r) stv(z + i0, z0);
    rr += vdot<T>(r0, r0);
    rz += vdot<T>(r0, z0);
    if (two) {
      stv(x + i1, x1);
      stv(r + i1, r1);
      if (z != r) stv(z + i1, z1);
      rr += vdot<T>(r1, r1);
      rz += vdot<T>(r1, z1);
    }
  }
  for (int i = o + nv * V + threadIdx.x; i < c.last; i += kBlock) {
    x[i] = x[i] + alpha * p[i];
    T ri = r[i] - alpha * q[i];
    r[i] = ri;
    T zi = d ? d[i] * ri : ri;
    if (z != r) z[i] = zi;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)zi;
  }
  double v[2] = {rr, rz};
  if (reduce_last<2>(v, partials, &st->counter[3]) && threadIdx.x == 0) {
    st->itx += 1;
    st->it += 1;
    stop_test(st, v[0], hist);
    if (!st->done) {
      st->rho_old = st->rho;
      st->rho = v[1];
      if (st->rho_old == 0.0) {
        st->itx += 1;  // next iteration counted, then breaks (psb_dcg.F90:253-258)
        st->breakdown = 1;
      } else {
        st->beta = st->rho / st->rho_old;
      }
    }
  }
}

// --------------------------------------------------------------- BiCGSTAB kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_a(int n, const T* __restrict__ r, T* __restrict__ p,
                                                 const T* __restrict__ v, const T* __restrict__ d,
                                                 T* __restrict__ f, KState* st) {
  if (halted(st)) return;
  const Chunk c = block_rows(n);
  const bool first = st->it == 0;
  const T beta = (T)st->beta, momega = (T)(-st->omega);
  constexpr int V = 16 / sizeof(T);
  const int nv = (c.last - c.first) / V;
  const int o = c.first;
  for (int k = threadIdx.x; k < nv; k += 2 * kBlock) {
    const int k1 = k + kBlock;
    const bool two = k1 < nv;
    const int i0 = o + V * k, i1 = o + V * (two ? k1 : k);
    v16<T> r0 = ldv(r + i0), r1 = ldv(r + i1);
    v16<T> p0 = r0, p1 = r1;
    if (!first) {
      v16<T> q0 = ldv(p + i0), q1 = ldv(p + i1), v0 = ldv(v + i0), v1 = ldv(v + i1);
      p0 = r0 + beta * (momega * v0 + q0);
      p1 = r1 + beta * (momega * v1 + q1);
    }
    v16<T> f0 = p0, f1 = p1;
    if (d) {
      f0 = ldv(d + i0) * p0;
      f1 = ldv(d + i1) * p1;
    }
    stv(p + i0, p0);
    if (f != p) stv(f + i0, f0);
    if (two) {
      stv(p + i1, p1);
      if (f != p) stv(f + i1, f1);
    }
  }
  for (int i = o + nv * V + threadIdx.x; i < c.last; i += kBlock) {
    T pi = first ? r[i] : r[i] + beta * (momega * v[i] + p[i]);
    p[i] = pi;
    if (f != p) f[i] = d ? d[i] * pi : pi;
  }
}

template <typename T, class M>
__global__ __launch_bounds__(kBlock) void k_bi_v(M m, const T* __restrict__ f, T* __restrict__ v,
                                                 const T* __restrict__ q, KState* st,
                                                 double* partials) {
  if (halted(st)) return;
  const Chunk c = block_rows(m.nr);
  double qv = 0.0;
  PSB_FUSED_ROWS(m, f, {
    v[row] = ax;
    qv += (double)q[row] * (double)ax;
  })
  double s[1] = {qv};
  if (reduce_last<1>(s, partials, &st->counter[4]) && threadIdx.x == 0) {
    st->sigma = s[0];
    if (s[0] == 0.0) {
      st->itx += 1;
      st->breakdown = 1;
    } else {
      st->alpha = st->rho / s[0];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_dot1(int n, const T* __restrict__ a,
                                                    const T* __restrict__ b, KState* st,
                                                    double* partials) {
  // sigma = a.b after a separate product (CSR path): same finalisation as k_bi_v
  if (halted(st)) return;
  const Chunk c = block_rows(n);
  double s0 = 0.0;
  for (int i = c.first + threadIdx.x; i < c.last; i += kBlock) s0 += (double)a[i] * (double)b[i];
  double s[1] = {s0};
  if (reduce_last<1>(s, partials, &st->counter[4]) && threadIdx.x == 0) {
    st->sigma = s[0];
    if (s[0] == 0.0) {
      st->itx += 1;
      st->breakdown = 1;
    } else {
      st->alpha = st->rho / s[0];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_s(int n, const T* __restrict__ r,
                                                 const T* __restrict__ v, const T* __restrict__ d,
                                                 T* __restrict__ s, T* __restrict__ z, KState* st) {
  if (halted(st)) return;
  const Chunk c = block_rows(n);
  const T malpha = (T)(-st->alpha);
  constexpr int V = 16 / sizeof(T);
  const int nv = (c.last - c.first) / V;
  const int o = c.first;
  for (int k = threadIdx.x; k < nv; k += 2 * kBlock) {
    const int k1 = k + kBlock;
    const bool two = k1 < nv;
    const int i0 = o + V * k, i1 = o + V * (two ? k1 : k);
    v16<T> s0 = malpha * ldv(v + i0) + ldv(r + i0);
    v16<T> s1 = malpha * ldv(v + i1) + ldv(r + i1);
    v16<T> z0 = s0, z1 = s1;
    if (d) {
      z0 = ldv(d + i0) * s0;
      z1 = ldv(d + i1) * s1;
    }
    stv(s + i0, s0);
    if (z != s) stv(z + i0, z0);
    if (two) {
      stv(s + i1, s1);
      if (z != s) stv(z + i1, z1);
    }
  }
  for (int i = o + nv * V + threadIdx.x; i < c.last; i += kBlock) {
    T si = malpha * v[i] + r[i];
    s[i] = si;
    if (z != s) z[i] = d ? d[i] * si : si;
  }
}

__device__ __forceinline__ void omega_final(KState* st, double tt, double ts) {
  if (tt == 0.0) {
    st->itx += 1;
    st->breakdown = 1;
    return;
  }
  st->omega = ts / tt;
  if (st->omega == 0.0) {
    st->itx += 1;
    st->breakdown = 1;
  }
}

template <typename T, class M>
__global__ __launch_bounds__(kBlock) void k_bi_t(M m, const T* __restrict__ z, T* __restrict__ t,
                                                 const T* __restrict__ s, KState* st,
                                                 double* partials) {
  if (halted(st)) return;
  const Chunk c = block_rows(m.nr);
  double tt = 0.0, ts = 0.0;
  PSB_FUSED_ROWS(m, z, {
    t[row] = ax;
    tt += (double)ax * (double)ax;
    ts += (double)ax * (double)s[row];
  })
  double v[2] = {tt, ts};
  if (reduce_last<2>(v, partials, &st->counter[5]) && threadIdx.x == 0) omega_final(st, v[0], v[1]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_dot2(int n, const T* __restrict__ t,
                                                    const T* __restrict__ s, KState* st,
                                                    double* partials) {
  if (halted(st)) return;
  const Chunk c = block_rows(n);
  double tt = 0.0, ts = 0.0;
  for (int i = c.first + threadIdx.x; i < c.last; i += kBlock) {
    tt += (double)t[i] * (double)t[i];
    ts += (double)t[i] * (double)s[i];
  }
  double v[2] = {tt, ts};
  if (reduce_last<2>(v, partials, &st->counter[5]) && threadIdx.x == 0) omega_final(st, v[0], v[1]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_x(int n, T* __restrict__ x, T* __restrict__ r,
                                                 const T* __restrict__ f, const T* __restrict__ z,
                                                 const T* __restrict__ s, const T* __restrict__ t,
                                                 const T* __restrict__ q, KState* st,
                                                 double* partials, double* hist) {
  if (halted(st)) return;
  const Chunk c = block_rows(n);
  const T alpha = (T)st->alpha, omega = (T)st->omega;
  double rr = 0.0, qr = 0.0;
  constexpr int V = 16 / sizeof(T);
  const int nv = (c.last - c.first) / V;
  const int o = c.first;
  for (int k = threadIdx.x; k < nv; k += 2 * kBlock) {
    const int k1 = k + kBlock;
    const bool two = k1 < nv;
    const int i0 = o + V * k, i1 = o + V * (two ? k1 : k);
    v16<T> x0 = ldv(x + i0), x1 = ldv(x + i1), f0 = ldv(f + i0), f1 = ldv(f + i1);
    v16<T> z0 = ldv(z + i0), z1 = ldv(z + i1), s0 = ldv(s + i0), s1 = ldv(s + i1);
    v16<T> t0 = ldv(t + i0), t1 = ldv(t + i1), q0 = ldv(q + i0), q1 = ldv(q + i1);
    x0 = omega * z0 + (alpha * f0 + x0);
    x1 = omega * z1 + (alpha * f1 + x1);
    v16<T> r0 = -omega * t0 + s0, r1 = -omega * t1 + s1;
    stv(x + i0, x0);
    stv(r + i0, r0);
    rr += vdot<T>(r0, r0);
    qr += vdot<T>(q0, r0);
    if (two) {
      stv(x + i1, x1);
      stv(r + i1, r1);
      rr += vdot<T>(r1, r1);
      qr += vdot<T>(q1, r1);
    }
  }
  for (int i = o + nv * V + threadIdx.x; i < c.last; i += kBlock) {
    T xi = alpha * f[i] + x[i];
    x[i] = omega * z[i] + xi;
    T ri = -omega * t[i] + s[i];
    r[i] = ri;
    rr += (double)ri * (double)ri;
    qr += (double)q[i] * (double)ri;
  }
  double v[2] = {rr, qr};
  if (reduce_last<2>(v, partials, &st->counter[6]) && threadIdx.x == 0) {
    st->itx += 1;
    st->it += 1;
    stop_test(st, v[0], hist);
    if (!st->done) {
      st->rho_old = st->rho;
      st->rho = v[1];
      if (st->rho == 0.0) {
        st->itx += 1;  // next iteration counted, then breaks (psb_dcgstab.f90:267-271)
        st->breakdown = 1;
      } else {
        st->beta = (st->rho / st->rho_old) * (st->alpha / st->omega);
      }
    }
  }
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

// --------------------------------------------------------------- reliable updates
// FP32 CG step with an FP64 residual-replacement trigger: as k_cg_x, but when
// the recursive residual has dropped by `delta` since the last replacement
// (or reached the absolute target, or the budget is spent) the iteration
// halts with breakdown = 2 so the host can replace r by the FP64 residual
// without restarting the recurrence (p and beta are kept).
__global__ __launch_bounds__(kBlock) void k_ru_x(int n, float* __restrict__ x, float* __restrict__ r,
                                                 const float* __restrict__ p,
                                                 const float* __restrict__ q,
                                                 const float* __restrict__ d,
                                                 float* __restrict__ z, KState* st,
                                                 double* partials) {
  if (halted(st)) return;
  const Chunk c = block_rows(n);
  const float alpha = (float)st->alpha;
  double rr = 0.0, rz = 0.0;
  constexpr int V = 4;
  const int nv = (c.last - c.first) / V;
  const int o = c.first;
  for (int k = threadIdx.x; k < nv; k += 2 * kBlock) {
    const int k1 = k + kBlock;
    const bool two = k1 < nv;
    const int i0 = o + V * k, i1 = o + V * (two ? k1 : k);
    v16<float> x0 = ldv(x + i0), p0 = ldv(p + i0), r0 = ldv(r + i0), q0 = ldv(q + i0);
    v16<float> x1 = ldv(x + i1), p1 = ldv(p + i1), r1 = ldv(r + i1), q1 = ldv(q + i1);
    v16<float> d0 = r0, d1 = r1;
    if (d) {
      d0 = ldv(d + i0);
      d1 = ldv(d + i1);
    }
    x0 = x0 + alpha * p0;
    r0 = r0 - alpha * q0;
    x1 = x1 + alpha * p1;
    r1 = r1 - alpha * q1;
    v16<float> z0 = d ? d0 * r0 : r0, z1 = d ? d1 * r1 : r1;
    stv(x + i0, x0);
    stv(r + i0, r0);
    if (z != r) stv(z + i0, z0);
    rr += vdot<float>(r0, r0);
    rz += vdot<float>(r0, z0);
    if (two) {
      stv(x + i1, x1);
      stv(r + i1, r1);
      if (z != r) stv(z + i1, z1);
      rr += vdot<float>(r1, r1);
      rz += vdot<float>(r1, z1);
    }
  }
  for (int i = o + nv * V + threadIdx.x; i < c.last; i += kBlock) {
    x[i] = x[i] + alpha * p[i];
    float ri = r[i] - alpha * q[i];
    r[i] = ri;
    float zi = d ? d[i] * ri : ri;
    if (z != r) z[i] = zi;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)zi;
  }
  double v[2] = {rr, rz};
  if (reduce_last<2>(v, partials, &st->counter[3]) && threadIdx.x == 0) {
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
    st->rho = v[1];
    if (st->rho_old == 0.0) {
      st->itx += 1;
      st->breakdown = 1;
    } else {
      st->beta = st->rho / st->rho_old;
    }
  }
}

// x64 += x32; x32 = 0
__global__ __launch_bounds__(kBlock) void k_ru_accum(int n, double* __restrict__ x,
                                                     float* __restrict__ x32) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] = x[i] + (double)x32[i];
    x32[i] = 0.0f;
  }
}

// r32 = (float)r64; z = d.*r32; rho = r32.z; beta = rho/rho_old (first: it = 0)
__global__ __launch_bounds__(kBlock) void k_ru_reset(int n, const double* __restrict__ r64,
                                                     const double* __restrict__ rnorm,
                                                     float* __restrict__ r,
                                                     const float* __restrict__ d,
                                                     float* __restrict__ z, KState* st,
                                                     double* partials, int first) {
  const Chunk c = block_rows(n);
  double rz = 0.0;
  for (int i = c.first + threadIdx.x; i < c.last; i += kBlock) {
    float ri = (float)r64[i];
    r[i] = ri;
    float zi = d ? d[i] * ri : ri;
    if (z != r) z[i] = zi;
    rz += (double)ri * (double)zi;
  }
  double v[1] = {rz};
  if (reduce_last<1>(v, partials, &st->counter[7]) && threadIdx.x == 0) {
    st->rref = *rnorm;
    st->breakdown = 0;
    st->rho = v[0];
    if (first) {
      st->it = 0;
    } else if (st->rho_old == 0.0) {
      st->breakdown = 1;
    } else {
      st->beta = st->rho / st->rho_old;
    }
  }
}

// --------------------------------------------------------------- host driver
enum { FMT_HLL = 0, FMT_ELL = 1, FMT_CSR = 2 };

struct Solver {
  int method = 0, fmt = 0, n = 0, etype = 0;
  void* mat = nullptr;
  bool use_single = false;       // HLL FP32 copy of a double matrix
  KState* st = nullptr;          // device
  KState* hst = nullptr;         // pinned host mirror
  double* partials = nullptr;
  double* hist = nullptr;
  std::vector<void*> bufs;
  hipStream_t stream = nullptr;

  int alloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return set_error(PSB_HIP_OUTOFMEM, "krylov", hipGetErrorString(e));
    bufs.push_back(*p);
    return PSB_HIP_SUCCESS;
  }
  ~Solver() {
    for (void* p : bufs) (void)hipFree(p);
    if (hst) (void)hipHostFree(hst);
  }
};

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
  } else if (s.fmt == FMT_ELL) {
    EllDev* e = (EllDev*)s.mat;
    m.ell.nr = e->rows;
    m.ell.pitch = e->pitch;
    m.ell.nzm = e->nzm;
    m.ell.ja = e->ja;
    m.ell.val = (const T*)e->val;
  }
  return m;
}

static int csr_product(Solver& s, const void* x, void* y) {
  // y = A*x through the CSR entry point (vector CSR + long rows)
  MultiVec mx{1, s.n, s.n, s.etype, const_cast<void*>(x)};
  MultiVec my{1, s.n, s.n, s.etype, y};
  return psb_hip_csr_spmv(s.mat, 1.0, &mx, 0.0, &my);
}

template <typename T>
static int launch_restart(Solver& s, const Mats<T>& m, const T* b, const T* x, T* r, T* q,
                          const T* d, T* z, int bicg) {
  dim3 g(fused_grid(s.n)), blk(kBlock);
  if (s.fmt == FMT_HLL)
    hipLaunchKernelGGL((k_restart<T, HllView<T>>), g, blk, 0, s.stream, m.hll, b, x, r, q, d, z,
                       s.st, s.partials, bicg);
  else if (s.fmt == FMT_ELL)
    hipLaunchKernelGGL((k_restart<T, EllView<T>>), g, blk, 0, s.stream, m.ell, b, x, r, q, d, z,
                       s.st, s.partials, bicg);
  else
    return PSB_HIP_UNSUPPORTED;
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

// One CG iteration (queued, no host synchronisation).
template <typename T>
static int cg_iter(Solver& s, const Mats<T>& m, T* x, T* r, T* p, T* q, const T* d, T* z) {
  dim3 g(fused_grid(s.n)), blk(kBlock);
  hipLaunchKernelGGL((k_cg_p<T>), g, blk, 0, s.stream, s.n, (const T*)z, p, s.st);
  if (s.fmt == FMT_HLL)
    hipLaunchKernelGGL((k_cg_q<T, HllView<T>>), g, blk, 0, s.stream, m.hll, (const T*)p, q, s.st,
                       s.partials);
  else if (s.fmt == FMT_ELL)
    hipLaunchKernelGGL((k_cg_q<T, EllView<T>>), g, blk, 0, s.stream, m.ell, (const T*)p, q, s.st,
                       s.partials);
  else {
    int rc = csr_product(s, p, q);
    if (rc) return rc;
    hipLaunchKernelGGL((k_cg_sigma<T>), g, blk, 0, s.stream, s.n, (const T*)p, (const T*)q, s.st,
                       s.partials);
  }
  hipLaunchKernelGGL((k_cg_x<T>), g, blk, 0, s.stream, s.n, x, r, (const T*)p, (const T*)q, d, z,
                     s.st, s.partials, s.hist);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

template <typename T>
static int bi_iter(Solver& s, const Mats<T>& m, T* x, T* r, T* q, T* p, T* v, T* f, T* sv, T* z,
                   T* t, const T* d) {
  dim3 g(fused_grid(s.n)), blk(kBlock);
  hipLaunchKernelGGL((k_bi_a<T>), g, blk, 0, s.stream, s.n, (const T*)r, p, (const T*)v, d, f,
                     s.st);
  if (s.fmt == FMT_HLL)
    hipLaunchKernelGGL((k_bi_v<T, HllView<T>>), g, blk, 0, s.stream, m.hll, (const T*)f, v,
                       (const T*)q, s.st, s.partials);
  else if (s.fmt == FMT_ELL)
    hipLaunchKernelGGL((k_bi_v<T, EllView<T>>), g, blk, 0, s.stream, m.ell, (const T*)f, v,
                       (const T*)q, s.st, s.partials);
  else {
    int rc = csr_product(s, f, v);
    if (rc) return rc;
    hipLaunchKernelGGL((k_bi_dot1<T>), g, blk, 0, s.stream, s.n, (const T*)q, (const T*)v, s.st,
                       s.partials);
  }
  hipLaunchKernelGGL((k_bi_s<T>), g, blk, 0, s.stream, s.n, (const T*)r, (const T*)v, d, sv, z,
                     s.st);
  if (s.fmt == FMT_HLL)
    hipLaunchKernelGGL((k_bi_t<T, HllView<T>>), g, blk, 0, s.stream, m.hll, (const T*)z, t,
                       (const T*)sv, s.st, s.partials);
  else if (s.fmt == FMT_ELL)
    hipLaunchKernelGGL((k_bi_t<T, EllView<T>>), g, blk, 0, s.stream, m.ell, (const T*)z, t,
                       (const T*)sv, s.st, s.partials);
  else {
    int rc = csr_product(s, z, t);
    if (rc) return rc;
    hipLaunchKernelGGL((k_bi_dot2<T>), g, blk, 0, s.stream, s.n, (const T*)t, (const T*)sv, s.st,
                       s.partials);
  }
  hipLaunchKernelGGL((k_bi_x<T>), g, blk, 0, s.stream, s.n, x, r, (const T*)f, (const T*)z,
                     (const T*)sv, (const T*)t, (const T*)q, s.st, s.partials, s.hist);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

static int read_state(Solver& s) {
  PSB_HIP_CHECK(hipMemcpyAsync(s.hst, s.st, sizeof(KState), hipMemcpyDeviceToHost, s.stream));
  PSB_HIP_CHECK(hipStreamSynchronize(s.stream));
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
  // vectors the preconditioner would only copy are aliased when d == NULL
  // the initial state goes through the pinned mirror (an async copy must not
  // read a stack object that may be gone when the copy runs)
  *s.hst = KState{};
  s.hst->eps = eps;
  s.hst->itmax = itmax;
  s.hst->istop = istop;
  s.hst->errden = 1.0;
  PSB_HIP_CHECK(hipMemcpyAsync(s.st, s.hst, sizeof(KState), hipMemcpyHostToDevice, s.stream));
  Mats<T> m = make_views<T>(s);
  dim3 g(fused_grid(n)), blk(kBlock);
  hipLaunchKernelGGL((k_bnorm<T>), g, blk, 0, s.stream, n, b, s.st, s.partials);
  PSB_HIP_LAUNCH_CHECK();
  if (check_every < 1) check_every = 8;
  int rc;
  for (;;) {
    // restart: r = b - A x (and the first z / rho)
    if (s.method == 0) {
      T *r = w[0], *z = d ? w[3] : w[0];
      if (s.fmt == FMT_CSR) {
        // r = b - A x via the product entry point, then the restart sums
        MultiVec mx{1, n, n, s.etype, x}, mr{1, n, n, s.etype, r}, mb{1, n, n, s.etype, (void*)b};
        rc = psb_hip_mv_axpby(n, 1.0, &mb, 0.0, &mr);
        if (rc) return rc;
        rc = psb_hip_csr_spmv(s.mat, -1.0, &mx, 1.0, &mr);
        if (rc) return rc;
        // identity "matrix" pass computing the sums: reuse k_restart on an
        // empty ELL view (row() = 0) with b := r
        EllView<T> zero{n, n, 0, nullptr, nullptr};
        hipLaunchKernelGGL((k_restart<T, EllView<T>>), g, blk, 0, s.stream, zero, (const T*)r,
                           (const T*)x, r, (T*)nullptr, d, z, s.st, s.partials, 0);
        PSB_HIP_LAUNCH_CHECK();
      } else {
        rc = launch_restart<T>(s, m, b, x, r, nullptr, d, z, 0);
        if (rc) return rc;
      }
    } else {
      T *r = w[0], *q = w[1];
      if (s.fmt == FMT_CSR) {
        MultiVec mx{1, n, n, s.etype, x}, mr{1, n, n, s.etype, r}, mb{1, n, n, s.etype, (void*)b};
        rc = psb_hip_mv_axpby(n, 1.0, &mb, 0.0, &mr);
        if (rc) return rc;
        rc = psb_hip_csr_spmv(s.mat, -1.0, &mx, 1.0, &mr);
        if (rc) return rc;
        EllView<T> zero{n, n, 0, nullptr, nullptr};
        hipLaunchKernelGGL((k_restart<T, EllView<T>>), g, blk, 0, s.stream, zero, (const T*)r,
                           (const T*)x, r, q, d, (T*)nullptr, s.st, s.partials, 1);
        PSB_HIP_LAUNCH_CHECK();
      } else {
        rc = launch_restart<T>(s, m, b, x, r, q, d, nullptr, 1);
        if (rc) return rc;
      }
    }
    // iterate
    bool stop = false;
    while (!stop) {
      for (int k = 0; k < check_every; k++) {
        if (s.method == 0) {
          T* z = d ? w[3] : w[0];
          rc = cg_iter<T>(s, m, x, w[0], w[1], w[2], d, z);
        } else {
          T* f = d ? w[4] : w[2];
          T* z = d ? w[6] : w[5];
          rc = bi_iter<T>(s, m, x, w[0], w[1], w[2], w[3], f, w[5], z, w[7], d);
        }
        if (rc) return rc;
      }
      rc = read_state(s);
      if (rc) return rc;
      if (s.hst->done || s.hst->breakdown) stop = true;
    }
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

static int setup(Solver& s, int method, int fmt, void* dmat, int itmax) {
  s.method = method;
  s.fmt = fmt;
  s.mat = dmat;
  s.stream = ctx().stream;
  if (fmt == FMT_HLL) {
    HllDev* h = (HllDev*)dmat;
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
  int rc = s.alloc((void**)&s.st, sizeof(KState));
  if (rc) return rc;
  rc = s.alloc((void**)&s.partials, sizeof(double) * 8 * kMaxGrid);
  if (rc) return rc;
  rc = s.alloc((void**)&s.hist, sizeof(double) * ((size_t)itmax + 2));
  if (rc) return rc;
  PSB_HIP_CHECK(hipHostMalloc((void**)&s.hst, sizeof(KState), hipHostMallocDefault));
  return PSB_HIP_SUCCESS;
}

static int mv_check(const MultiVec* v, int etype, int n) {
  if (!v || v->elem_type != etype || v->size < n || v->count != 1) return PSB_HIP_INVALID;
  return PSB_HIP_SUCCESS;
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
  dim3 g(fused_grid(n)), blk(kBlock), gs(stream_grid(n));
  if (d64) hipLaunchKernelGGL(k_d2s_vec, gs, blk, 0, stream, n, d64, d);
  hipLaunchKernelGGL((k_zero<float>), gs, blk, 0, stream, n, x32);
  // outer state: ||b||, and the FP64 residuals
  *outer.hst = KState{};
  outer.hst->itmax = 1 << 30;
  outer.hst->istop = 2;
  outer.hst->errden = 1.0;
  PSB_HIP_CHECK(hipMemcpyAsync(outer.st, outer.hst, sizeof(KState), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL((k_bnorm<double>), g, blk, 0, stream, n, b, outer.st, outer.partials);
  Mats<double> m64 = make_views<double>(outer);
  Mats<float> m32 = make_views<float>(inner);
  auto fp64_residual = [&]() -> int {
    hipLaunchKernelGGL((k_restart<double, HllView<double>>), g, blk, 0, stream, m64.hll, b,
                       (const double*)x, r64, (double*)nullptr, (const double*)nullptr, r64,
                       outer.st, outer.partials, 0);
    PSB_HIP_LAUNCH_CHECK();
    return read_state(outer);
  };
  if ((rc = fp64_residual())) return rc;
  const double bnorm = outer.hst->bnorm;
  // inner state: eps*errden is the absolute target, omega the replacement factor
  *inner.hst = KState{};
  inner.hst->eps = eps;
  inner.hst->errden = bnorm;
  inner.hst->delta = delta;
  inner.hst->itmax = inner_itmax;
  inner.hst->istop = 2;
  PSB_HIP_CHECK(hipMemcpyAsync(inner.st, inner.hst, sizeof(KState), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_ru_reset, g, blk, 0, stream, n, (const double*)r64,
                     (const double*)&outer.st->errnum, r, (const float*)d, z, inner.st,
                     inner.partials, 1);
  PSB_HIP_LAUNCH_CHECK();
  int updates = 0;
  double rnorm = outer.hst->errnum;
  bool conv = bnorm == 0.0 ? rnorm <= eps : rnorm <= eps * bnorm;
  if (check_every < 1) check_every = 8;
  while (!conv) {
    // queue FP32 iterations until a replacement (or breakdown) is requested
    bool stop = false;
    while (!stop) {
      for (int k = 0; k < check_every; k++) {
        hipLaunchKernelGGL((k_cg_p<float>), g, blk, 0, stream, n, (const float*)z, p, inner.st);
        hipLaunchKernelGGL((k_cg_q<float, HllView<float>>), g, blk, 0, stream, m32.hll,
                           (const float*)p, q, inner.st, inner.partials);
        hipLaunchKernelGGL(k_ru_x, g, blk, 0, stream, n, x32, r, (const float*)p,
                           (const float*)q, (const float*)d, z, inner.st, inner.partials);
      }
      PSB_HIP_LAUNCH_CHECK();
      if ((rc = read_state(inner))) return rc;
      stop = inner.hst->breakdown != 0;
    }
    // replacement: x += x32 in FP64, r = b - A x in FP64
    hipLaunchKernelGGL(k_ru_accum, gs, blk, 0, stream, n, x, x32);
    PSB_HIP_LAUNCH_CHECK();
    if ((rc = fp64_residual())) return rc;
    updates++;
    rnorm = outer.hst->errnum;
    conv = bnorm == 0.0 ? rnorm <= eps : rnorm <= eps * bnorm;
    if (conv || updates >= outer_itmax || inner.hst->itx >= inner_itmax) break;
    // a true breakdown (rho_old == 0) restarts the recurrence from the new residual
    int first = inner.hst->breakdown == 1 ? 1 : 0;
    hipLaunchKernelGGL(k_ru_reset, g, blk, 0, stream, n, (const double*)r64,
                       (const double*)&outer.st->errnum, r, (const float*)d, z, inner.st,
                       inner.partials, first);
    PSB_HIP_LAUNCH_CHECK();
  }
  *outer_its = updates;
  *inner_its = inner.hst->itx;
  *err = bnorm > 0.0 ? rnorm / bnorm : rnorm;
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
  if (istop != 2 && istop != 3)
    return set_error(PSB_HIP_UNSUPPORTED, __func__, "istop must be 2 or 3 in the fused solver");
  if (itmax < 0) return set_error(PSB_HIP_INVALID, __func__, "itmax");
  Solver s;
  rc = setup(s, method == 0 ? 0 : 1, fmt, dmat, itmax);
  if (rc) return rc;
  MultiVec *b = (MultiVec*)bv, *x = (MultiVec*)xv, *d = (MultiVec*)diag;
  if (mv_check(b, s.etype, s.n) || mv_check(x, s.etype, s.n) || (d && mv_check(d, s.etype, s.n)))
    return set_error(PSB_HIP_INVALID, __func__, "b/x/diag type or size");
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

int psb_hip_krylov_mixed(int method, void* dmat, void* diag, void* bv, void* xv, double eps,
                         int outer_itmax, double inner_eps, int inner_itmax, int check_every,
                         int variant, int* outer_its, int* inner_its, double* err) {
  int rc = ensure_init();
  if (rc) return rc;
  HllDev* h = (HllDev*)dmat;
  if (h->elem_type != PSB_HIP_TYPE_DOUBLE || !h->val32)
    return set_error(PSB_HIP_INVALID, __func__, "needs a double HLL matrix with a single copy");
  const int n = h->rows;
  MultiVec *b = (MultiVec*)bv, *x = (MultiVec*)xv, *d = (MultiVec*)diag;
  if (mv_check(b, PSB_HIP_TYPE_DOUBLE, n) || mv_check(x, PSB_HIP_TYPE_DOUBLE, n) ||
      (d && mv_check(d, PSB_HIP_TYPE_DOUBLE, n)))
    return set_error(PSB_HIP_INVALID, __func__, "b/x/diag type or size");
  if (variant == 1) {
    if (method != 0)
      return set_error(PSB_HIP_UNSUPPORTED, __func__, "reliable updates are implemented for CG");
    if (n == 0) {
      *outer_its = *inner_its = 0;
      *err = 0.0;
      return PSB_HIP_SUCCESS;
    }
    return mixed_reliable(h, (const double*)b->v, (double*)x->v,
                          d ? (const double*)d->v : nullptr, eps, outer_itmax, inner_eps,
                          inner_itmax, check_every, outer_its, inner_its, err);
  }
  hipStream_t stream = ctx().stream;
  // outer workspace (double): r, and the scalar state of the outer residual
  Solver outer;
  rc = setup(outer, 0, FMT_HLL, dmat, 1);
  if (rc) return rc;
  double* r = nullptr;
  float *r32 = nullptr, *d32 = nullptr, *x32 = nullptr;
  rc = outer.alloc((void**)&r, sizeof(double) * n);
  if (rc) return rc;
  rc = outer.alloc((void**)&r32, sizeof(float) * n);
  if (rc) return rc;
  rc = outer.alloc((void**)&x32, sizeof(float) * n);
  if (rc) return rc;
  if (d) {
    rc = outer.alloc((void**)&d32, sizeof(float) * n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_d2s_vec, dim3(stream_grid(n)), dim3(kBlock), 0, stream, n,
                       (const double*)d->v, d32);
    PSB_HIP_LAUNCH_CHECK();
  }
  Mats<double> m64 = make_views<double>(outer);
  dim3 g(fused_grid(n)), blk(kBlock);
  *outer.hst = KState{};
  outer.hst->itmax = 1 << 30;
  outer.hst->istop = 2;
  outer.hst->errden = 1.0;
  PSB_HIP_CHECK(hipMemcpyAsync(outer.st, outer.hst, sizeof(KState), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL((k_bnorm<double>), g, blk, 0, stream, n, (const double*)b->v, outer.st,
                     outer.partials);
  PSB_HIP_LAUNCH_CHECK();
  int total_inner = 0, outer_k = 0;
  double bnorm = 0.0, rnorm = 0.0;
  MultiVec mr32{1, n, n, PSB_HIP_TYPE_FLOAT, r32}, mx32{1, n, n, PSB_HIP_TYPE_FLOAT, x32};
  MultiVec md32{1, n, n, PSB_HIP_TYPE_FLOAT, d32};
  for (;;) {
    // r = b - A x in FP64 (CG restart kernel: also ||r||)
    hipLaunchKernelGGL((k_restart<double, HllView<double>>), g, blk, 0, stream, m64.hll,
                       (const double*)b->v, (const double*)x->v, r, (double*)nullptr,
                       (const double*)nullptr, r, outer.st, outer.partials, 0);
    PSB_HIP_LAUNCH_CHECK();
    rc = read_state(outer);
    if (rc) return rc;
    bnorm = outer.hst->bnorm;
    rnorm = outer.hst->errnum;
    bool conv = bnorm == 0.0 ? rnorm <= eps : rnorm <= eps * bnorm;
    if (conv || outer_k >= outer_itmax || rnorm == 0.0) break;
    outer_k++;
    // inner: A32 * d = r / ||r|| to inner_eps (no looser than needed at the end)
    double ie = inner_eps;
    if (bnorm > 0.0) {
      double need = 0.5 * eps * bnorm / rnorm;
      if (need > ie) ie = need < 0.5 ? need : 0.5;
    }
    hipLaunchKernelGGL(k_to_single, dim3(stream_grid(n)), dim3(kBlock), 0, stream, n,
                       (const double*)r, (const double*)&outer.st->errnum, r32);
    hipLaunchKernelGGL((k_zero<float>), dim3(stream_grid(n)), dim3(kBlock), 0, stream, n, x32);
    PSB_HIP_LAUNCH_CHECK();
    Solver inner;
    rc = setup(inner, method == 0 ? 0 : 1, FMT_HLL, dmat, inner_itmax);
    if (rc) return rc;
    inner.use_single = true;
    inner.etype = PSB_HIP_TYPE_FLOAT;
    int it = 0;
    double ierr = 0.0;
    rc = solve<float>(inner, r32, x32, d ? d32 : nullptr, ie, inner_itmax, 2, check_every, &it,
                      &ierr);
    if (rc) return rc;
    total_inner += it;
    // x += ||r|| * d  (FP64)
    hipLaunchKernelGGL(k_update_double, dim3(stream_grid(n)), dim3(kBlock), 0, stream, n,
                       (double*)x->v, (const float*)x32, (const double*)&outer.st->errnum);
    PSB_HIP_LAUNCH_CHECK();
    (void)mr32;
    (void)mx32;
    (void)md32;
  }
  *outer_its = outer_k;
  *inner_its = total_inner;
  *err = bnorm > 0.0 ? rnorm / bnorm : rnorm;
  return PSB_HIP_SUCCESS;
}

}  // extern "C"
