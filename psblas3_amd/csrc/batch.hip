// Batched mixed-precision Krylov engine: many independent systems of
// different sizes advance together, one HIP graph launch per step.
//
// Each live system owns a slot: a descriptor (its HLL matrix with the FP32
// value copy, its FP32 Jacobi diagonal, its vectors in the engine's arena)
// and a scalar state.  The rows of all live systems are cut into 1024-row
// tiles; the tile table lists (slot, first row) system after system.  Every
// streaming kernel walks the tile table grid-stride and skips the tiles of
// systems that are not in the phase it serves; every reduction stores one
// partial per tile, and a finalisation kernel with one block per slot adds
// its system's partials in tile order.  A system's arithmetic therefore does
// not depend on which other systems share the launch: its results are the
// same alone and in any mix.
//
// A step = check_every FP32 iterations of every iterating system, then one
// transition pass for the systems that asked for one:
//   CG (method 0): FP32 CG with FP64 residual replacement (reliable
//     updates): once the recursive residual has dropped by delta since the
//     last replacement (or reached the target), x64 += x32, r64 = b - A64 x64
//     in FP64, and the FP32 residual restarts from r64, keeping p.
//   BiCGSTAB (method 1): restarted refinement: FP32 BiCGSTAB solves
//     A d = r64/||r64|| to the inner tolerance, then x64 += ||r64|| d and
//     r64 = b - A64 x64 in FP64, and the inner solve restarts.
//   BiCGSTAB (method 2): reliable updates as for CG: the FP32 recurrence
//     keeps p, v and the shadow residual; at a replacement x64 += x32 and the
//     FP32 residual restarts from r64, with rho and beta recomputed from it.
// The transition pass also initialises newly admitted systems (r64 = b - A x0,
// ||b||) and decides convergence on the FP64 residual (||r64|| <= eps ||b||).
// The per-iteration operation sequences are those of the fused single-system
// solvers (krylov/psb_dcg.F90:241-294, krylov/psb_dcgstab.f90:248-364).
//
// The step's kernel sequence is captured once into a graph; admissions and
// releases only rewrite the device tables (stream-ordered copies between
// graph launches), so the graph never changes.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "psb_hip_internal.hpp"

namespace psbhip {

constexpr int kBT = kBlock;      // threads per block (128: serving 12% slower, batch 5%)
constexpr int kTileRows = 4 * kBT;  // rows per tile
constexpr int kBGrid = 2048;     // blocks of the tile kernels

enum BPhase { BP_RUN = 0, BP_BREAK = 1, BP_TRANS = 2, BP_INIT = 3 };

struct BState {
  double rho, rho_old, alpha, beta, omega, sigma;
  double rr;          // FP32 recursive residual norm after the last iteration
  double ref;         // CG: ||r64|| at the last replacement; BiCGSTAB: ||r64|| of the cycle
  double bnorm;       // ||b||
  double rnorm;       // ||b - A x64|| at the last transition
  double eps;         // target: rnorm <= eps * bnorm
  double delta;       // CG replacement factor
  double inner_eps;   // BiCGSTAB inner relative tolerance
  double inner_tol;   // BiCGSTAB inner absolute tolerance of the current cycle
  int it;             // iterations since the recurrence (re)started
  int itx;            // total FP32 iterations
  int itmax;
  int outer;          // transitions after the initial one
  int phase;          // BPhase
  int done;           // 1 converged, 2 iteration limit
  int pad[2];
};

struct BSys {
  int n, hksz;
  const int* hkoffs;
  const int* ja;
  const short* res16;   // 16-bit column offsets (HllDev::res16 / base16) or null
  const int* base16;
  const float* a32;
  const double* a64;
  const float* d32;   // FP32 Jacobi 1/a_ii, or null
  const double* b;
  double* x;
  double* r64;
  float *x32, *r, *p, *q, *z;   // q: CG q = A p; BiCGSTAB the shadow residual
  float *v, *f, *s, *t;         // BiCGSTAB only
};

struct BTables {
  const int2* __restrict__ tiles;  // (slot, first row)
  const int* ntiles;               // device scalar
  const int2* __restrict__ range;  // per slot: first tile, tile count
  const BSys* __restrict__ sys;
  BState* st;
  double* partials;                // two per tile
  int tcap;                        // tile table entries (>= kBGrid)
};

__device__ __forceinline__ bool b_iterating(const BState* s) {
  return s->phase == BP_RUN && s->done == 0;
}
__device__ __forceinline__ bool b_transiting(const BState* s) {
  return s->phase != BP_RUN && s->done == 0;
}

// (A x)[r] of a system's HLL matrix with values a (FP32 or FP64); HK > 0:
// the hack size as a compile-time constant
template <typename T, int HK>
__device__ __forceinline__ T b_row_hk(const BSys& sy, const T* a, int r,
                                      const T* __restrict__ x) {
  const int stride = HK > 0 ? HK : sy.hksz;
  int hk = r / stride;
  // tiles start on 64-row boundaries and rows past n clamp to n - 1: with the
  // hack size equal to the wavefront size a wavefront is one hack, so its
  // offsets come through scalar loads
  if (HK == kWave) hk = __builtin_amdgcn_readfirstlane(hk);
  const int lane = r - hk * stride;
  const int off0 = sy.hkoffs[hk];
  const int w = (sy.hkoffs[hk + 1] - off0) / stride;
  if (sy.res16)
    return slot_dot16<T, T>(sy.res16 + off0 + lane, sy.base16 + off0 / stride, a + off0 + lane,
                            stride, w, r, x);
  return slot_dot<T, T>(sy.ja + off0 + lane, a + off0 + lane, stride, w, x);
}

template <typename T>
__device__ __forceinline__ T b_row(const BSys& sy, const T* a, int r, const T* __restrict__ x) {
  return sy.hksz == 64 ? b_row_hk<T, 64>(sy, a, r, x) : b_row_hk<T, 0>(sy, a, r, x);
}

// The body runs for every tile of a system satisfying COND (st, sy, r0 =
// the tile's first row, t_ = tile index in scope).  COND is uniform over the
// block.  A tile is kTileRows rows: elementwise kernels give each of their
// kBT threads 4 consecutive rows (16-byte accesses), product kernels give
// each thread 4 rows kBT apart (a wavefront walks 4 whole hacks; 1024-thread
// blocks of one row per thread measured 6% slower on the 256-system CG
// batch).  The descriptor loads are kept short: the
// block's first tile entry is read together with the tile count (the table
// has at least gridDim.x entries), a tile's state and system descriptor
// together, and the next tile's entry before the body runs.
#define PSB_BATCH_TILES(TB, COND, ...)                                   \
  {                                                                      \
    const int nt_ = *(TB).ntiles;                                        \
    int2 tl_ = (TB).tiles[blockIdx.x];                                   \
    int2 tn_;                                                            \
    for (int t_ = blockIdx.x; t_ < nt_; t_ += gridDim.x, tl_ = tn_) {    \
      BState* st = (TB).st + tl_.x;                                      \
      const BSys sy = (TB).sys[tl_.x];                                   \
      tn_ = (TB).tiles[min(t_ + (int)gridDim.x, (TB).tcap - 1)];         \
      if (!(COND)) continue;                                             \
      const int r0 = tl_.y;                                              \
      __VA_ARGS__                                                        \
    }                                                                    \
  }

__device__ __forceinline__ void tile_partials(double a, double b, double* partials, int t) {
  __shared__ double lds[2][kBT / kWave];
  const double sa = block_sum<double, kBT>(a, lds[0]);
  const double sb = block_sum<double, kBT>(b, lds[1]);
  if (threadIdx.x == 0) {
    partials[2 * t] = sa;
    partials[2 * t + 1] = sb;
  }
}

// 4 consecutive entries p[0..3] (16-byte aligned); nv < 4: the first nv only
__device__ __forceinline__ void ld4(const float* p, int nv, float (&a)[4]) {
  if (nv == 4) {
    const v16<float> t = ldv(p);
    a[0] = t[0]; a[1] = t[1]; a[2] = t[2]; a[3] = t[3];
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) a[k] = k < nv ? p[k] : 0.0f;
  }
}
__device__ __forceinline__ void st4(float* p, int nv, const float (&a)[4]) {
  if (nv == 4) {
    v16<float> t;
    t[0] = a[0]; t[1] = a[1]; t[2] = a[2]; t[3] = a[3];
    stv(p, t);
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (k < nv) p[k] = a[k];
  }
}
__device__ __forceinline__ void ld4(const double* p, int nv, double (&a)[4]) {
  if (nv == 4) {
    const v16<double> t0 = ldv(p), t1 = ldv(p + 2);
    a[0] = t0[0]; a[1] = t0[1]; a[2] = t1[0]; a[3] = t1[1];
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) a[k] = k < nv ? p[k] : 0.0;
  }
}
__device__ __forceinline__ void st4(double* p, int nv, const double (&a)[4]) {
  if (nv == 4) {
    v16<double> t0, t1;
    t0[0] = a[0]; t0[1] = a[1]; t1[0] = a[2]; t1[1] = a[3];
    stv(p, t0);
    stv(p + 2, t1);
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (k < nv) p[k] = a[k];
  }
}

// first row and count (0..4) of this thread's 4 rows in an elementwise tile
#define PSB_QUAD(I, NV)                                                  \
  const int I = r0 + 4 * (int)threadIdx.x;                               \
  const int NV = sy.n - I < 4 ? (sy.n - I > 0 ? sy.n - I : 0) : 4;

// ------------------------------------------------------------------ CG iteration
__global__ __launch_bounds__(kBT) void kb_cg_p(BTables tb) {
  PSB_BATCH_TILES(tb, b_iterating(st), {
    PSB_QUAD(i, nv)
    if (nv > 0) {
      float z[4], p[4];
      ld4(sy.z + i, nv, z);
      if (st->it == 0) {
        st4(sy.p + i, nv, z);
      } else {
        ld4(sy.p + i, nv, p);
        const float beta = (float)st->beta;
_Pragma("unroll")
        for (int k = 0; k < 4; k++) p[k] = z[k] + beta * p[k];
        st4(sy.p + i, nv, p);
      }
    }
  })
}

// Product kernels: 4 rows per thread (rows r0 + k*kBT + thread, k = 0..3,
// each wavefront walking 4 whole hacks of the tile), so a block's descriptor
// loads and reduction serve the whole tile with kBT threads and 4 times as
// many blocks fit a CU as with one row per thread.
__global__ __launch_bounds__(kBT) void kb_cg_q(BTables tb) {
  PSB_BATCH_TILES(tb, b_iterating(st), {
    float ax[4];
_Pragma("unroll")
    for (int k = 0; k < 4; k++) {
      const int i = r0 + k * kBT + (int)threadIdx.x;
      ax[k] = b_row<float>(sy, sy.a32, i < sy.n ? i : sy.n - 1, sy.p);
    }
    double pq = 0.0;
_Pragma("unroll")
    for (int k = 0; k < 4; k++) {
      const int i = r0 + k * kBT + (int)threadIdx.x;
      if (i < sy.n) {
        sy.q[i] = ax[k];
        pq += (double)sy.p[i] * (double)ax[k];
      }
    }
    tile_partials(pq, 0.0, tb.partials, t_);
  })
}

__global__ __launch_bounds__(kBT) void kb_cg_x(BTables tb) {
  PSB_BATCH_TILES(tb, b_iterating(st), {
    PSB_QUAD(i, nv)
    double rr = 0.0, rz = 0.0;
    if (nv > 0) {
      const float alpha = (float)st->alpha;
      float x[4], p[4], r[4], q[4], z[4];
      ld4(sy.x32 + i, nv, x);
      ld4(sy.p + i, nv, p);
      ld4(sy.r + i, nv, r);
      ld4(sy.q + i, nv, q);
      if (sy.d32) ld4(sy.d32 + i, nv, z);
_Pragma("unroll")
      for (int k = 0; k < 4; k++) {
        x[k] = x[k] + alpha * p[k];
        r[k] = r[k] - alpha * q[k];
        z[k] = sy.d32 ? z[k] * r[k] : r[k];
        rr += (double)r[k] * (double)r[k];
        rz += (double)r[k] * (double)z[k];
      }
      st4(sy.x32 + i, nv, x);
      st4(sy.r + i, nv, r);
      st4(sy.z + i, nv, z);
    }
    tile_partials(rr, rz, tb.partials, t_);
  })
}

// ------------------------------------------------------------------ BiCGSTAB iteration
__global__ __launch_bounds__(kBT) void kb_bi_a(BTables tb) {
  PSB_BATCH_TILES(tb, b_iterating(st), {
    PSB_QUAD(i, nv)
    if (nv > 0) {
      float r[4], p[4], v[4], f[4];
      ld4(sy.r + i, nv, r);
      if (st->it != 0) {
        ld4(sy.p + i, nv, p);
        ld4(sy.v + i, nv, v);
        const float beta = (float)st->beta, momega = (float)(-st->omega);
_Pragma("unroll")
        for (int k = 0; k < 4; k++) p[k] = r[k] + beta * (momega * v[k] + p[k]);
      } else {
_Pragma("unroll")
        for (int k = 0; k < 4; k++) p[k] = r[k];
      }
      st4(sy.p + i, nv, p);
      if (sy.d32) {
        ld4(sy.d32 + i, nv, f);
_Pragma("unroll")
        for (int k = 0; k < 4; k++) f[k] = f[k] * p[k];
        st4(sy.f + i, nv, f);
      } else {
        st4(sy.f + i, nv, p);
      }
    }
  })
}

__global__ __launch_bounds__(kBT) void kb_bi_v(BTables tb) {
  PSB_BATCH_TILES(tb, b_iterating(st), {
    float ax[4];
_Pragma("unroll")
    for (int k = 0; k < 4; k++) {
      const int i = r0 + k * kBT + (int)threadIdx.x;
      ax[k] = b_row<float>(sy, sy.a32, i < sy.n ? i : sy.n - 1, sy.f);
    }
    double qv = 0.0;
_Pragma("unroll")
    for (int k = 0; k < 4; k++) {
      const int i = r0 + k * kBT + (int)threadIdx.x;
      if (i < sy.n) {
        sy.v[i] = ax[k];
        qv += (double)sy.q[i] * (double)ax[k];
      }
    }
    tile_partials(qv, 0.0, tb.partials, t_);
  })
}

__global__ __launch_bounds__(kBT) void kb_bi_s(BTables tb) {
  PSB_BATCH_TILES(tb, b_iterating(st), {
    PSB_QUAD(i, nv)
    if (nv > 0) {
      float v[4], r[4], z[4];
      ld4(sy.v + i, nv, v);
      ld4(sy.r + i, nv, r);
      const float malpha = (float)(-st->alpha);
_Pragma("unroll")
      for (int k = 0; k < 4; k++) r[k] = malpha * v[k] + r[k];   // s
      st4(sy.s + i, nv, r);
      if (sy.d32) {
        ld4(sy.d32 + i, nv, z);
_Pragma("unroll")
        for (int k = 0; k < 4; k++) z[k] = z[k] * r[k];
        st4(sy.z + i, nv, z);
      } else {
        st4(sy.z + i, nv, r);
      }
    }
  })
}

__global__ __launch_bounds__(kBT) void kb_bi_t(BTables tb) {
  PSB_BATCH_TILES(tb, b_iterating(st), {
    float ax[4];
_Pragma("unroll")
    for (int k = 0; k < 4; k++) {
      const int i = r0 + k * kBT + (int)threadIdx.x;
      ax[k] = b_row<float>(sy, sy.a32, i < sy.n ? i : sy.n - 1, sy.z);
    }
    double tt = 0.0, ts = 0.0;
_Pragma("unroll")
    for (int k = 0; k < 4; k++) {
      const int i = r0 + k * kBT + (int)threadIdx.x;
      if (i < sy.n) {
        sy.t[i] = ax[k];
        tt += (double)ax[k] * (double)ax[k];
        ts += (double)ax[k] * (double)sy.s[i];
      }
    }
    tile_partials(tt, ts, tb.partials, t_);
  })
}

__global__ __launch_bounds__(kBT) void kb_bi_x(BTables tb) {
  PSB_BATCH_TILES(tb, b_iterating(st), {
    PSB_QUAD(i, nv)
    double rr = 0.0, qr = 0.0;
    if (nv > 0) {
      const float alpha = (float)st->alpha, omega = (float)st->omega;
      float x[4], z[4], f[4], t[4], s4[4], q[4];
      ld4(sy.x32 + i, nv, x);
      ld4(sy.z + i, nv, z);
      ld4(sy.f + i, nv, f);
      ld4(sy.t + i, nv, t);
      ld4(sy.s + i, nv, s4);
      ld4(sy.q + i, nv, q);
_Pragma("unroll")
      for (int k = 0; k < 4; k++) {
        x[k] = omega * z[k] + (alpha * f[k] + x[k]);
        s4[k] = -omega * t[k] + s4[k];   // r
        rr += (double)s4[k] * (double)s4[k];
        qr += (double)q[k] * (double)s4[k];
      }
      st4(sy.x32 + i, nv, x);
      st4(sy.r + i, nv, s4);
    }
    tile_partials(rr, qr, tb.partials, t_);
  })
}

// ------------------------------------------------------------------ transitions
// x64 += scale * x32 (residual replacement: scale 1; restarted refinement:
// ||r64|| of the finished cycle), x32 = 0.  A newly admitted system only
// clears x32.
__global__ __launch_bounds__(kBT) void kb_accum(BTables tb, int unit_scale) {
  PSB_BATCH_TILES(tb, b_transiting(st), {
    PSB_QUAD(i, nv)
    if (nv > 0) {
      float x32[4];
      // a recurrence that ended on a non-finite residual norm leaves its
      // iterate out (st->rr: the norm after its last iteration)
      if (st->phase != BP_INIT && isfinite(st->rr)) {
        const double scale = unit_scale ? 1.0 : st->ref;
        double x[4];
        ld4(sy.x + i, nv, x);
        ld4(sy.x32 + i, nv, x32);
_Pragma("unroll")
        for (int k = 0; k < 4; k++) x[k] = x[k] + scale * (double)x32[k];
        st4(sy.x + i, nv, x);
      }
_Pragma("unroll")
      for (int k = 0; k < 4; k++) x32[k] = 0.0f;
      st4(sy.x32 + i, nv, x32);
    }
  })
}

// r64 = b - A64 x64 with partials r64.r64 and b.b (FP64)
__global__ __launch_bounds__(kBT) void kb_res(BTables tb) {
  PSB_BATCH_TILES(tb, b_transiting(st), {
    double ax[4];
_Pragma("unroll")
    for (int k = 0; k < 4; k++) {
      const int i = r0 + k * kBT + (int)threadIdx.x;
      ax[k] = b_row<double>(sy, sy.a64, i < sy.n ? i : sy.n - 1, sy.x);
    }
    double rr = 0.0, bb = 0.0;
_Pragma("unroll")
    for (int k = 0; k < 4; k++) {
      const int i = r0 + k * kBT + (int)threadIdx.x;
      if (i < sy.n) {
        const double bi = sy.b[i];
        const double ri = bi - ax[k];
        sy.r64[i] = ri;
        rr += ri * ri;
        bb += bi * bi;
      }
    }
    tile_partials(rr, bb, tb.partials, t_);
  })
}

// CG: r = (float) r64, z = d.*r, partial r.z
__global__ __launch_bounds__(kBT) void kb_cg_reset(BTables tb) {
  PSB_BATCH_TILES(tb, b_transiting(st), {
    PSB_QUAD(i, nv)
    double rz = 0.0;
    if (nv > 0) {
      double r64[4];
      float r[4], z[4];
      ld4(sy.r64 + i, nv, r64);
      if (sy.d32) ld4(sy.d32 + i, nv, z);
_Pragma("unroll")
      for (int k = 0; k < 4; k++) {
        r[k] = (float)r64[k];
        z[k] = sy.d32 ? z[k] * r[k] : r[k];
        rz += (double)r[k] * (double)z[k];
      }
      st4(sy.r + i, nv, r);
      st4(sy.z + i, nv, z);
    }
    tile_partials(rz, 0.0, tb.partials, t_);
  })
}

// BiCGSTAB: r = q = (float)(r64 / ||r64||), partial r.r
__global__ __launch_bounds__(kBT) void kb_bi_restart(BTables tb) {
  PSB_BATCH_TILES(tb, b_transiting(st), {
    PSB_QUAD(i, nv)
    double rr = 0.0;
    if (nv > 0) {
      const double inv = 1.0 / st->rnorm;
      double r64[4];
      float r[4];
      ld4(sy.r64 + i, nv, r64);
_Pragma("unroll")
      for (int k = 0; k < 4; k++) {
        r[k] = (float)(r64[k] * inv);
        rr += (double)r[k] * (double)r[k];
      }
      st4(sy.r + i, nv, r);
      st4(sy.q + i, nv, r);
    }
    tile_partials(rr, 0.0, tb.partials, t_);
  })
}

// BiCGSTAB with residual replacement: r = (float) r64; a newly admitted or
// broken-down system also takes q = r.  Partial q.r.
__global__ __launch_bounds__(kBT) void kb_biru_reset(BTables tb) {
  PSB_BATCH_TILES(tb, b_transiting(st), {
    PSB_QUAD(i, nv)
    double qr = 0.0;
    if (nv > 0) {
      const bool fresh = st->phase != BP_TRANS;
      double r64[4];
      float r[4], q[4];
      ld4(sy.r64 + i, nv, r64);
      if (!fresh) ld4(sy.q + i, nv, q);
_Pragma("unroll")
      for (int k = 0; k < 4; k++) {
        r[k] = (float)r64[k];
        if (fresh) q[k] = r[k];
        qr += (double)q[k] * (double)r[k];
      }
      st4(sy.r + i, nv, r);
      if (fresh) st4(sy.q + i, nv, q);
    }
    tile_partials(qr, 0.0, tb.partials, t_);
  })
}

// ------------------------------------------------------------------ per-system finalisation
enum BFin {
  BF_CG_Q, BF_CG_X, BF_BI_V, BF_BI_T, BF_BI_X, BF_BIRU_X, BF_RES, BF_CG_RESET, BF_BI_RESTART,
  BF_BIRU_RESET
};

template <int OP>
__global__ __launch_bounds__(kBT) void kb_fin(BTables tb) {
  BState* st = tb.st + blockIdx.x;
  const bool trans =
      OP == BF_RES || OP == BF_CG_RESET || OP == BF_BI_RESTART || OP == BF_BIRU_RESET;
  if (trans ? !b_transiting(st) : !b_iterating(st)) return;
  const int2 rg = tb.range[blockIdx.x];
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < rg.y; k += kBT) {
    a += tb.partials[2 * (rg.x + k)];
    b += tb.partials[2 * (rg.x + k) + 1];
  }
  __shared__ double lds[2][kBT / kWave];
  a = block_sum<double, kBT>(a, lds[0]);
  b = block_sum<double, kBT>(b, lds[1]);
  if (threadIdx.x != 0) return;
  // A zero or non-finite denominator or scalar is a breakdown: the FP32
  // recurrence stops before its vectors take the value, the transition
  // keeps the last finite iterate and restarts from the FP64 residual.  (An
  // FP32 inner BiCGSTAB on a convection-dominated request at L = 159
  // overflowed near a breakdown; the exact-zero tests alone let the NaN run
  // to the iteration limit.)
  if (OP == BF_CG_Q || OP == BF_BI_V) {
    st->sigma = a;
    const double alpha = a != 0.0 ? st->rho / a : 0.0;
    if (a == 0.0 || !isfinite(a) || !isfinite(alpha)) {
      st->itx += 1;
      st->phase = BP_BREAK;
    } else {
      st->alpha = alpha;
    }
  } else if (OP == BF_BI_T) {
    if (a == 0.0 || !isfinite(a) || !isfinite(b)) {
      st->itx += 1;
      st->phase = BP_BREAK;
      return;
    }
    st->omega = b / a;
    if (st->omega == 0.0 || !isfinite(st->omega)) {
      st->itx += 1;
      st->phase = BP_BREAK;
    }
  } else if (OP == BF_CG_X) {
    st->itx += 1;
    st->it += 1;
    const double rn = sqrt(a);
    st->rr = rn;
    if (!isfinite(rn)) {   // the iterate is dropped at the transition (kb_accum)
      st->phase = BP_BREAK;
      return;
    }
    st->rho_old = st->rho;
    if (rn <= st->delta * st->ref || rn <= st->eps * st->bnorm || st->itx >= st->itmax) {
      st->phase = BP_TRANS;
      return;
    }
    st->rho = b;
    if (st->rho_old == 0.0) {
      st->itx += 1;
      st->phase = BP_BREAK;
    } else {
      st->beta = st->rho / st->rho_old;
    }
  } else if (OP == BF_BI_X) {
    st->itx += 1;
    st->it += 1;
    const double rn = sqrt(a);
    st->rr = rn;
    if (!isfinite(rn)) {
      st->phase = BP_BREAK;
      return;
    }
    if (rn <= st->inner_tol || st->itx >= st->itmax) {
      st->phase = BP_TRANS;
      return;
    }
    st->rho_old = st->rho;
    st->rho = b;
    const double beta = (st->rho / st->rho_old) * (st->alpha / st->omega);
    if (st->rho == 0.0 || !isfinite(beta)) {
      st->itx += 1;
      st->phase = BP_BREAK;
    } else {
      st->beta = beta;
    }
  } else if (OP == BF_BIRU_X) {
    st->itx += 1;
    st->it += 1;
    const double rn = sqrt(a);
    st->rr = rn;
    st->rho_old = st->rho;
    if (!isfinite(rn)) {
      st->phase = BP_BREAK;
      return;
    }
    if (rn <= st->delta * st->ref || rn <= st->eps * st->bnorm || st->itx >= st->itmax) {
      st->phase = BP_TRANS;   // rho and beta follow from the replaced residual
      return;
    }
    st->rho = b;
    const double beta = (st->rho / st->rho_old) * (st->alpha / st->omega);
    if (st->rho == 0.0 || !isfinite(beta)) {
      st->itx += 1;
      st->phase = BP_BREAK;
    } else {
      st->beta = beta;
    }
  } else if (OP == BF_RES) {
    st->rnorm = sqrt(a);
    if (st->phase == BP_INIT) st->bnorm = sqrt(b);
    else st->outer += 1;
    const bool conv = st->bnorm == 0.0 ? st->rnorm <= st->eps : st->rnorm <= st->eps * st->bnorm;
    if (conv) st->done = 1;
    else if (st->itx >= st->itmax) st->done = 2;
  } else if (OP == BF_CG_RESET) {
    // the new residual's r.z; the recurrence keeps p unless it (re)starts
    const double rho = a;
    st->ref = st->rnorm;
    if (st->phase == BP_TRANS && st->rho_old != 0.0) {
      st->beta = rho / st->rho_old;
    } else {
      st->it = 0;
    }
    st->rho = rho;
    st->phase = BP_RUN;
  } else if (OP == BF_BIRU_RESET) {
    const double rho = a;   // q.r with the replaced residual
    st->ref = st->rnorm;
    if (st->phase == BP_TRANS && st->rho_old != 0.0 && rho != 0.0) {
      st->beta = (rho / st->rho_old) * (st->alpha / st->omega);
    } else {
      st->it = 0;           // (re)start: p = r, q = r was taken by the reset kernel
    }
    st->rho = rho;
    st->phase = BP_RUN;
  } else if (OP == BF_BI_RESTART) {
    // inner solve of A d = r64/||r64|| from d = 0: q = r, rho = q.r
    st->ref = st->rnorm;
    double tol = st->inner_eps;
    if (st->bnorm > 0.0) {
      const double need = 0.5 * st->eps * st->bnorm / st->rnorm;
      if (need > tol) tol = need < 0.5 ? need : 0.5;
    }
    st->inner_tol = tol * sqrt(a);
    st->it = 0;
    st->rho = a;
    st->rho_old = 0.0;
    st->phase = BP_RUN;
  }
}

__global__ __launch_bounds__(kBT) void kb_fill(double* __restrict__ x, int n, double v) {
  for (int i = blockIdx.x * kBT + threadIdx.x; i < n; i += gridDim.x * kBT) x[i] = v;
}

// ------------------------------------------------------------------ host engine
struct BMat {
  HllDev* h = nullptr;
  float* d32 = nullptr;   // owned
};

struct BSlot {
  bool live = false;
  bool reported = false;
  int mat = -1;
  long first_step = -1;   // first step whose state copy describes this system
  size_t off = 0, bytes = 0;
};

struct Stage {            // pinned staging of one step's table uploads
  int2* tiles = nullptr;
  int2* range = nullptr;
  BSys* sys = nullptr;
  int* ntiles = nullptr;
  BState* states = nullptr;
  hipEvent_t ev = nullptr;
  bool used = false;
};

struct Batch {
  int method = 0, nslots = 0, check_every = 8;
  double delta = 1e-2, inner_eps = 1e-4;
  bool use_graph = true;
  hipStream_t stream = nullptr;
  char* arena = nullptr;
  size_t arena_bytes = 0;
  int tmax = 0;
  // device tables
  int2* tiles = nullptr;
  int* ntiles = nullptr;
  int2* range = nullptr;
  BSys* sys = nullptr;
  BState* st = nullptr;
  double* partials = nullptr;
  // host
  std::vector<BMat> mats;
  std::vector<BSlot> slots;
  std::vector<BSys> sys_h;
  std::map<size_t, size_t> free_ranges;       // arena offset -> bytes
  std::vector<std::pair<int, BState>> pending;  // states to upload before the next step
  bool dirty = true;
  Stage stage[2];
  BState* ring[2] = {nullptr, nullptr};        // pinned state copies (pipelined poll)
  hipEvent_t ring_ev[2] = {nullptr, nullptr};
  long steps = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static BTables tables(const Batch& B) {
  BTables t;
  t.tiles = B.tiles;
  t.ntiles = B.ntiles;
  t.range = B.range;
  t.sys = B.sys;
  t.st = B.st;
  t.partials = B.partials;
  t.tcap = B.tmax;
  return t;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// arena bytes one system of n rows needs
static size_t sys_bytes(int method, int n) {
  const size_t d = align256(sizeof(double) * (size_t)n), f = align256(sizeof(float) * (size_t)n);
  return 3 * d + (method == 0 ? 5 : 9) * f;   // BiCGSTAB (1, 2): v, f, s, t too
}

static bool arena_take(Batch& B, size_t bytes, size_t* off) {
  for (auto it = B.free_ranges.begin(); it != B.free_ranges.end(); ++it) {
    if (it->second >= bytes) {
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t roff = it->first + bytes;
      B.free_ranges.erase(it);
      if (rest) B.free_ranges[roff] = rest;
      return true;
    }
  }
  return false;
}

static void arena_give(Batch& B, size_t off, size_t bytes) {
  auto it = B.free_ranges.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != B.free_ranges.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    B.free_ranges.erase(nx);
  }
  if (it != B.free_ranges.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      B.free_ranges.erase(it);
    }
  }
}

// The kernels of one step (check_every iterations + the transition pass).
static void enqueue_step(Batch& B) {
  const BTables tb = tables(B);
  dim3 g(kBGrid), blk(kBT), gs(B.nslots);
  hipStream_t s = B.stream;
  for (int k = 0; k < B.check_every; k++) {
    if (B.method == 0) {
      hipLaunchKernelGGL(kb_cg_p, g, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_cg_q, g, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_fin<BF_CG_Q>, gs, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_cg_x, g, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_fin<BF_CG_X>, gs, blk, 0, s, tb);
    } else {
      hipLaunchKernelGGL(kb_bi_a, g, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_bi_v, g, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_fin<BF_BI_V>, gs, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_bi_s, g, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_bi_t, g, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_fin<BF_BI_T>, gs, blk, 0, s, tb);
      hipLaunchKernelGGL(kb_bi_x, g, blk, 0, s, tb);
      if (B.method == 1) hipLaunchKernelGGL(kb_fin<BF_BI_X>, gs, blk, 0, s, tb);
      else hipLaunchKernelGGL(kb_fin<BF_BIRU_X>, gs, blk, 0, s, tb);
    }
  }
  hipLaunchKernelGGL(kb_accum, g, blk, 0, s, tb, B.method == 1 ? 0 : 1);
  hipLaunchKernelGGL(kb_res, g, blk, 0, s, tb);
  hipLaunchKernelGGL(kb_fin<BF_RES>, gs, blk, 0, s, tb);
  if (B.method == 0) {
    hipLaunchKernelGGL(kb_cg_reset, g, blk, 0, s, tb);
    hipLaunchKernelGGL(kb_fin<BF_CG_RESET>, gs, blk, 0, s, tb);
  } else if (B.method == 2) {
    hipLaunchKernelGGL(kb_biru_reset, g, blk, 0, s, tb);
    hipLaunchKernelGGL(kb_fin<BF_BIRU_RESET>, gs, blk, 0, s, tb);
  } else {
    hipLaunchKernelGGL(kb_bi_restart, g, blk, 0, s, tb);
    hipLaunchKernelGGL(kb_fin<BF_BI_RESTART>, gs, blk, 0, s, tb);
  }
}

// Upload the tables (when the live set changed) and the pending states
// through this step's pinned stage.
static int upload(Batch& B) {
  Stage& S = B.stage[B.steps & 1];
  if (S.used) PSB_HIP_CHECK(hipEventSynchronize(S.ev));
  if (B.dirty) {
    int nt = 0;
    for (int k = 0; k < B.nslots; k++) {
      const BSlot& sl = B.slots[k];
      const int nk = sl.live ? (B.sys_h[k].n + kTileRows - 1) / kTileRows : 0;
      if (nt + nk > B.tmax) return set_error(PSB_HIP_RUNTIME, "batch", "tile table overflow");
      S.range[k] = make_int2(nt, nk);
      for (int j = 0; j < nk; j++) S.tiles[nt + j] = make_int2(k, j * kTileRows);
      nt += nk;
    }
    *S.ntiles = nt;
    memcpy(S.sys, B.sys_h.data(), sizeof(BSys) * B.nslots);
    if (nt > 0)
      PSB_HIP_CHECK(hipMemcpyAsync(B.tiles, S.tiles, sizeof(int2) * nt, hipMemcpyHostToDevice,
                                   B.stream));
    PSB_HIP_CHECK(hipMemcpyAsync(B.ntiles, S.ntiles, sizeof(int), hipMemcpyHostToDevice, B.stream));
    PSB_HIP_CHECK(hipMemcpyAsync(B.range, S.range, sizeof(int2) * B.nslots,
                                 hipMemcpyHostToDevice, B.stream));
    PSB_HIP_CHECK(hipMemcpyAsync(B.sys, S.sys, sizeof(BSys) * B.nslots, hipMemcpyHostToDevice,
                                 B.stream));
    B.dirty = false;
  }
  for (auto& ps : B.pending) {
    B.slots[ps.first].first_step = B.steps;
    S.states[ps.first] = ps.second;
    PSB_HIP_CHECK(hipMemcpyAsync(B.st + ps.first, S.states + ps.first, sizeof(BState),
                                 hipMemcpyHostToDevice, B.stream));
  }
  B.pending.clear();
  PSB_HIP_CHECK(hipEventRecord(S.ev, B.stream));
  S.used = true;
  return PSB_HIP_SUCCESS;
}

static void destroy(Batch* B) {
  if (!B) return;
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  for (BMat& m : B->mats)
    if (m.d32) (void)hipFree(m.d32);
  void* dev[] = {B->arena, B->tiles, B->ntiles, B->range, B->sys, B->st, B->partials};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  for (Stage& S : B->stage) {
    void* hp[] = {S.tiles, S.range, S.sys, S.ntiles, S.states};
    for (void* p : hp)
      if (p) (void)hipHostFree(p);
    if (S.ev) (void)hipEventDestroy(S.ev);
  }
  for (int k = 0; k < 2; k++) {
    if (B->ring[k]) (void)hipHostFree(B->ring[k]);
    if (B->ring_ev[k]) (void)hipEventDestroy(B->ring_ev[k]);
  }
  delete B;
}

__global__ void kb_d2s(int n, const double* __restrict__ a, float* __restrict__ b) {
  for (int i = blockIdx.x * kBT + threadIdx.x; i < n; i += gridDim.x * kBT) b[i] = (float)a[i];
}

}  // namespace psbhip

using namespace psbhip;

extern "C" {

int psb_hip_batch_create(void** batch, int method, int max_systems, int64_t arena_bytes,
                         int check_every, double delta, double inner_eps, int use_graph) {
  PSB_ENSURE_INIT();
  if (!batch) return PSB_HIP_INVALID;
  *batch = nullptr;
  if (method < 0 || method > 2 || max_systems < 1 || arena_bytes <= 0 || check_every < 1)
    return set_error(PSB_HIP_INVALID, __func__, "method, max_systems, arena or check_every");
  Batch* B = new Batch();
  B->method = method;
  B->nslots = max_systems;
  B->check_every = check_every;
  B->delta = delta;
  B->inner_eps = inner_eps;
  B->use_graph = use_graph != 0;
  B->stream = ctx().stream;
  B->arena_bytes = align256((size_t)arena_bytes);
  // tiles: every arena row could be a tile row, plus one partial tile per slot
  const size_t row_bytes = sys_bytes(method, kTileRows) / kTileRows;
  B->tmax = (int)std::min<size_t>(B->arena_bytes / (row_bytes * kTileRows) + (size_t)max_systems,
                                  (size_t)1 << 30);
  B->tmax = std::max(B->tmax, kBGrid);   // the tile kernels read entry blockIdx.x unconditionally
  B->slots.resize(max_systems);
  B->sys_h.assign(max_systems, BSys{});
  B->free_ranges[0] = B->arena_bytes;
  hipError_t e = hipSuccess;
  auto dm = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  dm((void**)&B->arena, B->arena_bytes);
  dm((void**)&B->tiles, sizeof(int2) * B->tmax);
  dm((void**)&B->ntiles, sizeof(int));
  dm((void**)&B->range, sizeof(int2) * max_systems);
  dm((void**)&B->sys, sizeof(BSys) * max_systems);
  dm((void**)&B->st, sizeof(BState) * max_systems);
  dm((void**)&B->partials, sizeof(double) * 2 * (size_t)B->tmax);
  for (Stage& S : B->stage) {
    auto hm = [&](void** p, size_t bytes) {
      if (e == hipSuccess) e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    };
    hm((void**)&S.tiles, sizeof(int2) * B->tmax);
    hm((void**)&S.range, sizeof(int2) * max_systems);
    hm((void**)&S.sys, sizeof(BSys) * max_systems);
    hm((void**)&S.ntiles, sizeof(int));
    hm((void**)&S.states, sizeof(BState) * max_systems);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S.ev, hipEventDisableTiming);
  }
  for (int k = 0; k < 2; k++) {
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&B->ring[k], sizeof(BState) * max_systems, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&B->ring_ev[k], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    destroy(B);
    return set_error(e == hipErrorOutOfMemory ? PSB_HIP_OUTOFMEM : PSB_HIP_RUNTIME, __func__,
                     hipGetErrorString(e));
  }
  // every slot starts finished (no tiles, skipped by every kernel)
  std::vector<BState> idle(max_systems);
  for (BState& s : idle) {
    memset(&s, 0, sizeof(BState));
    s.done = 1;
  }
  if (hipMemcpy(B->st, idle.data(), sizeof(BState) * max_systems, hipMemcpyHostToDevice) !=
      hipSuccess) {
    destroy(B);
    return set_error(PSB_HIP_RUNTIME, __func__, "state upload");
  }
  *batch = B;
  return PSB_HIP_SUCCESS;
}

int psb_hip_batch_free(void* batch) {
  destroy((Batch*)batch);
  return PSB_HIP_SUCCESS;
}

int psb_hip_batch_register(void* batch, void* dmat, void* diag, int* mat_id) {
  Batch* B = (Batch*)batch;
  HllDev* h = (HllDev*)dmat;
  if (!B || !h || !mat_id) return PSB_HIP_INVALID;
  if (h->elem_type != PSB_HIP_TYPE_DOUBLE || h->rows < 1 || h->max_col > h->rows)
    return set_error(PSB_HIP_INVALID, __func__, "needs a square double HLL matrix");
  if (h->perm)
    return set_error(PSB_HIP_UNSUPPORTED, __func__, "row-sorted HLL: rows must be in order");
  if (!h->val32) {
    int rc = psb_hip_hll_make_single(h);
    if (rc) return rc;
  }
  BMat m;
  m.h = h;
  if (diag) {
    MultiVec* d = (MultiVec*)diag;
    if (d->elem_type != PSB_HIP_TYPE_DOUBLE || d->size < h->rows || d->count != 1)
      return set_error(PSB_HIP_INVALID, __func__, "diag");
    PSB_HIP_CHECK(hipMalloc((void**)&m.d32, sizeof(float) * (size_t)h->rows));
    hipLaunchKernelGGL(kb_d2s, dim3(stream_grid(h->rows)), dim3(kBT), 0, B->stream, h->rows,
                       (const double*)d->v, m.d32);
    PSB_HIP_LAUNCH_CHECK();
  }
  B->mats.push_back(m);
  *mat_id = (int)B->mats.size() - 1;
  return PSB_HIP_SUCCESS;
}

int64_t psb_hip_batch_footprint(void* batch, int mat_id) {
  Batch* B = (Batch*)batch;
  if (!B || mat_id < 0 || mat_id >= (int)B->mats.size()) return -1;
  return (int64_t)sys_bytes(B->method, B->mats[mat_id].h->rows);
}

int psb_hip_batch_submit(void* batch, int mat_id, void* bv, void* x0v, double eps, int itmax,
                         int* sys_id) {
  Batch* B = (Batch*)batch;
  if (!B || !sys_id || mat_id < 0 || mat_id >= (int)B->mats.size() || itmax < 0)
    return PSB_HIP_INVALID;
  *sys_id = -1;
  const BMat& m = B->mats[mat_id];
  const int n = m.h->rows;
  MultiVec* b = (MultiVec*)bv;
  MultiVec* x0 = (MultiVec*)x0v;
  if (!b || b->elem_type != PSB_HIP_TYPE_DOUBLE || b->size < n ||
      (x0 && (x0->elem_type != PSB_HIP_TYPE_DOUBLE || x0->size < n)))
    return set_error(PSB_HIP_INVALID, __func__, "b/x0 type or size");
  int slot = -1;
  for (int k = 0; k < B->nslots; k++)
    if (!B->slots[k].live) {
      slot = k;
      break;
    }
  const size_t bytes = sys_bytes(B->method, n);
  size_t off = 0;
  if (slot < 0 || !arena_take(*B, bytes, &off)) return PSB_HIP_BUSY;
  BSlot& sl = B->slots[slot];
  sl.live = true;
  sl.reported = false;
  sl.mat = mat_id;
  sl.off = off;
  sl.bytes = bytes;
  const size_t d = align256(sizeof(double) * (size_t)n), f = align256(sizeof(float) * (size_t)n);
  char* base = B->arena + off;
  BSys sy{};
  sy.n = n;
  sy.hksz = m.h->hksz;
  sy.hkoffs = m.h->hkoffs;
  sy.ja = m.h->ja;
  const bool i16 = hll_use_idx16(m.h);
  sy.res16 = i16 ? m.h->res16 : nullptr;
  sy.base16 = i16 ? m.h->base16 : nullptr;
  sy.a32 = m.h->val32;
  sy.a64 = (const double*)m.h->val;
  sy.d32 = m.d32;
  sy.b = (const double*)base;
  sy.x = (double*)(base + d);
  sy.r64 = (double*)(base + 2 * d);
  float* fv[9];
  for (int k = 0; k < (B->method == 0 ? 5 : 9); k++) fv[k] = (float*)(base + 3 * d + k * f);
  sy.x32 = fv[0];
  sy.r = fv[1];
  sy.p = fv[2];
  sy.q = fv[3];
  sy.z = fv[4];
  if (B->method != 0) {
    sy.v = fv[5];
    sy.f = fv[6];
    sy.s = fv[7];
    sy.t = fv[8];
  }
  B->sys_h[slot] = sy;
  PSB_HIP_CHECK(hipMemcpyAsync((void*)sy.b, b->v, sizeof(double) * (size_t)n,
                               hipMemcpyDeviceToDevice, B->stream));
  if (x0) {
    PSB_HIP_CHECK(hipMemcpyAsync(sy.x, x0->v, sizeof(double) * (size_t)n,
                                 hipMemcpyDeviceToDevice, B->stream));
  } else {
    hipLaunchKernelGGL(kb_fill, dim3(stream_grid(n)), dim3(kBT), 0, B->stream, sy.x, n, 0.0);
    PSB_HIP_LAUNCH_CHECK();
  }
  BState s;
  memset(&s, 0, sizeof(BState));
  s.phase = BP_INIT;
  s.eps = eps;
  s.itmax = itmax;
  s.delta = B->delta;
  s.inner_eps = B->inner_eps;
  B->pending.emplace_back(slot, s);
  B->dirty = true;
  *sys_id = slot;
  return PSB_HIP_SUCCESS;
}

int psb_hip_batch_can_admit(void* batch, int mat_id) {
  Batch* B = (Batch*)batch;
  if (!B || mat_id < 0 || mat_id >= (int)B->mats.size()) return 0;
  bool slot = false;
  for (const BSlot& sl : B->slots)
    if (!sl.live) {
      slot = true;
      break;
    }
  if (!slot) return 0;
  const size_t bytes = sys_bytes(B->method, B->mats[mat_id].h->rows);
  for (const auto& fr : B->free_ranges)
    if (fr.second >= bytes) return 1;
  return 0;
}

int psb_hip_batch_step(void* batch) {
  Batch* B = (Batch*)batch;
  if (!B) return PSB_HIP_INVALID;
  int rc = upload(*B);
  if (rc) return rc;
  if (B->use_graph) {
    if (!B->exec) {
      PSB_HIP_CHECK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
      enqueue_step(*B);
      PSB_HIP_CHECK(hipStreamEndCapture(B->stream, &B->graph));
      PSB_HIP_CHECK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    }
    PSB_HIP_CHECK(hipGraphLaunch(B->exec, B->stream));
  } else {
    enqueue_step(*B);
    PSB_HIP_LAUNCH_CHECK();
  }
  const int k = (int)(B->steps & 1);
  PSB_HIP_CHECK(hipMemcpyAsync(B->ring[k], B->st, sizeof(BState) * B->nslots,
                               hipMemcpyDeviceToHost, B->stream));
  PSB_HIP_CHECK(hipEventRecord(B->ring_ev[k], B->stream));
  B->steps++;
  return PSB_HIP_SUCCESS;
}

int psb_hip_batch_poll(void* batch, int wait_last, int max_out, int* ids, int* iters,
                       int* outer, double* relres, int* status, int* nout, int* nlive) {
  Batch* B = (Batch*)batch;
  if (!B || !nout) return PSB_HIP_INVALID;
  *nout = 0;
  if (nlive) {
    int c = 0;
    for (const BSlot& sl : B->slots) c += sl.live && !sl.reported;
    *nlive = c;
  }
  if (B->steps == 0) return PSB_HIP_SUCCESS;
  const long sidx = (wait_last || B->steps < 2) ? B->steps - 1 : B->steps - 2;
  const int k = (int)(sidx & 1);
  PSB_HIP_CHECK(hipEventSynchronize(B->ring_ev[k]));
  const BState* S = B->ring[k];
  for (int s = 0; s < B->nslots && *nout < max_out; s++) {
    BSlot& sl = B->slots[s];
    // copies queued before the system's first step still show the slot's
    // previous occupant (or the idle state)
    if (!sl.live || sl.reported || sl.first_step < 0 || sidx < sl.first_step) continue;
    if (S[s].done == 0) continue;
    sl.reported = true;
    const int j = (*nout)++;
    if (ids) ids[j] = s;
    if (iters) iters[j] = S[s].itx;
    if (outer) outer[j] = S[s].outer;
    if (relres) relres[j] = S[s].bnorm > 0.0 ? S[s].rnorm / S[s].bnorm : S[s].rnorm;
    if (status) status[j] = S[s].done;
  }
  if (nlive) {
    int c = 0;
    for (const BSlot& sl : B->slots) c += sl.live && !sl.reported;
    *nlive = c;
  }
  return PSB_HIP_SUCCESS;
}

int psb_hip_batch_result(void* batch, int sys_id, void* xv) {
  Batch* B = (Batch*)batch;
  if (!B || sys_id < 0 || sys_id >= B->nslots || !B->slots[sys_id].live) return PSB_HIP_INVALID;
  BSlot& sl = B->slots[sys_id];
  if (!sl.reported) return set_error(PSB_HIP_INVALID, __func__, "system not finished");
  if (xv) {
    MultiVec* x = (MultiVec*)xv;
    const int n = B->sys_h[sys_id].n;
    if (x->elem_type != PSB_HIP_TYPE_DOUBLE || x->size < n)
      return set_error(PSB_HIP_INVALID, __func__, "x type or size");
    PSB_HIP_CHECK(hipMemcpyAsync(x->v, B->sys_h[sys_id].x, sizeof(double) * (size_t)n,
                                 hipMemcpyDeviceToDevice, B->stream));
  }
  arena_give(*B, sl.off, sl.bytes);
  sl = BSlot{};
  B->sys_h[sys_id] = BSys{};
  B->dirty = true;
  return PSB_HIP_SUCCESS;
}

}  // extern "C"
