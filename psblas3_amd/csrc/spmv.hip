// Sparse matrix-vector products on gfx950: HLL (hacked ELLPACK), ELL and
// CSR device matrices, FP64 and FP32, plus the COO -> HLL/ELL device fills.
//
// Replaces cuda/hlldev.c, cuda/elldev.c, the SPGPU kernels
// cuda/spgpu/kernels/hell_spmv_base_template.cuh and ell_spmv_base_template.cuh,
// the device fills cuda/CUDA/psi_cuda_CopyCooToHlg.cuh / CopyCooToElg.cuh and
// the cuSPARSE CSR product wrapped by cuda/fcusparse_fct.h:108-124.
//
// HLL layout in HBM (the reference's, ext/psb_d_hll_mat_mod.f90:38-104):
// rows are grouped in hacks of hksz consecutive rows; hack h occupies
// val/ja[hkoffs[h] .. hkoffs[h+1]) stored column-major inside the hack
// (slot (r, j) at hkoffs[h] + j*hksz + r), so lane r of a wavefront reads
// consecutive addresses for every j: one 512-byte (FP64) / 256-byte (FP32 or
// index) coalesced load per column of the hack.  Kernel: one row per lane,
// 64 rows per wavefront (one hack at hksz = 64, two at hksz = 32); val and
// ja are streamed with non-temporal loads (they are read once), x is
// gathered through the XCD's L2 and the 256 MB Infinity Cache, y written
// once.  Blocks are remapped so each XCD walks a contiguous range of rows,
// which keeps the stencil's x neighbours (+-1, +-L, +-L^2) in one L2.
#include <string.h>

#include <algorithm>
#include <vector>

#include "psb_hip_internal.hpp"

namespace psbhip {

static int g_hll_variant = 0;
// multi-RHS HLL products over two rows per lane (k_hll_spmm_pair)
static bool g_spmm_pair = true;
// FP64 ELL products over two rows per lane (k_ell_spmv_pair)
static bool g_ell_pair = true;
// FP64 transposed products over two rows per lane (k_tmv_pair)
static bool g_tmv_pair = true;
// XCD order: chunks of blocks dealt to the XCDs in turn (0: one contiguous
// range per XCD; -1, the default: per matrix, xcd_chunk_for, and 16 blocks
// for the row-block CSR kernel).  On the 7-point stencils at L = 200 and 256
// chunks of 4-64 blocks measured 1-2% (FP64 pairs) and 2-5% (FP32) faster
// than launch order for HLL, contiguous ranges 3-5% slower; for ELL the
// chunks beat contiguous ranges by 9% (FP64) and 6% (FP32), for the
// row-block CSR kernel launch order by 1-2%.
static int g_hll_xchunk = -1;
constexpr int kXchunkDefault = 16;
// 0 auto (the row-block kernel, on the column slices when the matrix has
// them), 1 uniform lanes per row, 2 length-binned, 3 row-block on the whole
// matrix, 4 = 0
static int g_csr_variant = 0;
// Column slices built by psb_hip_csr_write: -1 automatic (matrices whose x is
// over kCsrSliceMinX bytes and whose entries mostly lie more than cols/16
// from the diagonal, in slices of about kCsrSliceBytes of x), 0 or 1 none,
// >= 2 that many.  On the n = 4M power-law matrix (uniform random columns, x
// 32 MB) 2, 3, 4, 5, 6 and 8 slices measured 942, 846, 779, 751, 807 and
// 916 us against 1041 us unsliced: a slice's x range stays in the XCDs' L2
// while its rows run, so fewer of the random gathers miss.
static int g_csr_slices = -1;
// Column slices of row-sorted HLL matrices (psb_hip_hll_copy_coo_sorted):
// the CSR rule for when to slice, slices of about kHllSliceBytes of x (FP64)
// or kHllSliceBytes32 (FP32); -1 automatic, 0 or 1 none, >= 2 that many.  On
// the n = 4M power-law matrix (x 32 MB FP64) 3, 4, 5, 6 and 8 slices
// measured 894, 866-869, 908, 1003 and 1241 us against 1008 us unsliced;
// FP32 (x 16 MB) 2, 3 and 4 slices 621, 571 and 633 us against 785 us.
static int g_hll_slices = -1;
constexpr size_t kHllSliceBytes = (size_t)8 << 20;
constexpr size_t kHllSliceBytes32 = 5600000;
constexpr size_t kCsrSliceMinX = (size_t)12 << 20;
constexpr size_t kCsrSliceBytes = 6400000;
constexpr int kCsrMaxSlices = 16;

constexpr int kCsrEntriesPerLane = 4;
constexpr int kCsrLongRow = 512;   // binned / row-block: longer rows take one wavefront each
constexpr int kCsrBlockNnz = 2048; // row-block kernel: entries of a block (products in LDS)

// lanes per row of the binned kernel: a power of two near len / 4, 1..64
inline int csr_class(int len) {
  int g = 1, c = 0;
  while (g < 64 && g * kCsrEntriesPerLane < len) {
    g <<= 1;
    c++;
  }
  return c;
}

// ------------------------------------------------------------------ HLL SpMV
// HK > 0: compile-time hack size (32 or 64); HK == 0: any hksz.  ord: the
// block order (launch, chunked XCD or plane-tile).
template <typename T, typename TV, int HK, bool IDX16>
__global__ __launch_bounds__(kBlock) void k_hll_spmv(int nr, int hksz, int nlog, BlockOrder ord,
                                                     const int* __restrict__ hkoffs,
                                                     const int* __restrict__ ja,
                                                     const short* __restrict__ res16,
                                                     const int* __restrict__ base16,
                                                     const TV* __restrict__ val,
                                                     const T* __restrict__ x, T alpha, T beta,
                                                     T* __restrict__ y) {
  const int lb = order_block(ord, nlog);
  if (lb >= nlog) return;
  const int row = lb * kBlock + threadIdx.x;
  const int r = row < nr ? row : nr - 1;
  const int stride = HK > 0 ? HK : hksz;
  int hk = r / stride;
  // hack size = wavefront size: the wavefront is one hack (rows past nr
  // clamp into the last one), so its offsets come through scalar loads
  if (HK == kWave) hk = __builtin_amdgcn_readfirstlane(hk);
  const int lane = r - hk * stride;
  const int off0 = hkoffs[hk];
  const int w = (hkoffs[hk + 1] - off0) / stride;
  T acc;
  if (IDX16)
    acc = slot_dot16<T, TV>(res16 + off0 + lane, base16 + off0 / stride, val + off0 + lane,
                            stride, w, r, x);
  else
    acc = slot_dot<T, TV>(ja + off0 + lane, val + off0 + lane, stride, w, x);
  if (row < nr) store_y(y, row, alpha, acc, beta);
}

// Two adjacent rows per lane (rows 2t and 2t+1 of a 2*kBlock-row tile, same
// hack since hksz is even): 8-byte index and 2-element value loads.  Rows
// past nr read the last hack's padding rows (the hack stores hksz rows).
template <typename T, typename TV, int HK, bool IDX16>
__global__ __launch_bounds__(kBlock) void k_hll_spmv_pair(int nr, int hksz, int nlog,
                                                          BlockOrder ord,
                                                          const int* __restrict__ hkoffs,
                                                          const int* __restrict__ ja,
                                                          const short* __restrict__ res16,
                                                          const int* __restrict__ base16,
                                                          const TV* __restrict__ val,
                                                          const T* __restrict__ x, T alpha,
                                                          T beta, T* __restrict__ y) {
  const int lb = order_block(ord, nlog);
  if (lb >= nlog) return;
  const int row = lb * (2 * kBlock) + 2 * (int)threadIdx.x;
  const int last = (nr - 1) & ~1;
  const int r = row < last ? row : last;
  const int stride = HK > 0 ? HK : hksz;
  const int hk = r / stride;
  const int lane = r - hk * stride;
  int off0, w;
  hll_pair_hack<HK>(hkoffs, nr, hk, stride, off0, w);
  T a0, a1;
  if (IDX16)
    slot_dot16_pair<T, TV>(res16 + off0 + lane, base16 + off0 / stride, val + off0 + lane, stride,
                           w, r, x, a0, a1);
  else
    slot_dot_pair<T, TV>(ja + off0 + lane, val + off0 + lane, stride, w, x, a0, a1);
  if (row + 1 < nr && (reinterpret_cast<uintptr_t>(y) & (2 * sizeof(T) - 1)) == 0) {
    // both rows in one 2-element store (two separate non-temporal stores,
    // 2 elements apart across the lanes, wrote 84 MB for the 64 MB of y at L=200)
    T r0 = alpha * a0, r1 = alpha * a1;
    if (beta != T(0)) {
      const vec2<T> yo = *reinterpret_cast<const vec2<T>*>(y + row);
      r0 = fma(beta, yo.x, r0);
      r1 = fma(beta, yo.y, r1);
    }
    vec2<T> yn;
    yn.x = r0;
    yn.y = r1;
    __builtin_nontemporal_store(yn, reinterpret_cast<vec2<T>*>(y + row));
  } else {
    if (row < nr) store_y(y, row, alpha, a0, beta);
    if (row + 1 < nr) store_y(y, row + 1, alpha, a1, beta);
  }
}

// Row-sorted HLL product (the permuted SPGPU kernel,
// cuda/spgpu/kernels/hell_spmv_base_template.cuh:227-252, which loops each
// row to its own length, :197-214): position p of hack h holds row perm[p]
// and irn[p] entries.  Threads take hksz-lane work units (hack, first slot,
// end slot, partial slot): lane p sums val * x over its row's slots inside
// the unit in slot order (slot_dot: loads in groups issued before the
// gathers), then writes y(perm[p]) = alpha*acc + beta*y, or, for a unit of a
// hack cut into several, its partial sum into part.  Padding slots past a
// row's length are not read.
template <typename T, typename TV>
__global__ __launch_bounds__(kBlock) void k_hll_spmv_units(int nr, int hksz, int nunits,
                                                           const int4* __restrict__ units,
                                                           const int* __restrict__ hkoffs,
                                                           const int* __restrict__ irn,
                                                           const int* __restrict__ perm,
                                                           const int* __restrict__ ja,
                                                           const TV* __restrict__ val,
                                                           const T* __restrict__ x, T alpha,
                                                           T beta, T* __restrict__ y,
                                                           T* __restrict__ part) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int u = (int)(t / hksz);
  if (u >= nunits) return;
  const int lane = (int)(t - (int64_t)u * hksz);
  const int4 un = units[u];
  const int p = un.x * hksz + lane;
  const int len = p < nr ? irn[p] : 0;
  const int e = len < un.z ? len : un.z;
  const int w = e > un.y ? e - un.y : 0;
  const size_t off = (size_t)hkoffs[un.x] + (size_t)un.y * hksz + lane;
  const T acc = slot_dot<T, TV>(ja + off, val + off, hksz, w, x);
  if (un.w >= 0) part[(size_t)un.w * hksz + lane] = acc;
  else if (p < nr) store_y(y, perm[p] - 1, alpha, acc, beta);
}

// The hacks cut into several units: lane p adds its units' partial sums in
// unit (slot) order and writes y(perm[p]).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_hll_units_combine(int nr, int hksz, int nsplit,
                                                              const int4* __restrict__ splits,
                                                              const int* __restrict__ perm,
                                                              const T* __restrict__ part,
                                                              T alpha, T beta,
                                                              T* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int s = (int)(t / hksz);
  if (s >= nsplit) return;
  const int lane = (int)(t - (int64_t)s * hksz);
  const int4 sp = splits[s];
  const int p = sp.x * hksz + lane;
  if (p >= nr) return;
  T acc = part[(size_t)sp.y * hksz + lane];
  for (int k = 1; k < sp.z; k++) acc += part[(size_t)(sp.y + k) * hksz + lane];
  store_y(y, perm[p] - 1, alpha, acc, beta);
}

// Multi-RHS HLL product (psb_d_hll_csmm, ext/impl/psb_d_hll_csmm.f90:33-235):
// one row per lane and up to NC columns of x/y (column-major, ldx/ldy
// elements apart) per launch.  Slots go in groups of S (S * NC = 32 x
// gathers per group, S <= 4) whose (ja, val) loads issue together, one group
// ahead of its use; every x gather of a group issues before its first fma.
// Slots past the row's width read its last slot, product dropped, as in
// slot_group; each slot is loaded once and used for every column, each
// column accumulating in slot order.
template <typename T, int NC, bool IDX16>
__global__ __launch_bounds__(kBlock) void k_hll_spmm(int nr, int hksz, int nlog,
                                                     const int* __restrict__ hkoffs,
                                                     const int* __restrict__ ja,
                                                     const short* __restrict__ res16,
                                                     const int* __restrict__ base16,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ x, int ldx,
                                                     T alpha, T beta, T* __restrict__ y,
                                                     int ldy, int ncols, BlockOrder ord) {
  constexpr int S = NC >= 32 ? 1 : NC >= 16 ? 2 : NC >= 8 ? 4 : 4;
  const int lb = order_block(ord, nlog);
  if (lb >= nlog) return;
  const int row = lb * kBlock + threadIdx.x;
  const int r = row < nr ? row : nr - 1;
  const int hk = r / hksz;
  const int lane = r - hk * hksz;
  const int off0 = hkoffs[hk];
  const int w = (hkoffs[hk + 1] - off0) / hksz;
  const int* pj = ja + off0 + lane;
  const short* pe = res16 + off0 + lane;
  const int* pb = base16 + off0 / hksz;
  const T* pv = val + off0 + lane;
  auto load_group = [&](int j, int (&c)[S], T (&v)[S]) {
#pragma unroll
    for (int k = 0; k < S; k++) {
      const int jj = j + k < w ? j + k : w - 1;
      if (IDX16)
        c[k] = r + pb[jj] + (int)ld_stream(pe + (size_t)jj * hksz);
      else
        c[k] = ld_stream(pj + (size_t)jj * hksz) - 1;
      v[k] = ld_stream(pv + (size_t)jj * hksz);
    }
  };
  T acc[NC];
#pragma unroll
  for (int q = 0; q < NC; q++) acc[q] = T(0);
  int c[S];
  T v[S];
  if (w > 0) load_group(0, c, v);
  for (int j = 0; j < w; j += S) {
    int cn[S];
    T vn[S];
#pragma unroll
    for (int k = 0; k < S; k++) {
      cn[k] = c[k];
      vn[k] = v[k];
    }
    if (j + S < w) load_group(j + S, cn, vn);
    T xv[NC][S];
#pragma unroll
    for (int q = 0; q < NC; q++) {
      const T* xq = x + (size_t)q * ldx;
#pragma unroll
      for (int k = 0; k < S; k++) xv[q][k] = q < ncols ? xq[c[k]] : T(0);
    }
#pragma unroll
    for (int q = 0; q < NC; q++) {
#pragma unroll
      for (int k = 0; k < S; k++) {
        const T t = fma(v[k], xv[q][k], acc[q]);
        acc[q] = j + k < w ? t : acc[q];
      }
    }
#pragma unroll
    for (int k = 0; k < S; k++) {
      c[k] = cn[k];
      v[k] = vn[k];
    }
  }
  if (row < nr) {
#pragma unroll
    for (int q = 0; q < NC; q++)
      if (q < ncols) store_y(y + (size_t)q * ldy, row, alpha, acc[q], beta);
  }
}

// The multi-RHS product over two adjacent rows per lane (rows 2t and 2t+1
// of a 2*kBlock-row tile, same hack since hksz is even), the columns in
// groups of NC <= 8: one offset-pair and one value-pair load per slot feed
// both rows, and over 16-bit offsets (banded matrices) each column's x
// entries of the two rows come in one 16-byte load where their columns are
// adjacent and even (gather_pair).  Per row and
// column the sum runs in slot order as in k_hll_spmm: the same bits.
template <typename T, int NC, bool IDX16>
__global__ __launch_bounds__(kBlock) void k_hll_spmm_pair(int nr, int hksz, int nlog,
                                                          const int* __restrict__ hkoffs,
                                                          const int* __restrict__ ja,
                                                          const short* __restrict__ res16,
                                                          const int* __restrict__ base16,
                                                          const T* __restrict__ val,
                                                          const T* __restrict__ x, int ldx,
                                                          T alpha, T beta, T* __restrict__ y,
                                                          int ldy, int ncols, BlockOrder ord) {
  static_assert(NC <= 8, "two rows per lane: up to 8 columns per group");
  constexpr int S = NC >= 8 ? 2 : 4;
  const int lb = order_block(ord, nlog);
  if (lb >= nlog) return;
  const int row = lb * (2 * kBlock) + 2 * (int)threadIdx.x;
  const int last = (nr - 1) & ~1;
  const int r = row < last ? row : last;
  const int hk = r / hksz;
  const int lane = r - hk * hksz;
  const int off0 = hkoffs[hk];
  const int w = (hkoffs[hk + 1] - off0) / hksz;
  const int* pj = ja + off0 + lane;
  const short* pe = res16 + off0 + lane;
  const int* pb = base16 + off0 / hksz;
  const T* pv = val + off0 + lane;
  const bool xal = (reinterpret_cast<uintptr_t>(x) & (2 * sizeof(T) - 1)) == 0 && (ldx & 1) == 0;
  auto load_group = [&](int j, int (&c0)[S], int (&c1)[S], vec2<T> (&v)[S]) {
#pragma unroll
    for (int k = 0; k < S; k++) {
      const int jj = j + k < w ? j + k : w - 1;
      if (IDX16) {
        const int b = r + pb[jj];
        const vec2<short> e =
            ld_stream(reinterpret_cast<const vec2<short>*>(pe + (size_t)jj * hksz));
        c0[k] = b + (int)e.x;
        c1[k] = b + 1 + (int)e.y;
      } else {
        const vec2<int> c = ld_stream(reinterpret_cast<const vec2<int>*>(pj + (size_t)jj * hksz));
        c0[k] = c.x - 1;
        c1[k] = c.y - 1;
      }
      v[k] = ld_stream(reinterpret_cast<const vec2<T>*>(pv + (size_t)jj * hksz));
    }
  };
  const bool yal = (reinterpret_cast<uintptr_t>(y) & (2 * sizeof(T) - 1)) == 0 && (ldy & 1) == 0;
  // columns in groups of NC, one after another: the group's slot loads come
  // again from the caches, and the group's pages are the ones in use
  for (int cg = 0; cg < ncols; cg += NC) {
    const int nc = ncols - cg < NC ? ncols - cg : NC;
    const T* xg = x + (size_t)cg * ldx;
    T* yg = y + (size_t)cg * ldy;
    T a0[NC], a1[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) a0[q] = a1[q] = T(0);
    int c0[S], c1[S];
    vec2<T> v[S];
    if (w > 0) load_group(0, c0, c1, v);
    for (int j = 0; j < w; j += S) {
      int n0[S], n1[S];
      vec2<T> nv[S];
#pragma unroll
      for (int k = 0; k < S; k++) {
        n0[k] = c0[k];
        n1[k] = c1[k];
        nv[k] = v[k];
      }
      if (j + S < w) load_group(j + S, n0, n1, nv);
      T x0[NC][S], x1[NC][S];
#pragma unroll
      for (int q = 0; q < NC; q++) {
        const T* xq = xg + (size_t)q * ldx;
#pragma unroll
        for (int k = 0; k < S; k++) {
          if (q >= nc) {
            x0[q][k] = x1[q][k] = T(0);
          } else if (IDX16) {
            gather_pair(xq, xal, c0[k], c1[k], x0[q][k], x1[q][k]);
          } else {
            x0[q][k] = xq[c0[k]];
            x1[q][k] = xq[c1[k]];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < NC; q++) {
#pragma unroll
        for (int k = 0; k < S; k++) {
          const T t0 = fma(v[k].x, x0[q][k], a0[q]);
          const T t1 = fma(v[k].y, x1[q][k], a1[q]);
          a0[q] = j + k < w ? t0 : a0[q];
          a1[q] = j + k < w ? t1 : a1[q];
        }
      }
#pragma unroll
      for (int k = 0; k < S; k++) {
        c0[k] = n0[k];
        c1[k] = n1[k];
        v[k] = nv[k];
      }
    }
#pragma unroll
    for (int q = 0; q < NC; q++) {
      if (q >= nc) continue;
      T* yq = yg + (size_t)q * ldy;
      if (row + 1 < nr && yal) {
        T r0 = alpha * a0[q], r1 = alpha * a1[q];
        if (beta != T(0)) {
          const vec2<T> yo = *reinterpret_cast<const vec2<T>*>(yq + row);
          r0 = fma(beta, yo.x, r0);
          r1 = fma(beta, yo.y, r1);
        }
        vec2<T> yn;
        yn.x = r0;
        yn.y = r1;
        __builtin_nontemporal_store(yn, reinterpret_cast<vec2<T>*>(yq + row));
      } else {
        if (row < nr) store_y(yq, row, alpha, a0[q], beta);
        if (row + 1 < nr) store_y(yq, row + 1, alpha, a1[q], beta);
      }
    }
  }
}

// ------------------------------------------------------------------ ELL SpMV
template <typename T, typename TV, bool IDX16>
__global__ __launch_bounds__(kBlock) void k_ell_spmv(int nr, int pitch, int nzm, int nlog,
                                                     int xchunk, const int* __restrict__ ja,
                                                     const short* __restrict__ res16,
                                                     const int* __restrict__ base16,
                                                     const TV* __restrict__ val,
                                                     const T* __restrict__ x, T alpha, T beta,
                                                     T* __restrict__ y) {
  const int lb = xcd_block_chunk(nlog, xchunk);
  if (lb >= nlog) return;
  const int row = lb * kBlock + threadIdx.x;
  const int r = row < nr ? row : nr - 1;
  T acc;
  if (IDX16)
    acc = slot_dot16<T, TV>(res16 + r, base16 + (size_t)(r >> 6) * nzm, val + r, pitch, nzm, r, x);
  else
    acc = slot_dot<T, TV>(ja + r, val + r, pitch, nzm, x);
  if (row < nr) store_y(y, row, alpha, acc, beta);
}

// Two adjacent rows per lane of an ELL matrix with an even pitch (rows 2t
// and 2t+1 of a 2*kBlock-row tile, one 64-row offset group): one 4-byte
// offset-pair (or 8-byte index-pair) and one 2-element value load per slot
// feed both rows, as in k_hll_spmv_pair; the two results leave in one 16-byte
// non-temporal store.  A row past nr reads the padding row the even pitch
// keeps after an odd nr (filled and encoded like the others).  Per row the
// sum runs in slot order from zero: the bits of k_ell_spmv.
template <typename T, typename TV, bool IDX16>
__global__ __launch_bounds__(kBlock) void k_ell_spmv_pair(int nr, int pitch, int nzm, int nlog,
                                                          int xchunk, const int* __restrict__ ja,
                                                          const short* __restrict__ res16,
                                                          const int* __restrict__ base16,
                                                          const TV* __restrict__ val,
                                                          const T* __restrict__ x, T alpha,
                                                          T beta, T* __restrict__ y) {
  const int lb = xcd_block_chunk(nlog, xchunk);
  if (lb >= nlog) return;
  const int row = lb * (2 * kBlock) + 2 * (int)threadIdx.x;
  const int last = (nr - 1) & ~1;
  const int r = row < last ? row : last;
  T a0, a1;
  if (IDX16)
    slot_dot16_pair<T, TV>(res16 + r, base16 + (size_t)(r >> 6) * nzm, val + r, pitch, nzm, r,
                           x, a0, a1);
  else
    slot_dot_pair<T, TV>(ja + r, val + r, pitch, nzm, x, a0, a1);
  if (row + 1 < nr && (reinterpret_cast<uintptr_t>(y) & (2 * sizeof(T) - 1)) == 0) {
    T r0 = alpha * a0, r1 = alpha * a1;
    if (beta != T(0)) {
      const vec2<T> yo = *reinterpret_cast<const vec2<T>*>(y + row);
      r0 = fma(beta, yo.x, r0);
      r1 = fma(beta, yo.y, r1);
    }
    vec2<T> yn;
    yn.x = r0;
    yn.y = r1;
    __builtin_nontemporal_store(yn, reinterpret_cast<vec2<T>*>(y + row));
  } else {
    if (row < nr) store_y(y, row, alpha, a0, beta);
    if (row + 1 < nr) store_y(y, row + 1, alpha, a1, beta);
  }
}

// 16-bit column offsets of an ELL matrix (EllDev::res16 / base16), one thread
// per group of 64 rows: each slot column's base is the ceiling midpoint of
// the group's offsets (column - row) over rows < nr; *bad is set when some
// residual does not fit.
__global__ __launch_bounds__(kBlock) void k_ell_idx16(int nr, int pitch, int nzm,
                                                      const int* __restrict__ ja,
                                                      short* __restrict__ res,
                                                      int* __restrict__ base, int* bad) {
  // bad[0]: some residual does not fit; bad[1]: max |column - row| (band)
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g * 64 >= nr) return;
  const int r0 = g * 64, r1 = r0 + 64 < nr ? r0 + 64 : nr;
  const int p1 = r0 + 64 < pitch ? r0 + 64 : pitch;
  for (int j = 0; j < nzm; j++) {
    const size_t q = (size_t)j * pitch;
    int lo = 0x7fffffff, hi = -0x7fffffff - 1;
    for (int r = r0; r < r1; r++) {
      const int d = ja[q + r] - 1 - r;
      lo = d < lo ? d : lo;
      hi = d > hi ? d : hi;
    }
    {
      const int64_t bw = -(int64_t)lo > (int64_t)hi ? -(int64_t)lo : (int64_t)hi;
      atomicMax(bad + 1, bw > 0x7fffffff ? 0x7fffffff : (int)bw);
    }
    const int b = (int)(((int64_t)lo + (int64_t)hi + 1) >> 1);
    bool fits = true;
    for (int r = r0; r < p1; r++) {
      const int64_t d = r < r1 ? (int64_t)ja[q + r] - 1 - r - b : 0;
      const bool ok = d >= -32768 && d <= 32767;
      fits = fits && ok;
      res[q + r] = (short)(ok ? d : 0);
    }
    if (!fits) atomicOr(bad, 1);
    base[(size_t)g * nzm + j] = b;
  }
}

// ------------------------------------------------------------------ CSR SpMV
// Vector CSR: G lanes per row (G = 1..64), the lanes stride the row and
// combine through a lane butterfly.  Rows longer than long_thr are skipped
// here and computed by k_csr_long (one wavefront per row).
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_csr_vector(int nr, int nlog,
                                                       const int* __restrict__ irp,
                                                       const int* __restrict__ ja,
                                                       const T* __restrict__ val,
                                                       const T* __restrict__ x, T alpha, T beta,
                                                       T* __restrict__ y, int long_thr) {
  const int lb = xcd_block(nlog);
  if (lb >= nlog) return;
  const int gid = lb * (kBlock / G) + (int)threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const int r = gid < nr ? gid : nr - 1;
  int b = irp[r] - 1, e = irp[r + 1] - 1;
  const bool skip = (e - b) > long_thr;
  if (skip) e = b;
  T acc = T(0);
  int p = b + lane;
  for (; p + 3 * G < e; p += 4 * G) {
    int c0 = ld_stream(ja + p), c1 = ld_stream(ja + p + G);
    int c2 = ld_stream(ja + p + 2 * G), c3 = ld_stream(ja + p + 3 * G);
    T v0 = ld_stream(val + p), v1 = ld_stream(val + p + G);
    T v2 = ld_stream(val + p + 2 * G), v3 = ld_stream(val + p + 3 * G);
    T x0 = x[c0 - 1], x1 = x[c1 - 1], x2 = x[c2 - 1], x3 = x[c3 - 1];
    acc = fma(v0, x0, acc);
    acc = fma(v1, x1, acc);
    acc = fma(v2, x2, acc);
    acc = fma(v3, x3, acc);
  }
  for (; p < e; p += G) acc = fma(ld_stream(val + p), x[ld_stream(ja + p) - 1], acc);
  if (G > 1) acc = group_sum<G>(acc);
  if (gid < nr && lane == 0 && !skip) store_y(y, gid, alpha, acc, beta);
}

// Row-block CSR (the CSR-stream scheme of Greathouse and Daga, SC14): a block
// takes a run of consecutive rows whose entries fit in LDS.  All threads
// stream the run's entries (coalesced index and value loads, eight per thread
// in flight, the x gathers after them) and store the products val * x in LDS;
// then one thread per row adds its row's products in CSR order from 0.  The
// products are separate multiplies and the sum separate adds, the order of
// the reference's CPU loop (psb_d_csr_csmv, base/serial/impl/
// psb_d_csr_impl.F90:46-417), so the row sums equal the CPU's bit for bit.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_csr_rowblock(int nblk, int xchunk,
                                                         const int2* __restrict__ blk,
                                                         const int* __restrict__ irp,
                                                         const int* __restrict__ ja,
                                                         const T* __restrict__ val,
                                                         const T* __restrict__ x, T alpha, T beta,
                                                         T* __restrict__ y) {
  constexpr int U = kCsrBlockNnz / kBlock;
  __shared__ T prod[kCsrBlockNnz];
  const int b = xchunk > 0 ? xcd_block_chunk(nblk, xchunk) : (int)blockIdx.x;
  if (b >= nblk) return;
  const int2 rr = blk[b];
  const int e0 = irp[rr.x] - 1, e1 = irp[rr.y] - 1;
  if (e1 > e0) {
    int c[U];
    T v[U], xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = min(e0 + (int)threadIdx.x + u * kBlock, e1 - 1);
      c[u] = ld_stream(ja + k);
      v[u] = ld_stream(val + k);
    }
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = x[c[u] - 1];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = (int)threadIdx.x + u * kBlock;
      if (e0 + k < e1) prod[k] = v[u] * xv[u];   // stored, so never fused into the sum
    }
  }
  __syncthreads();
  const int r = rr.x + (int)threadIdx.x;
  if (r < rr.y) {
    const int b = irp[r] - 1 - e0, e = irp[r + 1] - 1 - e0;
    T acc = T(0);
    for (int k = b; k < e; k++) acc = acc + prod[k];
    store_y(y, r, alpha, acc, beta);
  }
}

// Length-binned CSR: the rows of one class (G lanes per row, about four
// entries per lane) come from a list; same inner loop as k_csr_vector.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_csr_list(int count, const int* __restrict__ rows,
                                                     const int* __restrict__ irp,
                                                     const int* __restrict__ ja,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ x, T alpha, T beta,
                                                     T* __restrict__ y) {
  const int gid = (int)blockIdx.x * (kBlock / G) + (int)threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const int row = rows[gid < count ? gid : count - 1];
  const int b = irp[row] - 1, e = irp[row + 1] - 1;
  T acc = T(0);
  int p = b + lane;
  for (; p + 3 * G < e; p += 4 * G) {
    int c0 = ld_stream(ja + p), c1 = ld_stream(ja + p + G);
    int c2 = ld_stream(ja + p + 2 * G), c3 = ld_stream(ja + p + 3 * G);
    T v0 = ld_stream(val + p), v1 = ld_stream(val + p + G);
    T v2 = ld_stream(val + p + 2 * G), v3 = ld_stream(val + p + 3 * G);
    T x0 = x[c0 - 1], x1 = x[c1 - 1], x2 = x[c2 - 1], x3 = x[c3 - 1];
    acc = fma(v0, x0, acc);
    acc = fma(v1, x1, acc);
    acc = fma(v2, x2, acc);
    acc = fma(v3, x3, acc);
  }
  for (; p < e; p += G) acc = fma(ld_stream(val + p), x[ld_stream(ja + p) - 1], acc);
  if (G > 1) acc = group_sum<G>(acc);
  if (gid < count && lane == 0) store_y(y, row, alpha, acc, beta);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_csr_long(int nlong, const int* __restrict__ rows,
                                                     const int* __restrict__ irp,
                                                     const int* __restrict__ ja,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ x, T alpha, T beta,
                                                     T* __restrict__ y) {
  const int wid = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int wr = wid < nlong ? wid : nlong - 1;
  const int row = rows[wr];
  const int b = irp[row] - 1, e = irp[row + 1] - 1;
  T acc = T(0);
  int p = b + lane;
  for (; p + 3 * kWave < e; p += 4 * kWave) {
    int c0 = ld_stream(ja + p), c1 = ld_stream(ja + p + kWave);
    int c2 = ld_stream(ja + p + 2 * kWave), c3 = ld_stream(ja + p + 3 * kWave);
    T v0 = ld_stream(val + p), v1 = ld_stream(val + p + kWave);
    T v2 = ld_stream(val + p + 2 * kWave), v3 = ld_stream(val + p + 3 * kWave);
    T x0 = x[c0 - 1], x1 = x[c1 - 1], x2 = x[c2 - 1], x3 = x[c3 - 1];
    acc = fma(v0, x0, acc);
    acc = fma(v1, x1, acc);
    acc = fma(v2, x2, acc);
    acc = fma(v3, x3, acc);
  }
  for (; p < e; p += kWave) acc = fma(ld_stream(val + p), x[ld_stream(ja + p) - 1], acc);
  acc = wave_sum(acc);
  if (wid < nlong && lane == 0) store_y(y, row, alpha, acc, beta);
}

// ------------------------------------------------------------------ scaling
// scals: every stored value times s (psb_d_hll_scals and psb_d_ell_scals
// scale the whole value array, padding included; psb_d_csr_scals its nnz
// values).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale_vals(int64_t n, T s, T* __restrict__ val) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    val[i] = val[i] * s;
}

// scal with a vector d, one thread per row: left multiplies by d(row), right
// by d(column).  CSR (psb_d_csr_scal, base/serial/impl/psb_d_csr_impl.F90:
// 1607-1679): the row's entries.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_csr_scal(int rows, const int* __restrict__ irp,
                                                     const int* __restrict__ ja,
                                                     const T* __restrict__ d, int left,
                                                     T* __restrict__ val) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= rows) return;
  for (int k = irp[r] - 1; k < irp[r + 1] - 1; k++)
    val[k] = val[k] * (left ? d[r] : d[ja[k] - 1]);
}

// HLL (psb_d_hll_scal, ext/impl/psb_d_hll_scal.f90:83-130): the row's irn
// entries, its padding untouched.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_hll_scal(int rows, int hksz,
                                                     const int* __restrict__ hkoffs,
                                                     const int* __restrict__ irn,
                                                     const int* __restrict__ ja,
                                                     const int* __restrict__ perm,
                                                     const T* __restrict__ d, int left,
                                                     T* __restrict__ val) {
  const int r = blockIdx.x * kBlock + threadIdx.x;  // position (row when perm is null)
  if (r >= rows) return;
  const int hk = r / hksz;
  const size_t off = (size_t)hkoffs[hk] + (r - hk * hksz);
  const T dl = left ? d[perm ? perm[r] - 1 : r] : T(0);
  for (int j = 0; j < irn[r]; j++) {
    const size_t p = off + (size_t)j * hksz;
    val[p] = val[p] * (left ? dl : d[ja[p] - 1]);
  }
}

// ELL (psb_d_ell_scal, ext/impl/psb_d_ell_scal.f90:64-90): left scales the
// whole row, padding included; right the row's irn entries.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ell_scal(int rows, int pitch, int nzm,
                                                     const int* __restrict__ irn,
                                                     const int* __restrict__ ja,
                                                     const T* __restrict__ d, int left,
                                                     T* __restrict__ val) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= rows) return;
  const int w = left ? nzm : irn[r];
  for (int j = 0; j < w; j++) {
    const size_t p = (size_t)j * pitch + r;
    val[p] = val[p] * (left ? d[r] : d[ja[p] - 1]);
  }
}

// ------------------------------------------------------------------ transposed product
// y = alpha * A^T x + beta * y over the transpose of A stored like HLL: hacks
// of kTHack rows of A^T (columns of A), slot (c, j) at hkoffs[h] + j*kTHack +
// lane, one row per lane, so a wavefront's slot loads are coalesced.  Row c
// holds the rows i of A with an entry in column c, ascending, and is computed
// as the CPU's column-oriented loop updates y(c) (psb_d_csr_csmv trans 'T',
// base/serial/impl/psb_d_csr_impl.F90:308-370): y(c) scaled by beta first,
// then each entry added in that order, as y + v*x(i), y - v*x(i) (alpha = -1)
// or y + (alpha*v)*x(i), separately rounded.  Slots past the row's count
// (padding: source 0, value 0) are loaded and their update dropped.
constexpr int kTHack = 64;

template <typename T, bool IDX16>
__global__ __launch_bounds__(kBlock) void k_tmv_run(int rows, int cols,
                                                    const int* __restrict__ hkoffs,
                                                    const int* __restrict__ cnt,
                                                    const int* __restrict__ src,
                                                    const short* __restrict__ res16,
                                                    const int* __restrict__ base16,
                                                    const T* __restrict__ val,
                                                    const T* __restrict__ x, T alpha, T beta,
                                                    T* __restrict__ y) {
#pragma clang fp contract(off)   // separate multiply and add, as the CPU loop
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= rows) return;
  if (alpha == T(0)) {
    y[c] = beta == T(0) ? T(0) : beta * y[c];
    return;
  }
  T r;
  if (beta == T(0)) r = T(0);
  else if (beta == T(1)) r = y[c];
  else if (beta == T(-1)) r = -y[c];
  else r = beta * y[c];
  // a wavefront is one hack (64 columns of A): its offsets through scalar loads
  const int h = __builtin_amdgcn_readfirstlane(c / kTHack), lane = c - h * kTHack;
  const int h0 = hkoffs[h];
  const int off = h0 + lane;
  const int w = (hkoffs[h + 1] - h0) / kTHack;
  const int n = cnt[c];
  for (int j0 = 0; j0 < w; j0 += 4) {
    int s[4];
    T v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int j = j0 + k < w ? j0 + k : w - 1;
      if (IDX16) {
        const int sr = c + base16[h0 / kTHack + j] + (int)ld_stream(res16 + off + (size_t)j * kTHack);
        s[k] = sr < 0 ? 0 : (sr >= cols ? cols - 1 : sr);
      } else {
        s[k] = ld_stream(src + off + (size_t)j * kTHack);
      }
      v[k] = ld_stream(val + off + (size_t)j * kTHack);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const T xi = x[s[k]];
      T t;
      if (alpha == T(1)) t = r + v[k] * xi;
      else if (alpha == T(-1)) t = r - v[k] * xi;
      else t = r + alpha * v[k] * xi;
      r = j0 + k < n ? t : r;
    }
  }
  __builtin_nontemporal_store(r, y + c);
}

// The transposed product over 16-bit source offsets with two adjacent rows
// of A^T per lane (rows c and c+1, c even, one hack; a wavefront covers two
// hacks, their offsets through scalar loads): one offset-pair and one
// value-pair load per slot, the two rows' x entries in one 16-byte load
// where their source rows are adjacent and even, one 16-byte store.  Each
// row runs k_tmv_run's update sequence over its own count (a row past the
// last has count 0): the same bits.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tmv_pair(int rows, int cols,
                                                     const int* __restrict__ hkoffs,
                                                     const int* __restrict__ cnt,
                                                     const short* __restrict__ res16,
                                                     const int* __restrict__ base16,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ x, T alpha, T beta,
                                                     T* __restrict__ y) {
#pragma clang fp contract(off)   // separate multiply and add, as the CPU loop
  const int c = 2 * (int)(blockIdx.x * kBlock + threadIdx.x);
  if (c >= rows) return;
  const bool two = c + 1 < rows;
  if (alpha == T(0)) {
    y[c] = beta == T(0) ? T(0) : beta * y[c];
    if (two) y[c + 1] = beta == T(0) ? T(0) : beta * y[c + 1];
    return;
  }
  auto scaled = [&](int i) -> T {
    if (beta == T(0)) return T(0);
    if (beta == T(1)) return y[i];
    if (beta == T(-1)) return -y[i];
    return beta * y[i];
  };
  T r0 = scaled(c), r1 = two ? scaled(c + 1) : T(0);
  int off0, w;
  hll_pair_hack<kTHack>(hkoffs, rows, c / kTHack, kTHack, off0, w);
  const int lane = c % kTHack;
  const int n0 = cnt[c], n1 = two ? cnt[c + 1] : 0;
  const short* pe = res16 + off0 + lane;
  const int* pb = base16 + off0 / kTHack;
  const T* pv = val + off0 + lane;
  const bool xal = (reinterpret_cast<uintptr_t>(x) & (2 * sizeof(T) - 1)) == 0;
  for (int j0 = 0; j0 < w; j0 += 4) {
    int s0[4], s1[4];
    vec2<T> v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int j = j0 + k < w ? j0 + k : w - 1;
      const int b = c + pb[j];
      const vec2<short> e = ld_stream(reinterpret_cast<const vec2<short>*>(pe + (size_t)j * kTHack));
      const int a0 = b + (int)e.x, a1 = b + 1 + (int)e.y;
      s0[k] = a0 < 0 ? 0 : (a0 >= cols ? cols - 1 : a0);
      s1[k] = a1 < 0 ? 0 : (a1 >= cols ? cols - 1 : a1);
      v[k] = ld_stream(reinterpret_cast<const vec2<T>*>(pv + (size_t)j * kTHack));
    }
    T x0[4], x1[4];
#pragma unroll
    for (int k = 0; k < 4; k++) gather_pair(x, xal, s0[k], s1[k], x0[k], x1[k]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      T t0, t1;
      if (alpha == T(1)) {
        t0 = r0 + v[k].x * x0[k];
        t1 = r1 + v[k].y * x1[k];
      } else if (alpha == T(-1)) {
        t0 = r0 - v[k].x * x0[k];
        t1 = r1 - v[k].y * x1[k];
      } else {
        t0 = r0 + alpha * v[k].x * x0[k];
        t1 = r1 + alpha * v[k].y * x1[k];
      }
      r0 = j0 + k < n0 ? t0 : r0;
      r1 = j0 + k < n1 ? t1 : r1;
    }
  }
  if (two && (reinterpret_cast<uintptr_t>(y) & (2 * sizeof(T) - 1)) == 0) {
    vec2<T> yn;
    yn.x = r0;
    yn.y = r1;
    __builtin_nontemporal_store(yn, reinterpret_cast<vec2<T>*>(y + c));
  } else {
    __builtin_nontemporal_store(r0, y + c);
    if (two) __builtin_nontemporal_store(r1, y + c + 1);
  }
}

// 16-bit column offsets of a rows-in-order HLL matrix (HllDev::res16 /
// base16), one thread per hack: each slot column's base is the midpoint of
// its lanes' offsets (column - row), so the residuals fit in 16 bits when the
// offsets of one slot column span at most 65535; *bad is set otherwise (the
// matrix then keeps using ja).
__global__ __launch_bounds__(kBlock) void k_hll_idx16(int hksz, int noffs,
                                                      const int* __restrict__ hkoffs,
                                                      const int* __restrict__ ja,
                                                      short* __restrict__ res,
                                                      int* __restrict__ base, int* bad) {
  // bad[0]: some residual does not fit; bad[1]: max |column - row| (band)
  const int h = blockIdx.x * kBlock + threadIdx.x;
  if (h >= noffs) return;
  const int off0 = hkoffs[h];
  const int w = (hkoffs[h + 1] - off0) / hksz;
  const int r0 = h * hksz;
  for (int j = 0; j < w; j++) {
    const size_t q = (size_t)off0 + (size_t)j * hksz;
    int lo = 0x7fffffff, hi = -0x7fffffff - 1;
    for (int l = 0; l < hksz; l++) {
      const int d = ja[q + l] - 1 - (r0 + l);
      lo = d < lo ? d : lo;
      hi = d > hi ? d : hi;
    }
    {
      const int64_t bw = -(int64_t)lo > (int64_t)hi ? -(int64_t)lo : (int64_t)hi;
      atomicMax(bad + 1, bw > 0x7fffffff ? 0x7fffffff : (int)bw);
    }
    // ceiling midpoint: residuals in [-32768, 32767] for a span up to 65535
    const int b = (int)(((int64_t)lo + (int64_t)hi + 1) >> 1);
    bool fits = true;
    for (int l = 0; l < hksz; l++) {
      const int64_t d = (int64_t)ja[q + l] - 1 - (r0 + l) - b;
      const bool ok = d >= -32768 && d <= 32767;
      fits = fits && ok;
      res[q + l] = (short)(ok ? d : 0);
    }
    if (!fits) atomicOr(bad, 1);
    base[off0 / hksz + j] = b;
  }
}

// ------------------------------------------------------------------ fills
// psi_d_convert_hll_from_coo + psi_d_xtr_ell_from_coo, CPU padding rule:
// one thread per slot row; rows past nr in the last hack get ja = nr, val 0;
// an empty row is padded with the first column of its hack's entries.
// perm (row-sorted storage, null for rows in order): position i holds row
// perm[i] (1-based), irn is per position, idisp per row.
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_hll_fill(int nr, int hksz, int noffs,
                                                     const int* __restrict__ irn,
                                                     const int* __restrict__ hoffs,
                                                     const int* __restrict__ idisp,
                                                     const int* __restrict__ cja,
                                                     const TV* __restrict__ cval,
                                                     const int* __restrict__ perm,
                                                     int* __restrict__ ja, TV* __restrict__ val,
                                                     int* __restrict__ idiag) {
  const int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (t >= (int64_t)noffs * hksz) return;
  const int hk = (int)(t / hksz);
  const int lane = (int)(t - (int64_t)hk * hksz);
  const int i = hk * hksz + lane;  // 0-based position (the row when perm is null)
  const int off = hoffs[hk];
  const int w = (hoffs[hk + 1] - off) / hksz;
  if (i >= nr) {
    for (int j = 0; j < w; j++) {
      ja[off + (size_t)j * hksz + lane] = nr;
      val[off + (size_t)j * hksz + lane] = TV(0);
    }
    return;
  }
  const int row = perm ? perm[i] - 1 : i;
  const int rs = irn[i];
  const int p0 = idisp[row];
  int dg = 0;
  int ic = 0;
  for (int j = 0; j < rs; j++) {
    ic = cja[p0 + j];
    if (ic == row + 1) dg = j + 1;
    ja[off + (size_t)j * hksz + lane] = ic;
    val[off + (size_t)j * hksz + lane] = cval[p0 + j];
  }
  if (rs == 0 && w > 0) {
    // the first entry of the hack's first row (in bounds: some row of the
    // hack has entries); a sorted layout has its longest row first
    const int q = idisp[perm ? perm[hk * hksz] - 1 : hk * hksz];
    ic = q < idisp[nr] ? cja[q] : nr;
  }
  for (int j = rs; j < w; j++) {
    ja[off + (size_t)j * hksz + lane] = ic;
    val[off + (size_t)j * hksz + lane] = TV(0);
  }
  idiag[i] = dg;
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void k_ell_fill(int nr, int pitch, int nzm,
                                                     const int* __restrict__ irn,
                                                     const int* __restrict__ idisp,
                                                     const int* __restrict__ cja,
                                                     const TV* __restrict__ cval,
                                                     int* __restrict__ ja, TV* __restrict__ val,
                                                     int* __restrict__ idiag) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= pitch) return;
  if (i >= nr) {
    for (int j = 0; j < nzm; j++) {
      ja[(size_t)j * pitch + i] = 1;
      val[(size_t)j * pitch + i] = TV(0);
    }
    return;
  }
  const int rs = irn[i];
  const int p0 = idisp[i];
  int dg = 0, ic = 0;
  for (int j = 0; j < rs; j++) {
    ic = cja[p0 + j];
    if (ic == i + 1) dg = j + 1;
    ja[(size_t)j * pitch + i] = ic;
    val[(size_t)j * pitch + i] = cval[p0 + j];
  }
  if (rs == 0) ic = cja[0];
  for (int j = rs; j < nzm; j++) {
    ja[(size_t)j * pitch + i] = ic;
    val[(size_t)j * pitch + i] = TV(0);
  }
  idiag[i] = dg;
}

// diagonal extraction: d[i] = val(slot idiag[i]) or 0 when the row has none
// (row-sorted storage: position i holds row perm[i])
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_hll_diag(int nr, int hksz,
                                                     const int* __restrict__ hkoffs,
                                                     const int* __restrict__ idiag,
                                                     const int* __restrict__ perm,
                                                     const TV* __restrict__ val, TV* __restrict__ d) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nr) return;
  const int hk = i / hksz, lane = i - hk * hksz;
  const int s = idiag[i];
  d[perm ? perm[i] - 1 : i] = s > 0 ? val[hkoffs[hk] + (size_t)(s - 1) * hksz + lane] : TV(0);
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void k_ell_diag(int nr, int pitch, const int* __restrict__ idiag,
                                                     const TV* __restrict__ val, TV* __restrict__ d) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nr) return;
  const int s = idiag[i];
  d[i] = s > 0 ? val[(size_t)(s - 1) * pitch + i] : TV(0);
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void k_csr_diag(int nr, const int* __restrict__ irp,
                                                     const int* __restrict__ ja,
                                                     const TV* __restrict__ val, TV* __restrict__ d) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nr) return;
  TV a = TV(0);
  for (int p = irp[i] - 1; p < irp[i + 1] - 1; p++)
    if (ja[p] == i + 1) a = val[p];
  d[i] = a;
}

__global__ __launch_bounds__(kBlock) void k_d2s(int64_t n, const double* __restrict__ a,
                                                float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    b[i] = (float)a[i];
}

// ------------------------------------------------------------------ helpers
static bool scal_vec_ok(const MultiVec* d, int elem_type, int need) {
  return d && d->elem_type == elem_type && d->size >= need && d->count >= 1;
}

static void refresh_single(HllDev* m) {   // the FP32 copy follows the scaled values
  if (m->val32 && m->allocsize > 0)
    hipLaunchKernelGGL(k_d2s, dim3(stream_grid(m->allocsize)), dim3(kBlock), 0, ctx().stream,
                       (int64_t)m->allocsize, (const double*)m->val, m->val32);
}

template <typename T>
static void launch_scale(int64_t n, double s, void* val) {
  if (n > 0)
    hipLaunchKernelGGL((k_scale_vals<T>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream, n,
                       (T)s, (T*)val);
}

static int max_index(const int* ja, int64_t n) {
  int m = 0;
  for (int64_t k = 0; k < n; k++) m = ja[k] > m ? ja[k] : m;
  return m;
}

static int min_index(const int* ja, int64_t n) {
  int m = 0x7fffffff;
  for (int64_t k = 0; k < n; k++) m = ja[k] < m ? ja[k] : m;
  return n ? m : 1;
}

static int check_vectors(const MultiVec* x, const MultiVec* y, int elem_type, int need_x,
                         int need_y) {
  if (!x || !y) return PSB_HIP_INVALID;
  if (x->elem_type != elem_type || y->elem_type != elem_type) return PSB_HIP_INVALID;
  if (x->size < need_x || y->size < need_y) return PSB_HIP_UNSUPPORTED;
  if (x->count != y->count) return PSB_HIP_UNSUPPORTED;
  if (x->v == y->v) return PSB_HIP_UNSUPPORTED;  // the product is not in place
  return PSB_HIP_SUCCESS;
}

// 16-bit column offsets in the HLL products when the matrix has them
// (psb_hip_hll_set_index16 switches them off for comparisons)
static int g_hll_idx16 = 1;
bool hll_use_idx16(const HllDev* m) { return g_hll_idx16 && m->res16 && !m->perm; }

// Chunked XCD order whose chunk spans about an eighth of the band: the chunks
// an XCD takes in turn (c, c + 8, ...) are then about one band apart, so for a
// 3-D stencil (band L^2) the chunk before holds the z-1 neighbours of this
// one's rows in the same L2.  L=200 stencil: FP64 pairs (10 blocks) 116.5 us,
// FP32 and SpMM k=8 (20 blocks) 63.9 and 479.8 us, against 116.5-119.5, 64.6
// and 490.9-514.6 us with 16-block chunks.  Bands under 32 blocks (chunks
// below 4) and non-banded matrices keep kXchunkDefault; an explicit
// psb_hip_hll_set_xcd_chunk overrides.
int xcd_chunk_for(int band, int rows_per_block) {
  if (g_hll_xchunk >= 0) return g_hll_xchunk;
  const int64_t per = (int64_t)rows_per_block * kXcds;
  const int64_t c = ((int64_t)band + per / 2) / per;
  if (band <= 0 || c < 4 || c > 64) return kXchunkDefault;
  return (int)c;
}
bool ell_use_idx16(const EllDev* m) { return g_hll_idx16 && m->res16; }

static void ell_drop_idx16(EllDev* m) {
  hipFree(m->res16);
  hipFree(m->base16);
  m->res16 = nullptr;
  m->base16 = nullptr;
  m->band = 0;
}

// Build the 16-bit column offsets of an ELL matrix (kept only when every
// residual fits and memory allows).
static int ell_build_idx16(EllDev* m) {
  ell_drop_idx16(m);
  if (m->rows <= 0 || m->nzm <= 0) return PSB_HIP_SUCCESS;
  hipStream_t s = ctx().stream;
  const int ng = (m->rows + 63) / 64;
  int* bad = nullptr;
  hipError_t e = hipMalloc((void**)&m->res16, sizeof(short) * (size_t)m->pitch * m->nzm);
  if (e == hipSuccess) e = hipMalloc((void**)&m->base16, sizeof(int) * (size_t)ng * m->nzm);
  if (e == hipSuccess) e = hipMalloc((void**)&bad, 2 * sizeof(int));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    hipFree(bad);
    ell_drop_idx16(m);
    return PSB_HIP_SUCCESS;
  }
  int hbad[2] = {0, 0};
  PSB_HIP_CHECK(hipMemsetAsync(bad, 0, 2 * sizeof(int), s));
  hipLaunchKernelGGL(k_ell_idx16, dim3((unsigned)grid_for(ng, kBlock)), dim3(kBlock), 0, s, m->rows,
                     m->pitch, m->nzm, m->ja, m->res16, m->base16, bad);
  PSB_HIP_LAUNCH_CHECK();
  PSB_HIP_CHECK(hipMemcpyAsync(hbad, bad, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  hipFree(bad);
  if (hbad[0]) ell_drop_idx16(m);
  m->band = hbad[1];
  return PSB_HIP_SUCCESS;
}

// The block order of an HLL product over blocks of rows_per_block rows: the
// matrix's plane-tile order when it has one, else chunked XCD order (xcd) or
// launch order.
static BlockOrder hll_order(const HllDev* m, bool xcd, int rows_per_block, int nlog) {
  if (m->plane_rows > 0 && m->plane_tb > 0)
    return plane_order(nlog, rows_per_block, m->plane_rows, m->plane_tb);
  BlockOrder o;
  if (xcd) {
    o.mode = 1;
    o.xchunk = xcd_chunk_for(m->band, rows_per_block);
  }
  return o;
}

template <typename T, typename TV, int HK>
static void hll_launch_hk(const HllDev* m, bool pair, bool xcd, const TV* val, const T* x,
                          T alpha, T beta, T* y) {
  hipStream_t s = ctx().stream;
  const int rpb = pair ? 2 * kBlock : kBlock;
  const int nlog = grid_for(m->rows, rpb);
  const BlockOrder o = hll_order(m, xcd, rpb, nlog);
  dim3 grid(order_grid(o, nlog)), blk(kBlock);
  const bool i16 = hll_use_idx16(m);
  if (pair && i16)
    hipLaunchKernelGGL((k_hll_spmv_pair<T, TV, HK, true>), grid, blk, 0, s, m->rows, m->hksz,
                       nlog, o, m->hkoffs, m->ja, m->res16, m->base16, val, x, alpha, beta, y);
  else if (pair)
    hipLaunchKernelGGL((k_hll_spmv_pair<T, TV, HK, false>), grid, blk, 0, s, m->rows, m->hksz,
                       nlog, o, m->hkoffs, m->ja, m->res16, m->base16, val, x, alpha, beta, y);
  else if (i16)
    hipLaunchKernelGGL((k_hll_spmv<T, TV, HK, true>), grid, blk, 0, s, m->rows, m->hksz, nlog, o,
                       m->hkoffs, m->ja, m->res16, m->base16, val, x, alpha, beta, y);
  else
    hipLaunchKernelGGL((k_hll_spmv<T, TV, HK, false>), grid, blk, 0, s, m->rows, m->hksz, nlog, o,
                       m->hkoffs, m->ja, m->res16, m->base16, val, x, alpha, beta, y);
}

// Variant: 1 one row per lane, 2 two adjacent rows per lane (even hksz), 0
// chooses (pairs for 8-byte values, single rows for 4-byte values, both in
// XCD order); adding 4 deals the blocks to the XCDs in chunks of
// g_hll_xchunk blocks (xcd_block_chunk) instead of launch order.
// Row-sorted storage: the work units, then (when some hack is cut into
// several units) the partial-sum combination.
template <typename T, typename TV>
static int hll_units_launch(const HllDev* m, const TV* val, const T* x, T alpha, T beta, T* y) {
  hipStream_t s = ctx().stream;
  if (m->nunits > 0) {
    dim3 grid(grid_for((int64_t)m->nunits * m->hksz, kBlock));
    hipLaunchKernelGGL((k_hll_spmv_units<T, TV>), grid, dim3(kBlock), 0, s, m->rows, m->hksz,
                       m->nunits, (const int4*)m->units, (const int*)m->hkoffs,
                       (const int*)m->irn, (const int*)m->perm, (const int*)m->ja, val, x, alpha,
                       beta, y, (T*)m->part);
  }
  if (m->nsplit > 0) {
    dim3 grid(grid_for((int64_t)m->nsplit * m->hksz, kBlock));
    hipLaunchKernelGGL((k_hll_units_combine<T>), grid, dim3(kBlock), 0, s, m->rows, m->hksz,
                       m->nsplit, (const int4*)m->splits, (const int*)m->perm,
                       (const T*)m->part, alpha, beta, y);
  }
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

// Row-sorted storage with column slices: y = alpha A_0 x + beta y, then
// y = alpha A_s x + y slice after slice (the matrix's own values only; the
// FP32 copy of a double matrix runs unsliced).  Each slice is sorted by its
// own row lengths: slices sharing the whole matrix's row order and adding
// their sums by position measured 5-10% slower (their rows' lengths in a
// slice vary inside a hack, so the lanes of a wave diverge).
template <typename T, typename TV>
static int hll_units_sliced(const HllDev* m, const TV* val, const T* x, T alpha, T beta, T* y) {
  if (m->nslices > 0 && (const void*)val == m->val) {
    for (int s = 0; s < m->nslices; s++) {
      const HllDev* c = m->slices[s];
      const int rc =
          hll_units_launch<T, TV>(c, (const TV*)c->val, x, alpha, s == 0 ? beta : T(1), y);
      if (rc) return rc;
    }
    return PSB_HIP_SUCCESS;
  }
  return hll_units_launch<T, TV>(m, val, x, alpha, beta, y);
}

template <typename T, typename TV>
static int hll_launch(const HllDev* m, const TV* val, const T* x, T alpha, T beta, T* y) {
  if (m->perm) return hll_units_sliced<T, TV>(m, val, x, alpha, beta, y);
  const int v = m->variant != 0 ? m->variant : g_hll_variant;
  const int variant = v & 3;
  const bool xcd = (v & 4) != 0 || v == 0;
  bool pair = variant == 2 || (variant == 0 && sizeof(TV) == 8);
  if (m->hksz % 2 != 0) pair = false;
  if (m->hksz == 64) hll_launch_hk<T, TV, 64>(m, pair, xcd, val, x, alpha, beta, y);
  else if (m->hksz == 32) hll_launch_hk<T, TV, 32>(m, pair, xcd, val, x, alpha, beta, y);
  else hll_launch_hk<T, TV, 0>(m, pair, xcd, val, x, alpha, beta, y);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

template <typename T>
static int hll_spmm_launch(const HllDev* m, const T* val, const T* x, int ldx, T alpha, T beta,
                           T* y, int ldy, int ncols) {
  if (m->perm) {
    // row-sorted storage: one product per column (the matrix is read k times)
    for (int c = 0; c < ncols; c++) {
      int rc = hll_units_sliced<T, T>(m, val, x + (size_t)c * ldx, alpha, beta,
                                      y + (size_t)c * ldy);
      if (rc) return rc;
    }
    return PSB_HIP_SUCCESS;
  }
  hipStream_t s = ctx().stream;
  // two rows per lane, columns in groups of up to 8 inside the launch, up to
  // 16 columns (L=200 stencil against one row per lane: k = 4 270 -> 229 us,
  // k = 8 476 -> 433 us, k = 16 862 -> 812 us; k = 32 1605 against 1595 us,
  // so wider blocks stay one row per lane; 16 columns per group cut the
  // occupancy to 3 waves per SIMD: k = 16 1004 us)
  if (g_spmm_pair && m->hksz % 2 == 0 && ncols > 1 && ncols <= 16) {
    const int nlog = grid_for(m->rows, 2 * kBlock);
    const BlockOrder o = hll_order(m, true, 2 * kBlock, nlog);
    const bool i16 = hll_use_idx16(m);
    dim3 grid(order_grid(o, nlog)), blk(kBlock);
    {
      const int nc = ncols;
      const T* xc = x;
      T* yc = y;
#define PSB_SPMMP_CASE(NC)                                                                    \
  do {                                                                                        \
    if (i16)                                                                                  \
      hipLaunchKernelGGL((k_hll_spmm_pair<T, NC, true>), grid, blk, 0, s, m->rows, m->hksz,   \
                         nlog, m->hkoffs, m->ja, m->res16, m->base16, val, xc, ldx, alpha,    \
                         beta, yc, ldy, nc, o);                                               \
    else                                                                                      \
      hipLaunchKernelGGL((k_hll_spmm_pair<T, NC, false>), grid, blk, 0, s, m->rows, m->hksz,  \
                         nlog, m->hkoffs, m->ja, m->res16, m->base16, val, xc, ldx, alpha,    \
                         beta, yc, ldy, nc, o);                                               \
  } while (0)
      if (nc > 4) PSB_SPMMP_CASE(8);
      else if (nc > 2) PSB_SPMMP_CASE(4);
      else PSB_SPMMP_CASE(2);
#undef PSB_SPMMP_CASE
      PSB_HIP_LAUNCH_CHECK();
    }
    return PSB_HIP_SUCCESS;
  }
  // XCD block order, in chunks of g_hll_xchunk blocks: one contiguous range
  // per XCD measured 2-6% faster than launch order for k = 8, 16 and 32
  // columns on the L=200 stencil, chunks of 16 blocks another 2-5% (485.7 vs
  // 496.4-521.8 us for k = 8, 1782.8 vs 1873.4-1892.6 us for k = 32)
  const int nlog = grid_for(m->rows, kBlock);
  const BlockOrder o = hll_order(m, true, kBlock, nlog);
  const bool i16 = hll_use_idx16(m);
  dim3 grid(order_grid(o, nlog)), blk(kBlock);
  for (int c0 = 0; c0 < ncols; c0 += 32) {
    const int nc = ncols - c0 < 32 ? ncols - c0 : 32;
    const T* xc = x + (size_t)c0 * ldx;
    T* yc = y + (size_t)c0 * ldy;
#define PSB_SPMM_CASE(NC)                                                                     \
  do {                                                                                        \
    if (i16)                                                                                  \
      hipLaunchKernelGGL((k_hll_spmm<T, NC, true>), grid, blk, 0, s, m->rows, m->hksz, nlog,  \
                         m->hkoffs, m->ja, m->res16, m->base16, val, xc, ldx, alpha, beta, yc, \
                         ldy, nc, o);                                                         \
    else                                                                                      \
      hipLaunchKernelGGL((k_hll_spmm<T, NC, false>), grid, blk, 0, s, m->rows, m->hksz, nlog, \
                         m->hkoffs, m->ja, m->res16, m->base16, val, xc, ldx, alpha, beta, yc, \
                         ldy, nc, o);                                                         \
  } while (0)
    if (nc > 16) PSB_SPMM_CASE(32);
    else if (nc > 8) PSB_SPMM_CASE(16);
    else if (nc > 4) PSB_SPMM_CASE(8);
    else if (nc > 2) PSB_SPMM_CASE(4);
    else if (nc > 1) PSB_SPMM_CASE(2);
    else PSB_SPMM_CASE(1);
#undef PSB_SPMM_CASE
    PSB_HIP_LAUNCH_CHECK();
  }
  return PSB_HIP_SUCCESS;
}

template <typename T>
static int csr_launch_binned(CsrDev* m, const T* x, T alpha, T beta, T* y) {
  hipStream_t s = ctx().stream;
  const T* val = (const T*)m->val;
  for (int c = 0; c < 7; c++) {
    const int count = m->bin_start[c + 1] - m->bin_start[c];
    if (count == 0) continue;
    const int* rows = m->bin_rows + m->bin_start[c];
    const int G = 1 << c;
    dim3 grid(grid_for((int64_t)count * G, kBlock));
#define PSB_CSRB_CASE(GG)                                                                     \
  case GG:                                                                                    \
    hipLaunchKernelGGL((k_csr_list<T, GG>), grid, dim3(kBlock), 0, s, count, rows, m->irp,   \
                       m->ja, val, x, alpha, beta, y);                                        \
    break;
    switch (G) {
      PSB_CSRB_CASE(1)
      PSB_CSRB_CASE(2)
      PSB_CSRB_CASE(4)
      PSB_CSRB_CASE(8)
      PSB_CSRB_CASE(16)
      PSB_CSRB_CASE(32)
      PSB_CSRB_CASE(64)
    }
#undef PSB_CSRB_CASE
    PSB_HIP_LAUNCH_CHECK();
  }
  if (m->nlong_b > 0) {
    int nb = grid_for(m->nlong_b, kBlock / kWave);
    hipLaunchKernelGGL((k_csr_long<T>), dim3(nb), dim3(kBlock), 0, s, m->nlong_b,
                       m->long_rows_b, m->irp, m->ja, val, x, alpha, beta, y);
    PSB_HIP_LAUNCH_CHECK();
  }
  return PSB_HIP_SUCCESS;
}

template <typename T>
static int csr_launch_rowblock(CsrDev* m, const T* x, T alpha, T beta, T* y) {
  hipStream_t s = ctx().stream;
  const T* val = (const T*)m->val;
  if (m->nsblk > 0) {
    const int xc = g_hll_xchunk >= 0 ? g_hll_xchunk : kXchunkDefault;
    const int g = xc > 0 ? xcd_chunk_grid(m->nsblk, xc) : m->nsblk;
    hipLaunchKernelGGL((k_csr_rowblock<T>), dim3(g), dim3(kBlock), 0, s, m->nsblk, xc, m->sblk,
                       m->irp, m->ja, val, x, alpha, beta, y);
    PSB_HIP_LAUNCH_CHECK();
  }
  if (m->nlong_b > 0) {
    int nb = grid_for(m->nlong_b, kBlock / kWave);
    hipLaunchKernelGGL((k_csr_long<T>), dim3(nb), dim3(kBlock), 0, s, m->nlong_b,
                       m->long_rows_b, m->irp, m->ja, val, x, alpha, beta, y);
    PSB_HIP_LAUNCH_CHECK();
  }
  return PSB_HIP_SUCCESS;
}

template <typename T>
static int csr_launch(CsrDev* m, const T* x, T alpha, T beta, T* y) {
  if (m->nslices > 0 && (g_csr_variant == 0 || g_csr_variant == 4)) {
    // y = alpha A_0 x + beta y, then y = alpha A_s x + y slice after slice
    for (int s = 0; s < m->nslices; s++) {
      const int rc = csr_launch_rowblock<T>(m->slices[s], x, alpha, s == 0 ? beta : T(1), y);
      if (rc) return rc;
    }
    return PSB_HIP_SUCCESS;
  }
  const bool binned = g_csr_variant == 2;
  if (binned && m->bin_rows) return csr_launch_binned<T>(m, x, alpha, beta, y);
  if ((g_csr_variant == 0 || g_csr_variant == 3 || g_csr_variant == 4) && m->rows > 0)
    return csr_launch_rowblock<T>(m, x, alpha, beta, y);
  hipStream_t s = ctx().stream;
  const T* val = (const T*)m->val;
  int g = m->lanes;
  int nlog = grid_for((int64_t)m->rows * g, kBlock);
  dim3 grid(xcd_grid(nlog));
#define PSB_CSR_CASE(G)                                                                         \
  case G:                                                                                       \
    hipLaunchKernelGGL((k_csr_vector<T, G>), grid, dim3(kBlock), 0, s, m->rows, nlog, m->irp,  \
                       m->ja, val, x, alpha, beta, y, m->long_thr);                             \
    break;
  switch (g) {
    PSB_CSR_CASE(1)
    PSB_CSR_CASE(2)
    PSB_CSR_CASE(4)
    PSB_CSR_CASE(8)
    PSB_CSR_CASE(16)
    PSB_CSR_CASE(32)
    PSB_CSR_CASE(64)
    default: return PSB_HIP_INVALID;
  }
#undef PSB_CSR_CASE
  PSB_HIP_LAUNCH_CHECK();
  if (m->nlong > 0) {
    int nb = grid_for(m->nlong, kBlock / kWave);
    hipLaunchKernelGGL((k_csr_long<T>), dim3(nb), dim3(kBlock), 0, s, m->nlong, m->long_rows,
                       m->irp, m->ja, val, x, alpha, beta, y);
    PSB_HIP_LAUNCH_CHECK();
  }
  return PSB_HIP_SUCCESS;
}

}  // namespace psbhip

using namespace psbhip;

namespace psbhip {
static void free_csr_slices(CsrDev* m);

static int csr_write_impl(CsrDev* m, const int* irp, const int* ja, const void* val) {
  release_csr_sv(m);   // a prepared triangular solve describes the old values
  free_csr_slices(m);
  if (irp[0] != 1 || irp[m->rows] - 1 != m->nnz)
    return set_error(PSB_HIP_INVALID, __func__, "irp inconsistent with nnz");
  for (int i = 0; i < m->rows; i++)
    if (irp[i + 1] < irp[i]) return set_error(PSB_HIP_INVALID, __func__, "irp not monotone");
  int mn = min_index(ja, m->nnz), mx = max_index(ja, m->nnz);
  if (m->nnz > 0 && (mn < 1 || mx > m->cols))
    return set_error(PSB_HIP_INVALID, __func__, "ja out of range");
  m->max_col = mx;
  // analysis: lanes per row from the mean row length, long rows listed
  double mean = m->rows > 0 ? (double)m->nnz / m->rows : 0.0;
  int g = 1;
  while (g < 64 && g < mean) g <<= 1;
  m->lanes = g;
  m->long_thr = std::max(256, 8 * g);
  std::vector<int> longs;
  for (int i = 0; i < m->rows; i++)
    if (irp[i + 1] - irp[i] > m->long_thr) longs.push_back(i);
  m->nlong = (int)longs.size();
  // length-binned analysis
  int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<int> longs_b;
  for (int i = 0; i < m->rows; i++) {
    const int len = irp[i + 1] - irp[i];
    if (len > kCsrLongRow) {
      longs_b.push_back(i);
      continue;
    }
    cnt[csr_class(len)]++;
  }
  m->bin_start[0] = 0;
  for (int c = 0; c < 7; c++) m->bin_start[c + 1] = m->bin_start[c] + cnt[c];
  std::vector<int> bins((size_t)(m->bin_start[7] > 0 ? m->bin_start[7] : 1));
  {
    int fill[7];
    for (int c = 0; c < 7; c++) fill[c] = m->bin_start[c];
    for (int i = 0; i < m->rows; i++) {
      const int len = irp[i + 1] - irp[i];
      if (len <= kCsrLongRow) bins[(size_t)fill[csr_class(len)]++] = i;
    }
  }
  m->nlong_b = (int)longs_b.size();
  // row blocks: consecutive rows up to kBlock of them and kCsrBlockNnz
  // entries; a row longer than kCsrLongRow ends the block and is left out
  std::vector<int2> sblk;
  {
    int r0 = 0;
    while (r0 < m->rows) {
      if (irp[r0 + 1] - irp[r0] > kCsrLongRow) {
        r0++;
        continue;
      }
      int r1 = r0;
      while (r1 < m->rows && r1 - r0 < kBlock && irp[r1 + 1] - irp[r1] <= kCsrLongRow &&
             irp[r1 + 1] - irp[r0] <= kCsrBlockNnz)
        r1++;
      sblk.push_back(make_int2(r0, r1));
      r0 = r1;
    }
  }
  m->nsblk = (int)sblk.size();
  hipStream_t s = ctx().stream;
  hipFree(m->long_rows);
  m->long_rows = nullptr;
  hipFree(m->bin_rows);
  m->bin_rows = nullptr;
  hipFree(m->long_rows_b);
  m->long_rows_b = nullptr;
  if (m->nlong > 0) {
    PSB_HIP_CHECK(hipMalloc(&m->long_rows, sizeof(int) * m->nlong));
    PSB_HIP_CHECK(hipMemcpyAsync(m->long_rows, longs.data(), sizeof(int) * m->nlong,
                                 hipMemcpyHostToDevice, s));
  }
  if (m->rows > 0) {
    PSB_HIP_CHECK(hipMalloc(&m->bin_rows, sizeof(int) * bins.size()));
    PSB_HIP_CHECK(hipMemcpyAsync(m->bin_rows, bins.data(), sizeof(int) * bins.size(),
                                 hipMemcpyHostToDevice, s));
  }
  hipFree(m->sblk);
  m->sblk = nullptr;
  if (m->nsblk > 0) {
    PSB_HIP_CHECK(hipMalloc(&m->sblk, sizeof(int2) * m->nsblk));
    PSB_HIP_CHECK(hipMemcpyAsync(m->sblk, sblk.data(), sizeof(int2) * m->nsblk,
                                 hipMemcpyHostToDevice, s));
  }
  if (m->nlong_b > 0) {
    PSB_HIP_CHECK(hipMalloc(&m->long_rows_b, sizeof(int) * m->nlong_b));
    PSB_HIP_CHECK(hipMemcpyAsync(m->long_rows_b, longs_b.data(), sizeof(int) * m->nlong_b,
                                 hipMemcpyHostToDevice, s));
  }
  size_t es = elem_size(m->elem_type);
  PSB_HIP_CHECK(hipMemcpyAsync(m->irp, irp, sizeof(int) * (m->rows + 1), hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(m->ja, ja, sizeof(int) * (size_t)m->nnz, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(m->val, val, es * (size_t)m->nnz, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  return PSB_HIP_SUCCESS;
}

// Column slices (see g_csr_slices): slice s holds the entries of columns
// s*w+1 .. (s+1)*w of every row, in CSR order, as a CSR matrix of its own with
// the row-block analysis; the product runs the slices one after another,
// accumulating into y.
static void free_csr_slices(CsrDev* m) {
  if (!m->slices) return;
  for (int s = 0; s < m->nslices; s++) psb_hip_csr_free(m->slices[s]);
  delete[] m->slices;
  m->slices = nullptr;
  m->nslices = 0;
}

static int csr_slice_count(const CsrDev* m, const int* irp, const int* ja) {
  if (g_csr_slices >= 2) return std::min(g_csr_slices, std::max(m->cols, 1));
  if (g_csr_slices >= 0 || m->nnz == 0 || m->rows == 0) return 0;
  const size_t xb = (size_t)m->cols * elem_size(m->elem_type);
  if (xb <= kCsrSliceMinX) return 0;
  // scattered columns: most entries of about 65536 sampled rows at least
  // cols/16 columns away from the diagonal
  const int step = std::max(1, m->rows / 65536);
  const int64_t dist = m->cols / 16;
  int64_t far = 0, tot = 0;
  for (int i = 0; i < m->rows; i += step)
    for (int k = irp[i] - 1; k < irp[i + 1] - 1; k++) {
      const int64_t d = (int64_t)ja[k] - 1 - i;
      far += (d > dist || d < -dist) ? 1 : 0;
      tot++;
    }
  if (far * 2 <= tot) return 0;
  return (int)std::min<size_t>(kCsrMaxSlices, (xb + kCsrSliceBytes - 1) / kCsrSliceBytes);
}

static int csr_build_slices(CsrDev* m, const int* irp, const int* ja, const void* val) {
  const int S = csr_slice_count(m, irp, ja);
  if (S < 2) return PSB_HIP_SUCCESS;
  const size_t es = elem_size(m->elem_type);
  const int w = (m->cols + S - 1) / S;
  m->slices = new CsrDev*[S]();
  m->nslices = S;
  std::vector<int> sirp(m->rows + 1), sja;
  std::vector<char> sval;
  for (int s = 0; s < S; s++) {
    const int lo = s * w + 1, hi = std::min(m->cols, (s + 1) * w);
    sja.clear();
    sval.clear();
    sirp[0] = 1;
    for (int i = 0; i < m->rows; i++) {
      for (int k = irp[i] - 1; k < irp[i + 1] - 1; k++)
        if (ja[k] >= lo && ja[k] <= hi) {
          sja.push_back(ja[k]);
          const char* v = (const char*)val + (size_t)k * es;
          sval.insert(sval.end(), v, v + es);
        }
      sirp[i + 1] = (int)sja.size() + 1;
    }
    void* sub = nullptr;
    int rc = psb_hip_csr_alloc(&sub, m->rows, m->cols, (int)sja.size(), m->elem_type, 1);
    if (!rc) {
      m->slices[s] = (CsrDev*)sub;
      rc = csr_write_impl((CsrDev*)sub, sirp.data(), sja.data(), sval.data());
    }
    if (rc) {
      free_csr_slices(m);
      return rc;
    }
  }
  return PSB_HIP_SUCCESS;
}

// Back to rows in order (the arrays written next are in the reference's
// layout): drop the permutation and the work-unit plan.
static void hll_drop_idx16(HllDev* m) {
  hipFree(m->res16);
  hipFree(m->base16);
  m->res16 = nullptr;
  m->base16 = nullptr;
  m->band = 0;
}

static void hll_drop_slices(HllDev* m) {
  if (!m->slices) return;
  for (int s = 0; s < m->nslices; s++) psb_hip_hll_free(m->slices[s]);
  delete[] m->slices;
  m->slices = nullptr;
  m->nslices = 0;
}

static void hll_drop_sorted(HllDev* m) {
  hll_drop_idx16(m);   // ja is about to change
  hll_drop_slices(m);
  hipFree(m->perm);
  hipFree(m->units);
  hipFree(m->splits);
  hipFree(m->part);
  m->perm = nullptr;
  m->units = nullptr;
  m->splits = nullptr;
  m->part = nullptr;
  m->nunits = m->nsplit = 0;
}

// Build the 16-bit column offsets of a rows-in-order matrix; when a slot's
// offset does not fit (or memory is short) the matrix keeps only ja.
static int hll_build_idx16(HllDev* m) {
  hll_drop_idx16(m);
  if (m->perm || m->allocsize <= 0 || m->rows <= 0) return PSB_HIP_SUCCESS;
  hipStream_t s = ctx().stream;
  int* bad = nullptr;
  hipError_t e = hipMalloc((void**)&m->res16, sizeof(short) * (size_t)m->allocsize);
  if (e == hipSuccess)
    e = hipMalloc((void**)&m->base16, sizeof(int) * (size_t)(m->allocsize / m->hksz + 1));
  if (e == hipSuccess) e = hipMalloc((void**)&bad, 2 * sizeof(int));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    hipFree(bad);
    hll_drop_idx16(m);
    return PSB_HIP_SUCCESS;
  }
  int hbad[2] = {0, 0};
  PSB_HIP_CHECK(hipMemsetAsync(bad, 0, 2 * sizeof(int), s));
  hipLaunchKernelGGL(k_hll_idx16, dim3((unsigned)grid_for(m->noffs, kBlock)), dim3(kBlock), 0, s,
                     m->hksz, m->noffs, m->hkoffs, m->ja, m->res16, m->base16, bad);
  PSB_HIP_LAUNCH_CHECK();
  PSB_HIP_CHECK(hipMemcpyAsync(hbad, bad, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  hipFree(bad);
  if (hbad[0]) hll_drop_idx16(m);
  m->band = hbad[1];
  return PSB_HIP_SUCCESS;
}

// Row-sorted storage: copy perm (1-based row of each position, a bijection
// the caller has checked) and build the work-unit plan from the host hack
// offsets: hacks wider than kHllUnitSlots are cut, their units first (the
// widest hacks' units lead the launch), then every other hack in order.
static int hll_install_perm(HllDev* m, const int* perm, const int* hoffs) {
  const int hksz = m->hksz, noffs = m->noffs, nr = m->rows;
  std::vector<int4> units, splits;
  int npart = 0;
  for (int h = 0; h < noffs; h++) {
    const int w = (hoffs[h + 1] - hoffs[h]) / hksz;
    if (w <= kHllUnitSlots) continue;
    const int nu = (w + kHllUnitSlots - 1) / kHllUnitSlots;
    splits.push_back(make_int4(h, npart, nu, 0));
    for (int k = 0; k < nu; k++)
      units.push_back(make_int4(h, k * kHllUnitSlots, std::min(w, (k + 1) * kHllUnitSlots),
                                npart + k));
    npart += nu;
  }
  for (int h = 0; h < noffs; h++) {
    const int w = (hoffs[h + 1] - hoffs[h]) / hksz;
    if (w <= kHllUnitSlots) units.push_back(make_int4(h, 0, w, -1));
  }
  hipStream_t s = ctx().stream;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc((void**)&m->perm, sizeof(int) * (nr > 0 ? nr : 1));
  if (e == hipSuccess) e = hipMalloc((void**)&m->units, sizeof(int4) * (units.size() + 1));
  if (e == hipSuccess) e = hipMalloc((void**)&m->splits, sizeof(int4) * (splits.size() + 1));
  if (e == hipSuccess) e = hipMalloc(&m->part, (size_t)8 * hksz * (size_t)(npart > 0 ? npart : 1));
  if (e != hipSuccess) {
    hll_drop_sorted(m);
    return set_error(PSB_HIP_OUTOFMEM, "hll_install_perm", hipGetErrorString(e));
  }
  if (nr > 0)
    PSB_HIP_CHECK(hipMemcpyAsync(m->perm, perm, sizeof(int) * nr, hipMemcpyHostToDevice, s));
  if (!units.empty())
    PSB_HIP_CHECK(hipMemcpyAsync(m->units, units.data(), sizeof(int4) * units.size(),
                                 hipMemcpyHostToDevice, s));
  if (!splits.empty())
    PSB_HIP_CHECK(hipMemcpyAsync(m->splits, splits.data(), sizeof(int4) * splits.size(),
                                 hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));   // the vectors above are freed on return
  m->nunits = (int)units.size();
  m->nsplit = (int)splits.size();
  return PSB_HIP_SUCCESS;
}

static int check_perm(const int* perm, int nr) {
  std::vector<char> seen((size_t)(nr > 0 ? nr : 1), 0);
  for (int p = 0; p < nr; p++) {
    const int r = perm[p];
    if (r < 1 || r > nr || seen[r - 1]) return PSB_HIP_INVALID;
    seen[r - 1] = 1;
  }
  return PSB_HIP_SUCCESS;
}

}  // namespace psbhip

extern "C" {

// ================================================================== HLL
int psb_hip_hll_alloc(void** dmat, int hksz, int rows, int nzeros, int allocsize, int elem_type,
                      int first_index) {
  PSB_ENSURE_INIT();
  *dmat = nullptr;
  if (elem_type != PSB_HIP_TYPE_DOUBLE && elem_type != PSB_HIP_TYPE_FLOAT)
    return set_error(PSB_HIP_UNSUPPORTED, __func__, "element type");
  if (first_index != 1) return set_error(PSB_HIP_UNSUPPORTED, __func__, "first_index must be 1");
  if (hksz < 1 || rows < 0 || allocsize < 0) return set_error(PSB_HIP_INVALID, __func__, "sizes");
  HllDev* m = new HllDev();
  m->elem_type = elem_type;
  m->hksz = hksz;
  m->rows = rows;
  m->nzeros = nzeros;
  m->allocsize = allocsize;
  m->noffs = (rows + hksz - 1) / hksz;
  m->first_index = first_index;
  m->max_col = 0;
  size_t es = elem_size(elem_type);
  hipError_t e = hipSuccess;
  size_t asz = (size_t)(allocsize > 0 ? allocsize : 1);
  if (e == hipSuccess) e = hipMalloc(&m->val, asz * es);
  if (e == hipSuccess) e = hipMalloc(&m->ja, asz * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&m->hkoffs, sizeof(int) * (m->noffs + 1));
  if (e == hipSuccess) e = hipMalloc(&m->irn, sizeof(int) * (rows > 0 ? rows : 1));
  if (e == hipSuccess) e = hipMalloc(&m->idiag, sizeof(int) * (rows > 0 ? rows : 1));
  if (e != hipSuccess) {
    psb_hip_hll_free(m);
    return set_error(PSB_HIP_OUTOFMEM, __func__, hipGetErrorString(e));
  }
  *dmat = m;
  return PSB_HIP_SUCCESS;
}

void psb_hip_hll_free(void* dmat) {
  if (!dmat) return;
  HllDev* m = (HllDev*)dmat;
  if (ctx().initialized) hipStreamSynchronize(ctx().stream);
  hipFree(m->val);
  hipFree(m->ja);
  hipFree(m->hkoffs);
  hipFree(m->irn);
  hipFree(m->idiag);
  hipFree(m->val32);
  hipFree(m->perm);
  hipFree(m->units);
  hipFree(m->splits);
  hipFree(m->part);
  hipFree(m->res16);
  hipFree(m->base16);
  hll_drop_slices(m);
  delete m;
}

int psb_hip_hll_write(void* dmat, const void* val, const int* ja, const int* hkoffs,
                      const int* irn, const int* idiag) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  hipStream_t s = ctx().stream;
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  hll_drop_sorted(m);
  // the hack offsets must describe the allocation exactly
  if (hkoffs[0] != 0 || hkoffs[m->noffs] != m->allocsize)
    return set_error(PSB_HIP_INVALID, __func__, "hkoffs inconsistent with allocsize");
  for (int h = 0; h < m->noffs; h++)
    if (hkoffs[h + 1] < hkoffs[h] || (hkoffs[h + 1] - hkoffs[h]) % m->hksz != 0)
      return set_error(PSB_HIP_INVALID, __func__, "hkoffs not multiples of hksz");
  if (m->allocsize > 0 && min_index(ja, m->allocsize) < 1)
    return set_error(PSB_HIP_INVALID, __func__, "ja < 1");
  m->max_col = max_index(ja, m->allocsize);
  size_t es = elem_size(m->elem_type);
  PSB_HIP_CHECK(hipMemcpyAsync(m->val, val, (size_t)m->allocsize * es, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(m->ja, ja, (size_t)m->allocsize * 4, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(m->hkoffs, hkoffs, (size_t)(m->noffs + 1) * 4,
                               hipMemcpyHostToDevice, s));
  if (m->rows > 0) {
    PSB_HIP_CHECK(hipMemcpyAsync(m->irn, irn, (size_t)m->rows * 4, hipMemcpyHostToDevice, s));
    PSB_HIP_CHECK(hipMemcpyAsync(m->idiag, idiag, (size_t)m->rows * 4, hipMemcpyHostToDevice, s));
  }
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  return hll_build_idx16(m);
}

int psb_hip_hll_read(void* dmat, void* val, int* ja, int* hkoffs, int* irn, int* idiag) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  hipStream_t s = ctx().stream;
  size_t es = elem_size(m->elem_type);
  if (val)
    PSB_HIP_CHECK(hipMemcpyAsync(val, m->val, (size_t)m->allocsize * es, hipMemcpyDeviceToHost, s));
  if (ja)
    PSB_HIP_CHECK(hipMemcpyAsync(ja, m->ja, (size_t)m->allocsize * 4, hipMemcpyDeviceToHost, s));
  if (hkoffs)
    PSB_HIP_CHECK(hipMemcpyAsync(hkoffs, m->hkoffs, (size_t)(m->noffs + 1) * 4,
                                 hipMemcpyDeviceToHost, s));
  if (irn && m->rows > 0)
    PSB_HIP_CHECK(hipMemcpyAsync(irn, m->irn, (size_t)m->rows * 4, hipMemcpyDeviceToHost, s));
  if (idiag && m->rows > 0)
    PSB_HIP_CHECK(hipMemcpyAsync(idiag, m->idiag, (size_t)m->rows * 4, hipMemcpyDeviceToHost, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_get_params(void* dmat, int* hksz, int* rows, int* nzeros, int* allocsize,
                           int* hack_offs_len, int* first_index, int* elem_type) {
  HllDev* m = (HllDev*)dmat;
  if (!m) return PSB_HIP_INVALID;
  if (hksz) *hksz = m->hksz;
  if (rows) *rows = m->rows;
  if (nzeros) *nzeros = m->nzeros;
  if (allocsize) *allocsize = m->allocsize;
  if (hack_offs_len) *hack_offs_len = m->noffs + 1;
  if (first_index) *first_index = m->first_index;
  if (elem_type) *elem_type = m->elem_type;
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_copy_coo(int nr, int nc, int nza, int hksz, int noffs, int isz, const int* irn,
                         const int* hoffs, const int* idisp, const int* ja, const void* val,
                         void* dmat) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  if (m->rows != nr || m->hksz != hksz || m->noffs != noffs || m->allocsize != isz)
    return set_error(PSB_HIP_INVALID, __func__, "matrix shape mismatch");
  if (hoffs[0] != 0 || hoffs[noffs] != isz || idisp[0] != 0 || idisp[nr] != nza)
    return set_error(PSB_HIP_INVALID, __func__, "offsets inconsistent");
  for (int h = 0; h < noffs; h++) {
    int i0 = h * hksz, i1 = std::min(nr, i0 + hksz), mx = 0;
    for (int i = i0; i < i1; i++) mx = std::max(mx, irn[i]);
    if (hoffs[h + 1] - hoffs[h] != mx * hksz)
      return set_error(PSB_HIP_INVALID, __func__, "hack offsets do not match row counts");
  }
  for (int i = 0; i < nr; i++)
    if (idisp[i + 1] - idisp[i] != irn[i])
      return set_error(PSB_HIP_INVALID, __func__, "idisp does not match irn");
  int mn = min_index(ja, nza), mx = max_index(ja, nza);
  if (nza > 0 && (mn < 1 || mx > nc)) return set_error(PSB_HIP_INVALID, __func__, "ja out of range");
  m->max_col = std::max(mx, nr);  // padding of the last hack uses ja = nr
  hipStream_t s = ctx().stream;
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  hll_drop_sorted(m);
  size_t es = elem_size(m->elem_type);
  int *d_irn = nullptr, *d_hoffs = nullptr, *d_idisp = nullptr, *d_ja = nullptr;
  void* d_val = nullptr;
  hipError_t e = hipSuccess;
  size_t nzb = (size_t)(nza > 0 ? nza : 1);
  if (e == hipSuccess) e = hipMalloc(&d_irn, sizeof(int) * (nr > 0 ? nr : 1));
  if (e == hipSuccess) e = hipMalloc(&d_hoffs, sizeof(int) * (noffs + 1));
  if (e == hipSuccess) e = hipMalloc(&d_idisp, sizeof(int) * (nr + 1));
  if (e == hipSuccess) e = hipMalloc(&d_ja, sizeof(int) * nzb);
  if (e == hipSuccess) e = hipMalloc(&d_val, es * nzb);
  if (e != hipSuccess) {
    hipFree(d_irn); hipFree(d_hoffs); hipFree(d_idisp); hipFree(d_ja); hipFree(d_val);
    return set_error(PSB_HIP_OUTOFMEM, __func__, hipGetErrorString(e));
  }
  PSB_HIP_CHECK(hipMemcpyAsync(d_irn, irn, sizeof(int) * nr, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_hoffs, hoffs, sizeof(int) * (noffs + 1), hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_idisp, idisp, sizeof(int) * (nr + 1), hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_ja, ja, sizeof(int) * nza, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_val, val, es * nza, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(m->hkoffs, hoffs, sizeof(int) * (noffs + 1), hipMemcpyHostToDevice, s));
  if (nr > 0)
    PSB_HIP_CHECK(hipMemcpyAsync(m->irn, irn, sizeof(int) * nr, hipMemcpyHostToDevice, s));
  int64_t nthreads = (int64_t)noffs * hksz;
  if (nthreads > 0) {
    dim3 grid((unsigned)((nthreads + kBlock - 1) / kBlock));
    if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
      hipLaunchKernelGGL((k_hll_fill<double>), grid, dim3(kBlock), 0, s, nr, hksz, noffs, d_irn,
                         d_hoffs, d_idisp, d_ja, (const double*)d_val, (const int*)nullptr,
                         m->ja, (double*)m->val, m->idiag);
    else
      hipLaunchKernelGGL((k_hll_fill<float>), grid, dim3(kBlock), 0, s, nr, hksz, noffs, d_irn,
                         d_hoffs, d_idisp, d_ja, (const float*)d_val, (const int*)nullptr,
                         m->ja, (float*)m->val, m->idiag);
    PSB_HIP_LAUNCH_CHECK();
  }
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  hipFree(d_irn); hipFree(d_hoffs); hipFree(d_idisp); hipFree(d_ja); hipFree(d_val);
  m->nzeros = nza;
  return hll_build_idx16(m);
}

static int hll_copy_sorted_impl(int nr, int nc, int nza, int hksz, int noffs, int isz,
                                const int* perm, const int* irn_pos, const int* hoffs,
                                const int* idisp, const int* ja, const void* val, HllDev* m);
static int hll_build_slices(HllDev* m, int nc, const int* idisp, const int* ja, const void* val);

int psb_hip_hll_copy_coo_sorted(int nr, int nc, int nza, int hksz, int noffs, int isz,
                                const int* perm, const int* irn_pos, const int* hoffs,
                                const int* idisp, const int* ja, const void* val, void* dmat) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  if (!m || !perm || !irn_pos || !hoffs || !idisp) return PSB_HIP_INVALID;
  int rc = hll_copy_sorted_impl(nr, nc, nza, hksz, noffs, isz, perm, irn_pos, hoffs, idisp, ja,
                                val, m);
  if (rc) return rc;
  return hll_build_slices(m, nc, idisp, ja, val);
}

static int hll_copy_sorted_impl(int nr, int nc, int nza, int hksz, int noffs, int isz,
                                const int* perm, const int* irn_pos, const int* hoffs,
                                const int* idisp, const int* ja, const void* val, HllDev* m) {
  if (!m || !perm || !irn_pos || !hoffs || !idisp) return PSB_HIP_INVALID;
  if (m->rows != nr || m->hksz != hksz || m->noffs != noffs || m->allocsize != isz)
    return set_error(PSB_HIP_INVALID, __func__, "matrix shape mismatch");
  if (hoffs[0] != 0 || hoffs[noffs] != isz || idisp[0] != 0 || idisp[nr] != nza)
    return set_error(PSB_HIP_INVALID, __func__, "offsets inconsistent");
  if (check_perm(perm, nr))
    return set_error(PSB_HIP_INVALID, __func__, "perm is not a permutation of 1..nr");
  for (int p = 0; p < nr; p++)
    if (idisp[perm[p]] - idisp[perm[p] - 1] != irn_pos[p])
      return set_error(PSB_HIP_INVALID, __func__, "irn_pos does not match the rows' counts");
  for (int h = 0; h < noffs; h++) {
    int p0 = h * hksz, p1 = std::min(nr, p0 + hksz), mx = 0;
    for (int p = p0; p < p1; p++) mx = std::max(mx, irn_pos[p]);
    if (hoffs[h + 1] - hoffs[h] != mx * hksz)
      return set_error(PSB_HIP_INVALID, __func__, "hack offsets do not match row counts");
  }
  int mn = min_index(ja, nza), mx = max_index(ja, nza);
  if (nza > 0 && (mn < 1 || mx > nc)) return set_error(PSB_HIP_INVALID, __func__, "ja out of range");
  hipStream_t s = ctx().stream;
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  hll_drop_sorted(m);
  m->max_col = std::max(mx, nr);
  size_t es = elem_size(m->elem_type);
  int *d_irn = nullptr, *d_hoffs = nullptr, *d_idisp = nullptr, *d_ja = nullptr;
  void* d_val = nullptr;
  hipError_t e = hipSuccess;
  size_t nzb = (size_t)(nza > 0 ? nza : 1);
  if (e == hipSuccess) e = hipMalloc(&d_irn, sizeof(int) * (nr > 0 ? nr : 1));
  if (e == hipSuccess) e = hipMalloc(&d_hoffs, sizeof(int) * (noffs + 1));
  if (e == hipSuccess) e = hipMalloc(&d_idisp, sizeof(int) * (nr + 1));
  if (e == hipSuccess) e = hipMalloc(&d_ja, sizeof(int) * nzb);
  if (e == hipSuccess) e = hipMalloc(&d_val, es * nzb);
  if (e != hipSuccess) {
    hipFree(d_irn); hipFree(d_hoffs); hipFree(d_idisp); hipFree(d_ja); hipFree(d_val);
    return set_error(PSB_HIP_OUTOFMEM, __func__, hipGetErrorString(e));
  }
  int rc = hll_install_perm(m, perm, hoffs);
  if (rc) {
    hipFree(d_irn); hipFree(d_hoffs); hipFree(d_idisp); hipFree(d_ja); hipFree(d_val);
    return rc;
  }
  PSB_HIP_CHECK(hipMemcpyAsync(d_irn, irn_pos, sizeof(int) * nr, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_hoffs, hoffs, sizeof(int) * (noffs + 1), hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_idisp, idisp, sizeof(int) * (nr + 1), hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_ja, ja, sizeof(int) * nza, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_val, val, es * nza, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(m->hkoffs, hoffs, sizeof(int) * (noffs + 1), hipMemcpyHostToDevice, s));
  if (nr > 0)
    PSB_HIP_CHECK(hipMemcpyAsync(m->irn, irn_pos, sizeof(int) * nr, hipMemcpyHostToDevice, s));
  int64_t nthreads = (int64_t)noffs * hksz;
  if (nthreads > 0) {
    dim3 grid((unsigned)((nthreads + kBlock - 1) / kBlock));
    if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
      hipLaunchKernelGGL((k_hll_fill<double>), grid, dim3(kBlock), 0, s, nr, hksz, noffs, d_irn,
                         d_hoffs, d_idisp, d_ja, (const double*)d_val, (const int*)m->perm,
                         m->ja, (double*)m->val, m->idiag);
    else
      hipLaunchKernelGGL((k_hll_fill<float>), grid, dim3(kBlock), 0, s, nr, hksz, noffs, d_irn,
                         d_hoffs, d_idisp, d_ja, (const float*)d_val, (const int*)m->perm,
                         m->ja, (float*)m->val, m->idiag);
    PSB_HIP_LAUNCH_CHECK();
  }
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  hipFree(d_irn); hipFree(d_hoffs); hipFree(d_idisp); hipFree(d_ja); hipFree(d_val);
  m->nzeros = nza;
  if (m->val32) refresh_single(m);
  return PSB_HIP_SUCCESS;
}

// The rule of csr_slice_count on the host arrays of a row-sorted copy: x
// over kCsrSliceMinX bytes and most sampled entries farther than cols/16 from
// the diagonal; slices of about kCsrSliceBytes of x.
static int hll_slice_count(const HllDev* m, int nc, const int* idisp, const int* ja) {
  if (g_hll_slices >= 2) return std::min(g_hll_slices, std::max(nc, 1));
  if (g_hll_slices >= 0 || m->nzeros == 0 || m->rows == 0) return 0;
  const size_t xb = (size_t)nc * elem_size(m->elem_type);
  if (xb <= kCsrSliceMinX) return 0;
  const int step = std::max(1, m->rows / 65536);
  const int64_t dist = nc / 16;
  int64_t far = 0, tot = 0;
  for (int i = 0; i < m->rows; i += step)
    for (int k = idisp[i]; k < idisp[i + 1]; k++) {
      const int64_t d = (int64_t)ja[k] - 1 - i;
      far += (d > dist || d < -dist) ? 1 : 0;
      tot++;
    }
  if (far * 2 <= tot) return 0;
  const size_t sb = elem_size(m->elem_type) == 8 ? kHllSliceBytes : kHllSliceBytes32;
  return (int)std::min<size_t>(kCsrMaxSlices, (xb + sb - 1) / sb);
}

// Slice s: the entries of columns s*w+1 .. (s+1)*w of every row (row order
// and CSR order kept), sorted by its own row lengths over the whole matrix.
static int hll_build_slices(HllDev* m, int nc, const int* idisp, const int* ja, const void* val) {
  const int S = hll_slice_count(m, nc, idisp, ja);
  if (S < 2) return PSB_HIP_SUCCESS;
  const size_t es = elem_size(m->elem_type);
  const int nr = m->rows, hksz = m->hksz, noffs = m->noffs;
  const int w = (nc + S - 1) / S;
  m->slices = new HllDev*[S]();
  m->nslices = S;
  std::vector<int> sidisp(nr + 1), cnt(nr > 0 ? nr : 1), perm(nr > 0 ? nr : 1),
      cnt_pos(nr > 0 ? nr : 1), hoffs(noffs + 1), sja;
  std::vector<char> sval;
  for (int s = 0; s < S; s++) {
    const int lo = s * w + 1, hi = std::min(nc, (s + 1) * w);
    sja.clear();
    sval.clear();
    sidisp[0] = 0;
    for (int i = 0; i < nr; i++) {
      for (int k = idisp[i]; k < idisp[i + 1]; k++)
        if (ja[k] >= lo && ja[k] <= hi) {
          sja.push_back(ja[k]);
          const char* v = (const char*)val + (size_t)k * es;
          sval.insert(sval.end(), v, v + es);
        }
      sidisp[i + 1] = (int)sja.size();
      cnt[i] = sidisp[i + 1] - sidisp[i];
    }
    int64_t isz = 0;
    int rc = psb_hip_hll_sort_rows(nr, hksz, 0, cnt.data(), perm.data(), cnt_pos.data(),
                                   hoffs.data(), &isz);
    if (!rc && isz > 0x7fffffffLL) rc = set_error(PSB_HIP_UNSUPPORTED, __func__, "slice size");
    void* sub = nullptr;
    if (!rc) rc = psb_hip_hll_alloc(&sub, hksz, nr, (int)sja.size(), (int)isz, m->elem_type, 1);
    if (!rc) {
      m->slices[s] = (HllDev*)sub;
      rc = hll_copy_sorted_impl(nr, nc, (int)sja.size(), hksz, noffs, (int)isz, perm.data(),
                                cnt_pos.data(), hoffs.data(), sidisp.data(), sja.data(),
                                sval.data(), (HllDev*)sub);
    }
    if (rc) {
      hll_drop_slices(m);
      return rc;
    }
  }
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_set_perm(void* dmat, const int* perm, const int* hkoffs) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  if (!m || !perm || !hkoffs) return PSB_HIP_INVALID;
  if (check_perm(perm, m->rows))
    return set_error(PSB_HIP_INVALID, __func__, "perm is not a permutation of 1..rows");
  if (hkoffs[0] != 0 || hkoffs[m->noffs] != m->allocsize)
    return set_error(PSB_HIP_INVALID, __func__, "hkoffs inconsistent with allocsize");
  PSB_HIP_CHECK(hipStreamSynchronize(ctx().stream));
  hll_drop_sorted(m);
  return hll_install_perm(m, perm, hkoffs);
}

int psb_hip_hll_read_perm(void* dmat, int* perm) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  if (!m || !perm) return PSB_HIP_INVALID;
  if (!m->perm) {
    for (int p = 0; p < m->rows; p++) perm[p] = p + 1;
    return PSB_HIP_SUCCESS;
  }
  if (m->rows > 0)
    PSB_HIP_CHECK(hipMemcpy(perm, m->perm, sizeof(int) * (size_t)m->rows, hipMemcpyDeviceToHost));
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_units(void* dmat, int* nunits, int* nsplit) {
  HllDev* m = (HllDev*)dmat;
  if (!m) return PSB_HIP_INVALID;
  if (nunits) *nunits = m->nunits;
  if (nsplit) *nsplit = m->nsplit;
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_spmv(void* dmat, double alpha, void* xv, double beta, void* yv) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  MultiVec *x = (MultiVec*)xv, *y = (MultiVec*)yv;
  int rc = check_vectors(x, y, m->elem_type, m->max_col, m->rows);
  if (rc) return set_error(rc, __func__, "x/y type, size or aliasing");
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  if (x->count == 1) {
    if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
      return hll_launch<double, double>(m, (const double*)m->val, (const double*)x->v, alpha,
                                        beta, (double*)y->v);
    return hll_launch<float, float>(m, (const float*)m->val, (const float*)x->v, (float)alpha,
                                    (float)beta, (float*)y->v);
  }
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
    return hll_spmm_launch<double>(m, (const double*)m->val, (const double*)x->v, x->pitch, alpha,
                                   beta, (double*)y->v, y->pitch, x->count);
  return hll_spmm_launch<float>(m, (const float*)m->val, (const float*)x->v, x->pitch,
                                (float)alpha, (float)beta, (float*)y->v, y->pitch, x->count);
}

int psb_hip_hll_make_single(void* dmat) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  if (m->elem_type != PSB_HIP_TYPE_DOUBLE) return PSB_HIP_UNSUPPORTED;
  if (!m->val32) {
    PSB_HIP_CHECK(hipMalloc(&m->val32, sizeof(float) * (size_t)(m->allocsize > 0 ? m->allocsize : 1)));
  }
  if (m->allocsize > 0) {
    hipLaunchKernelGGL(k_d2s, dim3(stream_grid(m->allocsize)), dim3(kBlock), 0, ctx().stream,
                       (int64_t)m->allocsize, (const double*)m->val, m->val32);
    PSB_HIP_LAUNCH_CHECK();
  }
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_spmv_single(void* dmat, double alpha, void* xv, double beta, void* yv) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  if (!m->val32) return set_error(PSB_HIP_INVALID, __func__, "no single-precision copy");
  MultiVec *x = (MultiVec*)xv, *y = (MultiVec*)yv;
  int rc = check_vectors(x, y, PSB_HIP_TYPE_FLOAT, m->max_col, m->rows);
  if (rc) return set_error(rc, __func__, "x/y type, size or aliasing");
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  if (x->count != 1) return PSB_HIP_UNSUPPORTED;
  return hll_launch<float, float>(m, m->val32, (const float*)x->v, (float)alpha, (float)beta,
                                  (float*)y->v);
}

int psb_hip_hll_get_diag(void* dmat, void* dv) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  MultiVec* d = (MultiVec*)dv;
  if (d->elem_type != m->elem_type || d->size < m->rows) return PSB_HIP_INVALID;
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  dim3 grid(grid_for(m->rows, kBlock));
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_hll_diag<double>), grid, dim3(kBlock), 0, ctx().stream, m->rows,
                       m->hksz, m->hkoffs, m->idiag, (const int*)m->perm, (const double*)m->val,
                       (double*)d->v);
  else
    hipLaunchKernelGGL((k_hll_diag<float>), grid, dim3(kBlock), 0, ctx().stream, m->rows,
                       m->hksz, m->hkoffs, m->idiag, (const int*)m->perm, (const float*)m->val,
                       (float*)d->v);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_tune(void* dmat, void* xv, void* yv, int reps, int* best, float* us) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  if (!m || !best || reps < 1) return PSB_HIP_INVALID;
  static const int cand[4] = {1, 2, 5, 6};
  hipEvent_t e0, e1;
  PSB_HIP_CHECK(hipEventCreate(&e0));
  PSB_HIP_CHECK(hipEventCreate(&e1));
  const int saved = m->variant;
  float bt = 0.0f;
  int rc = PSB_HIP_SUCCESS;
  *best = 0;
  for (int c = 0; c < 4 && rc == PSB_HIP_SUCCESS; c++) {
    m->variant = cand[c];
    rc = psb_hip_hll_spmv(dmat, 1.0, xv, 0.0, yv);  // warm-up, argument checks
    if (rc) break;
    hipError_t e = hipEventRecord(e0, ctx().stream);
    for (int r = 0; r < reps && rc == PSB_HIP_SUCCESS; r++)
      rc = psb_hip_hll_spmv(dmat, 1.0, xv, 0.0, yv);
    if (e == hipSuccess) e = hipEventRecord(e1, ctx().stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess && rc == PSB_HIP_SUCCESS) rc = set_error(PSB_HIP_RUNTIME, __func__, hipGetErrorString(e));
    const float t = 1e3f * ms / (float)reps;
    if (us) us[c] = t;
    if (rc == PSB_HIP_SUCCESS && (*best == 0 || t < bt)) {
      *best = cand[c];
      bt = t;
    }
  }
  m->variant = rc == PSB_HIP_SUCCESS ? *best : saved;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int psb_hip_hll_set_slices(int slices) {
  const int old = g_hll_slices;
  if (slices >= -1 && slices <= kCsrMaxSlices) g_hll_slices = slices;
  return old;
}

int psb_hip_hll_slices(void* dmat) {
  const HllDev* m = (const HllDev*)dmat;
  return m ? m->nslices : -1;
}

int psb_hip_hll_set_index16(int on) {
  const int old = g_hll_idx16;
  if (on == 0 || on == 1) g_hll_idx16 = on;
  return old;
}

int psb_hip_ell_index16(void* dmat) {
  const EllDev* m = (const EllDev*)dmat;
  return m && m->res16 ? 1 : 0;
}

int psb_hip_hll_index16(void* dmat) {
  const HllDev* m = (const HllDev*)dmat;
  return m && m->res16 ? 1 : 0;
}

int psb_hip_hll_set_plane(void* dmat, int period_rows, int tile_blocks) {
  HllDev* m = (HllDev*)dmat;
  if (!m || period_rows < 0 || tile_blocks < 0)
    return set_error(PSB_HIP_INVALID, __func__, "matrix, period or tile");
  if (m->perm && period_rows > 0)
    return set_error(PSB_HIP_UNSUPPORTED, __func__, "row-sorted HLL");
  m->plane_rows = tile_blocks > 0 ? period_rows : 0;
  m->plane_tb = period_rows > 0 ? tile_blocks : 0;
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_set_variant(int variant) {
  int old = g_hll_variant;
  if ((variant & ~4) >= 0 && (variant & ~4) <= 2) g_hll_variant = variant;
  return old;
}

// ================================================================== ELL
int psb_hip_ell_alloc(void** dmat, int rows, int max_row_size, int nnzeros, int elem_type,
                      int first_index) {
  PSB_ENSURE_INIT();
  *dmat = nullptr;
  if (elem_type != PSB_HIP_TYPE_DOUBLE && elem_type != PSB_HIP_TYPE_FLOAT)
    return set_error(PSB_HIP_UNSUPPORTED, __func__, "element type");
  if (first_index != 1) return set_error(PSB_HIP_UNSUPPORTED, __func__, "first_index must be 1");
  if (rows < 0 || max_row_size < 0) return set_error(PSB_HIP_INVALID, __func__, "sizes");
  EllDev* m = new EllDev();
  m->elem_type = elem_type;
  m->rows = rows;
  m->pitch = ((rows + 31) / 32) * 32;  // cuda/spgpu/ell.c:33-37
  m->nzm = max_row_size;
  m->nnz = nnzeros;
  m->first_index = first_index;
  size_t slots = (size_t)m->pitch * m->nzm;
  if (slots == 0) slots = 1;
  size_t es = elem_size(elem_type);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&m->val, slots * es);
  if (e == hipSuccess) e = hipMalloc(&m->ja, slots * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&m->irn, sizeof(int) * (rows > 0 ? rows : 1));
  if (e == hipSuccess) e = hipMalloc(&m->idiag, sizeof(int) * (rows > 0 ? rows : 1));
  if (e != hipSuccess) {
    psb_hip_ell_free(m);
    return set_error(PSB_HIP_OUTOFMEM, __func__, hipGetErrorString(e));
  }
  *dmat = m;
  return PSB_HIP_SUCCESS;
}

void psb_hip_ell_free(void* dmat) {
  if (!dmat) return;
  EllDev* m = (EllDev*)dmat;
  if (ctx().initialized) hipStreamSynchronize(ctx().stream);
  hipFree(m->val);
  hipFree(m->ja);
  hipFree(m->irn);
  hipFree(m->idiag);
  ell_drop_idx16(m);
  delete m;
}

int psb_hip_ell_pitch(void* dmat) { return dmat ? ((EllDev*)dmat)->pitch : -1; }
int psb_hip_ell_max_row_size(void* dmat) { return dmat ? ((EllDev*)dmat)->nzm : -1; }

int psb_hip_ell_write(void* dmat, const void* val, const int* ja, int ldv, const int* irn,
                      const int* idiag) {
  PSB_ENSURE_INIT();
  EllDev* m = (EllDev*)dmat;
  if (ldv < m->rows) return set_error(PSB_HIP_INVALID, __func__, "ldv < rows");
  hipStream_t s = ctx().stream;
  size_t es = elem_size(m->elem_type);
  // validate the stored column indices of the real rows
  int mx = 0;
  for (int j = 0; j < m->nzm; j++)
    for (int i = 0; i < m->rows; i++) {
      int c = ja[(size_t)j * ldv + i];
      if (c < 1) return set_error(PSB_HIP_INVALID, __func__, "ja < 1");
      mx = c > mx ? c : mx;
    }
  m->max_col = mx;
  if (m->rows > 0 && m->nzm > 0) {
    // pad rows rows..pitch-1 with ja = 1, val = 0 through a zeroed staging copy
    std::vector<char> hv((size_t)m->pitch * m->nzm * es, 0);
    std::vector<int> hj((size_t)m->pitch * m->nzm, 1);
    for (int j = 0; j < m->nzm; j++) {
      memcpy(&hv[(size_t)j * m->pitch * es], (const char*)val + (size_t)j * ldv * es,
             (size_t)m->rows * es);
      memcpy(&hj[(size_t)j * m->pitch], ja + (size_t)j * ldv, (size_t)m->rows * 4);
    }
    PSB_HIP_CHECK(hipMemcpyAsync(m->val, hv.data(), hv.size(), hipMemcpyHostToDevice, s));
    PSB_HIP_CHECK(hipMemcpyAsync(m->ja, hj.data(), hj.size() * 4, hipMemcpyHostToDevice, s));
    PSB_HIP_CHECK(hipStreamSynchronize(s));
  }
  if (m->rows > 0) {
    PSB_HIP_CHECK(hipMemcpyAsync(m->irn, irn, (size_t)m->rows * 4, hipMemcpyHostToDevice, s));
    PSB_HIP_CHECK(hipMemcpyAsync(m->idiag, idiag, (size_t)m->rows * 4, hipMemcpyHostToDevice, s));
  }
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  return ell_build_idx16(m);
}

int psb_hip_ell_read(void* dmat, void* val, int* ja, int ldv, int* irn, int* idiag) {
  PSB_ENSURE_INIT();
  EllDev* m = (EllDev*)dmat;
  if (ldv < m->rows) return PSB_HIP_INVALID;
  hipStream_t s = ctx().stream;
  size_t es = elem_size(m->elem_type);
  if (m->nzm > 0 && m->rows > 0) {
    if (val)
      PSB_HIP_CHECK(hipMemcpy2DAsync(val, (size_t)ldv * es, m->val, (size_t)m->pitch * es,
                                     (size_t)m->rows * es, m->nzm, hipMemcpyDeviceToHost, s));
    if (ja)
      PSB_HIP_CHECK(hipMemcpy2DAsync(ja, (size_t)ldv * 4, m->ja, (size_t)m->pitch * 4,
                                     (size_t)m->rows * 4, m->nzm, hipMemcpyDeviceToHost, s));
  }
  if (irn && m->rows > 0)
    PSB_HIP_CHECK(hipMemcpyAsync(irn, m->irn, (size_t)m->rows * 4, hipMemcpyDeviceToHost, s));
  if (idiag && m->rows > 0)
    PSB_HIP_CHECK(hipMemcpyAsync(idiag, m->idiag, (size_t)m->rows * 4, hipMemcpyDeviceToHost, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  return PSB_HIP_SUCCESS;
}

int psb_hip_ell_copy_coo(int nr, int nc, int nza, const int* irn, const int* idisp,
                         const int* ja, const void* val, void* dmat) {
  PSB_ENSURE_INIT();
  EllDev* m = (EllDev*)dmat;
  if (m->rows != nr) return set_error(PSB_HIP_INVALID, __func__, "rows mismatch");
  int mxr = 0;
  for (int i = 0; i < nr; i++) {
    mxr = std::max(mxr, irn[i]);
    if (idisp[i + 1] - idisp[i] != irn[i])
      return set_error(PSB_HIP_INVALID, __func__, "idisp does not match irn");
  }
  if (idisp[0] != 0 || idisp[nr] != nza) return set_error(PSB_HIP_INVALID, __func__, "idisp");
  if (mxr != m->nzm) return set_error(PSB_HIP_INVALID, __func__, "max row size mismatch");
  int mn = min_index(ja, nza), mx = max_index(ja, nza);
  if (nza > 0 && (mn < 1 || mx > nc)) return set_error(PSB_HIP_INVALID, __func__, "ja out of range");
  m->max_col = std::max(mx, 1);
  hipStream_t s = ctx().stream;
  size_t es = elem_size(m->elem_type);
  int *d_irn = nullptr, *d_idisp = nullptr, *d_ja = nullptr;
  void* d_val = nullptr;
  size_t nzb = (size_t)(nza > 0 ? nza : 1);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&d_irn, sizeof(int) * (nr > 0 ? nr : 1));
  if (e == hipSuccess) e = hipMalloc(&d_idisp, sizeof(int) * (nr + 1));
  if (e == hipSuccess) e = hipMalloc(&d_ja, sizeof(int) * nzb);
  if (e == hipSuccess) e = hipMalloc(&d_val, es * nzb);
  if (e != hipSuccess) {
    hipFree(d_irn); hipFree(d_idisp); hipFree(d_ja); hipFree(d_val);
    return set_error(PSB_HIP_OUTOFMEM, __func__, hipGetErrorString(e));
  }
  PSB_HIP_CHECK(hipMemcpyAsync(d_irn, irn, sizeof(int) * nr, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_idisp, idisp, sizeof(int) * (nr + 1), hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_ja, ja, sizeof(int) * nza, hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipMemcpyAsync(d_val, val, es * nza, hipMemcpyHostToDevice, s));
  if (nr > 0) PSB_HIP_CHECK(hipMemcpyAsync(m->irn, irn, sizeof(int) * nr, hipMemcpyHostToDevice, s));
  if (m->pitch > 0 && m->nzm > 0) {
    dim3 grid(grid_for(m->pitch, kBlock));
    if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
      hipLaunchKernelGGL((k_ell_fill<double>), grid, dim3(kBlock), 0, s, nr, m->pitch, m->nzm,
                         d_irn, d_idisp, d_ja, (const double*)d_val, m->ja, (double*)m->val,
                         m->idiag);
    else
      hipLaunchKernelGGL((k_ell_fill<float>), grid, dim3(kBlock), 0, s, nr, m->pitch, m->nzm,
                         d_irn, d_idisp, d_ja, (const float*)d_val, m->ja, (float*)m->val,
                         m->idiag);
    PSB_HIP_LAUNCH_CHECK();
  } else if (nr > 0) {
    PSB_HIP_CHECK(hipMemsetAsync(m->idiag, 0, sizeof(int) * nr, s));
  }
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  hipFree(d_irn); hipFree(d_idisp); hipFree(d_ja); hipFree(d_val);
  m->nnz = nza;
  return ell_build_idx16(m);
}

int psb_hip_ell_get_diag(void* dmat, void* dv) {
  PSB_ENSURE_INIT();
  EllDev* m = (EllDev*)dmat;
  MultiVec* d = (MultiVec*)dv;
  if (d->elem_type != m->elem_type || d->size < m->rows) return PSB_HIP_INVALID;
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  if (m->nzm == 0) {
    PSB_HIP_CHECK(hipMemsetAsync(d->v, 0, elem_size(m->elem_type) * m->rows, ctx().stream));
    return PSB_HIP_SUCCESS;
  }
  dim3 grid(grid_for(m->rows, kBlock));
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_ell_diag<double>), grid, dim3(kBlock), 0, ctx().stream, m->rows,
                       m->pitch, m->idiag, (const double*)m->val, (double*)d->v);
  else
    hipLaunchKernelGGL((k_ell_diag<float>), grid, dim3(kBlock), 0, ctx().stream, m->rows,
                       m->pitch, m->idiag, (const float*)m->val, (float*)d->v);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_ell_spmv(void* dmat, double alpha, void* xv, double beta, void* yv) {
  PSB_ENSURE_INIT();
  EllDev* m = (EllDev*)dmat;
  MultiVec *x = (MultiVec*)xv, *y = (MultiVec*)yv;
  int rc = check_vectors(x, y, m->elem_type, m->max_col, m->rows);
  if (rc) return set_error(rc, __func__, "x/y type, size or aliasing");
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  hipStream_t s = ctx().stream;
  // two rows per lane for 8-byte values when the pitch is even (L=200
  // stencil FP64: 130.7 -> 111.8 us); FP32 one row per lane, as for HLL
  const bool pair = g_ell_pair && m->elem_type == PSB_HIP_TYPE_DOUBLE && m->pitch % 2 == 0;
  const int rpb = pair ? 2 * kBlock : kBlock;
  int nlog = grid_for(m->rows, rpb);
  const int xc = xcd_chunk_for(m->band, rpb);
  dim3 grid(xcd_chunk_grid(nlog, xc));
  const bool i16 = ell_use_idx16(m);
  // every column of x/y (cuda/elldev.c:195-215 loops count_)
  for (int c = 0; c < x->count; c++) {
#define PSB_ELL_CASE(K, T, I16)                                                                 \
  hipLaunchKernelGGL((K<T, T, I16>), grid, dim3(kBlock), 0, s, m->rows, m->pitch, m->nzm, nlog, \
                     xc, m->ja, m->res16, m->base16, (const T*)m->val,                          \
                     (const T*)x->v + (size_t)c * x->pitch, (T)alpha, (T)beta,                 \
                     (T*)y->v + (size_t)c * y->pitch)
    if (pair) {
      if (i16) PSB_ELL_CASE(k_ell_spmv_pair, double, true);
      else PSB_ELL_CASE(k_ell_spmv_pair, double, false);
    } else if (m->elem_type == PSB_HIP_TYPE_DOUBLE) {
      if (i16) PSB_ELL_CASE(k_ell_spmv, double, true);
      else PSB_ELL_CASE(k_ell_spmv, double, false);
    } else {
      if (i16) PSB_ELL_CASE(k_ell_spmv, float, true);
      else PSB_ELL_CASE(k_ell_spmv, float, false);
    }
#undef PSB_ELL_CASE
    PSB_HIP_LAUNCH_CHECK();
  }
  return PSB_HIP_SUCCESS;
}

// ================================================================== CSR
int psb_hip_csr_alloc(void** dmat, int rows, int cols, int nnz, int elem_type, int first_index) {
  PSB_ENSURE_INIT();
  *dmat = nullptr;
  if (elem_type != PSB_HIP_TYPE_DOUBLE && elem_type != PSB_HIP_TYPE_FLOAT)
    return set_error(PSB_HIP_UNSUPPORTED, __func__, "element type");
  if (first_index != 1) return set_error(PSB_HIP_UNSUPPORTED, __func__, "first_index must be 1");
  if (rows < 0 || cols < 0 || nnz < 0) return set_error(PSB_HIP_INVALID, __func__, "sizes");
  CsrDev* m = new CsrDev();
  m->elem_type = elem_type;
  m->rows = rows;
  m->cols = cols;
  m->nnz = nnz;
  m->first_index = first_index;
  m->lanes = 8;
  m->long_thr = 0x7fffffff;
  size_t es = elem_size(elem_type);
  size_t nzb = (size_t)(nnz > 0 ? nnz : 1);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&m->irp, sizeof(int) * (rows + 1));
  if (e == hipSuccess) e = hipMalloc(&m->ja, sizeof(int) * nzb);
  if (e == hipSuccess) e = hipMalloc(&m->val, es * nzb);
  if (e != hipSuccess) {
    psb_hip_csr_free(m);
    return set_error(PSB_HIP_OUTOFMEM, __func__, hipGetErrorString(e));
  }
  *dmat = m;
  return PSB_HIP_SUCCESS;
}

void psb_hip_csr_free(void* dmat) {
  if (!dmat) return;
  CsrDev* m = (CsrDev*)dmat;
  if (ctx().initialized) hipStreamSynchronize(ctx().stream);
  release_csr_sv(m);
  free_csr_slices(m);
  hipFree(m->irp);
  hipFree(m->ja);
  hipFree(m->val);
  hipFree(m->long_rows);
  hipFree(m->bin_rows);
  hipFree(m->long_rows_b);
  hipFree(m->sblk);
  delete m;
}

int psb_hip_csr_write(void* dmat, const int* irp, const int* ja, const void* val) {
  PSB_ENSURE_INIT();
  CsrDev* m = (CsrDev*)dmat;
  const int rc = csr_write_impl(m, irp, ja, val);
  if (rc) return rc;
  return csr_build_slices(m, irp, ja, val);
}


int psb_hip_csr_read(void* dmat, int* irp, int* ja, void* val) {
  PSB_ENSURE_INIT();
  CsrDev* m = (CsrDev*)dmat;
  hipStream_t s = ctx().stream;
  size_t es = elem_size(m->elem_type);
  if (irp)
    PSB_HIP_CHECK(hipMemcpyAsync(irp, m->irp, sizeof(int) * (m->rows + 1), hipMemcpyDeviceToHost, s));
  if (ja) PSB_HIP_CHECK(hipMemcpyAsync(ja, m->ja, sizeof(int) * (size_t)m->nnz, hipMemcpyDeviceToHost, s));
  if (val) PSB_HIP_CHECK(hipMemcpyAsync(val, m->val, es * (size_t)m->nnz, hipMemcpyDeviceToHost, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));
  return PSB_HIP_SUCCESS;
}

int psb_hip_csr_get_diag(void* dmat, void* dv) {
  PSB_ENSURE_INIT();
  CsrDev* m = (CsrDev*)dmat;
  MultiVec* d = (MultiVec*)dv;
  if (d->elem_type != m->elem_type || d->size < m->rows) return PSB_HIP_INVALID;
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  dim3 grid(grid_for(m->rows, kBlock));
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_csr_diag<double>), grid, dim3(kBlock), 0, ctx().stream, m->rows, m->irp,
                       m->ja, (const double*)m->val, (double*)d->v);
  else
    hipLaunchKernelGGL((k_csr_diag<float>), grid, dim3(kBlock), 0, ctx().stream, m->rows, m->irp,
                       m->ja, (const float*)m->val, (float*)d->v);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_set_xcd_chunk(int blocks) {
  const int old = g_hll_xchunk;
  if (blocks >= -1) g_hll_xchunk = blocks;
  return old;
}

int psb_hip_csr_set_variant(int variant) {
  const int old = g_csr_variant;
  if (variant >= 0 && variant <= 4) g_csr_variant = variant;
  return old;
}

int psb_hip_csr_set_slices(int slices) {
  const int old = g_csr_slices;
  if (slices >= -1 && slices <= kCsrMaxSlices) g_csr_slices = slices;
  return old;
}

int psb_hip_csr_slices(void* dmat) {
  CsrDev* m = (CsrDev*)dmat;
  return m ? m->nslices : -1;
}

int psb_hip_csr_get_analysis(void* dmat, int* lanes_per_row, int* long_rows, int* long_threshold) {
  CsrDev* m = (CsrDev*)dmat;
  if (!m) return PSB_HIP_INVALID;
  if (lanes_per_row) *lanes_per_row = m->lanes;
  if (long_rows) *long_rows = m->nlong;
  if (long_threshold) *long_threshold = m->long_thr;
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_scals(void* dmat, double s) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  if (!m) return PSB_HIP_INVALID;
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE) launch_scale<double>(m->allocsize, s, m->val);
  else launch_scale<float>(m->allocsize, s, m->val);
  refresh_single(m);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_ell_scals(void* dmat, double s) {
  PSB_ENSURE_INIT();
  EllDev* m = (EllDev*)dmat;
  if (!m) return PSB_HIP_INVALID;
  const int64_t n = (int64_t)m->pitch * m->nzm;
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE) launch_scale<double>(n, s, m->val);
  else launch_scale<float>(n, s, m->val);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_csr_scals(void* dmat, double s) {
  PSB_ENSURE_INIT();
  CsrDev* m = (CsrDev*)dmat;
  if (!m) return PSB_HIP_INVALID;
  release_csr_sv(m);
  for (int k = -1; k < m->nslices; k++) {   // the matrix, then its column slices
    CsrDev* c = k < 0 ? m : m->slices[k];
    if (m->elem_type == PSB_HIP_TYPE_DOUBLE) launch_scale<double>(c->nnz, s, c->val);
    else launch_scale<float>(c->nnz, s, c->val);
  }
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_scal(void* dmat, void* dv, int left) {
  PSB_ENSURE_INIT();
  HllDev* m = (HllDev*)dmat;
  const MultiVec* d = (const MultiVec*)dv;
  if (!m || !scal_vec_ok(d, m->elem_type, left ? m->rows : m->max_col))
    return set_error(PSB_HIP_INVALID, __func__, "d type or size");
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  dim3 grid(grid_for(m->rows, kBlock));
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_hll_scal<double>), grid, dim3(kBlock), 0, ctx().stream, m->rows, m->hksz,
                       m->hkoffs, m->irn, m->ja, (const int*)m->perm, (const double*)d->v, left,
                       (double*)m->val);
  else
    hipLaunchKernelGGL((k_hll_scal<float>), grid, dim3(kBlock), 0, ctx().stream, m->rows, m->hksz,
                       m->hkoffs, m->irn, m->ja, (const int*)m->perm, (const float*)d->v, left,
                       (float*)m->val);
  refresh_single(m);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_ell_scal(void* dmat, void* dv, int left) {
  PSB_ENSURE_INIT();
  EllDev* m = (EllDev*)dmat;
  const MultiVec* d = (const MultiVec*)dv;
  if (!m || !scal_vec_ok(d, m->elem_type, left ? m->rows : m->max_col))
    return set_error(PSB_HIP_INVALID, __func__, "d type or size");
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  dim3 grid(grid_for(m->rows, kBlock));
  if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_ell_scal<double>), grid, dim3(kBlock), 0, ctx().stream, m->rows, m->pitch,
                       m->nzm, m->irn, m->ja, (const double*)d->v, left, (double*)m->val);
  else
    hipLaunchKernelGGL((k_ell_scal<float>), grid, dim3(kBlock), 0, ctx().stream, m->rows, m->pitch,
                       m->nzm, m->irn, m->ja, (const float*)d->v, left, (float*)m->val);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_csr_scal(void* dmat, void* dv, int left) {
  PSB_ENSURE_INIT();
  CsrDev* m = (CsrDev*)dmat;
  const MultiVec* d = (const MultiVec*)dv;
  if (!m || !scal_vec_ok(d, m->elem_type, left ? m->rows : m->max_col))
    return set_error(PSB_HIP_INVALID, __func__, "d type or size");
  release_csr_sv(m);
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  dim3 grid(grid_for(m->rows, kBlock));
  for (int k = -1; k < m->nslices; k++) {   // the matrix, then its column slices
    CsrDev* c = k < 0 ? m : m->slices[k];
    if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
      hipLaunchKernelGGL((k_csr_scal<double>), grid, dim3(kBlock), 0, ctx().stream, c->rows,
                         c->irp, c->ja, (const double*)d->v, left, (double*)c->val);
    else
      hipLaunchKernelGGL((k_csr_scal<float>), grid, dim3(kBlock), 0, ctx().stream, c->rows,
                         c->irp, c->ja, (const float*)d->v, left, (float*)c->val);
  }
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

// ------------------------------------------------------------------ transposed product
int psb_hip_tmat_alloc(void** tmat, int nr, int nc, const int* irp, const int* ja,
                       const void* val, int elem_type) {
  PSB_ENSURE_INIT();
  if (!tmat) return PSB_HIP_INVALID;
  *tmat = nullptr;
  if (elem_type != PSB_HIP_TYPE_DOUBLE && elem_type != PSB_HIP_TYPE_FLOAT)
    return set_error(PSB_HIP_UNSUPPORTED, __func__, "element type");
  if (nr < 0 || nc < 0 || !irp || irp[0] != 1) return set_error(PSB_HIP_INVALID, __func__, "irp");
  const int nnz = irp[nr] - 1;
  for (int i = 0; i < nr; i++)
    if (irp[i + 1] < irp[i]) return set_error(PSB_HIP_INVALID, __func__, "irp not monotone");
  for (int k = 0; k < nnz; k++)
    if (ja[k] < 1 || ja[k] > nc) return set_error(PSB_HIP_INVALID, __func__, "ja out of range");
  // A^T row c = column c of A: its count, then hacks of kTHack rows whose
  // width is their longest row; entries placed rows of A ascending and, within
  // a row of A, in its CSR order
  const size_t es = elem_size(elem_type);
  const int nh = (nc + kTHack - 1) / kTHack;
  std::vector<int> cnt(nc > 0 ? nc : 1, 0), hk(nh + 1, 0);
  for (int k = 0; k < nnz; k++) cnt[ja[k] - 1]++;
  int64_t total = 0;
  for (int h = 0; h < nh; h++) {
    int w = 0;
    for (int c = h * kTHack; c < std::min(nc, (h + 1) * kTHack); c++) w = std::max(w, cnt[c]);
    total += (int64_t)w * kTHack;
    if (total > INT32_MAX)   // a hack is as wide as its fullest column of A
      return set_error(PSB_HIP_UNSUPPORTED, __func__, "transpose too large (padded slots)");
    hk[h + 1] = (int)total;
  }
  const size_t slots = (size_t)hk[nh];
  std::vector<int> src(slots > 0 ? slots : 1, 0), fill(nc > 0 ? nc : 1, 0);
  std::vector<char> tval((slots > 0 ? slots : 1) * es, 0);   // padding: source 0, value 0
  for (int i = 0; i < nr; i++)
    for (int k = irp[i] - 1; k < irp[i + 1] - 1; k++) {
      const int c = ja[k] - 1, h = c / kTHack;
      const size_t q = (size_t)hk[h] + (size_t)fill[c]++ * kTHack + (c - h * kTHack);
      src[q] = i;
      memcpy(tval.data() + q * es, (const char*)val + (size_t)k * es, es);
    }
  // 16-bit source offsets over each hack slot column's real entries
  std::vector<short> res(slots > 0 ? slots : 1, 0);
  std::vector<int> base(slots / kTHack + 1, 0);
  bool fits = true;
  for (int h = 0; h < nh && fits; h++) {
    const int w = (hk[h + 1] - hk[h]) / kTHack;
    const int c0 = h * kTHack, c1 = std::min(nc, c0 + kTHack);
    for (int j = 0; j < w && fits; j++) {
      const size_t q = (size_t)hk[h] + (size_t)j * kTHack;
      int64_t lo = INT64_MAX, hi = INT64_MIN;
      for (int c = c0; c < c1; c++)
        if (j < cnt[c]) {
          const int64_t d = (int64_t)src[q + (c - c0)] - c;
          lo = std::min(lo, d);
          hi = std::max(hi, d);
        }
      if (hi - lo > 65535) {
        fits = false;
        break;
      }
      const int64_t b = (lo + hi + 1) >> 1;
      base[(size_t)hk[h] / kTHack + j] = (int)b;
      for (int c = c0; c < c1; c++)
        if (j < cnt[c]) res[q + (c - c0)] = (short)((int64_t)src[q + (c - c0)] - c - b);
    }
  }
  TransDev* t = new TransDev();
  t->elem_type = elem_type;
  t->rows = nc;
  t->cols = nr;
  t->nnz = nnz;
  hipError_t e = hipMalloc(&t->hkoffs, sizeof(int) * hk.size());
  if (e == hipSuccess) e = hipMalloc(&t->cnt, sizeof(int) * cnt.size());
  if (e == hipSuccess) e = hipMalloc(&t->src, sizeof(int) * src.size());
  if (e == hipSuccess) e = hipMalloc(&t->val, tval.size());
  if (e == hipSuccess)
    e = hipMemcpy(t->hkoffs, hk.data(), sizeof(int) * hk.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(t->cnt, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(t->src, src.data(), sizeof(int) * src.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->val, tval.data(), tval.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && fits) {
    e = hipMalloc(&t->res16, sizeof(short) * res.size());
    if (e == hipSuccess) e = hipMalloc(&t->base16, sizeof(int) * base.size());
    if (e == hipSuccess)
      e = hipMemcpy(t->res16, res.data(), sizeof(short) * res.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(t->base16, base.data(), sizeof(int) * base.size(), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    psb_hip_tmat_free(t);
    return set_error(e == hipErrorOutOfMemory ? PSB_HIP_OUTOFMEM : PSB_HIP_RUNTIME, __func__,
                     hipGetErrorString(e));
  }
  *tmat = t;
  return PSB_HIP_SUCCESS;
}

void psb_hip_tmat_free(void* tmat) {
  if (!tmat) return;
  TransDev* t = (TransDev*)tmat;
  if (ctx().initialized) hipStreamSynchronize(ctx().stream);
  hipFree(t->hkoffs);
  hipFree(t->cnt);
  hipFree(t->src);
  hipFree(t->val);
  hipFree(t->res16);
  hipFree(t->base16);
  delete t;
}

int psb_hip_tmat_spmv(void* tmat, double alpha, void* xv, double beta, void* yv) {
  PSB_ENSURE_INIT();
  TransDev* t = (TransDev*)tmat;
  MultiVec *x = (MultiVec*)xv, *y = (MultiVec*)yv;
  if (!t) return PSB_HIP_INVALID;
  int rc = check_vectors(x, y, t->elem_type, t->cols, t->rows);
  if (rc) return set_error(rc, __func__, "x/y type, size or aliasing");
  if (t->rows == 0) return PSB_HIP_SUCCESS;
  dim3 grid(grid_for(t->rows, kBlock));
  const bool i16 = g_hll_idx16 && t->res16;
  // FP64 over 16-bit source offsets: two rows per lane (L=200 stencil:
  // 143-145 -> 127-128 us)
  const bool pair = g_tmv_pair && i16 && t->elem_type == PSB_HIP_TYPE_DOUBLE;
  if (pair) {
    dim3 gp(grid_for((t->rows + 1) / 2, kBlock));
    for (int c = 0; c < x->count; c++)
      hipLaunchKernelGGL((k_tmv_pair<double>), gp, dim3(kBlock), 0, ctx().stream, t->rows,
                         t->cols, t->hkoffs, t->cnt, t->res16, t->base16, (const double*)t->val,
                         (const double*)x->v + (size_t)c * x->pitch, alpha, beta,
                         (double*)y->v + (size_t)c * y->pitch);
    PSB_HIP_LAUNCH_CHECK();
    return PSB_HIP_SUCCESS;
  }
  for (int c = 0; c < x->count; c++) {
#define PSB_TMV_CASE(T, I16)                                                                     \
  hipLaunchKernelGGL((k_tmv_run<T, I16>), grid, dim3(kBlock), 0, ctx().stream, t->rows, t->cols, \
                     t->hkoffs, t->cnt, t->src, t->res16, t->base16, (const T*)t->val,           \
                     (const T*)x->v + (size_t)c * x->pitch, (T)alpha, (T)beta,                  \
                     (T*)y->v + (size_t)c * y->pitch)
    if (t->elem_type == PSB_HIP_TYPE_DOUBLE) {
      if (i16) PSB_TMV_CASE(double, true);
      else PSB_TMV_CASE(double, false);
    } else {
      if (i16) PSB_TMV_CASE(float, true);
      else PSB_TMV_CASE(float, false);
    }
#undef PSB_TMV_CASE
  }
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_csr_spmv(void* dmat, double alpha, void* xv, double beta, void* yv) {
  PSB_ENSURE_INIT();
  CsrDev* m = (CsrDev*)dmat;
  MultiVec *x = (MultiVec*)xv, *y = (MultiVec*)yv;
  int rc = check_vectors(x, y, m->elem_type, m->max_col, m->rows);
  if (rc) return set_error(rc, __func__, "x/y type, size or aliasing");
  if (m->rows == 0) return PSB_HIP_SUCCESS;
  for (int c = 0; c < x->count; c++) {
    if (m->elem_type == PSB_HIP_TYPE_DOUBLE)
      rc = csr_launch<double>(m, (const double*)x->v + (size_t)c * x->pitch, alpha, beta,
                              (double*)y->v + (size_t)c * y->pitch);
    else
      rc = csr_launch<float>(m, (const float*)x->v + (size_t)c * x->pitch, (float)alpha,
                             (float)beta, (float*)y->v + (size_t)c * y->pitch);
    if (rc) return rc;
  }
  return PSB_HIP_SUCCESS;
}

}  // extern "C"
