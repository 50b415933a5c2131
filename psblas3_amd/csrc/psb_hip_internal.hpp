// Internal declarations shared by the libpsb_hip translation units.
// Device code targets gfx950 (CDNA4): 64-lane wavefronts, DPP cross-lane
// moves, 160 KB LDS per CU, 8 XCDs each with its own L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "psb_hip.h"

namespace psbhip {

constexpr int kWave = 64;
constexpr int kBlock = 256;             // threads per block for streaming kernels
constexpr int kXcds = 8;                // MI355X: 8 XCDs, blocks dispatched round-robin
constexpr int kRedBlocks = 1024;        // max blocks of a grid reduction

// ------------------------------------------------------------------ errors
int set_error(int code, const char* where, const char* what);
#define PSB_HIP_CHECK(expr)                                                    \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::psbhip::set_error(e_ == hipErrorOutOfMemory ? PSB_HIP_OUTOFMEM  \
                                                           : PSB_HIP_RUNTIME,  \
                                 __func__, hipGetErrorString(e_));             \
  } while (0)
#define PSB_HIP_LAUNCH_CHECK() PSB_HIP_CHECK(hipGetLastError())

// ------------------------------------------------------------------ context
// One context per process (PSBLAS runs one process per device): the device,
// the stream every kernel is launched on, and the reduction workspace.
struct Context {
  int device = -1;
  bool initialized = false;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  // grid-reduction workspace: partials (double), completion counters and
  // result slots, all device memory; plus pinned host slots for read-back.
  double* red_partials = nullptr;       // kRedBlocks * 4 doubles
  unsigned int* red_counter = nullptr;  // 64 counters, kept at zero between uses
  double* red_result = nullptr;         // 64 result slots
  double* host_result = nullptr;        // pinned, 64 slots
  // the allocation-free solver's workspace: the device address of the
  // file-scope scratch array on this context's device (looked up at init),
  // a pinned host block for its state read-backs and two events
  char* scratch = nullptr;
  size_t scratch_bytes = 0;
  void* scratch_host = nullptr;         // pinned, kScratchHostBytes
  hipEvent_t scratch_ev[2] = {nullptr, nullptr};
};
// size of the device scratch array and of its pinned host companion
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr size_t kScratchHostBytes = 4096;
Context& ctx();
int ensure_init();                      // lazily initialises device 0
// early return with the error code when the library cannot initialise
#define PSB_ENSURE_INIT()          \
  do {                             \
    int rc_ = ::psbhip::ensure_init(); \
    if (rc_) return rc_;           \
  } while (0)

// ------------------------------------------------------------------ vectors
// Device multivector, the counterpart of cuda/vectordev.h:39-52.
struct MultiVec {
  int count;       // number of columns
  int size;        // rows per column
  int pitch;       // elements between columns
  int elem_type;   // PSB_HIP_TYPE_*
  void* v;         // device memory, pitch*count elements
};

inline size_t elem_size(int t) {
  switch (t) {
    case PSB_HIP_TYPE_INT: return 4;
    case PSB_HIP_TYPE_FLOAT: return 4;
    case PSB_HIP_TYPE_DOUBLE: return 8;
    default: return 0;
  }
}

// ------------------------------------------------------------------ launch helpers
inline int grid_for(int64_t n, int per_block) {
  int64_t g = (n + per_block - 1) / per_block;
  return (int)(g < 1 ? 1 : g);
}

// Grid of a streaming (grid-stride) kernel: enough waves to fill 256 CUs
// several times over without exceeding the work.
inline int stream_grid(int64_t n) {
  int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  int64_t cap = (int64_t)(ctx().num_cus > 0 ? ctx().num_cus : 256) * 8;
  if (g > cap) g = cap;
  return (int)(g < 1 ? 1 : g);
}

inline int red_grid(int64_t n) {
  int64_t g = (n + kBlock * 8 - 1) / (kBlock * 8);
  if (g > kRedBlocks) g = kRedBlocks;
  return (int)(g < 1 ? 1 : g);
}

// ------------------------------------------------------------------ device helpers
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) { __builtin_nontemporal_store(v, p); }

// DPP row rotate (within each 16-lane row) of a 32-bit value.  CTRL is the
// GFX9 dpp_ctrl encoding: row_ror:N = 0x120 + N.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  uint64_t u = __builtin_bit_cast(uint64_t, v);
  int lo = dpp_i32<CTRL>((int)(uint32_t)u);
  int hi = dpp_i32<CTRL>((int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __builtin_bit_cast(float, dpp_i32<CTRL>(__builtin_bit_cast(int, v)));
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp(T v) {
  if constexpr (sizeof(T) == 8) return dpp_f64<CTRL>(v);
  else return dpp_f32<CTRL>(v);
}

template <typename T>
__device__ __forceinline__ T readlane(T v, int lane) {
  if constexpr (sizeof(T) == 8) {
    uint64_t u = __builtin_bit_cast(uint64_t, v);
    uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
  }
}

// Sum over a full wavefront (all 64 lanes must be active).  Four DPP
// row_ror steps leave every lane of a 16-lane row holding the row sum; the
// four row sums are combined through readlane in a fixed order, so the
// result is wave-uniform and deterministic.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp<0x128>(v);  // row_ror:8
  v += dpp<0x124>(v);  // row_ror:4
  v += dpp<0x122>(v);  // row_ror:2
  v += dpp<0x121>(v);  // row_ror:1
  return (readlane(v, 0) + readlane(v, 16)) + (readlane(v, 32) + readlane(v, 48));
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  v = fmax(v, dpp<0x128>(v));
  v = fmax(v, dpp<0x124>(v));
  v = fmax(v, dpp<0x122>(v));
  v = fmax(v, dpp<0x121>(v));
  return fmax(fmax(readlane(v, 0), readlane(v, 16)), fmax(readlane(v, 32), readlane(v, 48)));
}

// Sum over aligned groups of G lanes (G = 2..16) inside a 16-lane row,
// result valid in every lane of the group.  Used by the vector-CSR kernel.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
  // xor butterfly inside the group (ds_swizzle / ds_bpermute lane moves)
  #pragma unroll
  for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
  return v;
}

// Block-wide sum over a block of NT threads (blockDim.x == NT, kBlock unless
// given); every thread gets the result.
template <typename T, int NT = kBlock>
__device__ __forceinline__ T block_sum(T v, T* lds /* NT/kWave */) {
  T w = wave_sum(v);
  int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[wid] = w;
  __syncthreads();
  T r = 0;
  #pragma unroll
  for (int k = 0; k < NT / kWave; k++) r += lds[k];
  __syncthreads();
  return r;
}

template <typename T>
__device__ __forceinline__ T block_max(T v, T* lds) {
  T w = wave_max(v);
  int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[wid] = w;
  __syncthreads();
  T r = lds[0];
  #pragma unroll
  for (int k = 1; k < kBlock / kWave; k++) r = fmax(r, lds[k]);
  __syncthreads();
  return r;
}

// Last-block-done grid reduction: every block stores its partial, the last
// block to arrive (device-scope counter) sums the partials in index order,
// stores the result and re-arms the counter.  Deterministic for a fixed
// grid size.  Returns true in the threads of the last block.
// NPART partial values per block (e.g. 2 for fused dot pairs).
template <int NPART>
__device__ __forceinline__ bool grid_reduce_sum(double (&v)[NPART], double* partials,
                                                unsigned int* counter, double* result) {
  __shared__ double lds[kBlock / kWave];
  __shared__ bool am_last;
  #pragma unroll
  for (int p = 0; p < NPART; p++) {
    double s = block_sum(v[p], lds);
    if (threadIdx.x == 0) partials[(size_t)p * kRedBlocks + blockIdx.x] = s;
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
  for (int p = 0; p < NPART; p++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
      s += __builtin_nontemporal_load(&partials[(size_t)p * kRedBlocks + b]);
    s = block_sum(s, lds);
    v[p] = s;
    if (threadIdx.x == 0) result[p] = s;
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

template <int NPART>
__device__ __forceinline__ bool grid_reduce_max(double (&v)[NPART], double* partials,
                                                unsigned int* counter, double* result) {
  __shared__ double lds[kBlock / kWave];
  __shared__ bool am_last;
  #pragma unroll
  for (int p = 0; p < NPART; p++) {
    double s = block_max(v[p], lds);
    if (threadIdx.x == 0) partials[(size_t)p * kRedBlocks + blockIdx.x] = s;
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
  for (int p = 0; p < NPART; p++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
      s = fmax(s, __builtin_nontemporal_load(&partials[(size_t)p * kRedBlocks + b]));
    s = block_max(s, lds);
    v[p] = s;
    if (threadIdx.x == 0) result[p] = s;
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

// XCD-aware block index: blocks are dispatched round-robin over the 8 XCDs;
// remap so that each XCD works on a contiguous range of the logical grid,
// keeping neighbouring rows (and the x entries they gather) in one L2.
// nlog = logical block count; the launched grid is ceil(nlog/8)*8.
__device__ __forceinline__ int xcd_block(int nlog) {
  int b = blockIdx.x;
  int per = (nlog + kXcds - 1) / kXcds;
  return (b % kXcds) * per + (b / kXcds);
}
inline int xcd_grid(int nlog) { return ((nlog + kXcds - 1) / kXcds) * kXcds; }

// Chunked XCD order: the logical blocks are cut into chunks of `chunk`
// blocks dealt round-robin to the XCDs (chunk c to XCD c % kXcds), each XCD
// walking its chunks in order; chunk <= 0 gives the contiguous order above.
// Launch xcd_chunk_grid(nlog, chunk) blocks; the mapping is a permutation,
// blocks mapped past nlog return.
__device__ __forceinline__ int xcd_block_chunk(int nlog, int chunk) {
  if (chunk <= 0) return xcd_block(nlog);
  const int b = blockIdx.x, j = b % kXcds, s = b / kXcds;
  return ((s / chunk) * kXcds + j) * chunk + (s % chunk);
}
inline int xcd_chunk_grid(int nlog, int chunk) {
  if (chunk <= 0) return xcd_grid(nlog);
  const int nc = (nlog + chunk - 1) / chunk;
  return ((nc + kXcds - 1) / kXcds) * kXcds * chunk;
}

// Block order of a product launch.  mode 0: launch order; 1: chunked XCD
// order (xcd_block_chunk, xchunk blocks per chunk); 2: plane-tile order for
// banded matrices whose couplings sit near multiples of a period of P rows
// (a 3-D stencil in natural order: P = L^2).  The row blocks are viewed as nz
// planes of pb blocks; tile columns of tb adjacent blocks are walked through
// every plane (block in the tile fastest, then plane), and whole tile
// columns are dealt to the XCDs in turn (column c to XCD c % kXcds), so the
// blocks that read the same x rows across planes run back to back on one
// XCD and find them in its L2.
struct BlockOrder {
  int mode = 0, xchunk = 0;
  int pb = 1, tb = 1, nz = 1, ntc = 1;
};

// Logical block of this workgroup under order o; nlog (past the matrix) for
// the workgroups the order leaves idle.
__device__ __forceinline__ int order_block(const BlockOrder& o, int nlog) {
  if (o.mode == 0) return (int)blockIdx.x;
  if (o.mode == 1) return xcd_block_chunk(nlog, o.xchunk);
  const int b = (int)blockIdx.x, j = b % kXcds, s = b / kXcds;
  const int per_col = o.nz * o.tb;
  const int k = s / per_col, rem = s - k * per_col;
  const int c = k * kXcds + j;
  const int z = rem / o.tb, sub = rem - z * o.tb;
  const int col = c * o.tb + sub;
  if (c >= o.ntc || col >= o.pb) return nlog;
  const int64_t pblk = (int64_t)z * o.pb + col;
  return pblk < nlog ? (int)pblk : nlog;
}

// Plane-tile order for nlog blocks of rows_per_block rows: planes of
// period_rows rows rounded to whole blocks, tiles of tile_blocks blocks.
inline BlockOrder plane_order(int nlog, int rows_per_block, int period_rows, int tile_blocks) {
  BlockOrder o;
  o.mode = 2;
  o.pb = (period_rows + rows_per_block / 2) / rows_per_block;
  if (o.pb < 1) o.pb = 1;
  o.tb = tile_blocks < 1 ? 1 : (tile_blocks > o.pb ? o.pb : tile_blocks);
  o.nz = (nlog + o.pb - 1) / o.pb;
  o.ntc = (o.pb + o.tb - 1) / o.tb;
  return o;
}

inline int order_grid(const BlockOrder& o, int nlog) {
  if (o.mode == 0) return nlog;
  if (o.mode == 1) return xcd_chunk_grid(nlog, o.xchunk);
  return ((o.ntc + kXcds - 1) / kXcds) * kXcds * o.nz * o.tb;
}

// ------------------------------------------------------------------ device matrices
struct HllDev {
  int elem_type;
  int hksz, rows, nzeros, allocsize, noffs, first_index;
  int max_col;          // largest column index stored (x must be this long)
  void* val;            // allocsize entries of elem_type
  int* ja;              // allocsize
  int* hkoffs;          // noffs + 1
  int* irn;             // rows
  int* idiag;           // rows
  float* val32;         // optional FP32 copy of a double matrix
  int variant;          // SpMV kernel variant chosen for this matrix (0: the global one)
  // Row-sorted storage (psb_hip_hll_copy_coo_sorted): position p of the hack
  // layout holds row perm[p] (1-based); irn and idiag are then per position.
  // The product runs work units (hack, first slot, end slot, partial slot or
  // -1): a hack wider than kHllUnitSlots is cut into several units whose
  // per-lane partial sums a second kernel adds in unit order (splits:
  // hack, first partial slot, unit count).  null perm: rows in order.
  int* perm = nullptr;
  int nunits = 0;
  int4* units = nullptr;
  int nsplit = 0;
  int4* splits = nullptr;
  void* part = nullptr;   // partial sums, (split units) x hksz entries of 8 bytes
  // plane-tile block order of the products (plane_order), 0: chunked XCD order
  int plane_rows = 0;
  int plane_tb = 0;
  // 16-bit column offsets (rows in order, built when every slot fits): the
  // 0-based column of slot (hack h, slot column j, lane l) is
  // h*hksz + l + base16[hkoffs[h]/hksz + j] + res16[hkoffs[h] + j*hksz + l],
  // i.e. ja - 1; base16 holds lane 0's offset of each slot column.  The
  // products read these 2-byte residuals instead of the 4-byte ja.
  short* res16 = nullptr;   // allocsize
  int* base16 = nullptr;    // allocsize / hksz
  int band = 0;             // max |column - row| over the stored slots (rows in order)
  // row-sorted matrices with scattered columns and a large x: the same rows
  // restricted to consecutive column ranges, each a row-sorted HLL matrix of
  // its own; the product runs them one after another accumulating into y
  int nslices = 0;
  HllDev** slices = nullptr;
};
// whether the products of m read its 16-bit column offsets (spmv.hip)
bool hll_use_idx16(const HllDev* m);
// slots of one work unit of the row-sorted HLL product
constexpr int kHllUnitSlots = 256;

struct EllDev {
  int elem_type;
  int rows, pitch, nzm, nnz, first_index;
  int max_col;
  void* val;            // pitch * nzm
  int* ja;
  int* irn;
  int* idiag;
  // 16-bit column offsets, as HllDev's over groups of 64 rows: the 0-based
  // column of slot j of row r is r + base16[(r/64)*nzm + j] + res16[j*pitch + r]
  // (rows past `rows` are never read and hold 0)
  short* res16 = nullptr;   // pitch * nzm
  int* base16 = nullptr;    // ceil(rows/64) * nzm
  int band = 0;             // max |column - row| over the stored slots of rows < rows
};
// XCD chunk (blocks) of a product over blocks of rows_per_block rows of a
// matrix of bandwidth band (spmv.hip)
int xcd_chunk_for(int band, int rows_per_block);
// whether the products of an ELL matrix read its 16-bit column offsets
bool ell_use_idx16(const EllDev* m);

struct CsrDev {
  int elem_type;
  int rows, cols, nnz, first_index;
  int max_col;
  int* irp;             // rows + 1, 1-based
  int* ja;
  void* val;
  int lanes;            // lanes per row of the uniform vector kernel
  int long_thr;         // rows longer than this go to the wave-per-row kernel
  int nlong;
  int* long_rows;       // 0-based row ids
  // length-binned kernel: rows grouped by lanes per row (1, 2, ..., 64), rows
  // longer than kCsrLongRow listed for the wave-per-row kernel
  int bin_start[8];     // class c: bin_rows[bin_start[c] .. bin_start[c+1])
  int* bin_rows;        // 0-based row ids, class after class, ascending
  int nlong_b;
  int* long_rows_b;
  // row-block kernel: blocks of consecutive rows (at most kBlock rows, their
  // entries fit the block's LDS), rows longer than kCsrLongRow excluded
  int nsblk;
  int2* sblk;           // (first row, one past the last row), 0-based
  // triangular solve prepared on first use (level analysis, see trsv.hip)
  void* sv = nullptr;
  // column slices: the same rows restricted to consecutive column ranges, each
  // a CSR matrix of its own, used by the product when x is large and the
  // columns scattered (the slices' x ranges stay in L2); null when unsliced
  int nslices = 0;
  CsrDev** slices = nullptr;
};

// The transpose of a matrix for y = alpha A^T x + beta y (spmv.hip), in hacks
// of 64 rows stored column-major like HLL: row c holds the entries (i, c) of
// A, rows i ascending, 0-based.
struct TransDev {
  int elem_type;
  int rows, cols, nnz;  // rows = columns of A, cols = rows of A
  int* hkoffs;          // hack offsets (slots), hacks + 1
  int* cnt;             // entries of each row
  int* src;             // the row i of A of each slot (0 in padding)
  void* val;            // each slot's value (0 in padding)
  // 16-bit source offsets, as HllDev's (built when every hack slot column's
  // real entries span at most 65535 rows of A): source row of slot (hack h,
  // slot column j, lane l) = h*64 + l + base16[hkoffs[h]/64 + j] + res16[slot],
  // clamped to [0, cols - 1] (padding slots hold 0; their product is dropped)
  short* res16 = nullptr;
  int* base16 = nullptr;
};

// frees the prepared triangular solve of a CSR matrix (trsv.hip)
void release_csr_sv(CsrDev* m);

// ------------------------------------------------------------------ inner products
// One group of G slots of a row (slot j at j*stride): acc += val[j] * x[ja[j]-1]
// in slot order.  Slots past the row's width w are read at slot w-1 (an
// in-bounds address) and their products dropped, so every index/value load of
// the group is unconditional and all of them issue before the first gather.
// (Masking each load instead makes the compiler wait for each index load in
// turn.)  Requires j < w.
// With xd / fd given (r the row, 0-based), the x entry gathered at column r
// -- the row's diagonal -- is kept in *xd and *fd set, for kernels that need
// x(r) beside the row's product (the fused CG's p.q) without reading x again.
template <int G, typename T, typename TV>
__device__ __forceinline__ T slot_group(const int* __restrict__ pj, const TV* __restrict__ pv,
                                        int stride, int j, int w, const T* __restrict__ x,
                                        T acc, int r = -1, T* xd = nullptr, int* fd = nullptr) {
  int c[G];
  TV v[G];
#pragma unroll
  for (int k = 0; k < G; k++) {
    const int jj = j + k < w ? j + k : w - 1;
    c[k] = ld_stream(pj + (size_t)jj * stride);
    v[k] = ld_stream(pv + (size_t)jj * stride);
  }
  T xv[G];
#pragma unroll
  for (int k = 0; k < G; k++) xv[k] = x[c[k] - 1];
#pragma unroll
  for (int k = 0; k < G; k++) {
    const T t = fma((T)v[k], xv[k], acc);
    acc = j + k < w ? t : acc;
    if (xd && c[k] - 1 == r) {
      *xd = xv[k];
      *fd = 1;
    }
  }
  return acc;
}

// sum_{j<w} val[j*stride] * x[ja[j*stride]-1], eight slots per group, the
// remainder in one group of 8 or 4.
template <typename T, typename TV>
__device__ __forceinline__ T slot_dot(const int* __restrict__ pj, const TV* __restrict__ pv,
                                      int stride, int w, const T* __restrict__ x, int r = -1,
                                      T* xd = nullptr, int* fd = nullptr) {
  T acc = T(0);
  int j = 0;
  for (; j + 8 <= w; j += 8) acc = slot_group<8, T, TV>(pj, pv, stride, j, w, x, acc, r, xd, fd);
  const int rem = w - j;
  if (rem > 4) acc = slot_group<8, T, TV>(pj, pv, stride, j, w, x, acc, r, xd, fd);
  else if (rem > 0) acc = slot_group<4, T, TV>(pj, pv, stride, j, w, x, acc, r, xd, fd);
  return acc;
}

template <typename T>
using vec2 = T __attribute__((ext_vector_type(2)));

// Offset and width of hack hk for a product with two adjacent rows per lane.
// With the hack size equal to the wavefront size a wavefront covers two whole
// hacks (the first lane's and the next; rows past nr clamp into the last
// hack), so the three offsets come through scalar loads.
template <int HK>
__device__ __forceinline__ void hll_pair_hack(const int* __restrict__ hkoffs, int nr, int hk,
                                              int stride, int& off0, int& w) {
  if (HK == kWave) {
    const int hka = __builtin_amdgcn_readfirstlane(hk);
    const int noffs = (nr + HK - 1) / HK;
    const int oa = hkoffs[hka], ob = hkoffs[hka + 1];
    const int oc = hka + 2 <= noffs ? hkoffs[hka + 2] : ob;
    const bool second = hk != hka;
    off0 = second ? ob : oa;
    w = ((second ? oc : ob) - off0) / stride;
  } else {
    off0 = hkoffs[hk];
    w = (hkoffs[hk + 1] - off0) / stride;
  }
}

// The same groups over 16-bit column offsets (HllDev::res16 / base16): pe
// and pv address the row's slot 0, pb the hack's slot-column bases, r is the
// row (0-based).  Column of slot j: r + pb[j] + pe[j*stride].
template <int G, typename T, typename TV>
__device__ __forceinline__ T slot_group16(const short* __restrict__ pe,
                                          const int* __restrict__ pb,
                                          const TV* __restrict__ pv, int stride, int j, int w,
                                          int r, const T* __restrict__ x, T acc,
                                          T* xd = nullptr, int* fd = nullptr) {
  int c[G];
  TV v[G];
#pragma unroll
  for (int k = 0; k < G; k++) {
    const int jj = j + k < w ? j + k : w - 1;
    c[k] = r + pb[jj] + (int)ld_stream(pe + (size_t)jj * stride);
    v[k] = ld_stream(pv + (size_t)jj * stride);
  }
  T xv[G];
#pragma unroll
  for (int k = 0; k < G; k++) xv[k] = x[c[k]];
#pragma unroll
  for (int k = 0; k < G; k++) {
    const T t = fma((T)v[k], xv[k], acc);
    acc = j + k < w ? t : acc;
    if (xd && c[k] == r) {
      *xd = xv[k];
      *fd = 1;
    }
  }
  return acc;
}

template <typename T, typename TV>
__device__ __forceinline__ T slot_dot16(const short* __restrict__ pe, const int* __restrict__ pb,
                                        const TV* __restrict__ pv, int stride, int w, int r,
                                        const T* __restrict__ x, T* xd = nullptr,
                                        int* fd = nullptr) {
  T acc = T(0);
  int j = 0;
  for (; j + 8 <= w; j += 8)
    acc = slot_group16<8, T, TV>(pe, pb, pv, stride, j, w, r, x, acc, xd, fd);
  const int rem = w - j;
  if (rem > 4) acc = slot_group16<8, T, TV>(pe, pb, pv, stride, j, w, r, x, acc, xd, fd);
  else if (rem > 0) acc = slot_group16<4, T, TV>(pe, pb, pv, stride, j, w, r, x, acc, xd, fd);
  return acc;
}

// x entries of columns c0 and c1 for the two rows of a lane: one 16-byte
// load when, on every active lane of the wavefront, c1 == c0 + 1 and c0 is
// even (x 16-byte aligned: xal), else two 8-byte loads.  Banded matrices
// take the single load on every slot whose column offset is even.  The
// choice is per wavefront (a per-lane choice runs both paths in wavefronts
// where only some lanes qualify).  Used over 16-bit offsets, i.e. on banded
// matrices.
template <typename T>
__device__ __forceinline__ void gather_pair(const T* __restrict__ x, bool xal, int c0, int c1,
                                            T& x0, T& x1) {
  if (__all(xal && c1 == c0 + 1 && (c0 & 1) == 0)) {
    const vec2<T> t = *reinterpret_cast<const vec2<T>*>(x + c0);
    x0 = t.x;
    x1 = t.y;
  } else {
    x0 = x[c0];
    x1 = x[c1];
  }
}

// Rows r (even) and r + 1 of one hack over 16-bit offsets: one 4-byte
// offset-pair load and one 2-element value load per slot.  xd / fd as in
// slot_group, two entries (rows r and r + 1).
template <int G, typename T, typename TV>
__device__ __forceinline__ void slot_group16_pair(const short* __restrict__ pe,
                                                  const int* __restrict__ pb,
                                                  const TV* __restrict__ pv, int stride, int j,
                                                  int w, int r, const T* __restrict__ x, T& a0,
                                                  T& a1, T* xd = nullptr, int* fd = nullptr) {
  int c0[G], c1[G];
  vec2<TV> v[G];
#pragma unroll
  for (int k = 0; k < G; k++) {
    const int jj = j + k < w ? j + k : w - 1;
    const int b = r + pb[jj];
    const vec2<short> e = ld_stream(reinterpret_cast<const vec2<short>*>(pe + (size_t)jj * stride));
    c0[k] = b + (int)e.x;
    c1[k] = b + 1 + (int)e.y;
    v[k] = ld_stream(reinterpret_cast<const vec2<TV>*>(pv + (size_t)jj * stride));
  }
  T x0[G], x1[G];
  const bool xal = (reinterpret_cast<uintptr_t>(x) & (2 * sizeof(T) - 1)) == 0;
#pragma unroll
  for (int k = 0; k < G; k++) gather_pair(x, xal, c0[k], c1[k], x0[k], x1[k]);
#pragma unroll
  for (int k = 0; k < G; k++) {
    const T t0 = fma((T)v[k].x, x0[k], a0);
    const T t1 = fma((T)v[k].y, x1[k], a1);
    a0 = j + k < w ? t0 : a0;
    a1 = j + k < w ? t1 : a1;
    if (xd) {
      if (c0[k] == r) {
        xd[0] = x0[k];
        fd[0] = 1;
      }
      if (c1[k] == r + 1) {
        xd[1] = x1[k];
        fd[1] = 1;
      }
    }
  }
}

template <typename T, typename TV>
__device__ __forceinline__ void slot_dot16_pair(const short* __restrict__ pe,
                                                const int* __restrict__ pb,
                                                const TV* __restrict__ pv, int stride, int w,
                                                int r, const T* __restrict__ x, T& a0, T& a1,
                                                T* xd = nullptr, int* fd = nullptr) {
  a0 = T(0);
  a1 = T(0);
  int j = 0;
  for (; j + 8 <= w; j += 8)
    slot_group16_pair<8, T, TV>(pe, pb, pv, stride, j, w, r, x, a0, a1, xd, fd);
  const int rem = w - j;
  if (rem > 4) slot_group16_pair<8, T, TV>(pe, pb, pv, stride, j, w, r, x, a0, a1, xd, fd);
  else if (rem > 0) slot_group16_pair<4, T, TV>(pe, pb, pv, stride, j, w, r, x, a0, a1, xd, fd);
}

// Two adjacent rows of one hack (same width w): one 8-byte index load and one
// 2-element value load per slot feed both rows.  pj/pv address the even row's
// slot 0 and are aligned to two elements.  Requires j < w.
template <int G, typename T, typename TV>
__device__ __forceinline__ void slot_group_pair(const int* __restrict__ pj,
                                                const TV* __restrict__ pv, int stride, int j,
                                                int w, const T* __restrict__ x, T& a0, T& a1,
                                                int r = -1, T* xd = nullptr, int* fd = nullptr) {
  vec2<int> c[G];
  vec2<TV> v[G];
#pragma unroll
  for (int k = 0; k < G; k++) {
    const int jj = j + k < w ? j + k : w - 1;
    c[k] = ld_stream(reinterpret_cast<const vec2<int>*>(pj + (size_t)jj * stride));
    v[k] = ld_stream(reinterpret_cast<const vec2<TV>*>(pv + (size_t)jj * stride));
  }
  // plain gathers: the 32-bit index path serves matrices with scattered
  // columns, where gather_pair's test rarely passes and only costs time
  // (rows-in-order HLL of a power-law matrix 8.9 -> 11.1 ms with it)
  T x0[G], x1[G];
#pragma unroll
  for (int k = 0; k < G; k++) {
    x0[k] = x[c[k].x - 1];
    x1[k] = x[c[k].y - 1];
  }
#pragma unroll
  for (int k = 0; k < G; k++) {
    const T t0 = fma((T)v[k].x, x0[k], a0);
    const T t1 = fma((T)v[k].y, x1[k], a1);
    a0 = j + k < w ? t0 : a0;
    a1 = j + k < w ? t1 : a1;
    if (xd) {
      if (c[k].x - 1 == r) {
        xd[0] = x0[k];
        fd[0] = 1;
      }
      if (c[k].y - 1 == r + 1) {
        xd[1] = x1[k];
        fd[1] = 1;
      }
    }
  }
}

template <typename T, typename TV>
__device__ __forceinline__ void slot_dot_pair(const int* __restrict__ pj,
                                              const TV* __restrict__ pv, int stride, int w,
                                              const T* __restrict__ x, T& a0, T& a1, int r = -1,
                                              T* xd = nullptr, int* fd = nullptr) {
  a0 = T(0);
  a1 = T(0);
  int j = 0;
  for (; j + 8 <= w; j += 8) slot_group_pair<8, T, TV>(pj, pv, stride, j, w, x, a0, a1, r, xd, fd);
  const int rem = w - j;
  if (rem > 4) slot_group_pair<8, T, TV>(pj, pv, stride, j, w, x, a0, a1, r, xd, fd);
  else if (rem > 0) slot_group_pair<4, T, TV>(pj, pv, stride, j, w, x, a0, a1, r, xd, fd);
}

// 16-byte vectors of T for the streaming vector kernels
template <typename T>
using v16 = T __attribute__((ext_vector_type(16 / sizeof(T))));
template <typename T>
__device__ __forceinline__ v16<T> ldv(const T* p) { return *reinterpret_cast<const v16<T>*>(p); }
template <typename T>
__device__ __forceinline__ void stv(T* p, v16<T> v) { *reinterpret_cast<v16<T>*>(p) = v; }
template <typename T>
__device__ __forceinline__ double vdot(v16<T> a, v16<T> b) {
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < (int)(16 / sizeof(T)); e++) s += (double)a[e] * (double)b[e];
  return s;
}

template <typename T>
__device__ __forceinline__ void store_y(T* __restrict__ y, int row, T alpha, T acc, T beta) {
  T r = alpha * acc;
  if (beta != T(0)) r = fma(beta, y[row], r);
  // y is written once per product: a non-temporal store leaves the L2 to x
  // (L=200 stencil: FP32 HLL 66.2 -> 63.3 us, ELL FP64 133 -> 130.5 us, FP64
  // pairs 120.5 -> 119.8 us)
  __builtin_nontemporal_store(r, y + row);
}

}  // namespace psbhip
