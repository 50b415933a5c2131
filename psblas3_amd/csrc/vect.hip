// Device multivectors and the BLAS-1 vector path.
// Replaces cuda/vectordev.c, cuda/dvectordev.c, cuda/svectordev.c,
// cuda/ivectordev.c and the SPGPU BLAS-1 kernels (cuda/spgpu/kernels/
// daxpby.cu, dupd_xyz.cu, dxyzw.cu, axy_base.cuh, ddot.cu, dnrm2.cu,
// amax_base.cuh, asum_base.cuh, abs_base.cuh, scal_base.cuh,
// setscal_base.cuh, gath_base.cuh, scat_base.cuh).
//
// Elementwise kernels evaluate exactly the expressions of the CPU routines
// (psi_daxpbyv, psi_d_upd_xyz, psi_dxyzw, d_base_mlt_a_2) with contraction
// off, so their results are bit-identical to the CPU path.  Reductions keep
// partial sums on the device (DPP wave sums, one last-block pass) and copy a
// single scalar back.
#include <math.h>
#include <string.h>

#include "psb_hip_internal.hpp"

#pragma clang fp contract(off)

namespace psbhip {

// ------------------------------------------------------------------ kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_setscal(T* __restrict__ x, int64_t first, int64_t n,
                                                    T val) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    x[first + i] = val;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scal(T* __restrict__ x, int64_t n, T alpha) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    x[i] = alpha * x[i];
}

// psi_daxpbyv: mode selects the branch of the CPU routine so that y is not
// read when beta == 0 and the single-term cases are a single product.
// z may alias y (axpby) or be a third vector (axpbyz).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpby(int n, int pitch_x, int pitch_y, int pitch_z,
                                                  T alpha, const T* __restrict__ x, T beta,
                                                  const T* y, T* z) {
  const int c = blockIdx.y;
  x += (int64_t)c * pitch_x;
  y += (int64_t)c * pitch_y;
  z += (int64_t)c * pitch_z;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    T r;
    if (alpha == T(0)) r = (beta == T(0)) ? T(0) : (beta == T(1) ? y[i] : beta * y[i]);
    else if (beta == T(0)) r = (alpha == T(1)) ? x[i] : alpha * x[i];
    else r = alpha * x[i] + beta * y[i];
    z[i] = r;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_upd_xyz(int n, T alpha, T beta, T gamma, T delta,
                                                    const T* __restrict__ x, T* __restrict__ y,
                                                    T* __restrict__ z) {
  const bool g = gamma != T(0) && !(alpha == T(0) && beta == T(0));
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    T yy;
    if (alpha != T(0) && beta != T(0)) yy = alpha * x[i] + beta * y[i];
    else if (alpha != T(0)) yy = alpha * x[i];
    else if (beta != T(0)) yy = beta * y[i];
    else yy = T(0);
    y[i] = yy;
    T zz;
    if (g && delta != T(0)) zz = gamma * yy + delta * z[i];
    else if (g) zz = gamma * yy;
    else if (delta != T(0)) zz = delta * z[i];
    else zz = T(0);
    z[i] = zz;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_xyzw(int n, T a, T b, T c, T d, T e, T f,
                                                 const T* __restrict__ x, T* __restrict__ y,
                                                 T* __restrict__ z, T* __restrict__ w) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    T yy = a * x[i] + b * y[i];
    T zz = c * yy + d * z[i];
    T ww = e * zz + f * w[i];
    y[i] = yy;
    z[i] = zz;
    w[i] = ww;
  }
}

// d_base_mlt_a_2: z = alpha*y*x + beta*z with the CPU's branch structure.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axybz(int n, int pitch, T alpha,
                                                  const T* __restrict__ x, const T* y, T beta,
                                                  T* z) {
  const int c = blockIdx.y;
  x += (int64_t)c * pitch;
  y += (int64_t)c * pitch;
  z += (int64_t)c * pitch;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if (alpha == T(0)) {
      if (beta != T(1)) z[i] = beta * z[i];
    } else if (alpha == T(1)) {
      if (beta == T(0)) z[i] = y[i] * x[i];
      else if (beta == T(1)) z[i] = z[i] + y[i] * x[i];
      else z[i] = beta * z[i] + y[i] * x[i];
    } else if (alpha == T(-1)) {
      if (beta == T(0)) z[i] = -y[i] * x[i];
      else if (beta == T(1)) z[i] = z[i] - y[i] * x[i];
      else z[i] = beta * z[i] - y[i] * x[i];
    } else {
      if (beta == T(0)) z[i] = alpha * y[i] * x[i];
      else if (beta == T(1)) z[i] = z[i] + alpha * y[i] * x[i];
      else z[i] = beta * z[i] + alpha * y[i] * x[i];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag_inv(int n, int size, T* d) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < size; i += gridDim.x * kBlock) {
    if (i < n) d[i] = (d[i] == T(0)) ? T(1) : T(1) / d[i];
    else d[i] = T(1);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_abs(int n, T alpha, const T* x, T* y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    y[i] = alpha * fabs(x[i]);
}

// Reductions.  op: 0 dot, 1 sum of squares, 2 max |x|, 3 sum |x|, 4 scaled
// sum of squares (x/scale)^2 with scale read from device memory.
template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_reduce(int n, const T* __restrict__ x,
                                                   const T* __restrict__ y,
                                                   const double* __restrict__ scale_p,
                                                   double* partials, unsigned int* counter,
                                                   double* result) {
  double acc[1] = {0.0};
  double inv = 1.0;
  if (OP == 4) inv = 1.0 / *scale_p;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double a = (double)x[i];
    if (OP == 0) acc[0] += a * (double)y[i];
    else if (OP == 1) acc[0] += a * a;
    else if (OP == 2) acc[0] = fmax(acc[0], fabs(a));
    else if (OP == 3) acc[0] += fabs(a);
    else { double t = a * inv; acc[0] += t * t; }
  }
  if (OP == 2) grid_reduce_max<1>(acc, partials, counter, result);
  else grid_reduce_sum<1>(acc, partials, counter, result);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gath(int n, const int* __restrict__ idx, int base,
                                                 const T* __restrict__ x, T* __restrict__ out) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    out[i] = x[idx[i] - base];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scat(int n, const int* __restrict__ idx, int base,
                                                 const T* __restrict__ in, T beta, T* x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    int j = idx[i] - base;
    x[j] = (beta == T(0)) ? in[i] : beta * x[j] + in[i];
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ins_add(int n, const int* __restrict__ idx, int base,
                                                    const T* __restrict__ in, T* x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    atomicAdd(&x[idx[i] - base], in[i]);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_fill_uniform(double* x, int64_t n, uint64_t key) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    uint64_t h = mix64(key ^ (uint64_t)i);
    x[i] = 2.0 * ((double)(h >> 11) * 0x1.0p-53) - 1.0;
  }
}

// ------------------------------------------------------------------ host helpers
static inline MultiVec* MV(void* p) { return static_cast<MultiVec*>(p); }

static int check_pair(const MultiVec* a, const MultiVec* b, int n) {
  if (!a || !b) return PSB_HIP_INVALID;
  if (n > a->size || n > b->size) return PSB_HIP_UNSUPPORTED;
  if (a->elem_type != b->elem_type) return PSB_HIP_INVALID;
  return PSB_HIP_SUCCESS;
}

// Run one grid reduction over column c and leave the scalar in
// red_result[slot]; returns the grid size used.
template <typename T, int OP>
static int launch_reduce(int n, const T* x, const T* y, const double* scale_p, int slot) {
  Context& c = ctx();
  int g = red_grid(n);
  hipLaunchKernelGGL((k_reduce<T, OP>), dim3(g), dim3(kBlock), 0, c.stream, n, x, y, scale_p,
                     c.red_partials, c.red_counter + slot, c.red_result + slot);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

static int fetch_results(int nslots, double* out) {
  Context& c = ctx();
  PSB_HIP_CHECK(hipMemcpyAsync(c.host_result, c.red_result, sizeof(double) * nslots,
                               hipMemcpyDeviceToHost, c.stream));
  PSB_HIP_CHECK(hipStreamSynchronize(c.stream));
  memcpy(out, c.host_result, sizeof(double) * nslots);
  return PSB_HIP_SUCCESS;
}

template <typename T>
static int nrm2_col(int n, const T* x, double* res) {
  int rc = launch_reduce<T, 1>(n, x, x, nullptr, 0);
  if (rc) return rc;
  double ss;
  rc = fetch_results(1, &ss);
  if (rc) return rc;
  // unscaled sum of squares is exact enough unless it over/underflowed;
  // then rescale by max|x| (the scaled dnrm2 of the CPU path)
  if (isfinite(ss) && (ss == 0.0 || ss > 1e-280)) {
    *res = sqrt(ss);
    return PSB_HIP_SUCCESS;
  }
  rc = launch_reduce<T, 2>(n, x, x, nullptr, 1);
  if (rc) return rc;
  rc = launch_reduce<T, 4>(n, x, x, ctx().red_result + 1, 2);
  if (rc) return rc;
  double r[3];
  rc = fetch_results(3, r);
  if (rc) return rc;
  *res = r[1] == 0.0 ? 0.0 : r[1] * sqrt(r[2]);
  return PSB_HIP_SUCCESS;
}

template <typename T>
static int setscal_impl(T val, int first, int last, int index_base, void* x) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(x);
  int64_t f = first - index_base, n = (int64_t)last - first + 1;
  if (n <= 0) return PSB_HIP_SUCCESS;
  if (f < 0 || f + n > (int64_t)v->pitch * v->count) return PSB_HIP_UNSUPPORTED;
  hipLaunchKernelGGL((k_setscal<T>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream,
                     (T*)v->v, f, n, val);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

template <int OP>
static int simple_reduce(double* res, int n, void* x) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(x);
  if (n > v->size) return PSB_HIP_UNSUPPORTED;
  for (int c = 0; c < v->count; c++) {
    if (n <= 0) { res[c] = 0.0; continue; }
    int rc;
    if (v->elem_type == PSB_HIP_TYPE_DOUBLE) {
      const double* p = (const double*)v->v + (int64_t)c * v->pitch;
      rc = launch_reduce<double, OP>(n, p, p, nullptr, 0);
    } else if (v->elem_type == PSB_HIP_TYPE_FLOAT) {
      const float* p = (const float*)v->v + (int64_t)c * v->pitch;
      rc = launch_reduce<float, OP>(n, p, p, nullptr, 0);
    } else {
      return PSB_HIP_UNSUPPORTED;
    }
    if (rc) return rc;
    rc = fetch_results(1, &res[c]);
    if (rc) return rc;
  }
  return PSB_HIP_SUCCESS;
}

}  // namespace psbhip

using namespace psbhip;


extern "C" {

int psb_hip_mv_alloc(void** mv, int count, int size, int elem_type) {
  PSB_ENSURE_INIT();
  *mv = nullptr;
  size_t es = elem_size(elem_type);
  if (es == 0 || count < 1 || size < 0) return set_error(PSB_HIP_UNSUPPORTED, __func__, "bad args");
  MultiVec* v = new MultiVec;
  v->count = count;
  v->size = size;
  v->elem_type = elem_type;
  // cuda/vectordev.c:83-145: pitch = size for one column, else 256-byte rows
  v->pitch = (count == 1) ? size : (int)((((size_t)size * es + 255) / 256) * 256 / es);
  size_t bytes = (size_t)v->pitch * count * es;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(&v->v, bytes);
  if (e != hipSuccess) {
    delete v;
    return set_error(PSB_HIP_OUTOFMEM, __func__, hipGetErrorString(e));
  }
  *mv = v;
  return PSB_HIP_SUCCESS;
}

void psb_hip_mv_free(void* mv) {
  if (!mv) return;
  MultiVec* v = MV(mv);
  if (ctx().initialized) hipStreamSynchronize(ctx().stream);
  hipFree(v->v);
  delete v;
}

int psb_hip_mv_size(void* mv) { return mv ? MV(mv)->size : -1; }
int psb_hip_mv_count(void* mv) { return mv ? MV(mv)->count : -1; }
int psb_hip_mv_pitch(void* mv) { return mv ? MV(mv)->pitch : -1; }
void* psb_hip_mv_data(void* mv) { return mv ? MV(mv)->v : nullptr; }

int psb_hip_mv_write_r2(void* mv, const void* host, int ld) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(mv);
  size_t es = elem_size(v->elem_type);
  if (v->size == 0) return PSB_HIP_SUCCESS;
  PSB_HIP_CHECK(hipMemcpy2DAsync(v->v, (size_t)v->pitch * es, host, (size_t)ld * es,
                                 (size_t)v->size * es, (size_t)v->count, hipMemcpyHostToDevice,
                                 ctx().stream));
  PSB_HIP_CHECK(hipStreamSynchronize(ctx().stream));
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_read_r2(void* mv, void* host, int ld) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(mv);
  size_t es = elem_size(v->elem_type);
  if (v->size == 0) return PSB_HIP_SUCCESS;
  PSB_HIP_CHECK(hipMemcpy2DAsync(host, (size_t)ld * es, v->v, (size_t)v->pitch * es,
                                 (size_t)v->size * es, (size_t)v->count, hipMemcpyDeviceToHost,
                                 ctx().stream));
  PSB_HIP_CHECK(hipStreamSynchronize(ctx().stream));
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_write(void* mv, const void* host) { return psb_hip_mv_write_r2(mv, host, MV(mv)->size); }
int psb_hip_mv_read(void* mv, void* host) { return psb_hip_mv_read_r2(mv, host, MV(mv)->size); }

int psb_hip_mv_write_first(void* mv, int first, int n, const void* host) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(mv);
  if (first < 1 || n < 0 || first - 1 + n > v->size) return PSB_HIP_UNSUPPORTED;
  size_t es = elem_size(v->elem_type);
  if (n == 0) return PSB_HIP_SUCCESS;
  PSB_HIP_CHECK(hipMemcpyAsync((char*)v->v + (size_t)(first - 1) * es, host, (size_t)n * es,
                               hipMemcpyHostToDevice, ctx().stream));
  PSB_HIP_CHECK(hipStreamSynchronize(ctx().stream));
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_read_first(void* mv, int first, int n, void* host) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(mv);
  if (first < 1 || n < 0 || first - 1 + n > v->size) return PSB_HIP_UNSUPPORTED;
  size_t es = elem_size(v->elem_type);
  if (n == 0) return PSB_HIP_SUCCESS;
  PSB_HIP_CHECK(hipMemcpyAsync(host, (char*)v->v + (size_t)(first - 1) * es, (size_t)n * es,
                               hipMemcpyDeviceToHost, ctx().stream));
  PSB_HIP_CHECK(hipStreamSynchronize(ctx().stream));
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_setscal_d(double val, int first, int last, int index_base, void* x) {
  if (MV(x)->elem_type != PSB_HIP_TYPE_DOUBLE) return PSB_HIP_INVALID;
  return setscal_impl<double>(val, first, last, index_base, x);
}
int psb_hip_mv_setscal_s(float val, int first, int last, int index_base, void* x) {
  if (MV(x)->elem_type != PSB_HIP_TYPE_FLOAT) return PSB_HIP_INVALID;
  return setscal_impl<float>(val, first, last, index_base, x);
}
int psb_hip_mv_setscal_i(int val, int first, int last, int index_base, void* x) {
  if (MV(x)->elem_type != PSB_HIP_TYPE_INT) return PSB_HIP_INVALID;
  return setscal_impl<int>(val, first, last, index_base, x);
}

int psb_hip_mv_geins(int n, void* irl, void* val, int dupl, int index_base, void* x) {
  PSB_ENSURE_INIT();
  MultiVec *vi = MV(irl), *vv = MV(val), *vx = MV(x);
  if (n > vi->size || n > vv->size) return PSB_HIP_UNSUPPORTED;
  if (vi->elem_type != PSB_HIP_TYPE_INT || vv->elem_type != vx->elem_type) return PSB_HIP_INVALID;
  if (n <= 0) return PSB_HIP_SUCCESS;
  int g = stream_grid(n);
  if (vx->elem_type == PSB_HIP_TYPE_DOUBLE) {
    if (dupl == PSB_HIP_INS_ADD)
      hipLaunchKernelGGL((k_ins_add<double>), dim3(g), dim3(kBlock), 0, ctx().stream, n,
                         (const int*)vi->v, index_base, (const double*)vv->v, (double*)vx->v);
    else
      hipLaunchKernelGGL((k_scat<double>), dim3(g), dim3(kBlock), 0, ctx().stream, n,
                         (const int*)vi->v, index_base, (const double*)vv->v, 0.0,
                         (double*)vx->v);
  } else if (vx->elem_type == PSB_HIP_TYPE_FLOAT) {
    if (dupl == PSB_HIP_INS_ADD)
      hipLaunchKernelGGL((k_ins_add<float>), dim3(g), dim3(kBlock), 0, ctx().stream, n,
                         (const int*)vi->v, index_base, (const float*)vv->v, (float*)vx->v);
    else
      hipLaunchKernelGGL((k_scat<float>), dim3(g), dim3(kBlock), 0, ctx().stream, n,
                         (const int*)vi->v, index_base, (const float*)vv->v, 0.0f,
                         (float*)vx->v);
  } else {
    return PSB_HIP_UNSUPPORTED;
  }
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_igath(void* x, int vector_id, int n, int first, const int* didx, int hfirst,
                     void* dhv, int index_base) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(x);
  if (n <= 0) return PSB_HIP_SUCCESS;
  if (vector_id < 0 || vector_id >= v->count) return PSB_HIP_UNSUPPORTED;
  int g = stream_grid(n);
  const int* idx = didx + (first - index_base);
  if (v->elem_type == PSB_HIP_TYPE_DOUBLE) {
    double* hv = (double*)dhv + (hfirst - index_base);
    hipLaunchKernelGGL((k_gath<double>), dim3(g), dim3(kBlock), 0, ctx().stream, n, idx,
                       index_base, (const double*)v->v + (int64_t)vector_id * v->pitch, hv);
  } else if (v->elem_type == PSB_HIP_TYPE_FLOAT) {
    float* hv = (float*)dhv + (hfirst - index_base);
    hipLaunchKernelGGL((k_gath<float>), dim3(g), dim3(kBlock), 0, ctx().stream, n, idx,
                       index_base, (const float*)v->v + (int64_t)vector_id * v->pitch, hv);
  } else {
    int* hv = (int*)dhv + (hfirst - index_base);
    hipLaunchKernelGGL((k_gath<int>), dim3(g), dim3(kBlock), 0, ctx().stream, n, idx,
                       index_base, (const int*)v->v + (int64_t)vector_id * v->pitch, hv);
  }
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_iscat(void* x, int vector_id, int n, int first, const int* didx, int hfirst,
                     const void* dhv, int index_base, double beta) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(x);
  if (n <= 0) return PSB_HIP_SUCCESS;
  if (vector_id < 0 || vector_id >= v->count) return PSB_HIP_UNSUPPORTED;
  int g = stream_grid(n);
  const int* idx = didx + (first - index_base);
  if (v->elem_type == PSB_HIP_TYPE_DOUBLE) {
    const double* hv = (const double*)dhv + (hfirst - index_base);
    hipLaunchKernelGGL((k_scat<double>), dim3(g), dim3(kBlock), 0, ctx().stream, n, idx,
                       index_base, hv, beta, (double*)v->v + (int64_t)vector_id * v->pitch);
  } else if (v->elem_type == PSB_HIP_TYPE_FLOAT) {
    const float* hv = (const float*)dhv + (hfirst - index_base);
    hipLaunchKernelGGL((k_scat<float>), dim3(g), dim3(kBlock), 0, ctx().stream, n, idx,
                       index_base, hv, (float)beta, (float*)v->v + (int64_t)vector_id * v->pitch);
  } else {
    const int* hv = (const int*)dhv + (hfirst - index_base);
    hipLaunchKernelGGL((k_scat<int>), dim3(g), dim3(kBlock), 0, ctx().stream, n, idx,
                       index_base, hv, (int)beta, (int*)v->v + (int64_t)vector_id * v->pitch);
  }
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_scal(double alpha, void* x) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(x);
  int64_t n = (int64_t)v->pitch * v->count;
  if (n <= 0) return PSB_HIP_SUCCESS;
  if (v->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_scal<double>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream,
                       (double*)v->v, n, alpha);
  else if (v->elem_type == PSB_HIP_TYPE_FLOAT)
    hipLaunchKernelGGL((k_scal<float>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream,
                       (float*)v->v, n, (float)alpha);
  else
    return PSB_HIP_UNSUPPORTED;
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_dot(double* res, int n, void* x, void* y) {
  PSB_ENSURE_INIT();
  MultiVec *vx = MV(x), *vy = MV(y);
  int rc = check_pair(vx, vy, n);
  if (rc) return rc;
  int cols = vx->count < vy->count ? vx->count : vy->count;
  for (int c = 0; c < cols; c++) {
    if (n <= 0) { res[c] = 0.0; continue; }
    if (vx->elem_type == PSB_HIP_TYPE_DOUBLE)
      rc = launch_reduce<double, 0>(n, (const double*)vx->v + (int64_t)c * vx->pitch,
                                    (const double*)vy->v + (int64_t)c * vy->pitch, nullptr, 0);
    else if (vx->elem_type == PSB_HIP_TYPE_FLOAT)
      rc = launch_reduce<float, 0>(n, (const float*)vx->v + (int64_t)c * vx->pitch,
                                   (const float*)vy->v + (int64_t)c * vy->pitch, nullptr, 0);
    else
      return PSB_HIP_UNSUPPORTED;
    if (rc) return rc;
    rc = fetch_results(1, &res[c]);
    if (rc) return rc;
  }
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_nrm2(double* res, int n, void* x) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(x);
  if (n > v->size) return PSB_HIP_UNSUPPORTED;
  for (int c = 0; c < v->count; c++) {
    if (n <= 0) { res[c] = 0.0; continue; }
    int rc;
    if (v->elem_type == PSB_HIP_TYPE_DOUBLE)
      rc = nrm2_col<double>(n, (const double*)v->v + (int64_t)c * v->pitch, &res[c]);
    else if (v->elem_type == PSB_HIP_TYPE_FLOAT)
      rc = nrm2_col<float>(n, (const float*)v->v + (int64_t)c * v->pitch, &res[c]);
    else
      return PSB_HIP_UNSUPPORTED;
    if (rc) return rc;
  }
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_amax(double* res, int n, void* x) { return simple_reduce<2>(res, n, x); }
int psb_hip_mv_asum(double* res, int n, void* x) { return simple_reduce<3>(res, n, x); }

int psb_hip_mv_axpbyz(int n, double alpha, void* x, double beta, void* y, void* z) {
  PSB_ENSURE_INIT();
  MultiVec *vx = MV(x), *vy = MV(y), *vz = MV(z);
  int rc = check_pair(vx, vy, n);
  if (rc) return rc;
  rc = check_pair(vy, vz, n);
  if (rc) return rc;
  if (n <= 0) return PSB_HIP_SUCCESS;
  int cols = vz->count;
  if (vx->count < cols || vy->count < cols) return PSB_HIP_UNSUPPORTED;
  dim3 grid(stream_grid(n), cols);
  if (vx->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_axpby<double>), grid, dim3(kBlock), 0, ctx().stream, n, vx->pitch,
                       vy->pitch, vz->pitch, alpha, (const double*)vx->v, beta,
                       (const double*)vy->v, (double*)vz->v);
  else if (vx->elem_type == PSB_HIP_TYPE_FLOAT)
    hipLaunchKernelGGL((k_axpby<float>), grid, dim3(kBlock), 0, ctx().stream, n, vx->pitch,
                       vy->pitch, vz->pitch, (float)alpha, (const float*)vx->v, (float)beta,
                       (const float*)vy->v, (float*)vz->v);
  else
    return PSB_HIP_UNSUPPORTED;
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_axpby(int n, double alpha, void* x, double beta, void* y) {
  return psb_hip_mv_axpbyz(n, alpha, x, beta, y, y);
}

int psb_hip_mv_upd_xyz(int n, double alpha, double beta, double gamma, double delta, void* x,
                       void* y, void* z) {
  PSB_ENSURE_INIT();
  MultiVec *vx = MV(x), *vy = MV(y), *vz = MV(z);
  int rc = check_pair(vx, vy, n);
  if (rc) return rc;
  rc = check_pair(vy, vz, n);
  if (rc) return rc;
  if (n <= 0) return PSB_HIP_SUCCESS;
  if (vx->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_upd_xyz<double>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream,
                       n, alpha, beta, gamma, delta, (const double*)vx->v, (double*)vy->v,
                       (double*)vz->v);
  else if (vx->elem_type == PSB_HIP_TYPE_FLOAT)
    hipLaunchKernelGGL((k_upd_xyz<float>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream,
                       n, (float)alpha, (float)beta, (float)gamma, (float)delta,
                       (const float*)vx->v, (float*)vy->v, (float*)vz->v);
  else
    return PSB_HIP_UNSUPPORTED;
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_xyzw(int n, double a, double b, double c, double d, double e, double f, void* x,
                    void* y, void* z, void* w) {
  PSB_ENSURE_INIT();
  MultiVec *vx = MV(x), *vy = MV(y), *vz = MV(z), *vw = MV(w);
  int rc = check_pair(vx, vy, n);
  if (rc) return rc;
  rc = check_pair(vz, vw, n);
  if (rc) return rc;
  // psi_dxyzw: all six coefficients must be non-zero, else nothing is done
  if (a == 0.0 || b == 0.0 || c == 0.0 || d == 0.0 || e == 0.0 || f == 0.0)
    return PSB_HIP_SUCCESS;
  if (n <= 0) return PSB_HIP_SUCCESS;
  if (vx->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_xyzw<double>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream, n,
                       a, b, c, d, e, f, (const double*)vx->v, (double*)vy->v, (double*)vz->v,
                       (double*)vw->v);
  else if (vx->elem_type == PSB_HIP_TYPE_FLOAT)
    hipLaunchKernelGGL((k_xyzw<float>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream, n,
                       (float)a, (float)b, (float)c, (float)d, (float)e, (float)f,
                       (const float*)vx->v, (float*)vy->v, (float*)vz->v, (float*)vw->v);
  else
    return PSB_HIP_UNSUPPORTED;
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_axybz(int n, double alpha, void* x, void* y, double beta, void* z) {
  PSB_ENSURE_INIT();
  MultiVec *vx = MV(x), *vy = MV(y), *vz = MV(z);
  int rc = check_pair(vx, vy, n);
  if (rc) return rc;
  rc = check_pair(vy, vz, n);
  if (rc) return rc;
  if (n <= 0) return PSB_HIP_SUCCESS;
  // one column of x (the diagonal) against every column of y and z
  int cols = vz->count;
  if (vx->count != 1 && vx->count != cols) return PSB_HIP_UNSUPPORTED;
  if (vy->count < cols || vy->pitch != vz->pitch) return PSB_HIP_UNSUPPORTED;
  for (int c = 0; c < cols; c++) {
    int64_t ox = (vx->count == 1) ? 0 : (int64_t)c * vx->pitch;
    int64_t oy = (int64_t)c * vy->pitch;
    if (vx->elem_type == PSB_HIP_TYPE_DOUBLE)
      hipLaunchKernelGGL((k_axybz<double>), dim3(stream_grid(n), 1), dim3(kBlock), 0,
                         ctx().stream, n, 0, alpha, (const double*)vx->v + ox,
                         (const double*)vy->v + oy, beta, (double*)vz->v + oy);
    else if (vx->elem_type == PSB_HIP_TYPE_FLOAT)
      hipLaunchKernelGGL((k_axybz<float>), dim3(stream_grid(n), 1), dim3(kBlock), 0,
                         ctx().stream, n, 0, (float)alpha, (const float*)vx->v + ox,
                         (const float*)vy->v + oy, (float)beta, (float*)vz->v + oy);
    else
      return PSB_HIP_UNSUPPORTED;
    PSB_HIP_LAUNCH_CHECK();
  }
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_axy(int n, double alpha, void* x, void* y) {
  return psb_hip_mv_axybz(n, alpha, x, y, 0.0, y);
}

int psb_hip_mv_diag_inv(int n, void* dv) {
  PSB_ENSURE_INIT();
  MultiVec* d = MV(dv);
  if (n > d->size) return PSB_HIP_UNSUPPORTED;
  if (d->size <= 0) return PSB_HIP_SUCCESS;
  if (d->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_diag_inv<double>), dim3(stream_grid(d->size)), dim3(kBlock), 0,
                       ctx().stream, n, d->size, (double*)d->v);
  else if (d->elem_type == PSB_HIP_TYPE_FLOAT)
    hipLaunchKernelGGL((k_diag_inv<float>), dim3(stream_grid(d->size)), dim3(kBlock), 0,
                       ctx().stream, n, d->size, (float*)d->v);
  else
    return PSB_HIP_UNSUPPORTED;
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_abs(int n, double alpha, void* x, void* y) {
  PSB_ENSURE_INIT();
  MultiVec *vx = MV(x), *vy = MV(y);
  int rc = check_pair(vx, vy, n);
  if (rc) return rc;
  if (n <= 0) return PSB_HIP_SUCCESS;
  if (vx->elem_type == PSB_HIP_TYPE_DOUBLE)
    hipLaunchKernelGGL((k_abs<double>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream, n,
                       alpha, (const double*)vx->v, (double*)vy->v);
  else if (vx->elem_type == PSB_HIP_TYPE_FLOAT)
    hipLaunchKernelGGL((k_abs<float>), dim3(stream_grid(n)), dim3(kBlock), 0, ctx().stream, n,
                       (float)alpha, (const float*)vx->v, (float*)vy->v);
  else
    return PSB_HIP_UNSUPPORTED;
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_mv_fill_uniform(void* x, uint64_t seed, uint32_t stream) {
  PSB_ENSURE_INIT();
  MultiVec* v = MV(x);
  if (v->elem_type != PSB_HIP_TYPE_DOUBLE) return PSB_HIP_UNSUPPORTED;
  if (v->size <= 0) return PSB_HIP_SUCCESS;
  for (int c = 0; c < v->count; c++) {
    // same key derivation as the host fill (host.cpp); column c uses stream + c
    uint64_t z = seed * 0xD1B54A32D192ED03ULL ^ ((uint64_t)(stream + (uint32_t)c) << 48);
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    const uint64_t key = z ^ (z >> 31);
    hipLaunchKernelGGL(k_fill_uniform, dim3(stream_grid(v->size)), dim3(kBlock), 0, ctx().stream,
                       (double*)v->v + (int64_t)c * v->pitch, (int64_t)v->size, key);
    PSB_HIP_LAUNCH_CHECK();
  }
  return PSB_HIP_SUCCESS;
}

}  // extern "C"
