// BJAC / ILU(0) preconditioner application on the device:
// y = alpha * U^-1 (d .* L^-1 x) + beta * y with L unit lower, U unit upper
// (psb_d_bjac_apply_vect, prec/impl/psb_d_bjacprec_impl.f90:166-173: psb_spsm
// with scale 'L' and the inverse pivots, then psb_spsm on U).
//
// Each triangle is solved by levels: the host analysis gives every row the
// length of its longest dependency chain, and one kernel per level solves all
// rows of that level, one thread per row.  A row's sum runs in its CSR order,
// y(i) = x(i) - sum_j val(j) y(ja(j)) with separate multiply and add, as the
// reference's inner_csrsv (base/serial/impl/psb_d_csr_impl.F90:890-941), so
// the result equals the sequential CPU solve bit for bit.  The per-level
// launches of both triangles (a stencil in natural order has about 3L
// levels) are captured once into a graph that works on internal buffers.
#pragma clang fp contract(off)

#include <string.h>

#include <algorithm>
#include <vector>

#include "psb_hip_internal.hpp"

namespace psbhip {

struct TriLevels {
  int n = 0, nnz = 0, nlev = 0;
  int* irp = nullptr;
  int* ja = nullptr;
  double* val = nullptr;
  int* rows = nullptr;          // rows grouped by level
  std::vector<int> lev_ptr;     // host: level l = rows[lev_ptr[l] .. lev_ptr[l+1])
};

struct IluDev {
  int n = 0;
  TriLevels L, U;
  double* d = nullptr;          // inverse pivots
  double* tin = nullptr;        // input copy
  double* t1 = nullptr;         // d .* L^-1 tin
  double* t2 = nullptr;         // U^-1 t1
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

__global__ __launch_bounds__(kBlock) void k_trsv_level(int count, const int* __restrict__ rows,
                                                       const int* __restrict__ irp,
                                                       const int* __restrict__ ja,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ x,
                                                       double* y) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int i = rows[t];
  double acc = 0.0;
  for (int p = irp[i] - 1; p < irp[i + 1] - 1; p++) acc = acc + val[p] * y[ja[p] - 1];
  y[i] = x[i] - acc;
}

__global__ __launch_bounds__(kBlock) void k_scale_by(int n, double* __restrict__ y,
                                                     const double* __restrict__ d) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = y[i] * d[i];
}

// Level of every row (longest chain of dependencies through the triangle),
// rows grouped by level in ascending row order within a level.
static void analyse(int n, bool lower, const int* irp, const int* ja, std::vector<int>& rows,
                    std::vector<int>& lev_ptr) {
  std::vector<int> lev(n, 0);
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

static int upload_tri(TriLevels& T, int n, bool lower, const int* irp, const int* ja,
                      const double* val) {
  T.n = n;
  T.nnz = irp[n] - 1;
  for (int i = 0; i < n; i++)
    for (int p = irp[i] - 1; p < irp[i + 1] - 1; p++) {
      const int c = ja[p];
      if (c < 1 || c > n || (lower ? c > i : c < i + 2))
        return set_error(PSB_HIP_INVALID, "ilu", "factor entry outside its triangle");
    }
  std::vector<int> rows;
  analyse(n, lower, irp, ja, rows, T.lev_ptr);
  T.nlev = (int)T.lev_ptr.size() - 1;
  const size_t nzb = (size_t)(T.nnz > 0 ? T.nnz : 1);
  PSB_HIP_CHECK(hipMalloc(&T.irp, sizeof(int) * (size_t)(n + 1)));
  PSB_HIP_CHECK(hipMalloc(&T.ja, sizeof(int) * nzb));
  PSB_HIP_CHECK(hipMalloc(&T.val, sizeof(double) * nzb));
  PSB_HIP_CHECK(hipMalloc(&T.rows, sizeof(int) * (size_t)(n > 0 ? n : 1)));
  hipStream_t s = ctx().stream;
  PSB_HIP_CHECK(hipMemcpyAsync(T.irp, irp, sizeof(int) * (size_t)(n + 1), hipMemcpyHostToDevice, s));
  if (T.nnz > 0) {
    PSB_HIP_CHECK(hipMemcpyAsync(T.ja, ja, sizeof(int) * (size_t)T.nnz, hipMemcpyHostToDevice, s));
    PSB_HIP_CHECK(hipMemcpyAsync(T.val, val, sizeof(double) * (size_t)T.nnz, hipMemcpyHostToDevice,
                                 s));
  }
  if (n > 0)
    PSB_HIP_CHECK(hipMemcpyAsync(T.rows, rows.data(), sizeof(int) * (size_t)n,
                                 hipMemcpyHostToDevice, s));
  PSB_HIP_CHECK(hipStreamSynchronize(s));   // host arrays are the caller's
  return PSB_HIP_SUCCESS;
}

static void enqueue_levels(const TriLevels& T, const double* x, double* y, hipStream_t s) {
  for (int l = 0; l < T.nlev; l++) {
    const int count = T.lev_ptr[l + 1] - T.lev_ptr[l];
    hipLaunchKernelGGL(k_trsv_level, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       count, T.rows + T.lev_ptr[l], T.irp, T.ja, T.val, x, y);
  }
}

static void free_tri(TriLevels& T) {
  void* p[] = {T.irp, T.ja, T.val, T.rows};
  for (void* q : p)
    if (q) (void)hipFree(q);
  T = TriLevels{};
}

static void destroy_ilu(IluDev* F) {
  if (!F) return;
  if (ctx().initialized) (void)hipStreamSynchronize(ctx().stream);
  if (F->exec) (void)hipGraphExecDestroy(F->exec);
  if (F->graph) (void)hipGraphDestroy(F->graph);
  free_tri(F->L);
  free_tri(F->U);
  void* p[] = {F->d, F->tin, F->t1, F->t2};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete F;
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
  int rc = upload_tri(F->L, n, true, lirp, lja, lval);
  if (!rc) rc = upload_tri(F->U, n, false, uirp, uja, uval);
  if (rc) {
    destroy_ilu(F);
    return rc;
  }
  const size_t nb = sizeof(double) * (size_t)(n > 0 ? n : 1);
  hipError_t e = hipSuccess;
  double** bufs[] = {&F->d, &F->tin, &F->t1, &F->t2};
  for (double** b : bufs)
    if (e == hipSuccess) e = hipMalloc((void**)b, nb);
  if (e == hipSuccess && n > 0)
    e = hipMemcpy(F->d, d, sizeof(double) * (size_t)n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    destroy_ilu(F);
    return set_error(e == hipErrorOutOfMemory ? PSB_HIP_OUTOFMEM : PSB_HIP_RUNTIME, __func__,
                     hipGetErrorString(e));
  }
  if (n > 0) {
    hipStream_t s = ctx().stream;
    PSB_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue_levels(F->L, F->tin, F->t1, s);
    hipLaunchKernelGGL(k_scale_by, dim3(stream_grid(n)), dim3(kBlock), 0, s, n, F->t1, F->d);
    enqueue_levels(F->U, F->t1, F->t2, s);
    e = hipStreamEndCapture(s, &F->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&F->exec, F->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      destroy_ilu(F);
      return set_error(PSB_HIP_RUNTIME, __func__, hipGetErrorString(e));
    }
  }
  *ilu = F;
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
  PSB_HIP_CHECK(hipMemcpyAsync(F->tin, x->v, sizeof(double) * (size_t)F->n,
                               hipMemcpyDeviceToDevice, s));
  PSB_HIP_CHECK(hipGraphLaunch(F->exec, s));
  MultiVec t2{1, F->n, F->n, PSB_HIP_TYPE_DOUBLE, F->t2};
  return psb_hip_mv_axpby(F->n, alpha, &t2, beta, y);
}

}  // extern "C"
