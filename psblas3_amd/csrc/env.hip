// Environment: device selection, library stream, reduction workspace,
// device properties, event timing and raw buffers.
// Replaces cuda/cuda_util.c, cuda/spgpu/core.c and the bind(c) set of
// cuda/psb_cuda_env_mod.F90:43-174.
#include <string.h>

#include <mutex>
#include <string>

#include "psb_hip_internal.hpp"

namespace psbhip {

static Context g_ctx;
static std::string g_last_error;
static std::mutex g_init_mutex;
static hipEvent_t g_ev_start = nullptr, g_ev_stop = nullptr;

// Workspace of psb_hip_krylov_fused_noalloc: a static device array, so that
// entry point needs no allocation.  Each device holds its own copy until the
// process ends.  In host code the name `scratch` is not a device address:
// init_device looks the address up (hipGetSymbolAddress, with the device
// current) into Context::scratch, and every kernel argument or copy uses
// that pointer.
static __device__ __align__(256) char scratch[kScratchBytes];

Context& ctx() { return g_ctx; }

int set_error(int code, const char* where, const char* what) {
  g_last_error = std::string(where) + ": " + (what ? what : "");
  return code;
}

static int init_device(int dev) {
  std::lock_guard<std::mutex> lock(g_init_mutex);
  Context& c = g_ctx;
  if (c.initialized) {
    if (c.device == dev) return PSB_HIP_SUCCESS;
    return set_error(PSB_HIP_INVALID, "psb_hip_init", "already initialised on another device");
  }
  int ndev = 0;
  PSB_HIP_CHECK(hipGetDeviceCount(&ndev));
  if (dev < 0 || dev >= ndev) return set_error(PSB_HIP_INVALID, "psb_hip_init", "no such device");
  PSB_HIP_CHECK(hipSetDevice(dev));
  hipDeviceProp_t prop;
  PSB_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  c.num_cus = prop.multiProcessorCount;
  // default priority stream: the library stream every kernel is launched on
  PSB_HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  PSB_HIP_CHECK(hipMalloc(&c.red_partials, sizeof(double) * kRedBlocks * 8));
  PSB_HIP_CHECK(hipMalloc(&c.red_counter, sizeof(unsigned int) * 64));
  PSB_HIP_CHECK(hipMalloc(&c.red_result, sizeof(double) * 64));
  PSB_HIP_CHECK(hipMemset(c.red_counter, 0, sizeof(unsigned int) * 64));
  PSB_HIP_CHECK(hipHostMalloc(&c.host_result, sizeof(double) * 64, hipHostMallocDefault));
  void* sp = nullptr;
  PSB_HIP_CHECK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  c.scratch = (char*)sp;
  c.scratch_bytes = kScratchBytes;
  PSB_HIP_CHECK(hipHostMalloc(&c.scratch_host, kScratchHostBytes, hipHostMallocDefault));
  for (hipEvent_t& e : c.scratch_ev)
    PSB_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  PSB_HIP_CHECK(hipEventCreate(&g_ev_start));
  PSB_HIP_CHECK(hipEventCreate(&g_ev_stop));
  PSB_HIP_CHECK(hipDeviceSynchronize());
  c.device = dev;
  c.initialized = true;
  return PSB_HIP_SUCCESS;
}

int ensure_init() {
  if (g_ctx.initialized) return PSB_HIP_SUCCESS;
  return init_device(0);
}

typedef double v2d __attribute__((ext_vector_type(2)));

// Each block copies contiguous 16 KB tiles (4 x 16 B per thread, loads
// issued before stores), grid-stride over the tiles.
__global__ __launch_bounds__(kBlock) void k_stream_copy(const v2d* __restrict__ src,
                                                       v2d* __restrict__ dst, int64_t n) {
  const int64_t tile = 4 * kBlock;
  const int64_t ntiles = n / tile;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const v2d* s = src + t * tile + threadIdx.x;
    v2d* d = dst + t * tile + threadIdx.x;
    v2d a0 = s[0], a1 = s[kBlock], a2 = s[2 * kBlock], a3 = s[3 * kBlock];
    __builtin_nontemporal_store(a0, d);
    __builtin_nontemporal_store(a1, d + kBlock);
    __builtin_nontemporal_store(a2, d + 2 * kBlock);
    __builtin_nontemporal_store(a3, d + 3 * kBlock);
  }
  for (int64_t i = ntiles * tile + blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

// STREAM triad a = b + s*c on 16-byte vectors, 4 per thread per tile.
// Read-only stream: per-block sums of a (4 x 16 bytes per thread in flight),
// one partial per block so the loads cannot be dropped.
__global__ __launch_bounds__(kBlock) void k_stream_read(const v2d* __restrict__ a, int64_t n,
                                                       double* __restrict__ partials) {
  __shared__ double lds[kBlock / kWave];
  const int64_t tile = 4 * kBlock;
  const int64_t ntiles = n / tile;
  v2d acc = {0.0, 0.0};
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const v2d* s = a + t * tile + threadIdx.x;
    const v2d a0 = __builtin_nontemporal_load(s), a1 = __builtin_nontemporal_load(s + kBlock);
    const v2d a2 = __builtin_nontemporal_load(s + 2 * kBlock);
    const v2d a3 = __builtin_nontemporal_load(s + 3 * kBlock);
    acc += (a0 + a1) + (a2 + a3);
  }
  for (int64_t i = ntiles * tile + blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    acc += a[i];
  const double s = block_sum(acc.x + acc.y, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// The traffic shape of a column-major sparse product without its gathers:
// y[i] = sum_k a[k*m + i] for k < ratio (ratio reads per write, coalesced).
__global__ __launch_bounds__(kBlock) void k_stream_mix(const v2d* __restrict__ a, int ratio,
                                                      int64_t m, v2d* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock) {
    v2d s = {0.0, 0.0};
    for (int k = 0; k < ratio; k++) s += __builtin_nontemporal_load(a + (int64_t)k * m + i);
    __builtin_nontemporal_store(s, y + i);
  }
}

__global__ __launch_bounds__(kBlock) void k_stream_triad(v2d* __restrict__ a,
                                                        const v2d* __restrict__ b,
                                                        const v2d* __restrict__ c, double s,
                                                        int64_t n) {
  const int64_t tile = 4 * kBlock;
  const int64_t ntiles = n / tile;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t o = t * tile + threadIdx.x;
    v2d b0 = b[o], b1 = b[o + kBlock], b2 = b[o + 2 * kBlock], b3 = b[o + 3 * kBlock];
    v2d c0 = c[o], c1 = c[o + kBlock], c2 = c[o + 2 * kBlock], c3 = c[o + 3 * kBlock];
    __builtin_nontemporal_store(b0 + s * c0, a + o);
    __builtin_nontemporal_store(b1 + s * c1, a + o + kBlock);
    __builtin_nontemporal_store(b2 + s * c2, a + o + 2 * kBlock);
    __builtin_nontemporal_store(b3 + s * c3, a + o + 3 * kBlock);
  }
  for (int64_t i = ntiles * tile + blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    a[i] = b[i] + s * c[i];
}

}  // namespace psbhip

using namespace psbhip;

extern "C" {

int psb_hip_init(int dev) { return init_device(dev); }

int psb_hip_exit(void) {
  std::lock_guard<std::mutex> lock(g_init_mutex);
  Context& c = g_ctx;
  if (!c.initialized) return PSB_HIP_SUCCESS;
  PSB_HIP_CHECK(hipStreamSynchronize(c.stream));
  hipFree(c.red_partials);
  hipFree(c.red_counter);
  hipFree(c.red_result);
  hipHostFree(c.host_result);
  hipHostFree(c.scratch_host);
  for (hipEvent_t e : c.scratch_ev) hipEventDestroy(e);
  hipEventDestroy(g_ev_start);
  hipEventDestroy(g_ev_stop);
  hipStreamDestroy(c.stream);
  c = Context();
  return PSB_HIP_SUCCESS;
}

int psb_hip_get_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int psb_hip_get_device(void) {
  int d = -1;
  if (hipGetDevice(&d) != hipSuccess) return -1;
  return d;
}

int psb_hip_set_device(int dev) {
  PSB_HIP_CHECK(hipSetDevice(dev));
  return PSB_HIP_SUCCESS;
}

static int prop_int(hipDeviceAttribute_t a) {
  if (ensure_init() != PSB_HIP_SUCCESS) return -1;
  int v = 0;
  if (hipDeviceGetAttribute(&v, a, g_ctx.device) != hipSuccess) return -1;
  return v;
}

int psb_hip_get_multiprocessors(void) { return prop_int(hipDeviceAttributeMultiprocessorCount); }
int psb_hip_get_memory_bus_width(void) { return prop_int(hipDeviceAttributeMemoryBusWidth); }
int psb_hip_get_memory_clock_rate(void) { return prop_int(hipDeviceAttributeMemoryClockRate); }
int psb_hip_get_warp_size(void) { return prop_int(hipDeviceAttributeWarpSize); }
int psb_hip_get_max_threads_per_mp(void) {
  return prop_int(hipDeviceAttributeMaxThreadsPerMultiProcessor);
}
int psb_hip_get_max_registers_per_block(void) {
  return prop_int(hipDeviceAttributeMaxRegistersPerBlock);
}
int psb_hip_device_has_uva(void) { return prop_int(hipDeviceAttributeUnifiedAddressing); }

int psb_hip_get_device_name(char* buf, int len) {
  int rc = ensure_init();
  if (rc) return rc;
  hipDeviceProp_t prop;
  PSB_HIP_CHECK(hipGetDeviceProperties(&prop, g_ctx.device));
  // the marketing name can be empty; the architecture name (gfx950...) is not
  const char* name = prop.name[0] ? prop.name : prop.gcnArchName;
  if (len > 0) {
    strncpy(buf, name, (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return PSB_HIP_SUCCESS;
}

double psb_hip_memory_peak_bandwidth(void) {
  // 2 * clock(kHz) * busWidth(bit) / 8 -> kB/s; / 1000 -> MB/s
  double clk = (double)psb_hip_get_memory_clock_rate();
  double bw = (double)psb_hip_get_memory_bus_width();
  return 2.0 * clk * bw / 8.0 / 1000.0;
}

int psb_hip_device_sync(void) {
  int rc = ensure_init();
  if (rc) return rc;
  PSB_HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
  return PSB_HIP_SUCCESS;
}

void* psb_hip_get_stream(void) {
  if (ensure_init() != PSB_HIP_SUCCESS) return nullptr;
  return (void*)g_ctx.stream;
}

const char* psb_hip_last_error(void) { return g_last_error.c_str(); }

int psb_hip_timer_start(void) {
  int rc = ensure_init();
  if (rc) return rc;
  PSB_HIP_CHECK(hipEventRecord(g_ev_start, g_ctx.stream));
  return PSB_HIP_SUCCESS;
}

double psb_hip_timer_stop(void) {
  if (ensure_init() != PSB_HIP_SUCCESS) return -1.0;
  if (hipEventRecord(g_ev_stop, g_ctx.stream) != hipSuccess) return -1.0;
  if (hipEventSynchronize(g_ev_stop) != hipSuccess) return -1.0;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, g_ev_start, g_ev_stop) != hipSuccess) return -1.0;
  return (double)ms;
}

int psb_hip_stream_copy(void* dst, const void* src, int64_t nbytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (nbytes % 16 != 0) return set_error(PSB_HIP_INVALID, __func__, "nbytes % 16");
  int64_t n = nbytes / 16;
  if (n == 0) return PSB_HIP_SUCCESS;
  int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  int64_t cap = (int64_t)(g_ctx.num_cus > 0 ? g_ctx.num_cus : 256) * 8;
  if (g > cap) g = cap;
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)g), dim3(kBlock), 0, g_ctx.stream,
                     (const v2d*)src, (v2d*)dst, n);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_stream_read(const void* a, int64_t nbytes, void* partials) {
  int rc = ensure_init();
  if (rc) return rc;
  if (nbytes % 16 != 0 || !partials) return set_error(PSB_HIP_INVALID, __func__, "nbytes % 16");
  int64_t n = nbytes / 16;
  if (n == 0) return PSB_HIP_SUCCESS;
  int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  int64_t cap = (int64_t)(g_ctx.num_cus > 0 ? g_ctx.num_cus : 256) * 8;
  if (g > cap) g = cap;
  hipLaunchKernelGGL(k_stream_read, dim3((unsigned)g), dim3(kBlock), 0, g_ctx.stream,
                     (const v2d*)a, n, (double*)partials);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_stream_mix(const void* a, int ratio, int64_t write_bytes, void* y) {
  int rc = ensure_init();
  if (rc) return rc;
  if (write_bytes % 16 != 0 || ratio < 1) return set_error(PSB_HIP_INVALID, __func__, "sizes");
  const int64_t m = write_bytes / 16;
  if (m == 0) return PSB_HIP_SUCCESS;
  int64_t g = (m + kBlock - 1) / kBlock;
  int64_t cap = (int64_t)(g_ctx.num_cus > 0 ? g_ctx.num_cus : 256) * 16;
  if (g > cap) g = cap;
  hipLaunchKernelGGL(k_stream_mix, dim3((unsigned)g), dim3(kBlock), 0, g_ctx.stream,
                     (const v2d*)a, ratio, m, (v2d*)y);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_stream_triad(void* a, const void* b, const void* c, double s, int64_t nbytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (nbytes % 16 != 0) return set_error(PSB_HIP_INVALID, __func__, "nbytes % 16");
  int64_t n = nbytes / 16;
  if (n == 0) return PSB_HIP_SUCCESS;
  int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  int64_t cap = (int64_t)(g_ctx.num_cus > 0 ? g_ctx.num_cus : 256) * 8;
  if (g > cap) g = cap;
  hipLaunchKernelGGL(k_stream_triad, dim3((unsigned)g), dim3(kBlock), 0, g_ctx.stream, (v2d*)a,
                     (const v2d*)b, (const v2d*)c, s, n);
  PSB_HIP_LAUNCH_CHECK();
  return PSB_HIP_SUCCESS;
}

int psb_hip_alloc_buffer(void** dptr, int64_t nbytes) {
  int rc = ensure_init();
  if (rc) return rc;
  *dptr = nullptr;
  if (nbytes <= 0) nbytes = 8;
  PSB_HIP_CHECK(hipMalloc(dptr, (size_t)nbytes));
  return PSB_HIP_SUCCESS;
}

int psb_hip_free_buffer(void* dptr) {
  if (!dptr) return PSB_HIP_SUCCESS;
  // buffers may still be in use by queued kernels
  if (g_ctx.initialized) PSB_HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
  PSB_HIP_CHECK(hipFree(dptr));
  return PSB_HIP_SUCCESS;
}

int psb_hip_write_buffer(void* dptr, const void* host, int64_t nbytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (nbytes <= 0) return PSB_HIP_SUCCESS;
  PSB_HIP_CHECK(hipMemcpyAsync(dptr, host, (size_t)nbytes, hipMemcpyHostToDevice, g_ctx.stream));
  PSB_HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
  return PSB_HIP_SUCCESS;
}

int psb_hip_read_buffer(void* host, const void* dptr, int64_t nbytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (nbytes <= 0) return PSB_HIP_SUCCESS;
  PSB_HIP_CHECK(hipMemcpyAsync(host, dptr, (size_t)nbytes, hipMemcpyDeviceToHost, g_ctx.stream));
  PSB_HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
  return PSB_HIP_SUCCESS;
}

}  // extern "C"
