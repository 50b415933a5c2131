/*
 * psb_hip.h -- C ABI of libpsb_hip, the MI355X (gfx950) device layer behind
 * PSBLAS's psb_[sd]_hip_{hlg,elg,csrg}_mat and psb_[sd]_hip_vect plug-in
 * types.  Plain C: pointers, sizes and opaque handles; no C++ or torch types.
 *
 * Every entry point names the reference interface it replaces.  Semantics
 * follow the reference CPU path where the CPU and the reference CUDA layer
 * disagree (HLL padding, xyzw, axybz with beta = 0).
 *
 * Conventions (as the reference's device layer):
 *   - indices are int32; matrix index arrays are 1-based (firstIndex = 1,
 *     cuda/impl/psb_d_cuda_hlg_to_gpu.F90:56);
 *   - every call is asynchronous on the library stream unless it returns a
 *     host scalar or copies to host memory (those synchronise);
 *   - return value 0 = success, else one of PSB_HIP_* below
 *     (cuda/spgpu/core.h:69-72 codes 0/1/3 keep their meaning).
 *   - device handles (void*) are owned by the caller and released with the
 *     matching free function (cuda/psb_d_cuda_hlg_mat_mod.F90:288-301).
 */
#ifndef PSB_HIP_H
#define PSB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes (cuda/spgpu/core.h:69-72) */
#define PSB_HIP_SUCCESS      0
#define PSB_HIP_UNSUPPORTED  1
#define PSB_HIP_INVALID      2
#define PSB_HIP_OUTOFMEM     3
#define PSB_HIP_RUNTIME      4
#define PSB_HIP_NOT_INIT     5
#define PSB_HIP_BUSY         6   /* batch engine: no free slot or arena room now */

/* element types (cuda/spgpu/core.h:77-79) */
#define PSB_HIP_TYPE_INT     0
#define PSB_HIP_TYPE_FLOAT   1
#define PSB_HIP_TYPE_DOUBLE  2

/* geins duplicate handling (cuda/vectordev.h:55-56) */
#define PSB_HIP_INS_OVERWRITE 0
#define PSB_HIP_INS_ADD       1

/* ===================================================================== */
/* Environment -- cuda/psb_cuda_env_mod.F90:43-174, cuda/cuda_util.h:54-69 */
/* ===================================================================== */

/* gpuInit(dev) + psb_cudaCreateHandle: select the device, create the library
   stream and the reduction workspace.  Idempotent. */
int psb_hip_init(int dev);
/* gpuClose + psb_cudaDestroyHandle */
int psb_hip_exit(void);
/* getDeviceCount / getDevice / setDevice */
int psb_hip_get_device_count(void);
int psb_hip_get_device(void);
int psb_hip_set_device(int dev);
/* getGPUMultiProcessors, getGPUMemoryBusWidth, getGPUMemoryClockRate (kHz),
   getGPUWarpSize, getGPUMaxThreadsPerMP, getGPUMaxRegistersPerBlock,
   DeviceHasUVA */
int psb_hip_get_multiprocessors(void);
int psb_hip_get_memory_bus_width(void);
int psb_hip_get_memory_clock_rate(void);
int psb_hip_get_warp_size(void);
int psb_hip_get_max_threads_per_mp(void);
int psb_hip_get_max_registers_per_block(void);
int psb_hip_device_has_uva(void);
/* cpyGPUNameString */
int psb_hip_get_device_name(char *buf, int len);
/* psb_cuda_MemoryPeakBandwidth (cuda/psb_cuda_env_mod.F90:294-300):
   2 * memClock(kHz) * busWidth(bit) / 8 / 1000, in MB/s */
double psb_hip_memory_peak_bandwidth(void);
/* cudaSync (cuda/cuda_util.h:69): wait for all work on the library stream */
int psb_hip_device_sync(void);
/* psb_cudaGetStream: the library stream as an opaque hipStream_t */
void *psb_hip_get_stream(void);
/* Last error text (for the host mirror's error messages). */
const char *psb_hip_last_error(void);

/* Event timing on the library stream (no reference counterpart; the
   reference times with MPI_Wtime around psb_cuda_DeviceSync). */
int psb_hip_timer_start(void);
/* milliseconds since psb_hip_timer_start, measured with HIP events on the
   library stream (synchronises) */
double psb_hip_timer_stop(void);

/* Device-to-device copy of nbytes (multiple of 16) with a streaming kernel
   on the library stream; used to measure achievable HBM bandwidth for the
   roofline denominator (no reference counterpart: the reference only
   reports the spec figure of psb_cuda_MemoryPeakBandwidth). */
int psb_hip_stream_copy(void *dst, const void *src, int64_t nbytes);
/* Read-only stream over nbytes of a (multiple of 16): one sum per block into
   partials (device, at least 8 * 8 * multiprocessors bytes); the bandwidth
   ceiling of read-dominated kernels such as SpMV. */
int psb_hip_stream_read(const void *a, int64_t nbytes, void *partials);
/* The traffic of a column-major sparse product without its gathers:
   y[i] = sum_{k<ratio} a[k*m + i] over 16-byte elements, m = write_bytes/16;
   reads ratio * write_bytes from a, writes write_bytes to y. */
int psb_hip_stream_mix(const void *a, int ratio, int64_t write_bytes, void *y);
/* STREAM triad a = b + s*c over nbytes per array (multiple of 16). */
int psb_hip_stream_triad(void *a, const void *b, const void *c, double s, int64_t nbytes);

/* ===================================================================== */
/* Raw device buffers -- cuda/psb_d_vectordev_mod.F90:85-160               */
/* (allocateDouble/writeDouble/readDouble/freeDouble and Float/Int)        */
/* ===================================================================== */
int psb_hip_alloc_buffer(void **dptr, int64_t nbytes);
int psb_hip_free_buffer(void *dptr);
int psb_hip_write_buffer(void *dptr, const void *host, int64_t nbytes);
int psb_hip_read_buffer(void *host, const void *dptr, int64_t nbytes);

/* ===================================================================== */
/* Multivectors -- cuda/vectordev.h:76-87, cuda/dvectordev.c:45-336        */
/* ===================================================================== */

/* FallocMultiVecDevice (cuda/vectordev.c:83-145): count columns of size
   entries; pitch = size if count == 1, else rounded up to 256 bytes. */
int psb_hip_mv_alloc(void **mv, int count, int size, int elem_type);
/* freeMultiVecDevice */
void psb_hip_mv_free(void *mv);
/* getMultiVecDeviceSize / Count / Pitch */
int psb_hip_mv_size(void *mv);
int psb_hip_mv_count(void *mv);
int psb_hip_mv_pitch(void *mv);
/* raw device pointer of column 0 (for interop and tests) */
void *psb_hip_mv_data(void *mv);
/* writeMultiVecDevice{Double,Float,Int} / readMultiVecDevice*: whole
   first column (count == 1) or count columns of size entries, host
   leading dimension = size. */
int psb_hip_mv_write(void *mv, const void *host);
int psb_hip_mv_read(void *mv, void *host);
/* writeMultiVecDeviceDoubleR2 / readMultiVecDeviceDoubleR2: host leading
   dimension ld */
int psb_hip_mv_write_r2(void *mv, const void *host, int ld);
int psb_hip_mv_read_r2(void *mv, void *host, int ld);
/* writeDoubleFirst / readDoubleFirst: n entries starting at 1-based
   position first of column 0 */
int psb_hip_mv_write_first(void *mv, int first, int n, const void *host);
int psb_hip_mv_read_first(void *mv, int first, int n, void *host);

/* setscalMultiVecDevice{Double,Float,Int} (cuda/dvectordev.c:85-96):
   x(first:last) = val, first/last relative to index_base */
int psb_hip_mv_setscal_d(double val, int first, int last, int index_base, void *x);
int psb_hip_mv_setscal_s(float val, int first, int last, int index_base, void *x);
int psb_hip_mv_setscal_i(int val, int first, int last, int index_base, void *x);

/* geinsMultiVecDeviceDouble (cuda/dvectordev.c:99-125): x(irl(i)) =
   val(i) (dupl = OVERWRITE) or x(irl(i)) += val(i) (dupl = ADD), n entries,
   irl an int multivector, val a multivector of x's type. */
int psb_hip_mv_geins(int n, void *irl, void *val, int dupl, int index_base, void *x);

/* igathMultiVecDeviceDouble / iscatMultiVecDeviceDouble (dvectordev.c:128-
   178): gather x(idx(first:first+n-1)) of column vector_id into the device
   buffer hv(hfirst:...), or scatter hv into x as x = beta*x + hv.  idx is a
   device int array (the VecIdx variants pass an int multivector handle). */
int psb_hip_mv_igath(void *x, int vector_id, int n, int first, const int *didx,
                     int hfirst, void *dhv, int index_base);
int psb_hip_mv_iscat(void *x, int vector_id, int n, int first, const int *didx,
                     int hfirst, const void *dhv, int index_base, double beta);

/* BLAS-1 over the first n entries of every column (d = double, s = float).
   References: dvectordev.c:180-336 (device), psi_d_serial_impl.F90 and
   psb_d_base_vect_mod.F90 (CPU semantics, which are followed). */
/* scalMultiVecDevice: x = alpha*x over the whole pitch*count */
int psb_hip_mv_scal(double alpha, void *x);
/* dotMultiVecDevice: res(c) = x(:,c).y(:,c); res is a host array of count
   entries; synchronises */
int psb_hip_mv_dot(double *res, int n, void *x, void *y);
/* nrm2MultiVecDevice / amaxMultiVecDevice / asumMultiVecDevice */
int psb_hip_mv_nrm2(double *res, int n, void *x);
int psb_hip_mv_amax(double *res, int n, void *x);
int psb_hip_mv_asum(double *res, int n, void *x);
/* axpbyMultiVecDevice (psi_daxpbyv semantics): y = alpha*x + beta*y */
int psb_hip_mv_axpby(int n, double alpha, void *x, double beta, void *y);
/* axpbyz (psi_daxpbyzv): z = alpha*x + beta*y */
int psb_hip_mv_axpbyz(int n, double alpha, void *x, double beta, void *y, void *z);
/* upd_xyzMultiVecDevice (psi_d_upd_xyz): y = alpha*x + beta*y;
   z = gamma*y + delta*z */
int psb_hip_mv_upd_xyz(int n, double alpha, double beta, double gamma, double delta,
                       void *x, void *y, void *z);
/* xyzwMultiVecDevice (psi_dxyzw; the CUDA kernel's a+x at dxyzw.cu:47 is
   not reproduced): y = a*x+b*y; z = c*y+d*z; w = e*z+f*w */
int psb_hip_mv_xyzw(int n, double a, double b, double c, double d, double e, double f,
                    void *x, void *y, void *z, void *w);
/* axyMultiVecDevice: y = alpha*x.*y */
int psb_hip_mv_axy(int n, double alpha, void *x, void *y);
/* axybzMultiVecDevice (d_base_mlt_a_2 semantics): z = alpha*y.*x + beta*z,
   z not read when beta == 0 */
int psb_hip_mv_axybz(int n, double alpha, void *x, void *y, double beta, void *z);
/* Jacobi build step of psb_d_diag_precbld (psb_d_diagprec_impl.f90:261-
   270): d(i) = 1/d(i), a zero replaced by 1, for i <= n; entries n+1..size
   set to 1 */
int psb_hip_mv_diag_inv(int n, void *d);
/* absMultiVecDevice2 / absMultiVecDevice: y = alpha*|x| */
int psb_hip_mv_abs(int n, double alpha, void *x, void *y);

/* ===================================================================== */
/* Host-side assembly helpers (no device work)                              */
/* ===================================================================== */

/* psb_d_fix_coo (base/serial/impl/psb_d_coo_impl.F90:4094-4896), row major:
   sorts the 1-based triplets by (row, column), stable, and merges
   duplicates (dupl: 0 overwrite, 1 add, 2 error -> returns
   PSB_HIP_INVALID).  In place; *nzout receives the new count. */
int psb_hip_fix_coo_d(int nr, int nc, int64_t nz, int *ia, int *ja, double *val, int dupl,
                      int64_t *nzout);
/* psi_compute_hckoff_from_coo (cuda/impl/psb_d_cuda_cp_hlg_from_coo.F90:117-
   188): from a row-sorted COO, irn(nr) row counts, idisp(nr+1) row starts
   (0-based), hkoffs(noffs+1) hack offsets; returns isz in *isz. */
int psb_hip_hll_offsets(int nr, int64_t nz, const int *ia, int hksz, int *irn, int *idisp,
                        int *hkoffs, int64_t *isz);

/* Host copies of the dual-memory plug-ins (cuda/psb_d_cuda_hlg_mat_mod.F90:
   40-84: host / sync / dev state; the host copy is the CPU format the GPU
   type extends).  Products on the host copy, for a product whose x or y is
   not a device vector (cuda/impl/psb_d_cuda_hlg_vect_mv.F90:77-108 falls
   back to the CPU format's product there): y = alpha*A*x + beta*y with the
   CPU products' arithmetic (ext/impl/psb_d_hll_csmv.f90:291-371,
   ext/impl/psb_d_ell_csmv.f90, base/serial/impl/psb_d_csr_impl.F90:46-417:
   each row's entries in stored order, products and sums separately rounded;
   y not read when beta = 0).  HLL: position p is row perm(p) when perm is not
   NULL.  ELL: column-major with leading dimension ld.  elem_type FLOAT or
   DOUBLE for val, x and y (host arrays). */
int psb_hip_host_hll_spmv(int nr, int hksz, const int *hkoffs, const int *irn, const int *ja,
                          const void *val, const int *perm, int elem_type, double alpha,
                          const void *x, double beta, void *y);
int psb_hip_host_ell_spmv(int nr, int ld, const int *irn, const int *ja, const void *val,
                          int elem_type, double alpha, const void *x, double beta, void *y);
int psb_hip_host_csr_spmv(int nr, const int *irp, const int *ja, const void *val, int elem_type,
                          double alpha, const void *x, double beta, void *y);
/* psi_d_convert_hll_from_coo on the host (ext/impl/psi_d_convert_hll_from_coo.f90:
   31-122 with psi_d_xtr_ell_from_coo's padding, :56-60): the HLL arrays ja /
   val (hkoffs(noffs+1) slots) and idiag(nr) from a row-sorted COO, irn and
   idisp as psb_hip_hll_offsets gives them (or, with perm, the per-position
   counts of psb_hip_hll_sort_rows and per-row idisp). */
int psb_hip_host_hll_fill(int nr, int hksz, const int *irn, const int *hkoffs, const int *idisp,
                          const int *cja, const void *cval, const int *perm, int elem_type,
                          int *ja, void *val, int *idiag);

/* Generators of the benchmark operators (host, COO, 1-based, entries in the
   insertion order of test/pargen/psb_d_pde3d.F90:478-560).  Call with
   NULL arrays to get the entry count.  a[3], b[3]: diffusion and
   convection coefficients; hscale = 0 drops h = 1/(L+1).  rhs (may be
   NULL) receives the boundary term of the sample g when gflag = 1. */
int64_t psb_hip_gen_pde3d(int L, const double *a, const double *b, double c, int hscale,
                          int gflag, int *ia, int *ja, double *val, double *rhs);
int64_t psb_hip_gen_aniso2d(int L, double eps, int *ia, int *ja, double *val);
int64_t psb_hip_gen_powerlaw(int n, int cap, uint64_t seed, int *ia, int *ja, double *val);
/* Counter-based uniform [-1,1) fill (host), the same stream the generators
   use: x(i) = u(seed, stream, i). */
void psb_hip_fill_uniform_host(uint64_t seed, uint32_t stream, int64_t n, double *x);
/* MatrixMarket I/O (host) -- util/psb_d_mmio_impl.f90.
   psb_hip_mm_read_coo: mm_mat_read (:315-440): a "matrix coordinate" file,
   type real or pattern (values 1), general or symmetric (the mirrored
   off-diagonal entries appended after the stored ones, in file order).
   With ia == NULL: *nr, *nc and *nz = the entry capacity needed.  Else the
   entries (1-based, unsorted: fix them with psb_hip_fix_coo_d) into arrays
   of capacity cap; *nz = the count.  PSB_HIP_RUNTIME: cannot open;
   PSB_HIP_UNSUPPORTED: other object/format/type/symmetry (reference info
   909 / 904); PSB_HIP_INVALID: malformed or short file (902 / 905). */
int psb_hip_mm_read_coo(const char *path, int64_t cap, int *nr, int *nc, int64_t *nz, int *ia,
                        int *ja, double *val);
/* mm_dvet_read (:32-102): "matrix array real general", first column; with
   b == NULL only *n. */
int psb_hip_mm_read_vec(const char *path, int64_t cap, int64_t *n, double *b);
/* the COO print (base/serial/impl/psb_d_coo_impl.F90:728-763) and
   mm_dvet1_write (:226-280): values with 19 significant digits */
int psb_hip_mm_write_coo(const char *path, const char *title, int nr, int nc, int64_t nz,
                         const int *ia, const int *ja, const double *val);
int psb_hip_mm_write_vec(const char *path, const char *title, int64_t n, const double *b);
/* The same values generated on the device into a double multivector (column c
   of a multi-column vector uses stream + c). */
int psb_hip_mv_fill_uniform(void *x, uint64_t seed, uint32_t stream);

/* ===================================================================== */
/* HLL (HLG) matrices -- cuda/hlldev.h:96-153, cuda/hlldev_mod.F90          */
/* ===================================================================== */

/* FallocHllDevice (cuda/hlldev.c:145-157). elem_type FLOAT or DOUBLE;
   allocsize = hkoffs(noffs+1) slots; first_index must be 1. */
int psb_hip_hll_alloc(void **dmat, int hksz, int rows, int nzeros, int allocsize,
                      int elem_type, int first_index);
/* freeHllDevice (also releases the hack offsets, which the reference
   leaks, cuda/hlldev.c:131-142) */
void psb_hip_hll_free(void *dmat);
/* writeHllDevice{Double,Float}: val/ja (allocsize), hkoffs (noffs+1),
   irn (rows), idiag (rows). val has the matrix element type. */
int psb_hip_hll_write(void *dmat, const void *val, const int *ja, const int *hkoffs,
                      const int *irn, const int *idiag);
/* readHllDevice{Double,Float} */
int psb_hip_hll_read(void *dmat, void *val, int *ja, int *hkoffs, int *irn, int *idiag);
/* getHllDeviceParams */
int psb_hip_hll_get_params(void *dmat, int *hksz, int *rows, int *nzeros, int *allocsize,
                           int *hack_offs_len, int *first_index, int *elem_type);
/* psiCopyCooToHlg{Double,Float} (cuda/hlldev.c:420-462): fill the device
   HLL from a row-sorted COO (ja, val of nza entries, 1-based) with the host
   irn/hoffs/idisp of psb_hip_hll_offsets.  Padding follows the CPU rule
   (ext/impl/psi_d_xtr_ell_from_coo.f90:56-60; last hack pre-filled with
   ja = nr, val = 0, psi_d_convert_hll_from_coo.f90:93-96).  idiag = 1-based
   slot of the diagonal, 0 if the row has none.  val has the element type
   of dmat. */
int psb_hip_hll_copy_coo(int nr, int nc, int nza, int hksz, int noffs, int isz,
                         const int *irn, const int *hoffs, const int *idisp, const int *ja,
                         const void *val, void *dmat);
/* Row-sorted HLL (the reorder array rIdx and per-row sizes rS of the SPGPU
   HLL kernel, cuda/spgpu/hell.h:37-101 and
   cuda/spgpu/kernels/hell_spmv_base_template.cuh:197-214,227-252, which the
   reference HLG type never fills: cuda/hlldev.c:202-204 passes NULL).
   psb_hip_hll_sort_rows (host): from the row counts irn(nr) of a row-sorted
   COO, perm(nr) = 1-based row at each position, rows sorted by descending
   count inside consecutive windows of `window` rows (rounded up to a
   multiple of hksz; <= 0: one window), ties in ascending row order;
   irn_pos(nr) the count at each position; hkoffs(noffs+1) the hack offsets
   over positions (hksz times each hack's widest count); *isz the slots.
   psb_hip_hll_copy_coo_sorted: as psb_hip_hll_copy_coo with position p
   holding row perm(p) (idisp per row, irn_pos per position); the matrix's
   irn and idiag are then per position.  The product writes y(perm(p)) from
   position p, each row summed over its own entries only (padding not read),
   in CSR order; a hack wider than 256 slots is computed in several work
   units whose partial sums are added in slot order (those rows then follow
   the per-row error bound, every other row equals the unsorted product's
   value).  Products of several columns run one column at a time.  A later
   psb_hip_hll_write / psb_hip_hll_copy_coo returns to rows in order.
   psb_hip_hll_read_perm: the permutation (1..nr when rows are in order).
   psb_hip_hll_units: work units and cut hacks of the product plan. */
int psb_hip_hll_sort_rows(int nr, int hksz, int window, const int *irn, int *perm, int *irn_pos,
                          int *hkoffs, int64_t *isz);
int psb_hip_hll_copy_coo_sorted(int nr, int nc, int nza, int hksz, int noffs, int isz,
                                const int *perm, const int *irn_pos, const int *hoffs,
                                const int *idisp, const int *ja, const void *val, void *dmat);
int psb_hip_hll_read_perm(void *dmat, int *perm);
int psb_hip_hll_units(void *dmat, int *nunits, int *nsplit);
/* After psb_hip_hll_write of arrays in a row-sorted layout (hkoffs the same
   host array written): install the permutation perm(rows) and the product's
   work-unit plan (the to_gpu of a row-sorted matrix built on the host). */
int psb_hip_hll_set_perm(void *dmat, const int *perm, const int *hkoffs);
/* Column slices of row-sorted matrices (no reference counterpart): when
   psb_hip_hll_copy_coo_sorted writes a matrix whose x is over 12 MiB and
   whose entries mostly lie more than cols/16 from the diagonal (scattered
   columns), the entries are also stored split into consecutive column
   ranges of about 8 MiB (FP64) or 5.6 MB (FP32) of x, each a row-sorted
   HLL matrix of its own, and
   the product runs them one after another accumulating into y, so each
   slice's x range stays in L2 (the rule of the CSR slices,
   psb_hip_csr_set_slices).  Rows then follow the per-row error bound.
   psb_hip_hll_set_slices: -1 automatic (default), 0 or 1 none, >= 2 that
   many, for matrices written afterwards; returns the previous value.
   psb_hip_hll_slices: the matrix's slice count (0 when unsliced). */
int psb_hip_hll_set_slices(int slices);
int psb_hip_hll_slices(void *dmat);
/* spmvHllDevice{Double,Float} (cuda/hlldev.c:185-207): y = alpha*A*x +
   beta*y for every column of x/y (the reference computes column 1 only),
   y not read when beta == 0.  x, y multivectors of the matrix type. */
int psb_hip_hll_spmv(void *dmat, double alpha, void *x, double beta, void *y);
/* Attach a single-precision copy of the values to a double HLL matrix so
   the mixed-precision solvers can run FP32 SpMVs on the same index arrays
   (no reference counterpart). */
int psb_hip_hll_make_single(void *dmat);
/* y = alpha*A*x + beta*y in single precision using that copy. */
int psb_hip_hll_spmv_single(void *dmat, double alpha, void *x, double beta, void *y);
/* Diagonal of the matrix into a vector of its type (0 where a row has no
   diagonal): the device side of get_diag used by psb_d_diag_precbld
   (prec/impl/psb_d_diagprec_impl.f90:254). */
int psb_hip_hll_get_diag(void *dmat, void *d);
/* psb_d_cuda_hlg_scals / psb_d_cuda_hlg_scal (cuda/impl/psb_d_cuda_hlg_scals.F90,
   psb_d_cuda_hlg_scal.F90: the CPU psb_d_hll_scals / psb_d_hll_scal, then
   to_gpu), done on the device copy: scals multiplies every stored value,
   padding included, by s; scal multiplies each row's entries (its irn
   entries, padding untouched) by d(row) when left, by d(column) otherwise
   (ext/impl/psb_d_hll_scal.f90:83-130).  d is a device vector of the
   matrix's type with rows entries (left) or the largest stored column
   (right).  The FP32 copy of psb_hip_hll_make_single follows. */
int psb_hip_hll_scals(void *dmat, double s);
int psb_hip_hll_scal(void *dmat, void *d, int left);
/* Kernel variant selection for the HLL SpMV (autotuner hook): 0 = auto
   (two rows per lane for double values, one for float, XCD block order),
   1 = one row per lane, 2 = two adjacent rows per lane (even hksz), both in
   launch order; + 4 = XCD block order.  Returns the previous value. */
int psb_hip_hll_set_variant(int variant);
/* Plane-tile block order of this matrix's products (single and multi-RHS;
   no reference counterpart): for banded matrices whose couplings sit near
   multiples of a period of period_rows rows (a 3-D stencil in natural order:
   L*L), the row blocks are walked in tile columns of tile_blocks blocks
   through every plane and whole tile columns are dealt to the XCDs, so the
   rows that read the same x entries across planes run on one XCD back to
   back.  The results are those of the default order.  period_rows = 0 or
   tile_blocks = 0 restores the default (chunked XCD order).  Not for
   row-sorted matrices. */
int psb_hip_hll_set_plane(void *dmat, int period_rows, int tile_blocks);
/* 16-bit column offsets (no reference counterpart): when a rows-in-order
   HLL matrix is written (psb_hip_hll_write, psb_hip_hll_copy_coo) and every
   slot's column lies within +-32767 of lane 0's column in the same hack
   slot column shifted by the row difference (banded matrices, the
   stencils), the library also keeps each slot's column as a 16-bit
   residual plus one 32-bit base per hack slot column, and the products
   (single and multi-RHS) read those instead of ja: 2 bytes of index traffic
   per slot instead of 4, the same columns, the same results bit for bit.
   ja stays as written (read back, transposed products, scaling).  ELL
   matrices (psb_hip_ell_write, psb_hip_ell_copy_coo) get the same, with a
   base per slot column of each group of 64 rows.
   psb_hip_hll_index16: 1 when the matrix has them (psb_hip_ell_index16 for
   ELL).  psb_hip_hll_set_index16
   (0 / 1, default 1) switches their use in products on or off for every
   matrix; returns the previous setting. */
int psb_hip_hll_index16(void *dmat);
int psb_hip_ell_index16(void *dmat);
int psb_hip_hll_set_index16(int on);
/* XCD block order of the HLL (variants 0 and + 4), ELL, multi-RHS HLL and
   row-block CSR products: logical blocks dealt to the XCDs in chunks of
   `blocks` (0: one contiguous range per XCD, for the row-block CSR kernel
   launch order).  -1, the default: per matrix, a chunk of about an eighth
   of the matrix's bandwidth (max |column - row|, recorded when an HLL or
   ELL matrix is written) when that is 4 to 64 blocks, so an XCD's
   successive chunks are about one band apart (the z-neighbours of a 3-D
   stencil); 16 otherwise and for CSR.  No counterpart in the reference.
   Returns the previous value. */
int psb_hip_hll_set_xcd_chunk(int blocks);
/* Autotuner (no counterpart in the reference): times y = A*x with each HLL
   kernel variant (1, 2, 5, 6 above; reps launches each, HIP events), stores
   the fastest in the matrix (it then overrides the global variant for this
   matrix) and returns it in *best; us[4] (may be NULL) receives the average
   microseconds per product of each candidate. */
int psb_hip_hll_tune(void *dmat, void *x, void *y, int reps, int *best, float *us);

/* ===================================================================== */
/* ELL (ELG) matrices -- cuda/elldev.h, cuda/elldev_mod.F90:37-160          */
/* ===================================================================== */

/* FallocEllDevice: pitch = ceil32(rows) (cuda/spgpu/ell.c:33-37);
   maxRowSize = ELL width nzm. */
int psb_hip_ell_alloc(void **dmat, int rows, int max_row_size, int nnzeros, int elem_type,
                      int first_index);
void psb_hip_ell_free(void *dmat);
/* getEllDevicePitch / getEllDeviceMaxRowSize */
int psb_hip_ell_pitch(void *dmat);
int psb_hip_ell_max_row_size(void *dmat);
/* writeEllDevice{Double,Float}: val/ja with host leading dimension ldv
   (>= rows) and nzm columns, irn (rows), idiag (rows) */
int psb_hip_ell_write(void *dmat, const void *val, const int *ja, int ldv, const int *irn,
                      const int *idiag);
int psb_hip_ell_read(void *dmat, void *val, int *ja, int ldv, int *irn, int *idiag);
/* psiCopyCooToElg{Double,Float}: fill from a row-sorted COO, CPU padding
   rule (empty row -> first column of the whole matrix); rows beyond nr in
   the pitch are ja = 1, val = 0. idisp(nr+1) row starts (0-based). */
int psb_hip_ell_copy_coo(int nr, int nc, int nza, const int *irn, const int *idisp,
                         const int *ja, const void *val, void *dmat);
int psb_hip_ell_get_diag(void *dmat, void *d);
/* psb_d_cuda_elg_scals / _scal (cuda/impl/psb_d_cuda_elg_scals.F90, _scal.F90):
   scals the whole value array; scal left the whole row (padding included),
   right the row's irn entries (ext/impl/psb_d_ell_scal.f90:64-90). */
int psb_hip_ell_scals(void *dmat, double s);
int psb_hip_ell_scal(void *dmat, void *d, int left);
/* spmvEllDevice{Double,Float} (cuda/elldev.c:195-215) */
int psb_hip_ell_spmv(void *dmat, double alpha, void *x, double beta, void *y);

/* ===================================================================== */
/* CSR (CSRG) matrices -- cuda/psb_d_cuda_csrg_mat_mod.F90, the cuSPARSE    */
/* wrapper cuda/fcusparse_fct.h:31,108-124 replaced by hand-written kernels */
/* ===================================================================== */

/* allocate for rows, nnz; elem_type FLOAT or DOUBLE; 1-based */
int psb_hip_csr_alloc(void **dmat, int rows, int cols, int nnz, int elem_type,
                      int first_index);
void psb_hip_csr_free(void *dmat);
/* write irp (rows+1, 1-based), ja, val (nnz); builds the row-length
   analysis (lanes per row, list of long rows) used by the SpMV. */
int psb_hip_csr_write(void *dmat, const int *irp, const int *ja, const void *val);
int psb_hip_csr_read(void *dmat, int *irp, int *ja, void *val);
/* spmvCSRGDevice: y = alpha*A*x + beta*y */
int psb_hip_csr_spmv(void *dmat, double alpha, void *x, double beta, void *y);
int psb_hip_csr_get_diag(void *dmat, void *d);
/* psb_d_cuda_csrg_scals / _scal (cuda/impl/psb_d_cuda_csrg_scals.F90, _scal.F90):
   the nnz values times s; scal left / right by d(row) / d(ja)
   (base/serial/impl/psb_d_csr_impl.F90:1607-1715).  A prepared triangular
   solve is dropped. */
int psb_hip_csr_scals(void *dmat, double s);
int psb_hip_csr_scal(void *dmat, void *d, int left);
/* the lanes-per-row and long-row threshold chosen by the analysis for the
   uniform kernel */
int psb_hip_csr_get_analysis(void *dmat, int *lanes_per_row, int *long_rows,
                             int *long_threshold);
/* CSR kernel choice (no counterpart in the reference, whose CSR product is
   cuSPARSE): 0 auto (the row-block kernel, fastest on every matrix
   measured, on the column slices when the matrix has them), 1 uniform lanes per row (a power of two >= the mean length),
   2 length-binned (rows grouped by lanes per row, about four entries per
   lane), 3 row-block (runs of consecutive rows whose entries fit in LDS:
   products streamed into LDS, then one thread per row sums them in CSR
   order, so those rows equal the CPU loop's bits when beta = 0; on the
   whole matrix even when it has column slices), 4 the same as 0.  In 2, 3
   and the row-block kernel of 0 / 4, rows over 512 entries take one
   wavefront each.
   Returns the previous value. */
int psb_hip_csr_set_variant(int variant);
/* Transposed product (the reference computes op 'T' / 'C' on the host:
   cuda/impl/psb_d_cuda_hlg_vect_mv.F90 and the CSRG / ELG ones call the CPU
   spmm for trans): tmat_alloc builds the transpose of A (nr x nc, CSR
   irp/ja/val on the host, 1-based) on the device; tmat_spmv computes
   y = alpha * A^T x + beta * y with the results of psb_d_csr_csmv trans
   'T' / 'C' (base/serial/impl/psb_d_csr_impl.F90:129-141,308-370) and of
   psb_d_hll_csmv (ext/impl/psb_d_hll_csmv.f90:70-110,300-330) bit for bit:
   each y(c) is scaled by beta, then updated entry by entry in the CPU's
   order.  x has nr entries, y nc; every column of x / y. */
int psb_hip_tmat_alloc(void **tmat, int nr, int nc, const int *irp, const int *ja,
                       const void *val, int elem_type);
void psb_hip_tmat_free(void *tmat);
int psb_hip_tmat_spmv(void *tmat, double alpha, void *x, double beta, void *y);
/* Column slices of a CSR matrix (no reference counterpart): for a matrix
   whose x does not fit the L2 caches and whose columns are scattered (more
   than half of the entries of sampled rows farther than cols/16 from the
   diagonal, x over 12 MiB), psb_hip_csr_write also stores the entries split
   by consecutive column ranges of about 6.4 MB of x each (at most 16), and
   the product (variant 0 or 4) runs the slices one after another,
   accumulating into y: each slice's random gathers hit an x range that stays
   in L2.  The row sums then add the slices' partial sums, so they follow the
   per-row error bound rather than the CPU loop's bits.  Uses twice the
   matrix's memory.  set_slices: -1 automatic (default), 0 or 1 none, >= 2
   that many, applied by the next psb_hip_csr_write; returns the previous
   value.  psb_hip_csr_slices: the matrix's slice count (0: none). */
int psb_hip_csr_set_slices(int slices);
int psb_hip_csr_slices(void *dmat);
/* spsvCSRGDevice (cuda/fcusparse_fct.h:132-200, called by
   psb_d_cuda_csrg_inner_vect_sv, cuda/impl/psb_d_cuda_csrg_inner_vect_sv.F90):
   y = alpha * op(T)^-1 x + beta * y for the triangular matrix T held in the
   CSR matrix (square).  lower / unit are the matrix's triangle flags (the
   reference reads them from a%is_lower() / a%is_unit()); a non-unit row
   keeps its diagonal as its last (lower) or first (upper) entry; trans 0
   'N', 1 'T', 2 'C'.  Results equal psb_d_csr_cssv
   (base/serial/impl/psb_d_csr_impl.F90:782-1039) bit for bit, for every
   trans and beta (the reference solves on the device only for 'N' with
   beta = 0 and copies to the host otherwise).  The level analysis is built
   from the device copy of the matrix on the first call for a (lower, unit,
   transposed) combination and kept until psb_hip_csr_write or
   psb_hip_csr_free; PSB_HIP_INVALID when an entry lies outside the triangle
   or a non-unit row is empty.  x and y may be the same vector; several
   columns are solved one after another. */
int psb_hip_csr_spsv(void *dmat, int lower, int unit, int trans, double alpha, void *x,
                     double beta, void *y);
/* levels and workgroup runs of the prepared solve (0, 0 before the first) */
int psb_hip_csr_spsv_levels(void *dmat, int *levels, int *segments);

/* ===================================================================== */
/* Device-resident Krylov solvers (no reference counterpart: the reference */
/* psb_krylov, krylov/psb_dkrylov.f90:81-227, returns every dot and nrm2    */
/* to the host).  Same recurrences as psb_dcg_vect / psb_dcgstab_vect with  */
/* the DIAG preconditioner, fused into 3 (CG) or 5 (BiCGSTAB) kernels per   */
/* iteration; scalars stay on the device and the host polls the stop flag   */
/* every check_every iterations.                                            */
/* ===================================================================== */

#define PSB_HIP_FMT_HLL 0
#define PSB_HIP_FMT_ELL 1
#define PSB_HIP_FMT_CSR 2

/* method 0 = CG, 1 = BiCGSTAB; fmt = PSB_HIP_FMT_*; diag = Jacobi inverse
   diagonal (psb_hip_mv_diag_inv) or NULL for no preconditioner; b, x
   vectors of the matrix type (x holds the initial guess); istop 2
   (||r||/||b||) or 3 (||r||/||r0||); *iter and *err as psb_krylov's iter
   and err; hist (may be NULL) receives err per iteration (itmax entries). */
int psb_hip_krylov_fused(int method, int fmt, void *dmat, void *diag, void *b, void *x,
                         double eps, int itmax, int istop, int check_every, int *iter,
                         double *err, double *hist);

/* psb_hip_krylov_fused with every device buffer it needs (state, partials,
   history, work vectors) carved from a static device array of 64 MiB that
   each GPU holds until the process ends, and the host read-backs through a
   pinned block made by psb_hip_init: it calls no hipMalloc / hipFree /
   hipHostMalloc.  No counterpart in the reference: it is for callers that
   cannot allocate (code under HIP graph capture of their own, serving loops
   that must not allocate per request).  Same arguments, results and errors
   as psb_hip_krylov_fused; needs psb_hip_init first (PSB_HIP_INVALID
   otherwise) and rejects with PSB_HIP_INVALID a solve whose buffers exceed
   the array: 4 (CG) or 8 (BiCGSTAB) vectors of n elements, 8 * (itmax + 2)
   bytes of history and 512 KiB of state and partials, i.e. n up to about
   2M rows for FP64 CG and 1M for FP64 BiCGSTAB.  Every call shares the
   array: calls that overlap on one GPU (several host threads) race. */
int psb_hip_krylov_fused_noalloc(int method, int fmt, void *dmat, void *diag, void *b, void *x,
                                 double eps, int itmax, int istop, int check_every, int *iter,
                                 double *err, double *hist);

/* FP32-inner / FP64-outer iterative refinement on a double HLL matrix with a
   single copy (psb_hip_hll_make_single).  FP64: r = b - A x, x += d and the
   convergence norms; FP32: the matrix values, the inner Krylov vectors and
   the Jacobi diagonal.  variant 0 (restarted refinement): each cycle solves
   A d = r/||r|| with FP32 CG (method 0) or BiCGSTAB (1) to inner_eps
   (loosened on the last cycle to what the target needs), then x += ||r|| d.
   variant 1 (reliable updates, CG only): one FP32 CG recurrence; whenever
   its residual has dropped by inner_eps since the last replacement, the
   FP32 iterate is added to x in FP64 and r is replaced by the FP64 b - A x,
   keeping the search direction.  Stops when ||b - A x||/||b|| <= eps
   (computed in FP64).  *outer_its: cycles / replacements, *inner_its: FP32
   iterations, *err: final FP64 relative residual. */
int psb_hip_krylov_mixed(int method, void *dmat, void *diag, void *b, void *x, double eps,
                         int outer_itmax, double inner_eps, int inner_itmax, int check_every,
                         int variant, int *outer_its, int *inner_its, double *err);

/* ===================================================================== */
/* BJAC / ILU(0) preconditioner -- prec/impl/psb_d_ilu0_fact.f90,           */
/* prec/impl/psb_d_bjacprec_impl.f90 (the reference's device triangular     */
/* solve is cuSPARSE SpSV, cuda/fcusparse_fct.h:145-215)                    */
/* ===================================================================== */
/* psb_dilu0_fact (ILU_N, upd 'F'): A in CSR (1-based, ascending columns) ->
   L strict lower (unit diagonal implied), U strict upper scaled by its row's
   inverse pivot, d inverse pivots; lja/lval and uja/uval need room for nnz.
   PSB_HIP_INVALID with *bad_row = i when |pivot_i| < 1.1e-16
   (psb_err_pivot_too_small_).  Host only. */
int psb_hip_ilu0_fact(int n, const int *irp, const int *ja, const double *val, int *lirp,
                      int *lja, double *lval, int *uirp, int *uja, double *uval, double *d,
                      int *bad_row);
/* psb_silu0_fact (prec/impl/psb_s_ilu0_fact.f90): the same in single
   precision, pivot threshold 5e-8 (s_epstol). */
int psb_hip_ilu0_fact_s(int n, const int *irp, const int *ja, const float *val, int *lirp,
                        int *lja, float *lval, int *uirp, int *uja, float *uval, float *d,
                        int *bad_row);
/* the factors on the device, levelled for the triangular solves */
int psb_hip_ilu_alloc(void **ilu, int n, const int *lirp, const int *lja, const double *lval,
                      const int *uirp, const int *uja, const double *uval, const double *d);
int psb_hip_ilu_free(void *ilu);
int psb_hip_ilu_levels(void *ilu, int *lower_levels, int *upper_levels);
/* runs of small levels solved by one workgroup (0 when the variant is 1) */
int psb_hip_ilu_segments(void *ilu, int *lower_segments, int *upper_segments);
/* triangular-solve kernels for factors allocated afterwards: 0 (default) runs
   of small levels by one workgroup each, other levels one launch each; 1 one
   launch per level.  Same results either way.  Returns the previous value. */
int psb_hip_ilu_set_variant(int variant);
/* psb_d_bjac_apply_vect (ILU, trans 'N'): y = alpha * U^-1 (d .* L^-1 x) + beta * y */
int psb_hip_ilu_apply(void *ilu, double alpha, void *x, double beta, void *y);

/* ===================================================================== */
/* Batched mixed-precision solves -- no counterpart in the reference.       */
/* Many independent systems (each an HLL double matrix registered once,     */
/* with its FP32 copy and optional Jacobi diagonal) advance together: a     */
/* step runs check_every FP32 iterations of every live system and one FP64  */
/* transition pass (method 0: CG with FP64 residual replacement when the    */
/* recursive residual drops by delta; method 1: BiCGSTAB restarted          */
/* refinement with inner tolerance inner_eps; method 2: BiCGSTAB with FP64  */
/* residual replacement as method 0), as one captured graph when            */
/* use_graph != 0.  Systems are admitted (submit) and retired (result)      */
/* between steps; the vectors of a system live in the engine's arena of     */
/* arena_bytes, and submit returns PSB_HIP_BUSY when no slot or arena room  */
/* is free.  A system's results do not depend on the other systems.         */
/* Per-iteration sequence: krylov/psb_dcg.F90:241-294 and                   */
/* krylov/psb_dcgstab.f90:248-364; stop test ||b - A x|| <= eps ||b|| in    */
/* FP64 (krylov/psb_d_krylov_conv_mod.f90:279-356, istop 2).                */
/* ===================================================================== */
int psb_hip_batch_create(void **batch, int method, int max_systems, int64_t arena_bytes,
                         int check_every, double delta, double inner_eps, int use_graph);
int psb_hip_batch_free(void *batch);
/* dmat: HLL double matrix (square); diag: double vector of 1/a_ii (as
   psb_hip_mv_diag_inv leaves it) or NULL.  The engine keeps pointers to
   both: they must outlive the engine. */
int psb_hip_batch_register(void *batch, void *dmat, void *diag, int *mat_id);
/* arena bytes one solve on matrix mat_id occupies (-1 for a bad id) */
int64_t psb_hip_batch_footprint(void *batch, int mat_id);
/* b (double vector, copied), x0 (double vector, copied) or NULL for zero;
   *sys_id receives the slot.  Takes effect at the next step. */
int psb_hip_batch_submit(void *batch, int mat_id, void *b, void *x0, double eps, int itmax,
                         int *sys_id);
/* 1 when a solve on matrix mat_id would be admitted now (a free slot and
   arena room), else 0 */
int psb_hip_batch_can_admit(void *batch, int mat_id);
/* queue one step (tables first when the live set changed) */
int psb_hip_batch_step(void *batch);
/* Systems finished as of the step before the last one queued (wait_last:
   the last one), up to max_out: slot ids, FP32 iterations, FP64
   transitions, final ||b - A x||/||b||, status (1 converged, 2 iteration
   limit).  *nlive: systems still running. */
int psb_hip_batch_poll(void *batch, int wait_last, int max_out, int *ids, int *iters,
                       int *outer, double *relres, int *status, int *nout, int *nlive);
/* copy a finished system's solution into x (double vector, may be NULL) and
   free its slot */
int psb_hip_batch_result(void *batch, int sys_id, void *x);

#ifdef __cplusplus
}
#endif

#endif /* PSB_HIP_H */
