/*
 * psb_oracle.c -- CPU restatement of PSBLAS 3.9's CPU path for the device
 * SpMV + Krylov vector path.  TEST INFRASTRUCTURE ONLY: it is the checker
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg.  Nothing in the product (psblas3_amd/) links, loads or calls it.
 *
 * Every function names the reference file:line it restates.  Arithmetic is
 * compiled with -ffp-contract=off so each operation is separately rounded,
 * as the reference's Fortran is on x86-64 (SSE2, no FMA contraction).
 *
 * Conventions: all index arrays are int32 and 1-based, as PSBLAS stores
 * them (psb_ipk_ = 4 bytes; device baseIndex = 1,
 * cuda/impl/psb_d_cuda_hlg_to_gpu.F90:56).  Dense 2-D arrays are
 * column-major (Fortran order).
 *
 * How each part is pinned (see tests/test_oracle_*.py):
 *   - conversions: closed forms on the stencil, brute-force dense rebuild,
 *     hand-derived tiny cases, and an independent numpy restatement;
 *   - SpMV: torture 2x2 goldens (test/torture/psb_d_mvsv_tester.f90),
 *     dense brute force, exact integer stencil cases, scipy.sparse;
 *   - BLAS-1: vecoperation known answers (test/kernel/vecoperation.f90),
 *     math.fsum error bounds;
 *   - Krylov: reference run of psb_d_pde3d L=40 BiCGSTAB+DIAG (80 its,
 *     err 9.197e-7), scipy.sparse.linalg solutions, CG invariants;
 *     FCG / GCR / RGMRES / BiCGSTAB(l): identities with CG and GMRES on
 *     cases where they coincide, brute-force minimal residual over the
 *     Krylov space for GCR / RGMRES, direct solves; CGS / BiCG: scipy's
 *     Jacobi-preconditioned cgs / bicg iteration counts, BiCG = CG on SPD;
 *   - transposed CSR product: the torture usmv cases with trans 't' / 'c',
 *     scipy A^T x, an exact integer case;
 *   - ILU(0) / BJAC: the defining property (L U equals A on A's pattern),
 *     exact LU on tridiagonal matrices, dense triangular solves (scipy);
 *     the library's host factorization equals it bit for bit; the FP32
 *     factorization by the same property at FP32 accuracy and against the
 *     FP64 factors;
 *   - CSR triangular solve (psb_d_csr_cssv): the torture SpSV known answers
 *     (test/torture/psb_{d,s}_mvsv_tester.f90), scipy dense triangular
 *     solves for every triangle / diagonal / op, exact integer solutions.
 *   - nrm2: netlib reference-BLAS scaled algorithm; the reference links an
 *     external BLAS (MKL here), whose summation order is not published:
 *     parity unpinned beyond the error bound.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Counter-based RNG shared by every generator (splitmix64 finaliser).  */
/* ------------------------------------------------------------------ */
static inline uint64_t orc_mix(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* uniform [0,1) with 53 random bits */
EXPORT double orc_uniform01(uint64_t seed, uint32_t stream, uint64_t i)
{
  uint64_t key = orc_mix(seed * 0xD1B54A32D192ED03ULL ^ ((uint64_t)stream << 48));
  uint64_t h = orc_mix(key ^ i);
  return (double)(h >> 11) * 0x1.0p-53;
}

/* uniform [-1,1) */
EXPORT double orc_uniform_pm1(uint64_t seed, uint32_t stream, uint64_t i)
{
  return 2.0 * orc_uniform01(seed, stream, i) - 1.0;
}

EXPORT void orc_fill_uniform_pm1(uint64_t seed, uint32_t stream, int64_t n, double *x)
{
  for (int64_t i = 0; i < n; i++) x[i] = orc_uniform_pm1(seed, stream, (uint64_t)i);
}

/* ------------------------------------------------------------------ */
/* Generators                                                          */
/* ------------------------------------------------------------------ */

/* ijk2idx, column-major (ix fastest): util/psb_partidx_mod.F90:377-410
   with col_major=.true. (:61). 1-based. */
static inline int orc_ijk2idx(int ix, int iy, int iz, int nx, int ny)
{
  return ((iz - 1) * ny + (iy - 1)) * nx + (ix - 1) + 1;
}

/* Boundary function g of the pde3d sample: test/pargen/psb_d_pde3d.F90:160-170 */
static double orc_g_sample(double x, double y, double z)
{
  double g = 0.0;
  if (x == 1.0) g = 1.0;
  else if (x == 0.0) g = exp(y * y - z * z);
  (void)z;
  return g;
}

/*
 * 7-point finite-difference operator on an L^3 grid with constant
 * coefficients, entries inserted in the order of the reference assembly
 * loop test/pargen/psb_d_pde3d.F90:478-560 (x-1, y-1, z-1, diag, z+1, y+1,
 * x+1), with deltah = 1/(L+1) (:249).  Off-diagonals:
 *   -a_i/h^2 - b_i/(2h) towards -i, -a_i/h^2 + b_i/(2h) towards +i,
 * diagonal (2*1)*(a1+a2+a3)/h^2 + c.
 * If hscale == 0 the h factors are dropped (h := 1), giving for instance the
 * SPD Poisson (6,-1) operator with a=1, b=0, c=0.
 * gflag: 0 -> rhs zero; 1 -> the sample boundary function g folded into
 * the rhs exactly as :490-556 do.
 * Returns the number of entries written (ia/ja/val may be NULL to count).
 */
EXPORT int64_t orc_gen_pde3d(int L, const double *a, const double *b, double c,
                             int hscale, int gflag,
                             int *ia, int *ja, double *val, double *rhs)
{
  double deltah = hscale ? 1.0 / (L + 1) : 1.0;
  double sqdeltah = deltah * deltah;
  double deltah2 = (2.0) * deltah;
  int64_t k = 0;
  double vxm = -a[0] / sqdeltah - b[0] / deltah2;
  double vym = -a[1] / sqdeltah - b[1] / deltah2;
  double vzm = -a[2] / sqdeltah - b[2] / deltah2;
  double vd = (2.0) * (a[0] + a[1] + a[2]) / sqdeltah + c;
  double vzp = -a[2] / sqdeltah + b[2] / deltah2;
  double vyp = -a[1] / sqdeltah + b[1] / deltah2;
  double vxp = -a[0] / sqdeltah + b[0] / deltah2;
  double gh = 1.0 / (L + 1); /* g is evaluated on the true grid */
  for (int iz = 1; iz <= L; iz++)
    for (int iy = 1; iy <= L; iy++)
      for (int ix = 1; ix <= L; ix++) {
        int row = orc_ijk2idx(ix, iy, iz, L, L);
        double x = (ix - 1) * gh, y = (iy - 1) * gh, z = (iz - 1) * gh;
        double zt = 0.0;
#define ORC_PUT(COL, V)                     \
  do {                                      \
    if (ia) { ia[k] = row; ja[k] = (COL); val[k] = (V); } \
    k++;                                    \
  } while (0)
        if (ix == 1) { if (gflag) zt = orc_g_sample(0.0, y, z) * (-vxm) + zt; }
        else ORC_PUT(orc_ijk2idx(ix - 1, iy, iz, L, L), vxm);
        if (iy == 1) { if (gflag) zt = orc_g_sample(x, 0.0, z) * (-vym) + zt; }
        else ORC_PUT(orc_ijk2idx(ix, iy - 1, iz, L, L), vym);
        if (iz == 1) { if (gflag) zt = orc_g_sample(x, y, 0.0) * (-vzm) + zt; }
        else ORC_PUT(orc_ijk2idx(ix, iy, iz - 1, L, L), vzm);
        ORC_PUT(row, vd);
        if (iz == L) { if (gflag) zt = orc_g_sample(x, y, 1.0) * (-vzp) + zt; }
        else ORC_PUT(orc_ijk2idx(ix, iy, iz + 1, L, L), vzp);
        if (iy == L) { if (gflag) zt = orc_g_sample(x, 1.0, z) * (-vyp) + zt; }
        else ORC_PUT(orc_ijk2idx(ix, iy + 1, iz, L, L), vyp);
        if (ix == L) { if (gflag) zt = orc_g_sample(1.0, y, z) * (-vxp) + zt; }
        else ORC_PUT(orc_ijk2idx(ix + 1, iy, iz, L, L), vxp);
        if (rhs) rhs[row - 1] = zt;
      }
#undef ORC_PUT
  return k;
}

/*
 * 2-D 5-point anisotropic operator -eps*u_xx - u_yy on an L^2 grid (ix
 * fastest), unscaled: diagonal 2*eps + 2, x-neighbours -eps, y-neighbours -1.
 * Insertion order x-1, y-1, diag, y+1, x+1 (the 2-D analogue of the pde3d
 * loop, test/pargen/psb_d_pde2d.F90).
 */
EXPORT int64_t orc_gen_aniso2d(int L, double eps, int *ia, int *ja, double *val)
{
  int64_t k = 0;
  double vd = 2.0 * eps + 2.0;
  for (int iy = 1; iy <= L; iy++)
    for (int ix = 1; ix <= L; ix++) {
      int row = (iy - 1) * L + ix;
#define ORC_PUT(COL, V) do { if (ia) { ia[k] = row; ja[k] = (COL); val[k] = (V); } k++; } while (0)
      if (ix > 1) ORC_PUT(row - 1, -eps);
      if (iy > 1) ORC_PUT(row - L, -1.0);
      ORC_PUT(row, vd);
      if (iy < L) ORC_PUT(row + L, -1.0);
      if (ix < L) ORC_PUT(row + 1, -eps);
#undef ORC_PUT
    }
  return k;
}

/*
 * Power-law sparse matrix: row i (1-based) has k_i = min(cap, n-1,
 * floor(c / v_i)) off-diagonal entries, v_i = 1 - U[0,1) in (0,1]
 * (P(K >= k) = c/k: a Zipf tail of exponent 2), columns drawn uniformly
 * without repetition from [1,n] \ {i}, plus the diagonal entry first.
 * Values U[-1,1).  c = 1.775, cap = 4095 gives a mean row length of 16.0.
 * Streams: 1 = row length, 2 = column draws, 3 = values.
 */
#define ORC_PL_C 1.775
EXPORT int orc_powerlaw_rowlen(int n, int cap, uint64_t seed, int i1)
{
  double v = 1.0 - orc_uniform01(seed, 1, (uint64_t)(i1 - 1));
  double kd = floor(ORC_PL_C / v);
  int kmax = cap < n - 1 ? cap : n - 1;
  int k = kd > (double)kmax ? kmax : (int)kd;
  return k + 1; /* + diagonal */
}

EXPORT int64_t orc_gen_powerlaw(int n, int cap, uint64_t seed, int *ia, int *ja, double *val)
{
  int64_t k = 0;
  int *seen = NULL;
  int seen_cap = 0;
  for (int i = 1; i <= n; i++) {
    int len = orc_powerlaw_rowlen(n, cap, seed, i);
    if (!ia) { k += len; continue; }
    /* simple open-addressing set for the drawn columns */
    int hs = 16;
    while (hs < 2 * len) hs <<= 1;
    if (hs > seen_cap) { free(seen); seen = (int *)malloc(sizeof(int) * hs); seen_cap = hs; }
    for (int t = 0; t < hs; t++) seen[t] = 0;
    uint64_t base = (uint64_t)(i - 1) << 20;
    ia[k] = i; ja[k] = i; val[k] = orc_uniform_pm1(seed, 3, base); k++;
    int got = 1;
    uint64_t attempt = 0;
    while (got < len) {
      double u = orc_uniform01(seed, 2, base + attempt);
      attempt++;
      int col = 1 + (int)(u * (double)n);
      if (col > n) col = n;
      if (col == i) continue;
      unsigned hsh = ((unsigned)col * 2654435761u) & (unsigned)(hs - 1);
      int dup = 0;
      while (seen[hsh]) {
        if (seen[hsh] == col) { dup = 1; break; }
        hsh = (hsh + 1) & (unsigned)(hs - 1);
      }
      if (dup) continue;
      seen[hsh] = col;
      ia[k] = i; ja[k] = col; val[k] = orc_uniform_pm1(seed, 3, base + (uint64_t)got); k++;
      got++;
    }
  }
  free(seen);
  return k;
}

/* ------------------------------------------------------------------ */
/* COO fix and format conversions                                      */
/* ------------------------------------------------------------------ */

typedef struct { int r, c; int64_t pos; } orc_key_t;

static int orc_key_cmp(const void *pa, const void *pb)
{
  const orc_key_t *a = (const orc_key_t *)pa, *b = (const orc_key_t *)pb;
  if (a->r != b->r) return a->r < b->r ? -1 : 1;
  if (a->c != b->c) return a->c < b->c ? -1 : 1;
  return a->pos < b->pos ? -1 : (a->pos > b->pos);
}

/*
 * psb_d_fix_coo_inner_rowmajor, serial build
 * (base/serial/impl/psb_d_coo_impl.F90:4252-4896): a stable bucket of the
 * entries by row (:4572-4588), a stable merge sort of the columns inside
 * each row (psi_msort_up, :4641-4645), then duplicates merged in that
 * order -- dupl=add (1) sums left to right (:4652-4663), dupl=ovwrt (0)
 * keeps the last (:4603-4617), dupl=err (2) returns -1.  Stability is
 * obtained here by sorting on (row, col, input position).  Returns nzout.
 */
EXPORT int64_t orc_fix_coo(int nr, int nc, int64_t nz, int *ia, int *ja, double *val, int dupl)
{
  (void)nc;
  if (nz < 2) return nz;
  orc_key_t *key = (orc_key_t *)malloc(sizeof(orc_key_t) * nz);
  double *vs = (double *)malloc(sizeof(double) * nz);
  int rows_ok = nr > 0;
  for (int64_t i = 0; i < nz && rows_ok; i++) rows_ok = ia[i] >= 1 && ia[i] <= nr;
  if (rows_ok) {
    /* the same total order (row, column, input position) as one sort of the
       keys: a stable counting sort by row, then each row sorted alone */
    int64_t *start = (int64_t *)calloc((size_t)nr + 1, sizeof(int64_t));
    for (int64_t i = 0; i < nz; i++) start[ia[i]]++;
    for (int r = 0; r < nr; r++) start[r + 1] += start[r];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nr + 1));
    memcpy(fill, start, sizeof(int64_t) * ((size_t)nr + 1));
    for (int64_t i = 0; i < nz; i++) {
      orc_key_t *k = &key[fill[ia[i] - 1]++];
      k->r = ia[i]; k->c = ja[i]; k->pos = i;
    }
    #pragma omp parallel for schedule(dynamic, 4096)
    for (int r = 0; r < nr; r++) {
      int64_t b = start[r], e = start[r + 1];
      if (e - b > 32) {
        qsort(key + b, (size_t)(e - b), sizeof(orc_key_t), orc_key_cmp);
      } else {
        for (int64_t i = b + 1; i < e; i++) {   /* insertion sort of a short row */
          orc_key_t t = key[i];
          int64_t j = i - 1;
          while (j >= b && orc_key_cmp(&key[j], &t) > 0) { key[j + 1] = key[j]; j--; }
          key[j + 1] = t;
        }
      }
    }
    free(start); free(fill);
  } else {
    for (int64_t i = 0; i < nz; i++) { key[i].r = ia[i]; key[i].c = ja[i]; key[i].pos = i; }
    qsort(key, (size_t)nz, sizeof(orc_key_t), orc_key_cmp);
  }
  for (int64_t i = 0; i < nz; i++) vs[i] = val[key[i].pos];
  int64_t k = 0;
  ia[0] = key[0].r; ja[0] = key[0].c; val[0] = vs[0];
  for (int64_t i = 1; i < nz; i++) {
    if (key[i].r == ia[k] && key[i].c == ja[k]) {
      if (dupl == 1) val[k] = val[k] + vs[i];
      else if (dupl == 0) val[k] = vs[i];
      else { free(key); free(vs); return -1; }
    } else {
      k++;
      ia[k] = key[i].r; ja[k] = key[i].c; val[k] = vs[i];
    }
  }
  free(key); free(vs);
  return k + 1;
}

/* Row counts of a row-sorted COO (shared by :4094 fix, the CSR, ELL and HLL
   conversions). */
EXPORT void orc_row_counts(int nr, int64_t nz, const int *ia, int *irn)
{
  for (int i = 0; i < nr; i++) irn[i] = 0;
  for (int64_t k = 0; k < nz; k++) irn[ia[k] - 1]++;
}

/* psb_d_cp_csr_from_coo (base/serial/impl/psb_d_csr_impl.F90:3154-3230):
   irp(1)=1, irp(i+1) = irp(i) + count(i). */
EXPORT void orc_csr_from_coo(int nr, int64_t nz, const int *ia, int *irp)
{
  int *cnt = (int *)calloc((size_t)nr, sizeof(int));
  orc_row_counts(nr, nz, ia, cnt);
  irp[0] = 1;
  for (int i = 0; i < nr; i++) irp[i + 1] = irp[i] + cnt[i];
  free(cnt);
}

/*
 * psi_d_xtr_ell_from_coo (ext/impl/psi_d_xtr_ell_from_coo.f90:31-62):
 * copies nrr rows (row-sorted COO slice iac/jac/valc) into a column-major
 * block ja/val with leading dimension ld and width mxrwl; pads each row
 * with its last column (or jac(1) of the slice if the row is empty) and a
 * zero value; diag(ii) = 1-based slot of the diagonal (left untouched when
 * the row has none).
 */
static void orc_xtr_ell(int i0, int nrr, int mxrwl, const int *iac, const int *jac,
                        const double *valc, int *ja, double *val, const int *irn,
                        int *diag, int ld)
{
  int64_t kc = 0;
  (void)i0;
  for (int ii = 0; ii < nrr; ii++) {
    int ncr = irn[ii];
    int ic = 0;
    for (int jj = 0; jj < ncr; jj++) {
      int ir = iac[kc];
      ic = jac[kc];
      if (ir == ic) diag[ii] = jj + 1;
      ja[ii + (int64_t)jj * ld] = ic;
      val[ii + (int64_t)jj * ld] = valc[kc];
      kc++;
    }
    if (ncr == 0) ic = jac[0];
    for (int jj = ncr; jj < mxrwl; jj++) {
      ja[ii + (int64_t)jj * ld] = ic;
      val[ii + (int64_t)jj * ld] = 0.0;
    }
  }
}

/*
 * HLL hack offsets: psi_d_convert_hll_from_coo
 * (ext/impl/psi_d_convert_hll_from_coo.f90:62-82).  hkoffs has
 * noffs+1 = ceil(nr/hksz)+1 entries, hkoffs[0] = 0 and
 * hkoffs[j+1] = hkoffs[j] + hksz * max(irn in hack j).  Returns isz.
 */
EXPORT int64_t orc_hll_offsets(int nr, int hksz, const int *irn, int *hkoffs)
{
  int noffs = (nr + hksz - 1) / hksz;
  hkoffs[0] = 0;
  int j = 0;
  for (int i = 0; i < nr; i += hksz) {
    int ir = hksz < nr - i ? hksz : nr - i;
    int mx = irn[i];
    for (int k = 1; k < ir; k++) if (irn[i + k] > mx) mx = irn[i + k];
    hkoffs[j + 1] = hkoffs[j] + mx * hksz;
    j++;
  }
  return noffs > 0 ? hkoffs[noffs] : 0;
}

/*
 * HLL fill: psi_d_convert_hll_from_coo :87-116.  The last hack is
 * pre-filled with ja = nr, val = 0 (:93-96), then every hack is extracted
 * with xtr_ell at ld = hksz.  idiag must be pre-initialised by the caller
 * (the reference leaves rows without a diagonal untouched).
 */
EXPORT void orc_hll_fill(int nr, int64_t nz, const int *ia, const int *jac, const double *valc,
                         int hksz, const int *hkoffs, const int *irn,
                         int *ja, double *val, int *idiag)
{
  int noffs = (nr + hksz - 1) / hksz;
  if (noffs == 0) return;
  int64_t isz = hkoffs[noffs];
  int64_t nzm = hkoffs[noffs] - hkoffs[noffs - 1];
  for (int64_t t = isz - nzm; t < isz; t++) { val[t] = 0.0; ja[t] = nr; }
  int64_t kc = 0;
  (void)nz;
  for (int i = 0; i < nr; i += hksz) {
    int ir = hksz < nr - i ? hksz : nr - i;
    int hk = i / hksz;
    int64_t sz = hkoffs[hk + 1] - hkoffs[hk];
    int mxrwl = (int)(sz / hksz);
    int64_t nza = 0;
    for (int r = 0; r < ir; r++) nza += irn[i + r];
    orc_xtr_ell(i + 1, ir, mxrwl, ia + kc, jac + kc, valc + kc,
                ja + hkoffs[hk], val + hkoffs[hk], irn + i, idiag + i, hksz);
    kc += nza;
  }
}

/*
 * ELL: psi_d_convert_ell_from_coo (ext/impl/psi_d_convert_ell_from_coo.f90:31-86)
 * ldv = ceil(nr/hsz)*hsz (hsz=1 on the CPU; the device pitch is ceil32,
 * cuda/spgpu/ell.c:33-37), nzm = max row length, one xtr_ell over all rows.
 * Rows nr+1..ldv are left untouched (the caller zero-initialises them).
 */
EXPORT int orc_ell_nzm(int nr, const int *irn)
{
  int nzm = 0;
  for (int i = 0; i < nr; i++) if (irn[i] > nzm) nzm = irn[i];
  return nzm;
}

EXPORT void orc_ell_fill(int nr, const int *ia, const int *jac, const double *valc,
                         int ldv, int nzm, const int *irn, int *ja, double *val, int *idiag)
{
  orc_xtr_ell(1, nr, nzm, ia, jac, valc, ja, val, irn, idiag, ldv);
}

/* ------------------------------------------------------------------ */
/* SpMV / SpMM, double and single precision                            */
/* ------------------------------------------------------------------ */

/*
 * psb_d_csr_csmv, no-transpose (base/serial/impl/psb_d_csr_impl.F90:46-417,
 * inner loops :133-300): acc = sum_j val(j)*x(ja(j)) in ascending j, then
 * y = beta*y + alpha*acc, with beta == 0 meaning y is not read and
 * alpha == 0 meaning y = beta*y.  The +-1 special cases of the reference
 * give the same values (multiplication by +-1 is exact).
 */
#define ORC_CSR_CSMV(NAME, T)                                                   \
EXPORT void NAME(int nr, const int *irp, const int *ja, const T *val,            \
                 T alpha, const T *x, T beta, T *y)                              \
{                                                                                \
  _Pragma("omp parallel for schedule(static)")                                   \
  for (int i = 0; i < nr; i++) {                                                 \
    if (alpha == (T)0) { y[i] = beta == (T)0 ? (T)0 : beta * y[i]; continue; }   \
    T acc = 0;                                                                   \
    for (int j = irp[i] - 1; j < irp[i + 1] - 1; j++) acc = acc + val[j] * x[ja[j] - 1]; \
    if (beta == (T)0) y[i] = alpha * acc;                                        \
    else y[i] = beta * y[i] + alpha * acc;                                       \
  }                                                                              \
}
ORC_CSR_CSMV(orc_d_csr_csmv, double)
ORC_CSR_CSMV(orc_s_csr_csmv, float)

/*
 * psb_d_csr_csmv, trans 'T' / 'C' (base/serial/impl/psb_d_csr_impl.F90:129-141,
 * 308-370; identical for real data): alpha == 0 gives y = beta*y; otherwise
 * y = beta*y (0 -> 0, 1 -> unchanged, -1 -> -y), then for the rows i in
 * ascending order and their entries in CSR order y(ja) = y(ja) + val*x(i)
 * (alpha == 1), y(ja) - val*x(i) (alpha == -1), else y(ja) + alpha*val*x(i).
 * The CPU HLL (ext/impl/psb_d_hll_csmv.f90:70-110,300-330) applies the same
 * updates in the same order (rows ascending, entries in stored order).
 * nr rows of A (x has nr entries), nc columns (y has nc).
 */
#define ORC_CSR_CSMV_T(NAME, T)                                                 \
EXPORT void NAME(int nr, int nc, const int *irp, const int *ja, const T *val,  \
                 T alpha, const T *x, T beta, T *y)                             \
{                                                                               \
  if (alpha == (T)0) {                                                          \
    for (int c = 0; c < nc; c++) y[c] = beta == (T)0 ? (T)0 : beta * y[c];      \
    return;                                                                     \
  }                                                                             \
  for (int c = 0; c < nc; c++)                                                  \
    y[c] = beta == (T)0 ? (T)0 : (beta == (T)1 ? y[c] : (beta == (T)-1 ? -y[c] : beta * y[c])); \
  for (int i = 0; i < nr; i++)                                                  \
    for (int j = irp[i] - 1; j < irp[i + 1] - 1; j++) {                         \
      const int c = ja[j] - 1;                                                  \
      if (alpha == (T)1) y[c] = y[c] + val[j] * x[i];                           \
      else if (alpha == (T)-1) y[c] = y[c] - val[j] * x[i];                     \
      else y[c] = y[c] + alpha * val[j] * x[i];                                 \
    }                                                                           \
}
ORC_CSR_CSMV_T(orc_d_csr_csmv_t, double)
ORC_CSR_CSMV_T(orc_s_csr_csmv_t, float)

/*
 * psb_d_ell_csmv, no-transpose (ext/impl/psb_d_ell_csmv.f90:33-433): the
 * loop runs over the full padded width nzm for every row (padding adds
 * 0*x), then y = alpha*acc + beta*y.
 */
#define ORC_ELL_CSMV(NAME, T)                                                   \
EXPORT void NAME(int nr, int ldv, int nzm, const int *ja, const T *val,          \
                 T alpha, const T *x, T beta, T *y)                              \
{                                                                                \
  _Pragma("omp parallel for schedule(static)")                                   \
  for (int i = 0; i < nr; i++) {                                                 \
    if (alpha == (T)0) { y[i] = beta == (T)0 ? (T)0 : beta * y[i]; continue; }   \
    T acc = 0;                                                                   \
    for (int j = 0; j < nzm; j++)                                                \
      acc = acc + val[i + (int64_t)j * ldv] * x[ja[i + (int64_t)j * ldv] - 1];   \
    if (beta == (T)0) y[i] = alpha * acc;                                        \
    else y[i] = alpha * acc + beta * y[i];                                       \
  }                                                                              \
}
ORC_ELL_CSMV(orc_d_ell_csmv, double)
ORC_ELL_CSMV(orc_s_ell_csmv, float)

/*
 * psb_d_hll_csmv, no-transpose (ext/impl/psb_d_hll_csmv.f90:110-275).
 * With psi_get_hll_vector() true (the default, ext/psi_i_ext_util_mod.f90:41)
 * and hksz in {4,8,16,24,32}, each full hack with mxrwl > 0 runs
 * notra_<hksz> (:487-523): tmp += val(:,j)*x(ja(:,j)) for j=1..mxrwl over
 * all hksz rows (padding included), y = alpha*tmp (+ beta*y).  A full hack
 * with mxrwl == 0 is skipped, leaving y untouched (reference quirk, kept
 * here when quirk_empty_hack != 0).  Other hack sizes and the remainder
 * hack run csmv_inner (:291-371): per row up to irn(i), no padding.
 */
#define ORC_HLL_CSMV(NAME, T)                                                    \
static void NAME##_inner(int i0, int m, const int *irn, T alpha, const int *ja,   \
                         const T *val, int ld, const T *x, T beta, T *y)          \
{                                                                                 \
  for (int i = 0; i < m; i++) {                                                   \
    if (alpha == (T)0) { y[i0 + i] = beta == (T)0 ? (T)0 : beta * y[i0 + i]; continue; } \
    T tmp = 0;                                                                    \
    for (int j = 0; j < irn[i]; j++) tmp = tmp + val[i + (int64_t)j * ld] * x[ja[i + (int64_t)j * ld] - 1]; \
    if (beta == (T)0) y[i0 + i] = alpha * tmp;                                    \
    else y[i0 + i] = alpha * tmp + beta * y[i0 + i];                              \
  }                                                                               \
}                                                                                 \
EXPORT void NAME(int nr, int hksz, const int *hkoffs, const int *irn,             \
                 const int *ja, const T *val, T alpha, const T *x, T beta, T *y,  \
                 int quirk_empty_hack)                                            \
{                                                                                 \
  int vec = (hksz == 4 || hksz == 8 || hksz == 16 || hksz == 24 || hksz == 32);  \
  int mmhk = (nr / hksz) * hksz;                                                  \
  int nfull = mmhk / hksz;                                                        \
  _Pragma("omp parallel for schedule(static)")                                    \
  for (int hk = 0; hk < nfull; hk++) {                                            \
    int i0 = hk * hksz;                                                           \
    int mxrwl = (hkoffs[hk + 1] - hkoffs[hk]) / hksz;                             \
    const int *hja = ja + hkoffs[hk];                                             \
    const T *hval = val + hkoffs[hk];                                             \
    if (mxrwl == 0 && quirk_empty_hack) continue;                                 \
    if (vec && mxrwl > 0) {                                                       \
      T tmp[32];                                                                  \
      for (int r = 0; r < hksz; r++) tmp[r] = 0;                                  \
      if (alpha != (T)0)                                                          \
        for (int j = 0; j < mxrwl; j++)                                           \
          for (int r = 0; r < hksz; r++)                                          \
            tmp[r] = tmp[r] + hval[r + (int64_t)j * hksz] * x[hja[r + (int64_t)j * hksz] - 1]; \
      for (int r = 0; r < hksz; r++) {                                            \
        if (beta == (T)0) y[i0 + r] = alpha * tmp[r];                             \
        else y[i0 + r] = alpha * tmp[r] + beta * y[i0 + r];                       \
      }                                                                           \
    } else {                                                                      \
      NAME##_inner(i0, hksz, irn + i0, alpha, hja, hval, hksz, x, beta, y);      \
    }                                                                             \
  }                                                                               \
  if (mmhk < nr) {                                                                \
    int hk = nfull;                                                               \
    int mxrwl = (hkoffs[hk + 1] - hkoffs[hk]) / hksz;                             \
    if (mxrwl > 0 || !quirk_empty_hack)                                           \
      NAME##_inner(mmhk, nr - mmhk, irn + mmhk, alpha, ja + hkoffs[hk],           \
                   val + hkoffs[hk], hksz, x, beta, y);                           \
  }                                                                               \
}
ORC_HLL_CSMV(orc_d_hll_csmv, double)
ORC_HLL_CSMV(orc_s_hll_csmv, float)

/*
 * psb_d_hll_csmm, no-transpose (ext/impl/psb_d_hll_csmm.f90:33-235):
 * per hack, per row, acc(1:nc) += val(i,j)*x(ja(i,j),1:nc) for j=1..irn(i)
 * (psb_d_hll_csmm_inner), y(i,:) = alpha*acc + beta*y(i,:).  x and y are
 * column-major with leading dimensions ldx, ldy.
 */
EXPORT void orc_d_hll_csmm(int nr, int hksz, const int *hkoffs, const int *irn,
                           const int *ja, const double *val, int k,
                           double alpha, const double *x, int ldx,
                           double beta, double *y, int ldy)
{
  int noffs = (nr + hksz - 1) / hksz;
  _Pragma("omp parallel for schedule(static)")
  for (int hk = 0; hk < noffs; hk++) {
    int i0 = hk * hksz;
    int m = hksz < nr - i0 ? hksz : nr - i0;
    const int *hja = ja + hkoffs[hk];
    const double *hval = val + hkoffs[hk];
    double acc[64];
    for (int i = 0; i < m; i++) {
      for (int c0 = 0; c0 < k; c0 += 64) {
        int cc = k - c0 < 64 ? k - c0 : 64;
        for (int c = 0; c < cc; c++) acc[c] = 0.0;
        for (int j = 0; j < irn[i0 + i]; j++) {
          double v = hval[i + (int64_t)j * hksz];
          int col = hja[i + (int64_t)j * hksz] - 1;
          for (int c = 0; c < cc; c++) acc[c] = acc[c] + v * x[col + (int64_t)(c0 + c) * ldx];
        }
        for (int c = 0; c < cc; c++) {
          double *yy = &y[i0 + i + (int64_t)(c0 + c) * ldy];
          if (beta == 0.0) *yy = alpha * acc[c];
          else *yy = alpha * acc[c] + beta * (*yy);
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* BLAS-1 (vector path)                                                */
/* ------------------------------------------------------------------ */

/* psi_daxpbyv (base/serial/psi_d_serial_impl.F90:593-745): y = alpha*x + beta*y
   with alpha == 0 -> y = beta*y (beta == 0 -> 0, y not read) and beta == 0 ->
   y = alpha*x.  The +-1 branches give the same values. */
#define ORC_AXPBY(NAME, T)                                                     \
EXPORT void NAME(int m, T alpha, const T *x, T beta, T *y)                      \
{                                                                               \
  _Pragma("omp parallel for schedule(static) if(m > 65536)")                    \
  for (int i = 0; i < m; i++) {                                                 \
    if (alpha == (T)0) y[i] = beta == (T)0 ? (T)0 : (beta == (T)1 ? y[i] : beta * y[i]); \
    else if (beta == (T)0) y[i] = alpha == (T)1 ? x[i] : alpha * x[i];          \
    else y[i] = alpha * x[i] + beta * y[i];                                     \
  }                                                                             \
}
ORC_AXPBY(orc_d_axpby, double)
ORC_AXPBY(orc_s_axpby, float)

/* psi_d_upd_xyz (psi_d_serial_impl.F90:1571-1766): y = alpha*x + beta*y;
   z = gamma*y + delta*z, every term with a zero coefficient omitted. */
EXPORT void orc_d_upd_xyz(int m, double alpha, double beta, double gamma, double delta,
                          const double *x, double *y, double *z)
{
  #pragma omp parallel for schedule(static) if(m > 65536)
  for (int i = 0; i < m; i++) {
    double yy;
    if (alpha != 0.0 && beta != 0.0) yy = alpha * x[i] + beta * y[i];
    else if (alpha != 0.0) yy = alpha * x[i];
    else if (beta != 0.0) yy = beta * y[i];
    else yy = 0.0;
    y[i] = yy;
    /* with alpha == beta == 0 the reference drops the gamma term (y is 0) */
    int g = gamma != 0.0 && !(alpha == 0.0 && beta == 0.0);
    double zz;
    if (g && delta != 0.0) zz = gamma * y[i] + delta * z[i];
    else if (g) zz = gamma * y[i];
    else if (delta != 0.0) zz = delta * z[i];
    else zz = 0.0;
    z[i] = zz;
  }
}

/* psi_dxyzw (psi_d_serial_impl.F90:1768-1866): y = a*x+b*y; z = c*y+d*z;
   w = e*z+f*w (all six coefficients must be non-zero, else nothing). */
EXPORT void orc_d_xyzw(int m, double a, double b, double c, double d, double e, double f,
                       const double *x, double *y, double *z, double *w)
{
  if (a == 0.0 || b == 0.0 || c == 0.0 || d == 0.0 || e == 0.0 || f == 0.0) return;
  #pragma omp parallel for schedule(static) if(m > 65536)
  for (int i = 0; i < m; i++) {
    y[i] = a * x[i] + b * y[i];
    z[i] = c * y[i] + d * z[i];
    w[i] = e * z[i] + f * w[i];
  }
}

/* d_base_mlt_a_2 (base/modules/serial/psb_d_base_vect_mod.F90:1265-1344):
   z = alpha*y*x + beta*z, i.e. for alpha == 1: z = y*x (+ beta*z). The
   evaluation order follows the reference branches. */
EXPORT void orc_d_mlt(int n, double alpha, const double *x, const double *y, double beta, double *z)
{
  #pragma omp parallel for schedule(static) if(n > 65536)
  for (int i = 0; i < n; i++) {
    if (alpha == 0.0) { if (beta != 1.0) z[i] = beta * z[i]; }
    else if (alpha == 1.0) {
      if (beta == 0.0) z[i] = y[i] * x[i];
      else if (beta == 1.0) z[i] = z[i] + y[i] * x[i];
      else z[i] = beta * z[i] + y[i] * x[i];
    } else if (alpha == -1.0) {
      if (beta == 0.0) z[i] = -y[i] * x[i];
      else if (beta == 1.0) z[i] = z[i] - y[i] * x[i];
      else z[i] = beta * z[i] - y[i] * x[i];
    } else {
      if (beta == 0.0) z[i] = alpha * y[i] * x[i];
      else if (beta == 1.0) z[i] = z[i] + alpha * y[i] * x[i];
      else z[i] = beta * z[i] + alpha * y[i] * x[i];
    }
  }
}

/* ddot, reference BLAS (called at psb_d_base_vect_mod.F90:991): sequential
   left-to-right sum (the netlib unroll by 5 keeps that order). */
EXPORT double orc_d_dot(int n, const double *x, const double *y)
{
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + x[i] * y[i];
  return s;
}

EXPORT float orc_s_dot(int n, const float *x, const float *y)
{
  float s = 0.0f;
  for (int i = 0; i < n; i++) s = s + x[i] * y[i];
  return s;
}

/* dnrm2, reference BLAS scaled sum of squares (called at
   psb_d_base_vect_mod.F90:1772). */
EXPORT double orc_d_nrm2(int n, const double *x)
{
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; i++) {
    if (x[i] != 0.0) {
      double a = fabs(x[i]);
      if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
      else ssq = ssq + (a / scale) * (a / scale);
    }
  }
  return scale * sqrt(ssq);
}

EXPORT float orc_s_nrm2(int n, const float *x)
{
  float scale = 0.0f, ssq = 1.0f;
  for (int i = 0; i < n; i++) {
    if (x[i] != 0.0f) {
      float a = fabsf(x[i]);
      if (scale < a) { ssq = 1.0f + ssq * (scale / a) * (scale / a); scale = a; }
      else ssq = ssq + (a / scale) * (a / scale);
    }
  }
  return scale * sqrtf(ssq);
}

/* d_base_amax (psb_d_base_vect_mod.F90:1781-1798) and d_base_asum (:1886-1903). */
EXPORT double orc_d_amax(int n, const double *x)
{
  double r = 0.0;
  for (int i = 0; i < n; i++) if (fabs(x[i]) > r) r = fabs(x[i]);
  return r;
}

EXPORT double orc_d_asum(int n, const double *x)
{
  double r = 0.0;
  for (int i = 0; i < n; i++) r = r + fabs(x[i]);
  return r;
}

/* ------------------------------------------------------------------ */
/* Jacobi (DIAG) preconditioner                                        */
/* ------------------------------------------------------------------ */

/* psb_d_diag_precbld (prec/impl/psb_d_diagprec_impl.f90:229-293):
   d(i) = 1/a_ii, a zero diagonal replaced by 1. */
EXPORT void orc_d_diag_build(int nr, const int *irp, const int *ja, const double *val, double *d)
{
  for (int i = 0; i < nr; i++) {
    double aii = 0.0;
    for (int j = irp[i] - 1; j < irp[i + 1] - 1; j++) if (ja[j] == i + 1) aii = val[j];
    d[i] = aii == 0.0 ? 1.0 : 1.0 / aii;
  }
}

/* ------------------------------------------------------------------ */
/* Krylov solvers over a CSR or HLL operator                           */
/* ------------------------------------------------------------------ */

typedef struct {
  int fmt; /* 0 = CSR, 1 = HLL */
  int nr;
  const int *irp, *ja;       /* CSR */
  const double *val;
  int hksz;                  /* HLL */
  const int *hkoffs, *irn, *hja;
  const double *hval;
} orc_mat_t;

static void orc_spmm(const orc_mat_t *a, double alpha, const double *x, double beta, double *y)
{
  if (a->fmt == 0) orc_d_csr_csmv(a->nr, a->irp, a->ja, a->val, alpha, x, beta, y);
  else orc_d_hll_csmv(a->nr, a->hksz, a->hkoffs, a->irn, a->hja, a->hval, alpha, x, beta, y, 1);
}

typedef struct {
  int istop;
  double eps, bn2, bni, ani, r0n2;
  int itmax;
} orc_conv_t;

/* psb_d_check_conv_vect (krylov/psb_d_krylov_conv_mod.f90:279-356), istop
   2 and 3 (istop 1 needs ||A||_inf: provided by the caller in ani). */
static int orc_check_conv(orc_conv_t *s, int it, int n, const double *x, const double *r,
                          double *errnum, double *errden)
{
  if (s->istop == 1) {
    *errnum = orc_d_amax(n, r);
    *errden = s->ani * orc_d_amax(n, x) + s->bni;
  } else if (s->istop == 2) {
    *errnum = orc_d_nrm2(n, r);
    *errden = s->bn2;
  } else {
    *errnum = orc_d_nrm2(n, r);
    *errden = s->r0n2;
  }
  int res;
  if (*errden == 0.0) res = (*errnum <= s->eps);
  else res = (*errnum <= s->eps * *errden);
  return res || (s->itmax <= it);
}

/* ------------------------------------------------------------------ */
/* BJAC with ILU(0) (prec/impl/psb_d_ilu0_fact.f90, psb_d_bjacprec_impl.f90) */
/* ------------------------------------------------------------------ */

/* psb_dilu0_factint (prec/impl/psb_d_ilu0_fact.f90:330-460), ialg = ILU_N,
   upd = 'F', no shift, on a CSR matrix with ascending columns.  L: strict
   lower part (unit diagonal implied), U: strict upper part scaled by the
   inverse pivot of its row, d: inverse pivots.  lirp/uirp (n+1), lja/lval
   and uja/uval with room for nnz entries.  Returns 0, or -i (1-based row)
   when the pivot of row i is below 1.1e-16 in magnitude
   (psb_err_pivot_too_small_). */
#define ORC_ILU0(NAME, T, EPS)                                                  \
EXPORT int NAME(int n, const int *irp, const int *ja, const T *val,                       \
                int *lirp, int *lja, T *lval, int *uirp, int *uja, T *uval, T *d)         \
{                                                                                         \
  int l1 = 0, l2 = 0;                                                                     \
  lirp[0] = 1;                                                                            \
  uirp[0] = 1;                                                                            \
  for (int i = 1; i <= n; i++) {                                                          \
    d[i - 1] = (T)0;                                                                      \
    /* ilu_copyin: L part (k < i), diagonal, U part (k > i) */                            \
    for (int j = irp[i - 1]; j < irp[i]; j++) {                                           \
      int k = ja[j - 1];                                                                  \
      if (k < i && k >= 1) { l1++; lval[l1 - 1] = val[j - 1]; lja[l1 - 1] = k; }          \
      else if (k == i) d[i - 1] = val[j - 1];                                             \
      else if (k > i && k <= n) { l2++; uval[l2 - 1] = val[j - 1]; uja[l2 - 1] = k; }     \
    }                                                                                     \
    lirp[i] = l1 + 1;                                                                     \
    uirp[i] = l2 + 1;                                                                     \
    T dia = d[i - 1];                                                                     \
    for (int kk = lirp[i - 1]; kk < lirp[i]; kk++) {                                      \
      T temp = lval[kk - 1];                                                              \
      int k = lja[kk - 1];                                                                \
      lval[kk - 1] = temp * d[k - 1];                                                     \
      int low1 = kk + 1, low2 = uirp[i - 1];                                              \
      for (int jj = uirp[k - 1]; jj < uirp[k]; jj++) {                                    \
        int j = uja[jj - 1];                                                              \
        if (j < i) {                                                                      \
          for (int ll = low1; ll < lirp[i]; ll++) {                                       \
            int l = lja[ll - 1];                                                          \
            if (l > j) { low1 = ll; break; }                                              \
            if (l == j) { lval[ll - 1] = lval[ll - 1] - temp * uval[jj - 1]; low1 = ll + 1; break; }\
          }                                                                               \
        } else if (j == i) {                                                              \
          dia = dia - temp * uval[jj - 1];                                                \
        } else {                                                                          \
          for (int ll = low2; ll < uirp[i]; ll++) {                                       \
            int l = uja[ll - 1];                                                          \
            if (l > j) { low2 = ll; break; }                                              \
            if (l == j) { uval[ll - 1] = uval[ll - 1] - temp * uval[jj - 1]; low2 = ll + 1; break; }\
          }                                                                               \
        }                                                                                 \
      }                                                                                   \
    }                                                                                     \
    if (fabs((double)dia) < (double)(EPS)) return -i;                                     \
    dia = (T)1 / dia;                                                                     \
    d[i - 1] = dia;                                                                       \
    for (int kk = uirp[i - 1]; kk < uirp[i]; kk++) uval[kk - 1] = uval[kk - 1] * dia;     \
  }                                                                                       \
  return 0;                                                                               \
}
ORC_ILU0(orc_dilu0_fact, double, 1.1e-16)
/* psb_silu0_factint (prec/impl/psb_s_ilu0_fact.f90): single precision, s_epstol */
ORC_ILU0(orc_silu0_fact, float, 5.e-8f)

/* inner_csrsv (base/serial/impl/psb_d_csr_impl.F90:890-941), no transpose,
   unit diagonal: lower (rows ascending) or upper (rows descending),
   y(i) = x(i) - sum_j val(j) y(ja(j)) accumulated in row order. */
EXPORT void orc_d_csr_trsv_unit(int n, int lower, const int *irp, const int *ja,
                                const double *val, const double *x, double *y)
{
  for (int t = 0; t < n; t++) {
    int i = lower ? t : n - 1 - t;
    double acc = 0.0;
    for (int j = irp[i] - 1; j < irp[i + 1] - 1; j++) acc = acc + val[j] * y[ja[j] - 1];
    y[i] = x[i] - acc;
  }
}

/* psb_d_csr_cssv (base/serial/impl/psb_d_csr_impl.F90:782-888) with its
   inner_csrsv (:890-1039): y = alpha * T^-1 x + beta * y for a triangular
   CSR matrix T.  lower/upper and unit/non-unit as the matrix is flagged; a
   non-unit row keeps its diagonal as its last (lower) or first (upper)
   entry and divides by it.  trans 0 is 'N' (row sums accumulated from 0 in
   CSR order, then x(i) - sum); 1 is 'T' and 'C' (identical for real data:
   y = x, then each solved y(i) is subtracted, times the row's entries, from
   the y of their columns, rows in solve order).  alpha == 0 returns beta*y
   without solving; beta == 0 scales the solution in place (alpha == -1 by
   negation); otherwise the solution goes through the library axpby. */
#define ORC_CSR_CSSV(NAME, T, AXPBY)                                            \
static void NAME##_inner(int tra, int lower, int unit, int n, const int *irp, \
                         const int *ja, const T *val, const T *x, T *y)       \
{                                                                              \
  if (!tra) {                                                                  \
    for (int t = 0; t < n; t++) {                                              \
      const int i = lower ? t : n - 1 - t;                                     \
      int b = irp[i] - 1, e = irp[i + 1] - 1;                                  \
      if (!unit) { if (lower) e--; else b++; }                                 \
      T acc = (T)0;                                                            \
      for (int j = b; j < e; j++) acc = acc + val[j] * y[ja[j] - 1];           \
      y[i] = unit ? x[i] - acc                                                 \
                  : (x[i] - acc) / val[lower ? irp[i + 1] - 2 : irp[i] - 1];   \
    }                                                                          \
    return;                                                                    \
  }                                                                            \
  for (int i = 0; i < n; i++) y[i] = x[i];                                     \
  for (int t = 0; t < n; t++) {                                                \
    const int i = lower ? n - 1 - t : t;                                       \
    int b = irp[i] - 1, e = irp[i + 1] - 1;                                    \
    if (!unit) {                                                               \
      if (lower) { y[i] = y[i] / val[e - 1]; e--; }                            \
      else { y[i] = y[i] / val[b]; b++; }                                      \
    }                                                                          \
    const T acc = y[i];                                                        \
    for (int j = b; j < e; j++) y[ja[j] - 1] = y[ja[j] - 1] - val[j] * acc;    \
  }                                                                            \
}                                                                              \
EXPORT void NAME(int n, int lower, int unit, int trans, T alpha, const int *irp, \
                 const int *ja, const T *val, const T *x, T beta, T *y)        \
{                                                                              \
  if (alpha == (T)0) {                                                         \
    for (int i = 0; i < n; i++) y[i] = beta == (T)0 ? (T)0 : beta * y[i];      \
    return;                                                                    \
  }                                                                            \
  if (beta == (T)0) {                                                          \
    NAME##_inner(trans, lower, unit, n, irp, ja, val, x, y);                   \
    if (alpha == (T)-1) for (int i = 0; i < n; i++) y[i] = -y[i];              \
    else if (alpha != (T)1) for (int i = 0; i < n; i++) y[i] = alpha * y[i];   \
    return;                                                                    \
  }                                                                            \
  T *tmp = calloc(n > 0 ? n : 1, sizeof(T));                                   \
  NAME##_inner(trans, lower, unit, n, irp, ja, val, x, tmp);                   \
  AXPBY(n, alpha, tmp, beta, y);                                               \
  free(tmp);                                                                   \
}
ORC_CSR_CSSV(orc_d_csr_cssv, double, orc_d_axpby)
ORC_CSR_CSSV(orc_s_csr_cssv, float, orc_s_axpby)

/* The preconditioner of the oracle solvers: DIAG (y = d .* x,
   psb_d_diag_apply_vect) or, when lirp is set, BJAC with ILU(0) factors
   (psb_d_bjac_apply_vect, prec/impl/psb_d_bjacprec_impl.f90:166-173):
   w = L^-1 x scaled by d (psb_spsm scale 'L'), y = U^-1 w.  x may alias y. */
typedef struct {
  const double *d;
  int n;
  const int *lirp, *lja;
  const double *lval;
  const int *uirp, *uja;
  const double *uval;
  const double *dd;
  double *tmp;
} orc_prec_t;

static void orc_papply(const orc_prec_t *p, const double *x, double *y)
{
  if (!p->lirp) {
    orc_d_mlt(p->n, 1.0, p->d, x, 0.0, y);
    return;
  }
  orc_d_csr_trsv_unit(p->n, 1, p->lirp, p->lja, p->lval, x, p->tmp);
  for (int i = 0; i < p->n; i++) p->tmp[i] = p->tmp[i] * p->dd[i];
  orc_d_csr_trsv_unit(p->n, 0, p->uirp, p->uja, p->uval, p->tmp, y);
}

EXPORT void orc_d_bjac_apply(int n, const int *lirp, const int *lja, const double *lval,
                             const int *uirp, const int *uja, const double *uval,
                             const double *dd, const double *x, double *y)
{
  double *tmp = calloc(n > 0 ? n : 1, sizeof(double));
  orc_prec_t p = {NULL, n, lirp, lja, lval, uirp, uja, uval, dd, tmp};
  orc_papply(&p, x, y);
  free(tmp);
}

/*
 * psb_dcg_vect (krylov/psb_dcg.F90:98-333) with the DIAG preconditioner
 * applied as z = d .* r (psb_d_diag_apply_vect, psb_d_diagprec_impl.f90:116).
 * hist (optional, length itmax+1) receives errnum/errden per iteration.
 */
static int orc_dcg(const orc_mat_t *a, const orc_prec_t *pc, const double *b, double *x,
                   double eps, int itmax, int istop, double ani,
                   int *iter, double *err, double *hist)
{
  int n = a->nr;
  double *p = calloc(n, sizeof(double)), *q = calloc(n, sizeof(double));
  double *r = calloc(n, sizeof(double)), *z = calloc(n, sizeof(double));
  orc_conv_t s = {istop, eps, 0, 0, ani, 0, itmax};
  double errnum = 0, errden = 1, alpha = 0, beta, rho, rho_old, sigma;
  int itx = 0;
  for (;;) { /* restart */
    if (itx >= itmax) break;
    int it = 0, done = 0;
    orc_d_axpby(n, 1.0, b, 0.0, r);
    orc_spmm(a, -1.0, x, 1.0, r);
    rho = 0.0;
    /* psb_init_conv runs inside the restart loop in CG (psb_dcg.F90:231) */
    if (istop == 2) s.bn2 = orc_d_nrm2(n, b);
    if (istop == 1) s.bni = orc_d_amax(n, b);
    if (istop == 3) s.r0n2 = orc_d_nrm2(n, r);
    for (;;) { /* iteration */
      it++; itx++;
      orc_papply(pc, r, z);
      rho_old = rho;
      rho = orc_d_dot(n, r, z);
      if (it == 1) orc_d_axpby(n, 1.0, z, 0.0, p);
      else {
        if (rho_old == 0.0) break;
        beta = rho / rho_old;
        orc_d_axpby(n, 1.0, z, beta, p);
      }
      orc_spmm(a, 1.0, p, 0.0, q);
      sigma = orc_d_dot(n, p, q);
      if (sigma == 0.0) break;
      alpha = rho / sigma;
      orc_d_axpby(n, alpha, p, 1.0, x);
      orc_d_axpby(n, -alpha, q, 1.0, r);
      int conv = orc_check_conv(&s, itx, n, x, r, &errnum, &errden);
      if (hist) hist[itx] = errden > 0 ? errnum / errden : errnum;
      if (conv) { done = 1; break; }
    }
    if (done) break;
  }
  (void)alpha;
  *iter = itx;
  *err = errden > 0 ? errnum / errden : errnum;
  free(p); free(q); free(r); free(z);
  return 0;
}

/* psb_dcgstab_vect (krylov/psb_dcgstab.f90:96-393), DIAG preconditioner. */
static int orc_dcgstab(const orc_mat_t *a, const orc_prec_t *pc, const double *b, double *x,
                       double eps, int itmax, int istop, double ani,
                       int *iter, double *err, double *hist)
{
  int n = a->nr;
  double *q = calloc(n, sizeof(double)), *r = calloc(n, sizeof(double));
  double *p = calloc(n, sizeof(double)), *v = calloc(n, sizeof(double));
  double *s = calloc(n, sizeof(double)), *t = calloc(n, sizeof(double));
  double *z = calloc(n, sizeof(double)), *f = calloc(n, sizeof(double));
  orc_conv_t st = {istop, eps, 0, 0, ani, 0, itmax};
  double errnum = 0, errden = 1;
  double alpha = 0, beta, omega = 0, rho = 0, rho_old, sigma, tau;
  int itx = 0;
  if (istop == 2) st.bn2 = orc_d_nrm2(n, b);
  if (istop == 1) st.bni = orc_d_amax(n, b);
  if (istop == 3) {
    memcpy(r, b, sizeof(double) * n);
    orc_spmm(a, -1.0, x, 1.0, r);
    st.r0n2 = orc_d_nrm2(n, r);
  }
  for (;;) { /* restart */
    if (itx >= itmax) break;
    int it = 0, done = 0;
    orc_d_axpby(n, 1.0, b, 0.0, r);
    orc_spmm(a, -1.0, x, 1.0, r);
    orc_d_axpby(n, 1.0, r, 0.0, q);
    if (orc_check_conv(&st, itx, n, x, r, &errnum, &errden)) break;
    rho = 0.0;
    for (;;) {
      it++; itx++;
      rho_old = rho;
      rho = orc_d_dot(n, q, r);
      if (rho == 0.0) break;
      if (it == 1) orc_d_axpby(n, 1.0, r, 0.0, p);
      else {
        beta = (rho / rho_old) * (alpha / omega);
        orc_d_axpby(n, -omega, v, 1.0, p);
        orc_d_axpby(n, 1.0, r, beta, p);
      }
      orc_papply(pc, p, f);
      orc_spmm(a, 1.0, f, 0.0, v);
      sigma = orc_d_dot(n, q, v);
      if (sigma == 0.0) break;
      alpha = rho / sigma;
      orc_d_axpby(n, 1.0, r, 0.0, s);
      orc_d_axpby(n, -alpha, v, 1.0, s);
      orc_papply(pc, s, z);
      orc_spmm(a, 1.0, z, 0.0, t);
      sigma = orc_d_dot(n, t, t);
      if (sigma == 0.0) break;
      tau = orc_d_dot(n, t, s);
      omega = tau / sigma;
      if (omega == 0.0) break;
      orc_d_axpby(n, alpha, f, 1.0, x);
      orc_d_axpby(n, omega, z, 1.0, x);
      orc_d_axpby(n, 1.0, s, 0.0, r);
      orc_d_axpby(n, -omega, t, 1.0, r);
      int conv = orc_check_conv(&st, itx, n, x, r, &errnum, &errden);
      if (hist) hist[itx] = errden > 0 ? errnum / errden : errnum;
      if (conv) { done = 1; break; }
    }
    if (done) break;
  }
  *iter = itx;
  *err = errden > 0 ? errnum / errden : errnum;
  free(q); free(r); free(p); free(v); free(s); free(t); free(z); free(f);
  return 0;
}

EXPORT int orc_d_krylov_csr(int method, int nr, const int *irp, const int *ja, const double *val,
                            const double *d, const double *b, double *x, double eps, int itmax,
                            int istop, double ani, int *iter, double *err, double *hist);

/* psb_d_init_conv_vect (krylov/psb_d_krylov_conv_mod.f90:205-277) */
static void orc_init_conv(orc_conv_t *s, const orc_mat_t *a, const double *x, const double *b)
{
  int n = a->nr;
  if (s->istop == 1) s->bni = orc_d_amax(n, b);
  if (s->istop == 2) s->bn2 = orc_d_nrm2(n, b);
  if (s->istop == 3) {
    double *r = malloc(sizeof(double) * (n > 0 ? n : 1));
    memcpy(r, b, sizeof(double) * n);
    orc_spmm(a, -1.0, x, 1.0, r);
    s->r0n2 = orc_d_nrm2(n, r);
    free(r);
  }
}

static double **orc_vecs(int k, int n)
{
  double **v = malloc(sizeof(double *) * k);
  for (int i = 0; i < k; i++) v[i] = calloc(n > 0 ? n : 1, sizeof(double));
  return v;
}

static void orc_free_vecs(double **v, int k)
{
  for (int i = 0; i < k; i++) free(v[i]);
  free(v);
}

/* dtrsm('l','u','n','n', m, 1, 1.0, h, ldh, y, ...) as the reference BLAS
   loop: column-oriented back substitution, H(1:m,1:m) upper, h[i][k] =
   H(i,k) with 1-based indices. */
static void orc_trsv_upper(int m, double **h, double *y)
{
  for (int k = m; k >= 1; k--) {
    if (y[k] != 0.0) {
      y[k] = y[k] / h[k][k];
      for (int i = 1; i <= k - 1; i++) y[i] = y[i] - y[k] * h[i][k];
    }
  }
}

/* reference BLAS drotg / drot (n = 1) */
static void orc_drotg(double *sa, double *sb, double *c, double *s)
{
  double roe = fabs(*sa) > fabs(*sb) ? *sa : *sb;
  double scale = fabs(*sa) + fabs(*sb);
  double r, z;
  if (scale == 0.0) {
    *c = 1.0; *s = 0.0; r = 0.0; z = 0.0;
  } else {
    double ta = *sa / scale, tb = *sb / scale;
    r = scale * sqrt(ta * ta + tb * tb);
    if (roe < 0.0) r = -r;
    *c = *sa / r;
    *s = *sb / r;
    z = 1.0;
    if (fabs(*sa) > fabs(*sb)) z = *s;
    if (fabs(*sb) >= fabs(*sa) && *c != 0.0) z = 1.0 / *c;
  }
  *sa = r;
  *sb = z;
}

static void orc_drot(double *x, double *y, double c, double s)
{
  double t = c * *x + s * *y;
  *y = c * *y - s * *x;
  *x = t;
}

/* psb_dfcg_vect (krylov/psb_dfcg.F90:150-336), DIAG preconditioner. */
static int orc_dfcg(const orc_mat_t *a, const orc_prec_t *pc, const double *b, double *x,
                    double eps, int itmax, int istop, double ani,
                    int *iter, double *err, double *hist)
{
  int n = a->nr;
  double **V = orc_vecs(5, n);
  double *w = V[0], *v = V[1], *d = V[2], *q = V[3], *r = V[4];
  orc_conv_t s = {istop, eps, 0, 0, ani, 0, itmax};
  orc_init_conv(&s, a, x, b);
  double errnum = 0, errden = 1, alpha, beta, gamma, delta, theta;
  int itx = 0, done = 0;
  for (;;) { /* restart */
    if (itx >= itmax) break;
    orc_d_axpby(n, 1.0, b, 0.0, r);
    orc_spmm(a, -1.0, x, 1.0, r);
    if (orc_check_conv(&s, itx, n, x, r, &errnum, &errden)) break;
    orc_papply(pc, r, v);
    orc_spmm(a, 1.0, v, 0.0, w);
    alpha = orc_d_dot(n, r, v);
    beta = orc_d_dot(n, w, v);
    if (beta == 0.0) break;
    orc_d_axpby(n, 1.0, v, 0.0, d);
    orc_d_axpby(n, 1.0, w, 0.0, q);
    delta = beta;
    theta = alpha / delta;
    orc_d_axpby(n, theta, d, 1.0, x);
    orc_d_axpby(n, -theta, q, 1.0, r);
    for (;;) { /* iteration */
      itx++;
      int conv = orc_check_conv(&s, itx, n, x, r, &errnum, &errden);
      if (hist) hist[itx] = errden > 0 ? errnum / errden : errnum;
      if (conv) { done = 1; break; }
      orc_papply(pc, r, v);
      orc_spmm(a, 1.0, v, 0.0, w);
      alpha = orc_d_dot(n, r, v);
      beta = orc_d_dot(n, w, v);
      gamma = orc_d_dot(n, q, v);
      theta = gamma / delta;
      orc_d_axpby(n, 1.0, v, -theta, d);
      orc_d_axpby(n, 1.0, w, -theta, q);
      delta = beta - (gamma * gamma) / delta;
      if (delta == 0.0) break;
      theta = alpha / delta;
      orc_d_axpby(n, theta, d, 1.0, x);
      orc_d_axpby(n, -theta, q, 1.0, r);
    }
    if (done) break;
  }
  *iter = itx;
  *err = errden > 0 ? errnum / errden : errnum;
  orc_free_vecs(V, 5);
  return 0;
}

/* psb_dgcr_vect (krylov/psb_dgcr.f90:246-345), DIAG preconditioner, restart
   length nl.  The recurrence (z, c, c_scale, h, alpha, r) and the stop
   tests are the reference's; x is assembled as x0 + sum of every cycle's
   correction (the reference sets x = 0 before the first cycle's
   correction and again before the last, which drops x0 and the earlier
   cycles whenever there is more than one). */
static int orc_dgcr(const orc_mat_t *a, const orc_prec_t *pc, const double *b, double *x,
                    double eps, int itmax, int istop, double ani, int nl,
                    int *iter, double *err, double *hist)
{
  int n = a->nr;
  double *r = calloc(n > 0 ? n : 1, sizeof(double));
  double **z = orc_vecs(nl + 2, n), **c = orc_vecs(nl + 2, n), **cs = orc_vecs(nl + 2, n);
  double **h = orc_vecs(nl + 2, nl + 2);
  double *alpha = calloc(nl + 2, sizeof(double));
  orc_conv_t s = {istop, eps, 0, 0, ani, 0, itmax};
  orc_init_conv(&s, a, x, b);
  double errnum = 0, errden = 1;
  int itx = 0, conv = 0;
  for (;;) { /* restart */
    if (itx >= itmax) break;
    for (int i = 0; i < nl + 2; i++) memset(h[i], 0, sizeof(double) * (nl + 2));
    int it = 0, j = 0;
    orc_d_axpby(n, 1.0, b, 0.0, r);
    orc_spmm(a, -1.0, x, 1.0, r);
    if (orc_check_conv(&s, itx, n, x, r, &errnum, &errden)) break;
    for (;;) { /* iteration */
      itx++; it++; j = it;
      orc_papply(pc, r, z[j]);
      orc_spmm(a, 1.0, z[j], 0.0, c[1]);
      for (int i = 1; i <= j - 1; i++) {
        h[i][j] = orc_d_dot(n, cs[i], c[i]);
        orc_d_axpby(n, 1.0, c[i], 0.0, c[i + 1]);
        orc_d_axpby(n, -h[i][j], cs[i], 1.0, c[i + 1]);
      }
      h[j][j] = orc_d_nrm2(n, c[j]);
      orc_d_axpby(n, 1.0 / h[j][j], c[j], 0.0, cs[j]);
      alpha[j] = orc_d_dot(n, cs[j], r);
      orc_d_axpby(n, -alpha[j], cs[j], 1.0, r);
      conv = orc_check_conv(&s, itx, n, x, r, &errnum, &errden);
      if (hist) hist[itx] = errden > 0 ? errnum / errden : errnum;
      if (conv || j >= nl) break;
    }
    orc_trsv_upper(j, h, alpha);
    for (int i = 1; i <= j; i++) orc_d_axpby(n, alpha[i], z[i], 1.0, x);
    if (conv) break;
  }
  *iter = itx;
  *err = errden > 0 ? errnum / errden : errnum;
  free(r); free(alpha);
  orc_free_vecs(z, nl + 2); orc_free_vecs(c, nl + 2); orc_free_vecs(cs, nl + 2);
  orc_free_vecs(h, nl + 2);
  return 0;
}

/* psb_drgmres_vect (krylov/psb_drgmres.f90:247-475), right DIAG
   preconditioning, restart nl.  istop 2 (||r||/||b||, the rotated
   least-squares residual) or 1 (normwise backward error, :387-405: at every
   inner iteration xt = x + M^-1 V y is rebuilt from the partial
   least-squares solve and ||b - A xt||_inf / (ani ||xt||_inf + ||b||_inf)
   tested; ani = ||A||_inf from the caller; xt is rebuilt from zero, where the
   reference accumulates into an uncleared xt).  At an invariant Krylov space
   (h(i+1,i) = 0) v(i+1) is not normalised and the solve ends. */
static int orc_drgmres(const orc_mat_t *a, const orc_prec_t *pc, const double *b, double *x,
                       double eps, int itmax, int istop, double ani, int nl,
                       int *iter, double *err, double *hist)
{
  int n = a->nr;
  double **v = orc_vecs(nl + 2, n), **h = orc_vecs(nl + 2, nl + 2);
  double *w = calloc(n > 0 ? n : 1, sizeof(double)), *w1 = calloc(n > 0 ? n : 1, sizeof(double));
  double *xt = calloc(n > 0 ? n : 1, sizeof(double));
  double *c = calloc(nl + 2, sizeof(double)), *sn = calloc(nl + 2, sizeof(double));
  double *rs = calloc(nl + 2, sizeof(double)), *rst = calloc(nl + 2, sizeof(double));
  double bn2 = 0, bni = 0, errnum = 0, errden = 1;
  if (istop == 1) bni = orc_d_amax(n, b);
  else bn2 = orc_d_nrm2(n, b);
  int itx = 0, done = 0;
  for (;;) { /* restart */
    orc_d_axpby(n, 1.0, b, 0.0, v[1]);
    orc_spmm(a, -1.0, x, 1.0, v[1]);
    rs[1] = orc_d_nrm2(n, v[1]);
    for (int k = 2; k < nl + 2; k++) rs[k] = 0.0;
    double scal = 1.0 / rs[1];
    if (istop == 1) {
      errnum = orc_d_amax(n, v[1]);
      errden = ani * orc_d_amax(n, x) + bni;
    } else {
      errnum = orc_d_nrm2(n, v[1]);
      errden = bn2;
    }
    if (errnum <= eps * errden) break;
    for (int i = 0; i < n; i++) v[1][i] = v[1][i] * scal;
    if (itx >= itmax) break;
    int i;
    for (i = 1; i <= nl; i++) { /* inner */
      itx++;
      orc_papply(pc, v[i], w1);
      orc_spmm(a, 1.0, w1, 0.0, w);
      for (int k = 1; k <= i; k++) {
        h[k][i] = orc_d_dot(n, v[k], w);
        orc_d_axpby(n, -h[k][i], v[k], 1.0, w);
      }
      h[i + 1][i] = orc_d_nrm2(n, w);
      int invariant = h[i + 1][i] == 0.0;
      if (!invariant) orc_d_axpby(n, 1.0 / h[i + 1][i], w, 0.0, v[i + 1]);
      for (int k = 2; k <= i; k++) orc_drot(&h[k - 1][i], &h[k][i], c[k - 1], sn[k - 1]);
      double rti = h[i][i], rti1 = h[i + 1][i];
      orc_drotg(&rti, &rti1, &c[i], &sn[i]);
      orc_drot(&h[i][i], &h[i + 1][i], c[i], sn[i]);
      h[i + 1][i] = 0.0;
      orc_drot(&rs[i], &rs[i + 1], c[i], sn[i]);
      if (istop == 1) {
        memcpy(rst, rs, sizeof(double) * (nl + 2));
        orc_trsv_upper(i, h, rst);
        memset(w1, 0, sizeof(double) * n);
        for (int k = 1; k <= i; k++) orc_d_axpby(n, rst[k], v[k], 1.0, w1);
        orc_papply(pc, w1, xt);
        orc_d_axpby(n, 1.0, x, 1.0, xt);
        orc_d_axpby(n, 1.0, b, 0.0, w1);
        orc_spmm(a, -1.0, xt, 1.0, w1);
        errnum = orc_d_amax(n, w1);
        errden = ani * orc_d_amax(n, xt) + bni;
      } else {
        errnum = fabs(rs[i + 1]);
      }
      if (hist) hist[itx] = errden > 0 ? errnum / errden : errnum;
      if (errnum <= eps * errden || invariant) { done = 1; break; }
    }
    if (istop == 1) {
      memcpy(x, xt, sizeof(double) * n);
    } else {
      int m = done ? i : nl;
      orc_trsv_upper(m, h, rs);
      memset(w1, 0, sizeof(double) * n);
      for (int k = 1; k <= m; k++) orc_d_axpby(n, rs[k], v[k], 1.0, w1);
      orc_papply(pc, w1, w);
      orc_d_axpby(n, 1.0, w, 1.0, x);
    }
    if (done) break;
  }
  *iter = itx;
  *err = errden > 0 ? errnum / errden : errnum;
  orc_free_vecs(v, nl + 2); orc_free_vecs(h, nl + 2);
  free(w); free(w1); free(xt); free(c); free(sn); free(rs); free(rst);
  return 0;
}

/* psb_dcgstabl_vect (krylov/psb_dcgstabl.f90:249-395): BiCGSTAB(l), left DIAG
   preconditioning (the stop test sees the preconditioned residual). */
static int orc_dcgstabl(const orc_mat_t *a, const orc_prec_t *pc, const double *b, double *x,
                        double eps, int itmax, int istop, double ani, int nl,
                        int *iter, double *err, double *hist)
{
  int n = a->nr;
  double *r = calloc(n > 0 ? n : 1, sizeof(double)), *rt0 = calloc(n > 0 ? n : 1, sizeof(double));
  double **rh = orc_vecs(nl + 1, n), **uh = orc_vecs(nl + 1, n);
  double **taum = orc_vecs(nl + 1, nl + 1);
  double *gamma = calloc(nl + 1, sizeof(double)), *gamma1 = calloc(nl + 1, sizeof(double));
  double *gamma2 = calloc(nl + 1, sizeof(double)), *sigma = calloc(nl + 1, sizeof(double));
  orc_conv_t s = {istop, eps, 0, 0, ani, 0, itmax};
  orc_init_conv(&s, a, x, b);
  double errnum = 0, errden = 1, rho, rho_old, alpha, omega, beta;
  int itx = 0, done = 0;
  for (;;) { /* restart */
    if (itx >= itmax) break;
    orc_d_axpby(n, 1.0, b, 0.0, r);
    orc_spmm(a, -1.0, x, 1.0, r);
    orc_papply(pc, r, r);
    orc_d_axpby(n, 1.0, r, 0.0, rt0);
    orc_d_axpby(n, 1.0, r, 0.0, rh[0]);
    orc_d_axpby(n, 0.0, r, 0.0, uh[0]);
    rho = 1.0; alpha = 0.0; omega = 1.0;
    if (orc_check_conv(&s, itx, n, x, r, &errnum, &errden)) break;
    for (;;) { /* iteration */
      itx += nl;
      rho = -omega * rho;
      int brk = 0;
      for (int j = 0; j < nl; j++) {
        rho_old = rho;
        rho = orc_d_dot(n, rh[j], rt0);
        if (rho == 0.0) { brk = 1; break; }
        beta = alpha * rho / rho_old;
        rho_old = rho;
        for (int k = 0; k <= j; k++) orc_d_axpby(n, 1.0, rh[k], -beta, uh[k]);
        orc_spmm(a, 1.0, uh[j], 0.0, uh[j + 1]);
        orc_papply(pc, uh[j + 1], uh[j + 1]);
        gamma[j] = orc_d_dot(n, uh[j + 1], rt0);
        if (gamma[j] == 0.0) { brk = 1; break; }
        alpha = rho / gamma[j];
        for (int k = 0; k <= j; k++) orc_d_axpby(n, -alpha, uh[k + 1], 1.0, rh[k]);
        orc_d_axpby(n, alpha, uh[0], 1.0, x);
        orc_spmm(a, 1.0, rh[j], 0.0, rh[j + 1]);
        orc_papply(pc, rh[j + 1], rh[j + 1]);
      }
      if (brk) break;
      for (int j = 1; j <= nl; j++) {
        for (int i = 1; i <= j - 1; i++) {
          taum[i][j] = orc_d_dot(n, rh[i], rh[j]);
          taum[i][j] = taum[i][j] / sigma[i];
          orc_d_axpby(n, -taum[i][j], rh[i], 1.0, rh[j]);
        }
        sigma[j] = orc_d_dot(n, rh[j], rh[j]);
        gamma1[j] = orc_d_dot(n, rh[0], rh[j]);
        gamma1[j] = gamma1[j] / sigma[j];
      }
      gamma[nl] = gamma1[nl];
      omega = gamma[nl];
      for (int j = nl - 1; j >= 1; j--) {
        gamma[j] = gamma1[j];
        for (int i = j + 1; i <= nl; i++) gamma[j] = gamma[j] - taum[j][i] * gamma[i];
      }
      for (int j = 1; j <= nl - 1; j++) {
        gamma2[j] = gamma[j + 1];
        for (int i = j + 1; i <= nl - 1; i++) gamma2[j] = gamma2[j] + taum[j][i] * gamma[i + 1];
      }
      orc_d_axpby(n, gamma[1], rh[0], 1.0, x);
      orc_d_axpby(n, -gamma1[nl], rh[nl], 1.0, rh[0]);
      orc_d_axpby(n, -gamma[nl], uh[nl], 1.0, uh[0]);
      for (int j = 1; j <= nl - 1; j++) {
        orc_d_axpby(n, -gamma[j], uh[j], 1.0, uh[0]);
        orc_d_axpby(n, gamma2[j], rh[j], 1.0, x);
        orc_d_axpby(n, -gamma1[j], rh[j], 1.0, rh[0]);
      }
      int conv = orc_check_conv(&s, itx, n, x, rh[0], &errnum, &errden);
      if (hist) hist[itx] = errden > 0 ? errnum / errden : errnum;
      if (conv) { done = 1; break; }
    }
    if (done) break;
  }
  *iter = itx;
  *err = errden > 0 ? errnum / errden : errnum;
  free(r); free(rt0); free(gamma); free(gamma1); free(gamma2); free(sigma);
  orc_free_vecs(rh, nl + 1); orc_free_vecs(uh, nl + 1); orc_free_vecs(taum, nl + 1);
  return 0;
}

/* psb_dcgs_vect (krylov/psb_dcgs.f90:100-400): CGS, a breakdown (rho or
   sigma zero) restarts from the true residual. */
static int orc_dcgs(const orc_mat_t *a, const orc_prec_t *pc, const double *b, double *x,
                    double eps, int itmax, int istop, double ani,
                    int *iter, double *err, double *hist)
{
  int n = a->nr;
  double **w = orc_vecs(10, n);
  double *q = w[0], *qt = w[1], *r = w[2], *rt = w[3], *p = w[4], *v = w[5], *uv = w[6];
  double *z = w[7], *f = w[8], *s = w[9];
  orc_conv_t st = {istop, eps, 0, 0, ani, 0, itmax};
  double errnum = 0, errden = 1, alpha, beta, rho = 0, rho_old, sigma;
  int itx = 0;
  if (istop == 2) st.bn2 = orc_d_nrm2(n, b);
  if (istop == 1) st.bni = orc_d_amax(n, b);
  if (istop == 3) {
    memcpy(r, b, sizeof(double) * n);
    orc_spmm(a, -1.0, x, 1.0, r);
    st.r0n2 = orc_d_nrm2(n, r);
  }
  for (;;) { /* restart */
    if (itx >= itmax) break;
    int it = 0, done = 0;
    orc_d_axpby(n, 1.0, b, 0.0, r);
    orc_spmm(a, -1.0, x, 1.0, r);
    orc_d_axpby(n, 1.0, r, 0.0, rt);
    if (orc_check_conv(&st, itx, n, x, r, &errnum, &errden)) break;
    rho = 0.0;
    for (;;) {
      it++; itx++;
      rho_old = rho;
      rho = orc_d_dot(n, rt, r);
      if (rho == 0.0) break;
      if (it == 1) {
        orc_d_axpby(n, 1.0, r, 0.0, uv);
        orc_d_axpby(n, 1.0, r, 0.0, p);
      } else {
        beta = rho / rho_old;
        orc_d_axpby(n, 1.0, r, 0.0, uv);
        orc_d_axpby(n, beta, q, 1.0, uv);
        orc_d_axpby(n, 1.0, q, beta, p);
        orc_d_axpby(n, 1.0, uv, beta, p);
      }
      orc_papply(pc, p, f);
      orc_spmm(a, 1.0, f, 0.0, v);
      sigma = orc_d_dot(n, rt, v);
      if (sigma == 0.0) break;
      alpha = rho / sigma;
      orc_d_axpby(n, 1.0, uv, 0.0, q);
      orc_d_axpby(n, -alpha, v, 1.0, q);
      orc_d_axpby(n, 1.0, uv, 0.0, s);
      orc_d_axpby(n, 1.0, q, 1.0, s);
      orc_papply(pc, s, z);
      orc_d_axpby(n, alpha, z, 1.0, x);
      orc_spmm(a, 1.0, z, 0.0, qt);
      orc_d_axpby(n, -alpha, qt, 1.0, r);
      int conv = orc_check_conv(&st, itx, n, x, r, &errnum, &errden);
      if (hist) hist[itx] = errden > 0 ? errnum / errden : errnum;
      if (conv) { done = 1; break; }
    }
    if (done) break;
  }
  *iter = itx;
  *err = errden > 0 ? errnum / errden : errnum;
  orc_free_vecs(w, 10);
  return 0;
}

/* psb_dbicg_vect (krylov/psb_dbicg.f90:100-330): BiCG with the shadow
   residual rt = r0; the products with A^T are psb_d_csr_csmv's transposed
   branch (CSR matrices); the preconditioner's transposed apply is DIAG's,
   the same as its apply. */
static int orc_dbicg(const orc_mat_t *a, const orc_prec_t *pc, const double *b, double *x,
                     double eps, int itmax, int istop, double ani,
                     int *iter, double *err, double *hist)
{
  int n = a->nr;
  double **w = orc_vecs(8, n);
  double *q = w[0], *qt = w[1], *r = w[2], *rt = w[3], *p = w[4], *pt = w[5], *z = w[6];
  double *zt = w[7];
  orc_conv_t st = {istop, eps, 0, 0, ani, 0, itmax};
  double errnum = 0, errden = 1, alpha, beta, rho = 0, rho_old, sigma;
  int itx = 0;
  if (istop == 2) st.bn2 = orc_d_nrm2(n, b);
  if (istop == 1) st.bni = orc_d_amax(n, b);
  if (istop == 3) {
    memcpy(r, b, sizeof(double) * n);
    orc_spmm(a, -1.0, x, 1.0, r);
    st.r0n2 = orc_d_nrm2(n, r);
  }
  for (;;) { /* restart */
    if (itx >= itmax) break;
    int it = 0, done = 0;
    orc_d_axpby(n, 1.0, b, 0.0, r);
    orc_spmm(a, -1.0, x, 1.0, r);
    orc_d_axpby(n, 1.0, r, 0.0, rt);
    rho = 0.0;
    if (orc_check_conv(&st, itx, n, x, r, &errnum, &errden)) break;
    for (;;) {
      it++; itx++;
      orc_papply(pc, r, z);
      orc_papply(pc, rt, zt);
      rho_old = rho;
      rho = orc_d_dot(n, rt, z);
      if (rho == 0.0) break;
      if (it == 1) {
        orc_d_axpby(n, 1.0, z, 0.0, p);
        orc_d_axpby(n, 1.0, zt, 0.0, pt);
      } else {
        beta = rho / rho_old;
        orc_d_axpby(n, 1.0, z, beta, p);
        orc_d_axpby(n, 1.0, zt, beta, pt);
      }
      orc_spmm(a, 1.0, p, 0.0, q);
      orc_d_csr_csmv_t(n, n, a->irp, a->ja, a->val, 1.0, pt, 0.0, qt);
      sigma = orc_d_dot(n, pt, q);
      if (sigma == 0.0) break;
      alpha = rho / sigma;
      orc_d_axpby(n, alpha, p, 1.0, x);
      orc_d_axpby(n, -alpha, q, 1.0, r);
      orc_d_axpby(n, -alpha, qt, 1.0, rt);
      int conv = orc_check_conv(&st, itx, n, x, r, &errnum, &errden);
      if (hist) hist[itx] = errden > 0 ? errnum / errden : errnum;
      if (conv) { done = 1; break; }
    }
    if (done) break;
  }
  *iter = itx;
  *err = errden > 0 ? errnum / errden : errnum;
  orc_free_vecs(w, 8);
  return 0;
}

/* psb_dkrylov_vect with the further methods (krylov/psb_dkrylov.f90:178-206):
   2 FCG, 3 GCR, 4 RGMRES, 5 BiCGSTAB(l), 6 CGS, 7 BiCG; irst = restart
   length / l. */
EXPORT int orc_d_krylov2_csr(int method, int irst, int nr, const int *irp, const int *ja,
                             const double *val, const double *d, const double *b, double *x,
                             double eps, int itmax, int istop, double ani, int *iter,
                             double *err, double *hist)
{
  orc_mat_t a = {0, nr, irp, ja, val, 0, NULL, NULL, NULL, NULL};
  orc_prec_t pcv = {d, nr, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL};
  const orc_prec_t *pc = &pcv;
  switch (method) {
    case 2: return orc_dfcg(&a, pc, b, x, eps, itmax, istop, ani, iter, err, hist);
    case 3: return orc_dgcr(&a, pc, b, x, eps, itmax, istop, ani, irst, iter, err, hist);
    case 4: return orc_drgmres(&a, pc, b, x, eps, itmax, istop, ani, irst, iter, err, hist);
    case 5: return orc_dcgstabl(&a, pc, b, x, eps, itmax, istop, ani, irst, iter, err, hist);
    case 6: return orc_dcgs(&a, pc, b, x, eps, itmax, istop, ani, iter, err, hist);
    case 7: return orc_dbicg(&a, pc, b, x, eps, itmax, istop, ani, iter, err, hist);
    default: return orc_d_krylov_csr(method, nr, irp, ja, val, d, b, x, eps, itmax, istop, ani,
                                     iter, err, hist);
  }
}

/* psb_dkrylov_vect dispatcher (krylov/psb_dkrylov.f90:81-227): method 0 =
   CG, anything else = BiCGSTAB (the reference's default branch). */
EXPORT int orc_d_krylov_csr(int method, int nr, const int *irp, const int *ja, const double *val,
                            const double *d, const double *b, double *x, double eps, int itmax,
                            int istop, double ani, int *iter, double *err, double *hist)
{
  orc_mat_t a = {0, nr, irp, ja, val, 0, NULL, NULL, NULL, NULL};
  orc_prec_t pc = {d, nr, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL};
  if (method == 0) return orc_dcg(&a, &pc, b, x, eps, itmax, istop, ani, iter, err, hist);
  return orc_dcgstab(&a, &pc, b, x, eps, itmax, istop, ani, iter, err, hist);
}

EXPORT int orc_d_krylov_hll(int method, int nr, int hksz, const int *hkoffs, const int *irn,
                            const int *hja, const double *hval,
                            const double *d, const double *b, double *x, double eps, int itmax,
                            int istop, double ani, int *iter, double *err, double *hist)
{
  orc_mat_t a = {1, nr, NULL, NULL, NULL, hksz, hkoffs, irn, hja, hval};
  orc_prec_t pc = {d, nr, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL};
  if (method == 0) return orc_dcg(&a, &pc, b, x, eps, itmax, istop, ani, iter, err, hist);
  return orc_dcgstab(&a, &pc, b, x, eps, itmax, istop, ani, iter, err, hist);
}

/* Any method (0 CG, 1 BiCGSTAB, 2-5 as orc_d_krylov2_csr) with the BJAC /
   ILU(0) preconditioner of factors from orc_dilu0_fact. */
EXPORT int orc_d_krylov_ilu_csr(int method, int irst, int nr, const int *irp, const int *ja,
                                const double *val, const int *lirp, const int *lja,
                                const double *lval, const int *uirp, const int *uja,
                                const double *uval, const double *dd, const double *b,
                                double *x, double eps, int itmax, int istop, double ani,
                                int *iter, double *err, double *hist)
{
  orc_mat_t a = {0, nr, irp, ja, val, 0, NULL, NULL, NULL, NULL};
  double *tmp = calloc(nr > 0 ? nr : 1, sizeof(double));
  orc_prec_t pcv = {NULL, nr, lirp, lja, lval, uirp, uja, uval, dd, tmp};
  const orc_prec_t *pc = &pcv;
  int rc;
  switch (method) {
    case 0: rc = orc_dcg(&a, pc, b, x, eps, itmax, istop, ani, iter, err, hist); break;
    case 2: rc = orc_dfcg(&a, pc, b, x, eps, itmax, istop, ani, iter, err, hist); break;
    case 3: rc = orc_dgcr(&a, pc, b, x, eps, itmax, istop, ani, irst, iter, err, hist); break;
    case 4: rc = orc_drgmres(&a, pc, b, x, eps, itmax, istop, ani, irst, iter, err, hist); break;
    case 5: rc = orc_dcgstabl(&a, pc, b, x, eps, itmax, istop, ani, irst, iter, err, hist); break;
    default: rc = orc_dcgstab(&a, pc, b, x, eps, itmax, istop, ani, iter, err, hist); break;
  }
  free(tmp);
  return rc;
}

EXPORT int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
