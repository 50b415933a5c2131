// Host-side assembly: COO fix (sort + duplicate merge), HLL hack offsets,
// and the generators of the benchmark operators.
//
// psb_hip_fix_coo_d restates psb_d_fix_coo_inner_rowmajor
// (base/serial/impl/psb_d_coo_impl.F90:4252-4896) with a counting sort by
// row and a per-row stable column sort spread over host threads; the
// result is the reference's: rows ascending, columns ascending within a
// row, duplicates merged left to right in input order.
#include <ctype.h>
#include <math.h>

#include <cmath>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "psb_hip.h"

namespace {

int host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  if (n > 16) n = 16;
  return (int)n;
}

template <typename F>
void parallel_for(int64_t n, F f) {
  int nt = host_threads();
  if (n < 65536 || nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    int64_t lo = t * chunk, hi = std::min<int64_t>(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto& t : th) t.join();
}

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct Uniform {
  uint64_t key;
  Uniform(uint64_t seed, uint32_t stream)
      : key(mix64(seed * 0xD1B54A32D192ED03ULL ^ ((uint64_t)stream << 48))) {}
  double u01(uint64_t i) const { return (double)(mix64(key ^ i) >> 11) * 0x1.0p-53; }
  double pm1(uint64_t i) const { return 2.0 * u01(i) - 1.0; }
};

// ---------------------------------------------------------------- ILU(0)
// psb_dilu0_factint / psb_silu0_factint (prec/impl/psb_d_ilu0_fact.f90:330-460,
// psb_s_ilu0_fact.f90: the same in single precision), ialg ILU_N,
// upd 'F', no shift: IKJ elimination restricted to the pattern of A, L
// scaled by the inverse pivots, each U row scaled by its row's inverse pivot.

template <typename T>
int ilu0_fact(int n, const int* irp, const int* ja, const T* val, int* lirp, int* lja, T* lval,
              int* uirp, int* uja, T* uval, T* d, int* bad_row) {
  // pivot threshold: d_epstol / s_epstol (base/modules/psb_const_mod.F90:165-166)
  const T epstol = sizeof(T) == 8 ? T(1.1e-16) : T(5.e-8f);
  if (n < 0 || !irp || !lirp || !uirp || !d) return PSB_HIP_INVALID;
  if (bad_row) *bad_row = 0;
  int nl = 0, nu = 0;
  lirp[0] = 1;
  uirp[0] = 1;
  for (int i = 1; i <= n; i++) {
    T dia = T(0);
    for (int p = irp[i - 1]; p < irp[i]; p++) {
      const int k = ja[p - 1];
      if (k >= 1 && k < i) {
        lval[nl] = val[p - 1];
        lja[nl++] = k;
      } else if (k == i) {
        dia = val[p - 1];
      } else if (k > i && k <= n) {
        uval[nu] = val[p - 1];
        uja[nu++] = k;
      }
    }
    lirp[i] = nl + 1;
    uirp[i] = nu + 1;
    for (int kk = lirp[i - 1]; kk < lirp[i]; kk++) {
      const T lik = lval[kk - 1];
      const int k = lja[kk - 1];
      lval[kk - 1] = lik * d[k - 1];
      int lpos = kk + 1, upos = uirp[i - 1];
      for (int jj = uirp[k - 1]; jj < uirp[k]; jj++) {
        const int j = uja[jj - 1];
        const T u = uval[jj - 1];
        if (j == i) {
          dia = dia - lik * u;
        } else if (j < i) {
          for (; lpos < lirp[i]; lpos++) {
            if (lja[lpos - 1] > j) break;
            if (lja[lpos - 1] == j) {
              lval[lpos - 1] = lval[lpos - 1] - lik * u;
              lpos++;
              break;
            }
          }
        } else {
          for (; upos < uirp[i]; upos++) {
            if (uja[upos - 1] > j) break;
            if (uja[upos - 1] == j) {
              uval[upos - 1] = uval[upos - 1] - lik * u;
              upos++;
              break;
            }
          }
        }
      }
    }
    if (std::fabs(dia) < epstol) {   // psb_err_pivot_too_small_
      if (bad_row) *bad_row = i;
      return PSB_HIP_INVALID;
    }
    dia = T(1) / dia;
    d[i - 1] = dia;
    for (int kk = uirp[i - 1]; kk < uirp[i]; kk++) uval[kk - 1] = uval[kk - 1] * dia;
  }
  return PSB_HIP_SUCCESS;
}

// ---------------------------------------------------------------- host copies
// The CPU side of the dual-memory plug-ins: the matrix formats' own products
// (the csmv the GPU types inherit from the CPU HLL / ELL / CSR types) on the
// host copy, used when an operand is not a device vector
// (cuda/impl/psb_d_cuda_hlg_vect_mv.F90:77-108), and the host fills of the
// HLL / ELL layouts.  Each row sums its own entries in stored order with
// separately rounded products and sums (this file is compiled without
// contraction), then y = alpha*acc + beta*y (y not read when beta = 0; y =
// beta*y when alpha = 0), as psb_d_hll_csmv (ext/impl/psb_d_hll_csmv.f90:
// 291-371), psb_d_ell_csmv and psb_d_csr_csmv
// (base/serial/impl/psb_d_csr_impl.F90:46-417) compute them.

template <typename T>
inline void host_row_out(T* y, int64_t row, T alpha, T acc, T beta) {
  if (alpha == T(0)) y[row] = beta == T(0) ? T(0) : beta * y[row];
  else if (beta == T(0)) y[row] = alpha * acc;
  else y[row] = alpha * acc + beta * y[row];
}

template <typename T>
void host_hll_spmv(int nr, int hksz, const int* hkoffs, const int* irn, const int* ja,
                   const T* val, const int* perm, T alpha, const T* x, T beta, T* y) {
  const int noffs = (nr + hksz - 1) / hksz;
  parallel_for(noffs, [&](int64_t lo, int64_t hi) {
    for (int64_t h = lo; h < hi; h++) {
      const int64_t p0 = h * hksz, p1 = std::min<int64_t>(nr, p0 + hksz);
      for (int64_t p = p0; p < p1; p++) {
        const int64_t base = hkoffs[h] + (p - p0);
        T acc = T(0);
        for (int j = 0; j < irn[p]; j++) {
          const int64_t s = base + (int64_t)j * hksz;
          acc = acc + val[s] * x[ja[s] - 1];
        }
        host_row_out(y, perm ? perm[p] - 1 : p, alpha, acc, beta);
      }
    }
  });
}

template <typename T>
void host_ell_spmv(int nr, int ld, const int* irn, const int* ja, const T* val, T alpha,
                   const T* x, T beta, T* y) {
  parallel_for(nr, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) {
      T acc = T(0);
      for (int j = 0; j < irn[i]; j++) {
        const int64_t s = i + (int64_t)j * ld;
        acc = acc + val[s] * x[ja[s] - 1];
      }
      host_row_out(y, i, alpha, acc, beta);
    }
  });
}

template <typename T>
void host_csr_spmv(int nr, const int* irp, const int* ja, const T* val, T alpha, const T* x,
                   T beta, T* y) {
  parallel_for(nr, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) {
      T acc = T(0);
      for (int64_t k = irp[i] - 1; k < irp[i + 1] - 1; k++) acc = acc + val[k] * x[ja[k] - 1];
      host_row_out(y, i, alpha, acc, beta);
    }
  });
}

// HLL layout on the host from a row-sorted COO: the rule of the device fill
// (CPU padding: a row's last column, an empty row the hack's first entry,
// positions past nr ja = nr, val = 0; position p holds row perm[p] when perm
// is given).
template <typename T>
void host_hll_fill(int nr, int hksz, const int* irn, const int* hkoffs, const int* idisp,
                   const int* cja, const T* cval, const int* perm, int* ja, T* val, int* idiag) {
  const int noffs = (nr + hksz - 1) / hksz;
  parallel_for(noffs, [&](int64_t lo, int64_t hi) {
    for (int64_t h = lo; h < hi; h++) {
      const int off = hkoffs[h];
      const int w = (hkoffs[h + 1] - off) / hksz;
      for (int lane = 0; lane < hksz; lane++) {
        const int64_t i = h * hksz + lane;
        if (i >= nr) {
          for (int j = 0; j < w; j++) {
            ja[off + (int64_t)j * hksz + lane] = nr;
            val[off + (int64_t)j * hksz + lane] = T(0);
          }
          continue;
        }
        const int row = perm ? perm[i] - 1 : (int)i;
        const int rs = irn[i];
        const int p0 = idisp[row];
        int dg = 0, ic = 0;
        for (int j = 0; j < rs; j++) {
          ic = cja[p0 + j];
          if (ic == row + 1) dg = j + 1;
          ja[off + (int64_t)j * hksz + lane] = ic;
          val[off + (int64_t)j * hksz + lane] = cval[p0 + j];
        }
        if (rs == 0 && w > 0) {
          const int q = idisp[perm ? perm[h * hksz] - 1 : h * hksz];
          ic = q < idisp[nr] ? cja[q] : nr;
        }
        for (int j = rs; j < w; j++) {
          ja[off + (int64_t)j * hksz + lane] = ic;
          val[off + (int64_t)j * hksz + lane] = T(0);
        }
        idiag[i] = dg;
      }
    }
  });
}

}  // namespace

extern "C" {

int psb_hip_host_hll_spmv(int nr, int hksz, const int* hkoffs, const int* irn, const int* ja,
                          const void* val, const int* perm, int elem_type, double alpha,
                          const void* x, double beta, void* y) {
  if (nr < 0 || hksz < 1 || (nr > 0 && (!hkoffs || !irn || !x || !y))) return PSB_HIP_INVALID;
  if (elem_type == PSB_HIP_TYPE_DOUBLE)
    host_hll_spmv<double>(nr, hksz, hkoffs, irn, ja, (const double*)val, perm, alpha,
                          (const double*)x, beta, (double*)y);
  else if (elem_type == PSB_HIP_TYPE_FLOAT)
    host_hll_spmv<float>(nr, hksz, hkoffs, irn, ja, (const float*)val, perm, (float)alpha,
                         (const float*)x, (float)beta, (float*)y);
  else
    return PSB_HIP_UNSUPPORTED;
  return PSB_HIP_SUCCESS;
}

int psb_hip_host_ell_spmv(int nr, int ld, const int* irn, const int* ja, const void* val,
                          int elem_type, double alpha, const void* x, double beta, void* y) {
  if (nr < 0 || ld < nr || (nr > 0 && (!irn || !x || !y))) return PSB_HIP_INVALID;
  if (elem_type == PSB_HIP_TYPE_DOUBLE)
    host_ell_spmv<double>(nr, ld, irn, ja, (const double*)val, alpha, (const double*)x, beta,
                          (double*)y);
  else if (elem_type == PSB_HIP_TYPE_FLOAT)
    host_ell_spmv<float>(nr, ld, irn, ja, (const float*)val, (float)alpha, (const float*)x,
                         (float)beta, (float*)y);
  else
    return PSB_HIP_UNSUPPORTED;
  return PSB_HIP_SUCCESS;
}

int psb_hip_host_csr_spmv(int nr, const int* irp, const int* ja, const void* val, int elem_type,
                          double alpha, const void* x, double beta, void* y) {
  if (nr < 0 || (nr > 0 && (!irp || !x || !y))) return PSB_HIP_INVALID;
  if (elem_type == PSB_HIP_TYPE_DOUBLE)
    host_csr_spmv<double>(nr, irp, ja, (const double*)val, alpha, (const double*)x, beta,
                          (double*)y);
  else if (elem_type == PSB_HIP_TYPE_FLOAT)
    host_csr_spmv<float>(nr, irp, ja, (const float*)val, (float)alpha, (const float*)x,
                         (float)beta, (float*)y);
  else
    return PSB_HIP_UNSUPPORTED;
  return PSB_HIP_SUCCESS;
}

int psb_hip_host_hll_fill(int nr, int hksz, const int* irn, const int* hkoffs, const int* idisp,
                          const int* cja, const void* cval, const int* perm, int elem_type,
                          int* ja, void* val, int* idiag) {
  if (nr < 0 || hksz < 1) return PSB_HIP_INVALID;
  if (elem_type == PSB_HIP_TYPE_DOUBLE)
    host_hll_fill<double>(nr, hksz, irn, hkoffs, idisp, cja, (const double*)cval, perm, ja,
                          (double*)val, idiag);
  else if (elem_type == PSB_HIP_TYPE_FLOAT)
    host_hll_fill<float>(nr, hksz, irn, hkoffs, idisp, cja, (const float*)cval, perm, ja,
                         (float*)val, idiag);
  else
    return PSB_HIP_UNSUPPORTED;
  return PSB_HIP_SUCCESS;
}


int psb_hip_fix_coo_d(int nr, int nc, int64_t nz, int* ia, int* ja, double* val, int dupl,
                      int64_t* nzout) {
  (void)nc;
  if (nz < 2) {
    *nzout = nz;
    return PSB_HIP_SUCCESS;
  }
  if (dupl < 0 || dupl > 2) return PSB_HIP_INVALID;
  // 1. counts per row
  std::vector<int64_t> start((size_t)nr + 2, 0);
  for (int64_t k = 0; k < nz; k++) {
    int r = ia[k];
    if (r < 1 || r > nr) return PSB_HIP_INVALID;
    start[(size_t)r]++;
  }
  // 2. exclusive scan: rowptr[r] = first slot of row r+1
  std::vector<int64_t> rowptr((size_t)nr + 1);
  int64_t acc = 0;
  for (int r = 0; r < nr; r++) {
    rowptr[r] = acc;
    acc += start[(size_t)r + 1];
  }
  rowptr[nr] = acc;
  // 3. stable scatter by row
  std::vector<int> jas((size_t)nz);
  std::vector<double> vs((size_t)nz);
  {
    std::vector<int64_t> fill(rowptr.begin(), rowptr.end() - 1);
    for (int64_t k = 0; k < nz; k++) {
      int64_t p = fill[(size_t)ia[k] - 1]++;
      jas[(size_t)p] = ja[k];
      vs[(size_t)p] = val[k];
    }
  }
  // 4. per-row stable sort by column
  parallel_for(nr, [&](int64_t lo, int64_t hi) {
    std::vector<std::pair<int, int64_t>> tmp;
    std::vector<double> vtmp;
    for (int64_t r = lo; r < hi; r++) {
      int64_t b = rowptr[r], e = rowptr[r + 1];
      int64_t len = e - b;
      if (len < 2) continue;
      if (len <= 32) {  // insertion sort, stable
        for (int64_t i = b + 1; i < e; i++) {
          int c = jas[i];
          double v = vs[i];
          int64_t j = i - 1;
          while (j >= b && jas[j] > c) {
            jas[j + 1] = jas[j];
            vs[j + 1] = vs[j];
            j--;
          }
          jas[j + 1] = c;
          vs[j + 1] = v;
        }
      } else {
        tmp.resize(len);
        for (int64_t i = 0; i < len; i++) tmp[i] = {jas[b + i], i};
        std::stable_sort(tmp.begin(), tmp.end(),
                         [](const std::pair<int, int64_t>& x, const std::pair<int, int64_t>& y) {
                           return x.first < y.first;
                         });
        vtmp.assign(vs.begin() + b, vs.begin() + e);
        for (int64_t i = 0; i < len; i++) {
          jas[b + i] = tmp[i].first;
          vs[b + i] = vtmp[tmp[i].second];
        }
      }
    }
  });
  // 5. merge duplicates, left to right
  int64_t k = -1;
  for (int r = 0; r < nr; r++) {
    for (int64_t p = rowptr[r]; p < rowptr[r + 1]; p++) {
      if (k >= 0 && ia[k] == r + 1 && ja[k] == jas[p]) {
        if (dupl == 1) val[k] = val[k] + vs[p];
        else if (dupl == 0) val[k] = vs[p];
        else return PSB_HIP_INVALID;
      } else {
        k++;
        ia[k] = r + 1;
        ja[k] = jas[p];
        val[k] = vs[p];
      }
    }
  }
  *nzout = k + 1;
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_offsets(int nr, int64_t nz, const int* ia, int hksz, int* irn, int* idisp,
                        int* hkoffs, int64_t* isz) {
  if (hksz < 1 || nr < 0) return PSB_HIP_INVALID;
  memset(irn, 0, sizeof(int) * (size_t)nr);
  for (int64_t k = 0; k < nz; k++) {
    int r = ia[k];
    if (r < 1 || r > nr) return PSB_HIP_INVALID;
    irn[r - 1]++;
  }
  idisp[0] = 0;
  int64_t acc = 0;
  for (int i = 0; i < nr; i++) {
    acc += irn[i];
    if (acc > 0x7fffffffLL) return PSB_HIP_UNSUPPORTED;
    idisp[i + 1] = (int)acc;
  }
  int noffs = (nr + hksz - 1) / hksz;
  int64_t off = 0;
  hkoffs[0] = 0;
  for (int h = 0; h < noffs; h++) {
    int i0 = h * hksz;
    int i1 = std::min(nr, i0 + hksz);
    int mx = 0;
    for (int i = i0; i < i1; i++) mx = std::max(mx, irn[i]);
    off += (int64_t)mx * hksz;
    if (off > 0x7fffffffLL) return PSB_HIP_UNSUPPORTED;
    hkoffs[h + 1] = (int)off;
  }
  *isz = off;
  return PSB_HIP_SUCCESS;
}

int psb_hip_hll_sort_rows(int nr, int hksz, int window, const int* irn, int* perm, int* irn_pos,
                          int* hkoffs, int64_t* isz) {
  if (hksz < 1 || nr < 0 || !irn || !perm || !irn_pos || !hkoffs || !isz) return PSB_HIP_INVALID;
  int64_t win = window <= 0 ? (int64_t)nr : ((int64_t)window + hksz - 1) / hksz * hksz;
  if (win < 1) win = 1;
  for (int i = 0; i < nr; i++) {
    if (irn[i] < 0) return PSB_HIP_INVALID;
    perm[i] = i;
  }
  // descending length inside each window, ties in ascending row order
  const int64_t nwin = ((int64_t)nr + win - 1) / win;
  parallel_for(nwin, [&](int64_t lo, int64_t hi) {
    for (int64_t w = lo; w < hi; w++) {
      int* b = perm + w * win;
      int* e = perm + std::min<int64_t>((w + 1) * win, nr);
      std::stable_sort(b, e, [&](int a, int c) { return irn[a] > irn[c]; });
    }
  });
  for (int p = 0; p < nr; p++) {
    irn_pos[p] = irn[perm[p]];
    perm[p] += 1;  // 1-based row of position p
  }
  const int noffs = (nr + hksz - 1) / hksz;
  int64_t off = 0;
  hkoffs[0] = 0;
  for (int h = 0; h < noffs; h++) {
    const int p0 = h * hksz, p1 = std::min(nr, p0 + hksz);
    int mx = 0;
    for (int p = p0; p < p1; p++) mx = std::max(mx, irn_pos[p]);
    off += (int64_t)mx * hksz;
    if (off > 0x7fffffffLL) return PSB_HIP_UNSUPPORTED;
    hkoffs[h + 1] = (int)off;
  }
  *isz = off;
  return PSB_HIP_SUCCESS;
}

// ------------------------------------------------------------------ generators

static double g_sample(double x, double y, double z) {
  // boundary function of the pde3d sample, test/pargen/psb_d_pde3d.F90:160-170
  if (x == 1.0) return 1.0;
  if (x == 0.0) return exp(y * y - z * z);
  return 0.0;
}

int64_t psb_hip_gen_pde3d(int L, const double* a, const double* b, double c, int hscale,
                          int gflag, int* ia, int* ja, double* val, double* rhs) {
  const int64_t n = (int64_t)L * L * L;
  if (!ia) return 7 * n - 6 * (int64_t)L * L;
  const double h = hscale ? 1.0 / (L + 1) : 1.0;
  const double h2 = h * h, two_h = 2.0 * h;
  const double vxm = -a[0] / h2 - b[0] / two_h, vxp = -a[0] / h2 + b[0] / two_h;
  const double vym = -a[1] / h2 - b[1] / two_h, vyp = -a[1] / h2 + b[1] / two_h;
  const double vzm = -a[2] / h2 - b[2] / two_h, vzp = -a[2] / h2 + b[2] / two_h;
  const double vd = 2.0 * (a[0] + a[1] + a[2]) / h2 + c;
  const double gh = 1.0 / (L + 1);
  const int64_t plane = (int64_t)L * L;
  // one z-plane per task; entries before plane iz: interior planes have
  // 7L^2 - 4L entries (x and y boundaries), the two end planes L^2 fewer
  const int64_t per_plane = 7 * plane - 4 * (int64_t)L;
  parallel_for(L, [&](int64_t z0, int64_t z1) {
    for (int64_t izm = z0; izm < z1; izm++) {
      int iz = (int)izm + 1;
      int64_t k = izm * per_plane - (izm > 0 ? plane : 0);
      for (int iy = 1; iy <= L; iy++)
        for (int ix = 1; ix <= L; ix++) {
          int row = (int)(((int64_t)(iz - 1) * L + (iy - 1)) * L + ix);
          double x = (ix - 1) * gh, y = (iy - 1) * gh, z = (iz - 1) * gh;
          double zt = 0.0;
          if (ix == 1) {
            if (gflag) zt = g_sample(0.0, y, z) * (-vxm) + zt;
          } else {
            ia[k] = row; ja[k] = row - 1; val[k] = vxm; k++;
          }
          if (iy == 1) {
            if (gflag) zt = g_sample(x, 0.0, z) * (-vym) + zt;
          } else {
            ia[k] = row; ja[k] = row - L; val[k] = vym; k++;
          }
          if (iz == 1) {
            if (gflag) zt = g_sample(x, y, 0.0) * (-vzm) + zt;
          } else {
            ia[k] = row; ja[k] = (int)(row - plane); val[k] = vzm; k++;
          }
          ia[k] = row; ja[k] = row; val[k] = vd; k++;
          if (iz == L) {
            if (gflag) zt = g_sample(x, y, 1.0) * (-vzp) + zt;
          } else {
            ia[k] = row; ja[k] = (int)(row + plane); val[k] = vzp; k++;
          }
          if (iy == L) {
            if (gflag) zt = g_sample(x, 1.0, z) * (-vyp) + zt;
          } else {
            ia[k] = row; ja[k] = row + L; val[k] = vyp; k++;
          }
          if (ix == L) {
            if (gflag) zt = g_sample(1.0, y, z) * (-vxp) + zt;
          } else {
            ia[k] = row; ja[k] = row + 1; val[k] = vxp; k++;
          }
          if (rhs) rhs[row - 1] = zt;
        }
    }
  });
  return 7 * n - 6 * plane;
}

int64_t psb_hip_gen_aniso2d(int L, double eps, int* ia, int* ja, double* val) {
  const int64_t n = (int64_t)L * L;
  const int64_t total = 5 * n - 4 * (int64_t)L;
  if (!ia) return total;
  const double vd = 2.0 * eps + 2.0;
  int64_t k = 0;
  for (int iy = 1; iy <= L; iy++)
    for (int ix = 1; ix <= L; ix++) {
      int row = (iy - 1) * L + ix;
      if (ix > 1) { ia[k] = row; ja[k] = row - 1; val[k] = -eps; k++; }
      if (iy > 1) { ia[k] = row; ja[k] = row - L; val[k] = -1.0; k++; }
      ia[k] = row; ja[k] = row; val[k] = vd; k++;
      if (iy < L) { ia[k] = row; ja[k] = row + L; val[k] = -1.0; k++; }
      if (ix < L) { ia[k] = row; ja[k] = row + 1; val[k] = -eps; k++; }
    }
  return k;
}

// Power-law rows: length 1 + min(cap, n-1, floor(1.775 / v)), v in (0,1];
// columns uniform without repetition, diagonal first; values U[-1,1).
static int powerlaw_len(const Uniform& ul, int n, int cap, int i1) {
  double v = 1.0 - ul.u01((uint64_t)(i1 - 1));
  double kd = floor(1.775 / v);
  int kmax = std::min(cap, n - 1);
  int k = kd > (double)kmax ? kmax : (int)kd;
  return k + 1;
}

int64_t psb_hip_gen_powerlaw(int n, int cap, uint64_t seed, int* ia, int* ja, double* val) {
  Uniform ul(seed, 1), uc(seed, 2), uv(seed, 3);
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int i = 1; i <= n; i++) off[i] = off[i - 1] + powerlaw_len(ul, n, cap, i);
  if (!ia) return off[n];
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    std::vector<int> seen;
    for (int64_t i0 = lo; i0 < hi; i0++) {
      int i = (int)i0 + 1;
      int len = (int)(off[i] - off[i - 1]);
      int64_t k = off[i - 1];
      int hs = 16;
      while (hs < 2 * len) hs <<= 1;
      seen.assign((size_t)hs, 0);
      uint64_t base = (uint64_t)(i - 1) << 20;
      ia[k] = i; ja[k] = i; val[k] = uv.pm1(base); k++;
      int got = 1;
      uint64_t attempt = 0;
      while (got < len) {
        double u = uc.u01(base + attempt);
        attempt++;
        int col = 1 + (int)(u * (double)n);
        if (col > n) col = n;
        if (col == i) continue;
        unsigned h = ((unsigned)col * 2654435761u) & (unsigned)(hs - 1);
        bool dup = false;
        while (seen[h]) {
          if (seen[h] == col) { dup = true; break; }
          h = (h + 1) & (unsigned)(hs - 1);
        }
        if (dup) continue;
        seen[h] = col;
        ia[k] = i; ja[k] = col; val[k] = uv.pm1(base + (uint64_t)got); k++;
        got++;
      }
    }
  });
  return off[n];
}

void psb_hip_fill_uniform_host(uint64_t seed, uint32_t stream, int64_t n, double* x) {
  Uniform u(seed, stream);
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) x[i] = u.pm1((uint64_t)i);
  });
}

// ---------------------------------------------------------------- ILU(0)
int psb_hip_ilu0_fact(int n, const int* irp, const int* ja, const double* val, int* lirp,
                      int* lja, double* lval, int* uirp, int* uja, double* uval, double* d,
                      int* bad_row) {
  return ilu0_fact<double>(n, irp, ja, val, lirp, lja, lval, uirp, uja, uval, d, bad_row);
}

int psb_hip_ilu0_fact_s(int n, const int* irp, const int* ja, const float* val, int* lirp,
                        int* lja, float* lval, int* uirp, int* uja, float* uval, float* d,
                        int* bad_row) {
  return ilu0_fact<float>(n, irp, ja, val, lirp, lja, lval, uirp, uja, uval, d, bad_row);
}

// ---------------------------------------------------------------- MatrixMarket
// mm_mat_read / mm_dvet_read / the COO print and mm_dvet1_write
// (util/psb_d_mmio_impl.f90:32-102,226-280,315-486;
// base/serial/impl/psb_d_coo_impl.F90:728-763).

struct MMFile {
  FILE* f = nullptr;
  char object[32] = {0}, format[32] = {0}, type[32] = {0}, sym[32] = {0};
  long long nrow = 0, ncol = 0, nnz = 0;
  ~MMFile() {
    if (f) fclose(f);
  }
};

static void mm_lower(char* s) {
  for (; *s; s++) *s = (char)tolower((unsigned char)*s);
}

// Opens the file, reads the banner and the size line (nnz only for
// coordinate files).  0, or a PSB_HIP_* code.
static int mm_open(const char* path, bool coordinate, MMFile& m) {
  m.f = fopen(path, "r");
  if (!m.f) return PSB_HIP_RUNTIME;
  char line[1100];
  if (!fgets(line, sizeof line, m.f)) return PSB_HIP_INVALID;
  char banner[32];
  if (sscanf(line, "%31s %31s %31s %31s %31s", banner, m.object, m.format, m.type, m.sym) != 5)
    return PSB_HIP_INVALID;
  mm_lower(m.object);
  mm_lower(m.format);
  mm_lower(m.type);
  mm_lower(m.sym);
  if (strcmp(m.object, "matrix") != 0 || strcmp(m.format, coordinate ? "coordinate" : "array"))
    return PSB_HIP_UNSUPPORTED;
  do {
    if (!fgets(line, sizeof line, m.f)) return PSB_HIP_INVALID;
  } while (line[0] == '%');
  const int want = coordinate ? 3 : 2;
  if (sscanf(line, "%lld %lld %lld", &m.nrow, &m.ncol, &m.nnz) < want) return PSB_HIP_INVALID;
  if (m.nrow < 0 || m.ncol < 0 || m.nnz < 0 || m.nrow > 0x7fffffff || m.ncol > 0x7fffffff)
    return PSB_HIP_INVALID;
  return PSB_HIP_SUCCESS;
}

// The rest of the file in memory, NUL-terminated.
static bool mm_slurp(FILE* f, std::vector<char>& buf) {
  const long pos = ftell(f);
  if (fseek(f, 0, SEEK_END) != 0) return false;
  const long end = ftell(f);
  if (fseek(f, pos, SEEK_SET) != 0 || end < pos) return false;
  buf.resize((size_t)(end - pos) + 1);
  const size_t got = fread(buf.data(), 1, (size_t)(end - pos), f);
  buf[got] = 0;
  return true;
}

int psb_hip_mm_read_coo(const char* path, int64_t cap, int* nr, int* nc, int64_t* nz, int* ia,
                        int* ja, double* val) {
  MMFile m;
  int rc = mm_open(path, true, m);
  if (rc) return rc;
  const bool pattern = strcmp(m.type, "pattern") == 0;
  const bool symmetric = strcmp(m.sym, "symmetric") == 0;
  if ((!pattern && strcmp(m.type, "real") != 0) || (!symmetric && strcmp(m.sym, "general") != 0))
    return PSB_HIP_UNSUPPORTED;   // the reference reads real and pattern, general and symmetric
  *nr = (int)m.nrow;
  *nc = (int)m.ncol;
  const int64_t need = symmetric ? 2 * m.nnz : m.nnz;
  if (!ia) {                      // size query
    *nz = need;
    return PSB_HIP_SUCCESS;
  }
  if (cap < need) {
    *nz = need;
    return PSB_HIP_INVALID;
  }
  std::vector<char> buf;
  if (!mm_slurp(m.f, buf)) return PSB_HIP_RUNTIME;
  char* p = buf.data();
  for (int64_t k = 0; k < m.nnz; k++) {
    char* e;
    const long r = strtol(p, &e, 10);
    if (e == p) return PSB_HIP_INVALID;   // unexpected end of file or bad entry
    p = e;
    const long c = strtol(p, &e, 10);
    if (e == p) return PSB_HIP_INVALID;
    p = e;
    double v = 1.0;
    if (!pattern) {
      v = strtod(p, &e);
      if (e == p) return PSB_HIP_INVALID;
      p = e;
    }
    if (r < 1 || r > m.nrow || c < 1 || c > m.ncol) return PSB_HIP_INVALID;
    ia[k] = (int)r;
    ja[k] = (int)c;
    val[k] = v;
  }
  int64_t n = m.nnz;
  if (symmetric) {                // mirrored off-diagonal entries, appended in read order
    for (int64_t k = 0; k < m.nnz; k++) {
      if (ia[k] != ja[k]) {
        ia[n] = ja[k];
        ja[n] = ia[k];
        val[n] = val[k];
        n++;
      }
    }
  }
  *nz = n;
  return PSB_HIP_SUCCESS;
}

int psb_hip_mm_read_vec(const char* path, int64_t cap, int64_t* n, double* b) {
  MMFile m;
  int rc = mm_open(path, false, m);
  if (rc) return rc;
  if (strcmp(m.type, "real") != 0 || strcmp(m.sym, "general") != 0) return PSB_HIP_UNSUPPORTED;
  *n = m.nrow;
  if (!b) return PSB_HIP_SUCCESS;
  if (cap < m.nrow) return PSB_HIP_INVALID;
  std::vector<char> buf;
  if (!mm_slurp(m.f, buf)) return PSB_HIP_RUNTIME;
  char* p = buf.data();
  for (int64_t i = 0; i < m.nrow; i++) {   // the first column
    char* e;
    b[i] = strtod(p, &e);
    if (e == p) return PSB_HIP_INVALID;
    p = e;
  }
  return PSB_HIP_SUCCESS;
}

int psb_hip_mm_write_coo(const char* path, const char* title, int nr, int nc, int64_t nz,
                         const int* ia, const int* ja, const double* val) {
  FILE* f = fopen(path, "w");
  if (!f) return PSB_HIP_RUNTIME;
  fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
  if (title) fprintf(f, "%% %s\n", title);
  fprintf(f, "%%\n%% COO\n");
  fprintf(f, "%d %d %lld\n", nr, nc, (long long)nz);
  for (int64_t k = 0; k < nz; k++) fprintf(f, "%d %d %.18e\n", ia[k], ja[k], val[k]);
  const bool ok = ferror(f) == 0;
  fclose(f);
  return ok ? PSB_HIP_SUCCESS : PSB_HIP_RUNTIME;
}

int psb_hip_mm_write_vec(const char* path, const char* title, int64_t n, const double* b) {
  FILE* f = fopen(path, "w");
  if (!f) return PSB_HIP_RUNTIME;
  fprintf(f, "%%%%MatrixMarket matrix array real general\n");
  fprintf(f, "%% %s\n%% \n", title ? title : "");
  fprintf(f, "%lld 1\n", (long long)n);
  for (int64_t i = 0; i < n; i++) fprintf(f, "%.18e\n", b[i]);
  const bool ok = ferror(f) == 0;
  fclose(f);
  return ok ? PSB_HIP_SUCCESS : PSB_HIP_RUNTIME;
}

}  // extern "C"
