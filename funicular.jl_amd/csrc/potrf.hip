// Upper Cholesky factor of a k x k Hermitian (or symmetric) matrix on the
// device: cholesky_upper's plain factorization (src/panelblas.jl:148-155),
// kept next to the Gram accumulator it factors so that a CholQR pass does not
// round-trip G and R through the host.
//
// Blocked right-looking Cholesky, PB = 32 columns per block, four launches
// per block column J:
//   potrf_diag_kernel    factors the diagonal block R_JJ in LDS (one workgroup),
//   potrf_panel_kernel   solves R_JJ^H X = A_JL for each block L > J,
//   potrf_update_kernel  A_IL -= X_I^H X_L for every upper tile J < I <= L.
// A pivot that is not positive (or not finite) stops the factorization and
// records its 1-based column in *info, as LAPACK's potrf reports it; the
// kernels after it return at once. R's lower triangle is zero.
#include "fhip_internal.hpp"

#include <cmath>

namespace fhip {

constexpr int PB = 32;
constexpr int PT = 256;

template <typename TG, typename T, bool CPLX>
__global__ __launch_bounds__(PT) void potrf_prepare_kernel(T *__restrict__ R, int64_t ldr,
                                                           const TG *__restrict__ G, int64_t ldg, int k,
                                                           int *__restrict__ info) {
  const int64_t n = (int64_t)k * k;
  if (blockIdx.x == 0 && threadIdx.x == 0) *info = 0;
  for (int64_t e = (int64_t)blockIdx.x * PT + threadIdx.x; e < n; e += (int64_t)gridDim.x * PT) {
    const int64_t i = e % k;
    const int64_t j = e / k;
    TG re = TG(0), im = TG(0);
    if (i <= j) load_elem<TG, TG, CPLX>(G, i + j * ldg, true, re, im);
    store_elem<T, CPLX>(R, i + j * ldr, (T)re, (T)im);
  }
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(PT) void potrf_diag_kernel(T *__restrict__ R, int64_t ldr, int k, int J,
                                                        int *__restrict__ info) {
  if (*info != 0) return;
  __shared__ T are[PB][PB + 1];
  __shared__ T aim[PB][PB + 1];
  __shared__ int failed;
  const int t = threadIdx.x;
  const int c0 = J * PB;
  const int nb = (k - c0) < PB ? (k - c0) : PB;
  for (int e = t; e < PB * PB; e += PT) {
    const int i = e % PB, j = e / PB;
    T re = T(0), im = T(0);
    if (i < nb && j < nb && i <= j) load_elem<T, T, CPLX>(R, (c0 + i) + (int64_t)(c0 + j) * ldr, true, re, im);
    are[i][j] = re;
    aim[i][j] = im;
  }
  if (t == 0) failed = 0;
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (t == 0) {
      const T d = are[j][j];
      if (!(d > T(0)) || !isfinite(d)) {
        failed = c0 + j + 1;
      } else {
        are[j][j] = sqrt(d);
        aim[j][j] = T(0);
      }
    }
    __syncthreads();
    if (failed) break;
    const T r = are[j][j];
    for (int l = j + 1 + t; l < nb; l += PT) {
      are[j][l] /= r;
      aim[j][l] /= r;
    }
    __syncthreads();
    const int m = nb - j - 1;
    for (int e = t; e < m * m; e += PT) {
      const int i = j + 1 + e % m;
      const int l = j + 1 + e / m;
      if (i <= l) {
        // a_il -= conj(a_ji) a_jl
        const T xr = are[j][i], xi = aim[j][i], yr = are[j][l], yi = aim[j][l];
        are[i][l] -= xr * yr + xi * yi;
        if (CPLX) aim[i][l] -= xr * yi - xi * yr;
      }
    }
    __syncthreads();
  }
  if (failed) {
    if (t == 0) *info = failed;
    return;
  }
  for (int e = t; e < PB * PB; e += PT) {
    const int i = e % PB, j = e / PB;
    if (i < nb && j < nb && i <= j) store_elem<T, CPLX>(R, (c0 + i) + (int64_t)(c0 + j) * ldr, are[i][j], aim[i][j]);
  }
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(PT) void potrf_panel_kernel(T *__restrict__ R, int64_t ldr, int k, int J,
                                                         const int *__restrict__ info) {
  if (*info != 0) return;
  __shared__ T rre[PB][PB + 1];
  __shared__ T rim[PB][PB + 1];
  __shared__ T xre[PB][PB + 1];
  __shared__ T xim[PB][PB + 1];
  const int t = threadIdx.x;
  const int c0 = J * PB;  // a full block: J is not the last block column
  const int l0 = (J + 1 + blockIdx.x) * PB;
  const int nl = (k - l0) < PB ? (k - l0) : PB;
  for (int e = t; e < PB * PB; e += PT) {
    const int i = e % PB, j = e / PB;
    T re = T(0), im = T(0);
    if (i <= j) load_elem<T, T, CPLX>(R, (c0 + i) + (int64_t)(c0 + j) * ldr, true, re, im);
    rre[i][j] = re;
    rim[i][j] = im;
    re = T(0);
    im = T(0);
    if (j < nl) load_elem<T, T, CPLX>(R, (c0 + i) + (int64_t)(l0 + j) * ldr, true, re, im);
    xre[i][j] = re;
    xim[i][j] = im;
  }
  __syncthreads();
  // R_JJ^H X = A by forward substitution, one column of X per thread:
  // x_i = (a_i - sum_{m<i} conj(r_mi) x_m) / r_ii, with r_ii real.
  if (t < nl) {
    for (int i = 0; i < PB; ++i) {
      T sr = xre[i][t], si = xim[i][t];
      for (int m = 0; m < i; ++m) {
        const T ar = rre[m][i], ai = rim[m][i], br = xre[m][t], bi = xim[m][t];
        sr -= ar * br + ai * bi;
        if (CPLX) si -= ar * bi - ai * br;
      }
      xre[i][t] = sr / rre[i][i];
      xim[i][t] = si / rre[i][i];
    }
  }
  __syncthreads();
  for (int e = t; e < PB * PB; e += PT) {
    const int i = e % PB, j = e / PB;
    if (j < nl) store_elem<T, CPLX>(R, (c0 + i) + (int64_t)(l0 + j) * ldr, xre[i][j], xim[i][j]);
  }
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(PT) void potrf_update_kernel(T *__restrict__ R, int64_t ldr, int k, int J, int m,
                                                          const int *__restrict__ info) {
  if (*info != 0) return;
  __shared__ T pre[PB][PB + 1];
  __shared__ T pim[PB][PB + 1];
  __shared__ T qre[PB][PB + 1];
  __shared__ T qim[PB][PB + 1];
  // Upper tile t of the m x m trailing block grid -> (I, L), I <= L, numbered
  // column by column.
  int L = 0;
  int rem = blockIdx.x;
  while (rem > L) {
    rem -= L + 1;
    ++L;
  }
  const int I = rem;
  const int c0 = J * PB;
  const int i0 = (J + 1 + I) * PB;
  const int l0 = (J + 1 + L) * PB;
  const int ni = (k - i0) < PB ? (k - i0) : PB;
  const int nl = (k - l0) < PB ? (k - l0) : PB;
  const int t = threadIdx.x;
  for (int e = t; e < PB * PB; e += PT) {
    const int r = e % PB, c = e / PB;
    T re = T(0), im = T(0);
    if (c < ni) load_elem<T, T, CPLX>(R, (c0 + r) + (int64_t)(i0 + c) * ldr, true, re, im);
    pre[r][c] = re;
    pim[r][c] = im;
    re = T(0);
    im = T(0);
    if (c < nl) load_elem<T, T, CPLX>(R, (c0 + r) + (int64_t)(l0 + c) * ldr, true, re, im);
    qre[r][c] = re;
    qim[r][c] = im;
  }
  __syncthreads();
  for (int e = t; e < PB * PB; e += PT) {
    const int i = e % PB, l = e / PB;
    if (i >= ni || l >= nl || (I == L && i > l)) continue;
    T sr = T(0), si = T(0);
    for (int q = 0; q < PB; ++q) {
      // conj(x_qi) x_ql
      const T ar = pre[q][i], ai = pim[q][i], br = qre[q][l], bi = qim[q][l];
      sr += ar * br + ai * bi;
      if (CPLX) si += ar * bi - ai * br;
    }
    T dr, di;
    load_elem<T, T, CPLX>(R, (i0 + i) + (int64_t)(l0 + l) * ldr, true, dr, di);
    store_elem<T, CPLX>(R, (i0 + i) + (int64_t)(l0 + l) * ldr, dr - sr, di - si);
  }
}

template <typename TG, typename T, bool CPLX>
static int launch_potrf(T *R, int64_t ldr, const TG *G, int64_t ldg, int k, int *info, hipStream_t s) {
  int64_t g = cdiv64((int64_t)k * k, PT);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL((potrf_prepare_kernel<TG, T, CPLX>), dim3((unsigned)g), dim3(PT), 0, s, R, ldr, G, ldg, k, info);
  int st = check_launch("potrf_prepare_kernel");
  if (st) return st;
  const int nbk = (k + PB - 1) / PB;
  for (int J = 0; J < nbk; ++J) {
    hipLaunchKernelGGL((potrf_diag_kernel<T, CPLX>), dim3(1), dim3(PT), 0, s, R, ldr, k, J, info);
    if ((st = check_launch("potrf_diag_kernel"))) return st;
    const int m = nbk - J - 1;
    if (m == 0) break;
    hipLaunchKernelGGL((potrf_panel_kernel<T, CPLX>), dim3(m), dim3(PT), 0, s, R, ldr, k, J, (const int *)info);
    if ((st = check_launch("potrf_panel_kernel"))) return st;
    hipLaunchKernelGGL((potrf_update_kernel<T, CPLX>), dim3(m * (m + 1) / 2), dim3(PT), 0, s, R, ldr, k, J, m,
                       (const int *)info);
    if ((st = check_launch("potrf_update_kernel"))) return st;
  }
  return FHIP_OK;
}

}  // namespace fhip

using namespace fhip;

extern "C" int fhip_potrf_upper(void *R, int64_t ldr, int rtype, const void *G, int64_t ldg, int gtype, int64_t k,
                                int *info, void *stream) {
  if (!valid_type(rtype) || !valid_type(gtype))
    return fail(FHIP_E_UNSUPPORTED, "fhip_potrf_upper: unknown type %d / %d", rtype, gtype);
  const bool same = rtype == gtype;
  const bool narrowing = (gtype == FHIP_C64 && rtype == FHIP_C32) || (gtype == FHIP_F64 && rtype == FHIP_F32);
  if (!same && !narrowing)
    return fail(FHIP_E_UNSUPPORTED, "fhip_potrf_upper: cannot factor a type %d matrix into type %d", gtype, rtype);
  if (k < 0) return fail(FHIP_E_ARGUMENT, "fhip_potrf_upper: negative dimension");
  if (k > (1 << 16)) return fail(FHIP_E_ARGUMENT, "fhip_potrf_upper: too many columns (%lld)", (long long)k);
  if (!info) return fail(FHIP_E_ARGUMENT, "fhip_potrf_upper: null info pointer");
  if (k > 0 && (!R || !G)) return fail(FHIP_E_ARGUMENT, "fhip_potrf_upper: null matrix pointer");
  if (k > 0 && (ldr < k || ldg < k))
    return fail(FHIP_E_ARGUMENT, "fhip_potrf_upper: leading dimension smaller than the matrix");
  hipStream_t s = (hipStream_t)stream;
  const int ik = (int)k;
  if (k == 0) {
    hipLaunchKernelGGL((potrf_prepare_kernel<double, double, false>), dim3(1), dim3(PT), 0, s, (double *)nullptr, 1,
                       (const double *)nullptr, 1, 0, info);
    return check_launch("potrf_prepare_kernel");
  }
  switch (rtype) {
    case FHIP_C64:
      return launch_potrf<double, double, true>((double *)R, ldr, (const double *)G, ldg, ik, info, s);
    case FHIP_F64:
      return launch_potrf<double, double, false>((double *)R, ldr, (const double *)G, ldg, ik, info, s);
    case FHIP_C32:
      if (same) return launch_potrf<float, float, true>((float *)R, ldr, (const float *)G, ldg, ik, info, s);
      return launch_potrf<double, float, true>((float *)R, ldr, (const double *)G, ldg, ik, info, s);
    case FHIP_F32:
      if (same) return launch_potrf<float, float, false>((float *)R, ldr, (const float *)G, ldg, ik, info, s);
      return launch_potrf<double, float, false>((float *)R, ldr, (const double *)G, ldg, ik, info, s);
  }
  return fail(FHIP_E_UNSUPPORTED, "fhip_potrf_upper: unknown type %d", rtype);
}
