// Kernels (b) and (b'): the right-upper triangular solve Y <- Y R^-1 that
// replaces rdiv_upper! (src/kernels.jl:108, semantics :90-99, called from
// rdiv_rows! src/panelblas.jl:157-169), and the row-block product D <- A C
// that replaces rightmul_gemm! (src/kernels.jl:73, semantics :57-64, called
// from rightmul! src/panelblas.jl:203-245).
//
// Both are "rows are independent" products: a workgroup owns RW rows. The
// solve is blocked by CB columns and left looking,
//     X_J = (Y_J - sum_{I<J} X_I R_IJ) * inv(R_JJ),
// with the CB x CB diagonal inverses formed once per call by a small kernel.
// X_J overwrites Y_J in place, and later blocks read it back. The products
// run on v_mfma_f64_16x16x4 (f32 form for single precision), with operand
// slices staged through LDS.
#include "fhip_internal.hpp"

namespace fhip {

constexpr int NN_THREADS = 256;
constexpr int NN_RW = 64;   // rows per workgroup
constexpr int NN_CB = 32;   // output columns per block
constexpr int NN_KC = 32;   // inner dimension staged per LDS pass (== NN_CB)
constexpr int NN_LDA = NN_RW + 4;
constexpr int NN_LDB = NN_CB + 4;

// As[kk][r] <- A[row0 + r, kc + kk] for kk < NN_KC, r < NN_RW (zero outside
// rows < m and columns < kmax).
template <typename TIn, typename T, bool CPLX>
__device__ inline void stage_rows(T *__restrict__ As_re, T *__restrict__ As_im, const TIn *__restrict__ A,
                                  int64_t lda, int64_t row0, int64_t m, int kc, int kmax) {
  const int t = threadIdx.x;
#pragma unroll
  for (int e = 0; e < NN_KC * NN_RW / NN_THREADS; ++e) {
    const int idx = t + NN_THREADS * e;
    const int r = idx % NN_RW;
    const int kk = idx / NN_RW;
    const int64_t row = row0 + r;
    const int col = kc + kk;
    T re, im;
    load_elem<TIn, T, CPLX>(A, row + (int64_t)col * lda, row < m && col < kmax, re, im);
    As_re[kk * NN_LDA + r] = re;
    if (CPLX) As_im[kk * NN_LDA + r] = im;
  }
}

// Bs[kk][c] <- B[kc + kk, c0 + c] for kk < NN_KC, c < NN_CB (zero outside
// rows < kmax and columns < ncols).
template <typename TIn, typename T, bool CPLX>
__device__ inline void stage_cols(T *__restrict__ Bs_re, T *__restrict__ Bs_im, const TIn *__restrict__ B,
                                  int64_t ldb, int kc, int kmax, int c0, int ncols) {
  const int t = threadIdx.x;
#pragma unroll
  for (int e = 0; e < NN_KC * NN_CB / NN_THREADS; ++e) {
    const int idx = t + NN_THREADS * e;
    const int kk = idx % NN_KC;
    const int c = idx / NN_KC;
    const int krow = kc + kk;
    const int col = c0 + c;
    T re, im;
    load_elem<TIn, T, CPLX>(B, (int64_t)krow + (int64_t)col * ldb, krow < kmax && col < ncols, re, im);
    Bs_re[kk * NN_LDB + c] = re;
    if (CPLX) Bs_im[kk * NN_LDB + c] = im;
  }
}

// Inverse of each NN_CB x NN_CB diagonal block of R (the last one cut to what
// is left of k), zero padded, column major, one block after another. One
// thread per column, back substitution down the column.
template <typename TS, typename T, bool CPLX>
__global__ __launch_bounds__(NN_CB) void trsm_diag_inv_kernel(const TS *__restrict__ R, int64_t ldr, int k,
                                                               T *__restrict__ Dinv) {
  constexpr int NP = CPLX ? 2 : 1;
  const int J = blockIdx.x;
  const int j = threadIdx.x;
  const int c0 = J * NN_CB;
  const int nb = (k - c0) < NN_CB ? (k - c0) : NN_CB;
  T *Z = Dinv + (int64_t)J * NN_CB * NN_CB * NP;
  for (int i = 0; i < NN_CB; ++i)
    for (int p = 0; p < NP; ++p) Z[(i + j * NN_CB) * NP + p] = T(0);
  if (j >= nb) return;
  auto rd = [&](int i, int l, double &re, double &im) {
    const int64_t idx = (int64_t)(c0 + i) + (int64_t)(c0 + l) * ldr;
    if (CPLX) {
      re = (double)R[2 * idx];
      im = (double)R[2 * idx + 1];
    } else {
      re = (double)R[idx];
      im = 0.0;
    }
  };
  // Column j of the inverse, held in double while it is formed.
  double zr[NN_CB], zi[NN_CB];
  {
    double tr, ti;
    rd(j, j, tr, ti);
    const double d = tr * tr + ti * ti;
    zr[j] = tr / d;
    zi[j] = -ti / d;
  }
  for (int i = j - 1; i >= 0; --i) {
    double sr = 0.0, si = 0.0;
    for (int l = i + 1; l <= j; ++l) {
      double tr, ti;
      rd(i, l, tr, ti);
      sr += tr * zr[l] - ti * zi[l];
      si += tr * zi[l] + ti * zr[l];
    }
    double tr, ti;
    rd(i, i, tr, ti);
    const double d = tr * tr + ti * ti;
    // z = -s / t
    zr[i] = -(sr * tr + si * ti) / d;
    zi[i] = -(si * tr - sr * ti) / d;
  }
  for (int i = 0; i <= j; ++i) {
    Z[(i + j * NN_CB) * NP] = (T)zr[i];
    if (CPLX) Z[(i + j * NN_CB) * NP + 1] = (T)zi[i];
  }
}

// The solve proper. A workgroup owns TR_RW rows; wave w holds rows
// [32 w, 32 w + 32) x the 32 columns of the current block as 2 x 2 blocks of
// the 16x16 matrix-core tile. The update streams TR_KC columns of the solved
// rows and of R at a time, the next slice's loads in flight during the
// current slice's products.
#ifndef FHIP_TRSM_KC
#define FHIP_TRSM_KC 16
#endif
constexpr int TR_RW = 128;
constexpr int TR_KC = FHIP_TRSM_KC;
constexpr int TR_LDA = TR_RW + 4;
constexpr int TR_LDB = NN_CB + 4;
constexpr int TR_XPT = TR_KC * TR_RW / NN_THREADS;  // X elements per thread per slice (8)
constexpr int TR_RPT = TR_KC * NN_CB / NN_THREADS;  // R elements per thread per slice (2)

// CT = false solves X R = Y (left to right over column blocks); CT = true
// solves X R^H = Y, whose factor R^H is lower triangular, right to left.
template <typename TS, typename T, bool CPLX, bool CT>
__global__ __launch_bounds__(NN_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void trsm_runn_kernel(
    TS *__restrict__ Y, int64_t ldy, int64_t m, int k, const TS *__restrict__ R, int64_t ldr,
    const T *__restrict__ Dinv) {
  constexpr int NP = CPLX ? 2 : 1;
  __shared__ T As[NP][TR_KC * TR_LDA];
  __shared__ T Bs[NP][TR_KC * TR_LDB];
  typedef typename Mfma<T>::acc acc_t;

  const int64_t row0 = (int64_t)blockIdx.x * TR_RW;
  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  const int li = lane & 15;
  const int lk = lane >> 4;
  const int mi0 = 32 * w;
  const int nbk = (k + NN_CB - 1) / NN_CB;

  T xr[TR_XPT], xi[TR_XPT], rr[TR_RPT], ri[TR_RPT];
  // Slice kc of the update: X[row0 + r, kc + kk] and R[kc + kk, c0 + c].
  auto fetch = [&](int kc, int c0) {
#pragma unroll
    for (int e = 0; e < TR_XPT; ++e) {
      const int idx = t + NN_THREADS * e;
      const int r = idx % TR_RW;
      const int kk = idx / TR_RW;
      const int64_t row = row0 + r;
      load_elem<TS, T, CPLX>(Y, row + (int64_t)(kc + kk) * ldy, row < m && kc + kk < k, xr[e], xi[e]);
    }
#pragma unroll
    for (int e = 0; e < TR_RPT; ++e) {
      const int idx = t + NN_THREADS * e;
      const int kk = idx % TR_KC;
      const int c = idx / TR_KC;
      if (CT) {
        // Element (kc + kk, c0 + c) of R^H is conj(R[c0 + c, kc + kk]).
        load_elem<TS, T, CPLX>(R, (int64_t)(c0 + c) + (int64_t)(kc + kk) * ldr, kc + kk < k && c0 + c < k,
                               rr[e], ri[e]);
        ri[e] = -ri[e];
      } else {
        load_elem<TS, T, CPLX>(R, (int64_t)(kc + kk) + (int64_t)(c0 + c) * ldr, c0 + c < k, rr[e], ri[e]);
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int e = 0; e < TR_XPT; ++e) {
      const int idx = t + NN_THREADS * e;
      const int r = idx % TR_RW;
      const int kk = idx / TR_RW;
      As[0][kk * TR_LDA + r] = xr[e];
      if (CPLX) As[NP - 1][kk * TR_LDA + r] = xi[e];
    }
#pragma unroll
    for (int e = 0; e < TR_RPT; ++e) {
      const int idx = t + NN_THREADS * e;
      const int kk = idx % TR_KC;
      const int c = idx / TR_KC;
      Bs[0][kk * TR_LDB + c] = rr[e];
      if (CPLX) Bs[NP - 1][kk * TR_LDB + c] = ri[e];
    }
  };

  for (int step = 0; step < nbk; ++step) {
    const int J = CT ? nbk - 1 - step : step;
    const int c0 = J * NN_CB;
    // The block's right-hand side seeds the update sums. Complex blocks keep
    // the three 3M sums: p1 = Y_re, p2 = 0, p3 = Y_re + Y_im, counted down by
    // the products, so that combine3m yields Y - sum X_I R_IJ directly.
    acc_t p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = row0 + mi0 + 16 * mb + Mfma<T>::row(lk, r);
          const int col = c0 + 16 * nb + li;
          T re, im;
          load_elem<TS, T, CPLX>(Y, row + (int64_t)col * ldy, row < m && col < k, re, im);
          p1[mb][nb][r] = re;
          p2[mb][nb][r] = CPLX ? T(0) : im;
          p3[mb][nb][r] = re + im;
        }
    // Update from the blocks already solved: those left of J for R, right of
    // J for R^H.
    const int kbeg = CT ? c0 + NN_CB : 0;
    const int kend = CT ? k : c0;
    if (kbeg < kend) fetch(kbeg, c0);
    for (int kc = kbeg; kc < kend; kc += TR_KC) {
      __syncthreads();
      stash();
      __syncthreads();
      if (kc + TR_KC < kend) fetch(kc + TR_KC, c0);
      if constexpr (CPLX)
        mma_steps_3m<T, 2, 2, PROD_SUB>(As[0], As[NP - 1], TR_LDA, mi0, Bs[0], Bs[NP - 1], TR_LDB, 0, 0,
                                        TR_KC / 4, 1, p1, p2, p3);
      else
        mma_steps<T, false, 2, 2, PROD_SUB>(As[0], As[0], TR_LDA, mi0, Bs[0], Bs[0], TR_LDB, 0, 0, TR_KC / 4,
                                            1, p1, p2);
    }
    acc_t acc_re[2][2], acc_im[2][2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        acc_re[mb][nb] = p1[mb][nb];
        acc_im[mb][nb] = p2[mb][nb];
        if constexpr (CPLX) combine3m<PROD_ADD>(p1[mb][nb], p2[mb][nb], p3[mb][nb], acc_re[mb][nb], acc_im[mb][nb]);
      }
    // Multiply by the inverse of the diagonal block, one 16-column half at a
    // time. This product is an eighth of the work at k = 512; it keeps the
    // four-product form, which needs two accumulator sets instead of three.
    acc_t x_re[2][2], x_im[2][2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        x_re[mb][nb] = Mfma<T>::zero();
        x_im[mb][nb] = Mfma<T>::zero();
      }
    const T *Dj = Dinv + (int64_t)J * NN_CB * NN_CB * NP;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      __syncthreads();
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = mi0 + 16 * mb + Mfma<T>::row(lk, r);
          As[0][li * TR_LDA + row] = acc_re[mb][h][r];
          if (CPLX) As[NP - 1][li * TR_LDA + row] = acc_im[mb][h][r];
        }
#pragma unroll
      for (int e = 0; e < TR_RPT; ++e) {
        const int idx = t + NN_THREADS * e;
        const int kk = idx % TR_KC;
        const int c = idx / TR_KC;
        T re, im;
        if (CT) {
          // inv(R_JJ^H) = inv(R_JJ)^H.
          load_elem<T, T, CPLX>(Dj, (int64_t)c + (int64_t)(16 * h + kk) * NN_CB, true, re, im);
          im = -im;
        } else {
          load_elem<T, T, CPLX>(Dj, (int64_t)(16 * h + kk) + (int64_t)c * NN_CB, true, re, im);
        }
        Bs[0][kk * TR_LDB + c] = re;
        if (CPLX) Bs[NP - 1][kk * TR_LDB + c] = im;
      }
      __syncthreads();
      mma_steps<T, CPLX, 2, 2, PROD_ADD>(As[0], As[NP - 1], TR_LDA, mi0, Bs[0], Bs[NP - 1], TR_LDB, 0, 0,
                                         TR_KC / 4, 1, x_re, x_im);
    }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = row0 + mi0 + 16 * mb + Mfma<T>::row(lk, r);
          const int col = c0 + 16 * nb + li;
          if (row < m && col < k)
            store_elem<TS, CPLX>(Y, row + (int64_t)col * ldy, (TS)x_re[mb][nb][r], (TS)x_im[mb][nb][r]);
        }
    // The stores above are read back by the next block's update: the barrier
    // is a workgroup-scope release/acquire.
    __syncthreads();
  }
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(NN_THREADS) void gemm_nn_kernel(T *__restrict__ D, int64_t ldd,
                                                              const T *__restrict__ A, int64_t lda,
                                                              const T *__restrict__ C, int64_t ldc,
                                                              int64_t m, int s, int r, int ncb,
                                                              int64_t nrc) {
  constexpr int NP = CPLX ? 2 : 1;
  __shared__ T As[NP][NN_KC * NN_LDA];
  __shared__ T Bs[NP][NN_KC * NN_LDB];
  typedef typename Mfma<T>::acc acc_t;

  // Column blocks of one row chunk share an XCD (ids go round-robin to XCDs).
  const int id = blockIdx.x;
  const int xcd = id & 7;
  const int local = id >> 3;
  const int cb = local % ncb;
  const int64_t rc = (int64_t)(local / ncb) * 8 + xcd;
  if (rc >= nrc) return;
  const int64_t row0 = rc * NN_RW;
  const int c0 = cb * NN_CB;

  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  const int li = lane & 15;
  const int lk = lane >> 4;
  const int mi0 = 16 * w;

  acc_t p1[1][2], p2[1][2], p3[1][2];
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    p1[0][nb] = Mfma<T>::zero();
    p2[0][nb] = Mfma<T>::zero();
    p3[0][nb] = Mfma<T>::zero();
  }
  for (int kc = 0; kc < s; kc += NN_KC) {
    __syncthreads();
    stage_rows<T, T, CPLX>(As[0], As[NP - 1], A, lda, row0, m, kc, s);
    stage_cols<T, T, CPLX>(Bs[0], Bs[NP - 1], C, ldc, kc, s, c0, r);
    __syncthreads();
    if constexpr (CPLX)
      mma_steps_3m<T, 1, 2, PROD_ADD>(As[0], As[NP - 1], NN_LDA, mi0, Bs[0], Bs[NP - 1], NN_LDB, 0, 0, NN_KC / 4,
                                      1, p1, p2, p3);
    else
      mma_steps<T, false, 1, 2, PROD_ADD>(As[0], As[0], NN_LDA, mi0, Bs[0], Bs[0], NN_LDB, 0, 0, NN_KC / 4, 1,
                                          p1, p2);
  }
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    acc_t re = p1[0][nb], im = p2[0][nb];
    if constexpr (CPLX) combine3m<PROD_ADD>(p1[0][nb], p2[0][nb], p3[0][nb], re, im);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t row = row0 + mi0 + Mfma<T>::row(lk, q);
      const int col = c0 + 16 * nb + li;
      if (row < m && col < r) store_elem<T, CPLX>(D, row + (int64_t)col * ldd, re[q], im[q]);
    }
  }
}

template <typename TS, typename T, bool CPLX, bool CT = false>
static int launch_trsm(TS *Y, int64_t ldy, int64_t m, int k, const TS *R, int64_t ldr, T *Dinv,
                       hipStream_t stream) {
  const int nbk = (k + NN_CB - 1) / NN_CB;
  hipLaunchKernelGGL((trsm_diag_inv_kernel<TS, T, CPLX>), dim3(nbk), dim3(NN_CB), 0, stream, R, ldr, k, Dinv);
  int st = check_launch("trsm_diag_inv_kernel");
  if (st || m == 0) return st;
  const int64_t grid = cdiv64(m, TR_RW);
  hipLaunchKernelGGL((trsm_runn_kernel<TS, T, CPLX, CT>), dim3((unsigned)grid), dim3(NN_THREADS), 0, stream, Y, ldy,
                     m, k, R, ldr, (const T *)Dinv);
  return check_launch("trsm_runn_kernel");
}

template <typename T, bool CPLX>
static int launch_gemm(T *D, int64_t ldd, const T *A, int64_t lda, const T *C, int64_t ldc, int64_t m,
                       int s, int r, hipStream_t stream) {
  const int ncb = (r + NN_CB - 1) / NN_CB;
  const int64_t nrc = cdiv64(m, NN_RW);
  const int64_t grid = 8 * cdiv64(nrc, 8) * ncb;
  hipLaunchKernelGGL((gemm_nn_kernel<T, CPLX>), dim3((unsigned)grid), dim3(NN_THREADS), 0, stream, D, ldd, A, lda,
                     C, ldc, m, s, r, ncb, nrc);
  return check_launch("gemm_nn_kernel");
}

}  // namespace fhip

using namespace fhip;

extern "C" int fhip_trsm_workspace_bytes(size_t *nbytes, int64_t k, int compute) {
  if (!nbytes) return fail(FHIP_E_ARGUMENT, "fhip_trsm_workspace_bytes: null output pointer");
  if (!valid_type(compute)) return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_runn: unknown type %d", compute);
  if (k < 0) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runn: negative dimension");
  *nbytes = (size_t)cdiv64(k, NN_CB) * NN_CB * NN_CB * eltype_size(compute);
  return FHIP_OK;
}

extern "C" int fhip_trsm_runn(void *Y, int64_t ldy, int64_t m, int64_t k, const void *R, int64_t ldr,
                              int type, int compute, void *work, size_t work_bytes, void *stream) {
  if (!valid_type(type) || !valid_type(compute))
    return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_runn: unknown type %d / %d", type, compute);
  if (!(compute == type || (type == FHIP_C64 && compute == FHIP_C32)))
    return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_runn: no kernel for storage type %d with compute type %d", type,
                compute);
  if (m < 0 || k < 0) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runn: negative dimension");
  if (k == 0) return FHIP_OK;
  if (k > (1 << 20)) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runn: too many columns");
  if (!Y || !R) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runn: null matrix pointer");
  if (ldr < k || (m > 0 && ldy < m))
    return fail(FHIP_E_ARGUMENT, "fhip_trsm_runn: leading dimension smaller than the matrix");
  size_t need = 0;
  fhip_trsm_workspace_bytes(&need, k, compute);
  if (!work || work_bytes < need)
    return fail(FHIP_E_WORKSPACE, "fhip_trsm_runn: workspace of %zu bytes is smaller than the %zu needed",
                work_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  const int ik = (int)k;
  if (type == FHIP_C64 && compute == FHIP_C32)
    return launch_trsm<double, float, true>((double *)Y, ldy, m, ik, (const double *)R, ldr, (float *)work, s);
  switch (type) {
    case FHIP_C64:
      return launch_trsm<double, double, true>((double *)Y, ldy, m, ik, (const double *)R, ldr, (double *)work, s);
    case FHIP_C32:
      return launch_trsm<float, float, true>((float *)Y, ldy, m, ik, (const float *)R, ldr, (float *)work, s);
    case FHIP_F64:
      return launch_trsm<double, double, false>((double *)Y, ldy, m, ik, (const double *)R, ldr, (double *)work, s);
    case FHIP_F32:
      return launch_trsm<float, float, false>((float *)Y, ldy, m, ik, (const float *)R, ldr, (float *)work, s);
  }
  return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_runn: unknown type %d", type);
}

extern "C" int fhip_trsm_runc(void *Y, int64_t ldy, int64_t m, int64_t k, const void *R, int64_t ldr, int type,
                              void *work, size_t work_bytes, void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_runc: unknown type %d", type);
  if (m < 0 || k < 0) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runc: negative dimension");
  if (k == 0) return FHIP_OK;
  if (k > (1 << 20)) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runc: too many columns");
  if (!Y || !R) return fail(FHIP_E_ARGUMENT, "fhip_trsm_runc: null matrix pointer");
  if (ldr < k || (m > 0 && ldy < m))
    return fail(FHIP_E_ARGUMENT, "fhip_trsm_runc: leading dimension smaller than the matrix");
  size_t need = 0;
  fhip_trsm_workspace_bytes(&need, k, type);
  if (!work || work_bytes < need)
    return fail(FHIP_E_WORKSPACE, "fhip_trsm_runc: workspace of %zu bytes is smaller than the %zu needed",
                work_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  const int ik = (int)k;
  switch (type) {
    case FHIP_C64:
      return launch_trsm<double, double, true, true>((double *)Y, ldy, m, ik, (const double *)R, ldr, (double *)work, s);
    case FHIP_C32:
      return launch_trsm<float, float, true, true>((float *)Y, ldy, m, ik, (const float *)R, ldr, (float *)work, s);
    case FHIP_F64:
      return launch_trsm<double, double, false, true>((double *)Y, ldy, m, ik, (const double *)R, ldr, (double *)work, s);
    case FHIP_F32:
      return launch_trsm<float, float, false, true>((float *)Y, ldy, m, ik, (const float *)R, ldr, (float *)work, s);
  }
  return fail(FHIP_E_UNSUPPORTED, "fhip_trsm_runc: unknown type %d", type);
}

extern "C" int fhip_gemm_nn(void *D, int64_t ldd, const void *A, int64_t lda, const void *C, int64_t ldc,
                            int64_t m, int64_t s, int64_t r, int type, void *stream) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "fhip_gemm_nn: unknown type %d", type);
  if (m < 0 || s < 0 || r < 0) return fail(FHIP_E_ARGUMENT, "fhip_gemm_nn: negative dimension");
  if (m == 0 || r == 0) return FHIP_OK;
  if (s > (1 << 20) || r > (1 << 20)) return fail(FHIP_E_ARGUMENT, "fhip_gemm_nn: too many columns");
  if (!D || (s > 0 && (!A || !C))) return fail(FHIP_E_ARGUMENT, "fhip_gemm_nn: null matrix pointer");
  if (ldd < m || (s > 0 && (lda < m || ldc < s)))
    return fail(FHIP_E_ARGUMENT, "fhip_gemm_nn: leading dimension smaller than the matrix");
  hipStream_t st = (hipStream_t)stream;
  const int is = (int)s, ir = (int)r;
  switch (type) {
    case FHIP_C64: return launch_gemm<double, true>((double *)D, ldd, (const double *)A, lda, (const double *)C, ldc, m, is, ir, st);
    case FHIP_C32: return launch_gemm<float, true>((float *)D, ldd, (const float *)A, lda, (const float *)C, ldc, m, is, ir, st);
    case FHIP_F64: return launch_gemm<double, false>((double *)D, ldd, (const double *)A, lda, (const double *)C, ldc, m, is, ir, st);
    case FHIP_F32: return launch_gemm<float, false>((float *)D, ldd, (const float *)A, lda, (const float *)C, ldc, m, is, ir, st);
  }
  return fail(FHIP_E_UNSUPPORTED, "fhip_gemm_nn: unknown type %d", type);
}
