// Kernel (b'): the row-block product D <- A C that replaces rightmul_gemm!
// (src/kernels.jl:73, semantics :57-64, called from rightmul!
// src/panelblas.jl:203-245), the panel update of the QR adjoint.
//
// Rows are independent: a workgroup owns RW rows and CB output columns. The
// products run on v_mfma_f64_16x16x4 (f32 form for single precision) in the
// three-product complex form, with operand slices staged through LDS.
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
