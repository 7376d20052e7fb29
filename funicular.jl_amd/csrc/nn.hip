// Kernel (b'): the row-block product D <- A C that replaces rightmul_gemm!
// (src/kernels.jl:73, semantics :57-64, called from rightmul!
// src/panelblas.jl:203-245), the panel update of the QR adjoint; and its
// scaled form D <- alpha A C + beta D, the update step of the panel-by-panel
// (left-looking) triangular solve.
//
// Rows are independent: a workgroup owns GW_ROWS rows and GW_CB output
// columns, its four waves a 32 x 32 tile each. The inner dimension streams
// through LDS in 16-column slices, the next slice's loads in flight during
// the current slice's products, which run on v_mfma_f64_16x16x4 (the f32 form
// for single precision) in the three-product complex form. Workgroup ids go
// round-robin to the 8 XCDs, so the id is split as (xcd, local) and the
// output-column blocks of one row chunk run on one XCD: the rows of A come
// from HBM once and the other column blocks read them from that XCD's L2.
#include "fhip_internal.hpp"

namespace fhip {

constexpr int GW_THREADS = 256;
constexpr int GW_ROWS = 64;  // rows per workgroup
constexpr int GW_CB = 64;    // output columns per workgroup
constexpr int GW_KC = 16;    // inner dimension per LDS slice
constexpr int GW_APT = GW_KC * GW_ROWS / GW_THREADS;  // A elements per thread per slice
constexpr int GW_BPT = GW_KC * GW_CB / GW_THREADS;    // C elements per thread per slice

// SCALED: D <- alpha A C + beta D (beta = 0 does not read D); otherwise D <- A C.
template <typename T, bool CPLX, bool SCALED>
__global__ __launch_bounds__(GW_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void gemm_nn_kernel(
    T *__restrict__ D, int64_t ldd, const T *__restrict__ A, int64_t lda, const T *__restrict__ C, int64_t ldc,
    int64_t m, int s, int r, int ncb, int64_t nrc, T alr, T ali, T ber, T bei) {
  constexpr int NP = CPLX ? 2 : 1;
  // Complex FP64 planes use the swizzled LDS layout (laid out for 8-byte
  // elements); the others pad each row by 4.
  constexpr bool SWZ = CPLX && sizeof(T) == 8;
  constexpr int LDA_S = SWZ ? GW_ROWS : GW_ROWS + 4;
  constexpr int LDB_S = SWZ ? GW_CB : GW_CB + 4;
  __shared__ T As[NP][GW_KC * LDA_S];
  __shared__ T Bs[NP][GW_KC * LDB_S];
  typedef typename Mfma<T>::acc acc_t;

  const int id = blockIdx.x;
  const int xcd = id & 7;
  const int local = id >> 3;
  const int cb = local % ncb;
  const int64_t rc = (int64_t)(local / ncb) * 8 + xcd;
  if (rc >= nrc) return;  // the whole workgroup: the grid is padded to full rounds of 8
  const int64_t row0 = rc * GW_ROWS;
  const int c0 = cb * GW_CB;

  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  const int li = lane & 15;
  const int lk = lane >> 4;
  const int mi0 = 32 * (w & 1);
  const int nj0 = 32 * (w >> 1);
  // Waves whose columns lie past r only stage slices.
  const bool active = c0 + nj0 < r;

  T ar[GW_APT], ai[GW_APT], br[GW_BPT], bi[GW_BPT];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int e = 0; e < GW_APT; ++e) {
      const int idx = t + GW_THREADS * e;
      const int rr = idx % GW_ROWS;
      const int kk = idx / GW_ROWS;
      const int64_t row = row0 + rr;
      load_elem<T, T, CPLX>(A, row + (int64_t)(kc + kk) * lda, row < m && kc + kk < s, ar[e], ai[e]);
    }
#pragma unroll
    for (int e = 0; e < GW_BPT; ++e) {
      const int idx = t + GW_THREADS * e;
      const int kk = idx % GW_KC;
      const int c = idx / GW_KC;
      load_elem<T, T, CPLX>(C, (int64_t)(kc + kk) + (int64_t)(c0 + c) * ldc, kc + kk < s && c0 + c < r, br[e],
                            bi[e]);
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int e = 0; e < GW_APT; ++e) {
      const int idx = t + GW_THREADS * e;
      const int rr = idx % GW_ROWS;
      const int kk = idx / GW_ROWS;
      const int at = kk * LDA_S + (SWZ ? (rr ^ lds_swizzle(kk)) : rr);
      As[0][at] = ar[e];
      if (CPLX) As[NP - 1][at] = ai[e];
    }
#pragma unroll
    for (int e = 0; e < GW_BPT; ++e) {
      const int idx = t + GW_THREADS * e;
      const int kk = idx % GW_KC;
      const int c = idx / GW_KC;
      const int at = kk * LDB_S + (SWZ ? (c ^ lds_swizzle(kk)) : c);
      Bs[0][at] = br[e];
      if (CPLX) Bs[NP - 1][at] = bi[e];
    }
  };

  acc_t p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      p1[mb][nb] = Mfma<T>::zero();
      p2[mb][nb] = Mfma<T>::zero();
      p3[mb][nb] = Mfma<T>::zero();
    }

  if (s > 0) fetch(0);
  for (int kc = 0; kc < s; kc += GW_KC) {
    __syncthreads();
    stash();
    __syncthreads();
    if (kc + GW_KC < s) fetch(kc + GW_KC);
    if (!active) continue;
    if constexpr (CPLX)
      mma_steps_3m<T, 2, 2, PROD_ADD, SWZ>(As[0], As[NP - 1], LDA_S, mi0, Bs[0], Bs[NP - 1], LDB_S, nj0, 0,
                                           GW_KC / 4, 1, p1, p2, p3);
    else
      mma_steps<T, false, 2, 2, PROD_ADD>(As[0], As[0], LDA_S, mi0, Bs[0], Bs[0], LDB_S, nj0, 0, GW_KC / 4, 1,
                                          p1, p2);
  }
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      acc_t re = p1[mb][nb], im = p2[mb][nb];
      if constexpr (CPLX) combine3m<PROD_ADD>(p1[mb][nb], p2[mb][nb], p3[mb][nb], re, im);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t row = row0 + mi0 + 16 * mb + Mfma<T>::row(lk, q);
        const int col = c0 + nj0 + 16 * nb + li;
        if (row < m && col < r) {
          if constexpr (SCALED) {
            T outr = alr * re[q] - ali * im[q];
            T outi = alr * im[q] + ali * re[q];
            if (ber != T(0) || bei != T(0)) {
              T dr, di;
              load_elem<T, T, CPLX>(D, row + (int64_t)col * ldd, true, dr, di);
              outr += ber * dr - bei * di;
              outi += ber * di + bei * dr;
            }
            store_elem<T, CPLX>(D, row + (int64_t)col * ldd, outr, outi);
          } else {
            store_elem<T, CPLX>(D, row + (int64_t)col * ldd, re[q], im[q]);
          }
        }
      }
    }
}

template <typename T, bool CPLX, bool SCALED>
static int launch_gemm(T *D, int64_t ldd, const T *A, int64_t lda, const T *C, int64_t ldc, int64_t m,
                       int s, int r, hipStream_t stream, double alr = 1, double ali = 0, double ber = 0,
                       double bei = 0) {
  const int ncb = (r + GW_CB - 1) / GW_CB;
  const int64_t nrc = cdiv64(m, GW_ROWS);
  const int64_t grid = 8 * cdiv64(nrc, 8) * ncb;
  if (grid > 0x7fffffff) return fail(FHIP_E_ARGUMENT, "fhip_gemm_nn: too many rows (%lld)", (long long)m);
  hipLaunchKernelGGL((gemm_nn_kernel<T, CPLX, SCALED>), dim3((unsigned)grid), dim3(GW_THREADS), 0, stream, D, ldd, A,
                     lda, C, ldc, m, s, r, ncb, nrc, (T)alr, (T)ali, (T)ber, (T)bei);
  return check_launch("gemm_nn_kernel");
}

static int gemm_checks(const char *name, void *D, int64_t ldd, const void *A, int64_t lda, const void *C,
                       int64_t ldc, int64_t m, int64_t s, int64_t r, int type) {
  if (!valid_type(type)) return fail(FHIP_E_UNSUPPORTED, "%s: unknown type %d", name, type);
  if (m < 0 || s < 0 || r < 0) return fail(FHIP_E_ARGUMENT, "%s: negative dimension", name);
  if (s > (1 << 20) || r > (1 << 20)) return fail(FHIP_E_ARGUMENT, "%s: too many columns", name);
  if (m == 0 || r == 0) return FHIP_OK;
  if (!D || (s > 0 && (!A || !C))) return fail(FHIP_E_ARGUMENT, "%s: null matrix pointer", name);
  if (ldd < m || (s > 0 && (lda < m || ldc < s)))
    return fail(FHIP_E_ARGUMENT, "%s: leading dimension smaller than the matrix", name);
  return FHIP_OK;
}

}  // namespace fhip

using namespace fhip;

extern "C" int fhip_gemm_nn(void *D, int64_t ldd, const void *A, int64_t lda, const void *C, int64_t ldc,
                            int64_t m, int64_t s, int64_t r, int type, void *stream) {
  int st = gemm_checks("fhip_gemm_nn", D, ldd, A, lda, C, ldc, m, s, r, type);
  if (st) return st;
  if (m == 0 || r == 0) return FHIP_OK;
  hipStream_t sm = (hipStream_t)stream;
  const int is = (int)s, ir = (int)r;
  switch (type) {
    case FHIP_C64: return launch_gemm<double, true, false>((double *)D, ldd, (const double *)A, lda, (const double *)C, ldc, m, is, ir, sm);
    case FHIP_C32: return launch_gemm<float, true, false>((float *)D, ldd, (const float *)A, lda, (const float *)C, ldc, m, is, ir, sm);
    case FHIP_F64: return launch_gemm<double, false, false>((double *)D, ldd, (const double *)A, lda, (const double *)C, ldc, m, is, ir, sm);
    case FHIP_F32: return launch_gemm<float, false, false>((float *)D, ldd, (const float *)A, lda, (const float *)C, ldc, m, is, ir, sm);
  }
  return fail(FHIP_E_UNSUPPORTED, "fhip_gemm_nn: unknown type %d", type);
}

extern "C" int fhip_gemm_nn_scaled(void *D, int64_t ldd, const void *A, int64_t lda, const void *C, int64_t ldc,
                                   int64_t m, int64_t s, int64_t r, double alpha_re, double alpha_im,
                                   double beta_re, double beta_im, int type, void *stream) {
  int st = gemm_checks("fhip_gemm_nn_scaled", D, ldd, A, lda, C, ldc, m, s, r, type);
  if (st) return st;
  if (m == 0 || r == 0) return FHIP_OK;
  hipStream_t sm = (hipStream_t)stream;
  const int is = (int)s, ir = (int)r;
  const bool cplx = type == FHIP_C64 || type == FHIP_C32;
  const double ai = cplx ? alpha_im : 0.0, bi = cplx ? beta_im : 0.0;
  switch (type) {
    case FHIP_C64: return launch_gemm<double, true, true>((double *)D, ldd, (const double *)A, lda, (const double *)C, ldc, m, is, ir, sm, alpha_re, ai, beta_re, bi);
    case FHIP_C32: return launch_gemm<float, true, true>((float *)D, ldd, (const float *)A, lda, (const float *)C, ldc, m, is, ir, sm, alpha_re, ai, beta_re, bi);
    case FHIP_F64: return launch_gemm<double, false, true>((double *)D, ldd, (const double *)A, lda, (const double *)C, ldc, m, is, ir, sm, alpha_re, ai, beta_re, bi);
    case FHIP_F32: return launch_gemm<float, false, true>((float *)D, ldd, (const float *)A, lda, (const float *)C, ldc, m, is, ir, sm, alpha_re, ai, beta_re, bi);
  }
  return fail(FHIP_E_UNSUPPORTED, "fhip_gemm_nn_scaled: unknown type %d", type);
}
